"""Deterministic gradients (drh_set_deterministic): fixed-point integer sums make the gradient
independent of the order the waves finish in.  Repeated runs agree bit for bit; against the
float-atomic path the gradient differs by the float atomics' own rounding (their sums round
at every step, in whatever order the waves arrive; the fixed-point sums round only terms below
2^-9): within 1e-4 relative L2, and
against the oracle it keeps the parity bar (1e-3).  Cases cover the interior pass through its
LDS, hash-table and uniform-texel paths (now replaced by fixed-point atomics), the edge pass
with one and with two differentiable meshes, a RESTRICTED_TRANSFORM mesh, and the fused
optimisation step replayed from a hipGraph."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import drh, scenes as S
from golden_cases import rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _moved(m, off):
    out = m.copy()
    out.pos = (m.pos + np.asarray(off, np.float32)).astype(np.float32)
    return out


def _case(name):
    lam = S.lambert_config2()
    cam = S.camera_look((0.5, 1.0, 2.8), 128, 128)
    if name == "vcol_two_meshes":
        box = _moved(S.with_random_colors(S.cube(0.3, 1), 10), (0.9, 0.5, 0.6))
        return [S.with_random_colors(S.icosphere(2), 7), box], drh.VERTEX_COLOR, None, cam
    if name == "vcol_large_hash":          # G > the interior's LDS range: hash-table path
        return [S.with_random_colors(S.icosphere(4), 3)], drh.VERTEX_COLOR, None, cam
    if name == "texture_lds":
        return [S.with_uv_texture(S.icosphere(3), 8, 8, 5)], drh.TEXTURE_COLOR, None, cam
    if name == "lambert_uniform":          # 1x1 albedo: the interior's uniform-texel path
        return [S.with_albedo(S.icosphere(3), 0.8)], drh.LAMBERT, lam, cam
    if name == "silhouette_restricted":
        m = S.icosphere(2)
        m.restricted = np.array([0.02, -0.01, 0.0, 0.1, 0.05, -0.05, 1.0], np.float32)
        return [m], drh.SILHOUETTE, None, cam
    raise KeyError(name)


CASES = ["vcol_two_meshes", "vcol_large_hash", "texture_lds", "lambert_uniform", "silhouette_restricted"]


def _grad(name, det, N=1 << 15):
    meshes, mode, lam, cam = _case(name)
    tg = [_moved(m, (0.03, -0.02, 0.01)) for m in meshes]
    for t in tg:
        t.restricted = None
    tgt = O.Scene(tg, O.TRACER_BVH).render(mode, 4, [cam], lam=lam)[0]
    r = drh.DiffRender(mode=mode, spp=4, light=lam)
    try:
        r.commit(meshes)
        r.set_deterministic(det)
        loss, g, _ = r.d_render_and_compare([cam], [tgt], N, images=False)
    finally:
        r.close()
    return g, loss


@pytest.mark.parametrize("name", CASES)
def test_deterministic_gradient_repeats_bit_for_bit(name):
    g1, l1 = _grad(name, 1)
    g2, l2 = _grad(name, 1)
    gf, lf = _grad(name, 0)
    e = rel_l2(g1, gf)
    say(name, "G", g1.size, "rel L2 deterministic vs float atomics", e)
    assert np.abs(gf).sum() > 0
    assert np.array_equal(g1, g2)
    assert l1 == l2 == lf
    assert e < 1e-4


@pytest.mark.parametrize("order", [drh.EDGE_ORDER_AUTO, drh.EDGE_ORDER_BINNED])
def test_deterministic_gradient_repeats_with_fewer_samples_than_edges(order):
    """N < E makes the automatic edge order pick the pixel-bin order, whose in-bin ranks come
    from atomic counters; deterministic mode keeps the rnd0 slot order instead (for an explicit
    EDGE_ORDER_BINNED too), so three runs agree bit for bit."""
    m = S.with_random_colors(S.icosphere(4), 3)            # E = 7680 edges
    cam = S.camera_look((0.5, 1.0, 2.8), 128, 128)
    tgt = O.Scene([_moved(m, (0.03, -0.02, 0.01))], O.TRACER_BVH).render(drh.VERTEX_COLOR, 1, [cam])[0]
    N = 1 << 12
    gs = []
    for _ in range(3):
        r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=1)
        try:
            r.commit([m])
            assert N < len(r.edges()[0])
            r.set_edge_order(order)
            r.set_deterministic(1)
            gs.append(r.d_render_and_compare([cam], [tgt], N, images=False)[1])
        finally:
            r.close()
    assert np.abs(gs[0][:3 * m.V]).sum() > 0
    assert np.array_equal(gs[0], gs[1]) and np.array_equal(gs[0], gs[2])


def test_deterministic_gradient_matches_oracle():
    meshes, mode, lam, cam = _case("vcol_two_meshes")
    tg = [_moved(m, (0.03, -0.02, 0.01)) for m in meshes]
    sc = O.Scene(meshes, O.TRACER_BVH)
    img = sc.render(mode, 4, [cam])[0]
    tgt = O.Scene(tg, O.TRACER_BVH).render(mode, 4, [cam])[0]
    _, adj = O.loss(img, tgt)
    N = 1 << 15
    go = sc.d_render(mode, 4, [cam], [adj], N)
    r = drh.DiffRender(mode=mode, spp=4)
    try:
        r.commit(meshes)
        r.set_deterministic(1)
        g = r.d_render([cam], [adj], N)
    finally:
        r.close()
    e = rel_l2(g, go)
    say("deterministic vs oracle rel L2", e)
    assert e < 1e-3


def test_deterministic_fits_repeat_bit_for_bit():
    """Ten Adam steps (the later ones replayed from a hipGraph) twice: identical parameters."""
    meshes, mode, lam, cam = _case("vcol_two_meshes")
    tg = [_moved(m, (0.03, -0.02, 0.01)) for m in meshes]
    tgt = O.Scene(tg, O.TRACER_BVH).render(mode, 4, [cam])[0]
    out = []
    for _ in range(2):
        r = drh.DiffRender(mode=mode, spp=4)
        try:
            r.commit(meshes)
            r.set_deterministic(1)
            r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
            dt = r.device_array(tgt.reshape(-1))
            dl = r.device_array(np.zeros(1, np.float32))
            cs = drh.cameras_array([cam])
            for _ in range(10):
                r.opt_step_dev(cs, 1, dt.ptr, 1 << 14, dl.ptr)
            out.append(r.get_params())
        finally:
            r.close()
    (p1, c1), (p2, c2) = out
    assert np.abs(p1 - np.concatenate([m.pos for m in meshes])).max() > 1e-4   # the fit moved
    assert np.array_equal(p1, p2) and np.array_equal(c1, c2)


def test_set_deterministic_rejects_bad_values():
    r = drh.DiffRender(mode=drh.SILHOUETTE, spp=1)
    try:
        with pytest.raises(drh.DrhError):
            r.set_deterministic(3)
    finally:
        r.close()
