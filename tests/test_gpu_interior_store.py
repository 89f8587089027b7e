"""Interior samples aligned with the render's grid, and the hit buffer stored by the render
versus recomputed by the interior pass (multi-view adjoint).  The aligned interior estimate of
the colour gradient is the exact derivative of the rendered image (which is linear in the
vertex colours at fixed samples), so finite differences of GPU renders pin it; storing the
hits must change nothing but the summation order."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(np.asarray(b, np.float64)), 1e-30))


def grad(meshes, cams, targets, mode, spp, N, aligned, store, light=None):
    r = drh.DiffRender(mode=mode, spp=spp, light=light)
    r.commit(meshes)
    r.set_interior_samples(aligned, store)
    loss, g, imgs = r.d_render_and_compare(cams, targets, N)
    r.close()
    return loss, g, imgs


CASES = [
    ("vcol", drh.VERTEX_COLOR, lambda m: S.with_random_colors(m, 3)),
    ("texture", drh.TEXTURE_COLOR, lambda m: S.with_uv_texture(m, 8, 8, seed=5)),
    ("lambert albedo", drh.LAMBERT, lambda m: S.with_albedo(m, 0.8)),
]


@pytest.mark.parametrize("name,mode,dress", CASES, ids=[c[0] for c in CASES])
def test_stored_hits_equal_recomputed_hits(name, mode, dress):
    lam = S.lambert_config2() if mode == drh.LAMBERT else None
    m = dress(S.icosphere(3))
    t = dress(S.radial_displace(S.icosphere(3), 0.08, seed=9))
    cams = S.cameras_ring(3, 96, 96)
    r = drh.DiffRender(mode=mode, spp=4, light=lam)
    r.commit([t])
    targets = r.render(cams)
    r.close()
    l0, g0, i0 = grad([m], cams, targets, mode, 4, 1 << 14, aligned=True, store=False, light=lam)
    l1, g1, i1 = grad([m], cams, targets, mode, 4, 1 << 14, aligned=True, store=True, light=lam)
    lh, gh, _ = grad([m], cams, targets, mode, 4, 1 << 14, aligned=False, store=False, light=lam)
    for a, b in zip(i0, i1):
        assert np.array_equal(a, b)
    assert l0 == l1
    e = rel_l2(g1, g0)
    say(name, "store vs recompute", e, "aligned vs Hammersley", rel_l2(g0, gh))
    assert e < 1e-5
    assert not np.array_equal(g0, gh)   # a different (equally valid) estimator


def test_aligned_colour_gradient_is_the_derivative_of_the_render():
    """VCOL with no edge samples: the aligned interior colour gradient must equal the change of
    sum_p adj_p . I_p when one vertex colour component moves (the image is linear in it)."""
    m = S.with_random_colors(S.icosphere(2), 3)
    cam = S.camera_look((0.5, 0.7, 2.8), 64, 64)
    rng = np.random.default_rng(1)
    adj = rng.uniform(-1, 1, size=(64, 64, 3)).astype(np.float32)
    r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=4)
    r.commit([m])
    r.set_interior_samples(True, False)
    g = r.d_render([cam], [adj], 0)
    base = r.render([cam])[0].astype(np.float64)
    coords = rng.choice(3 * m.V, 30, replace=False)
    fd = []
    for k in coords:
        mm = m.copy()
        mm.col = m.col.copy()
        mm.col.reshape(-1)[k] += 1.0
        r.commit([mm])
        d = r.render([cam])[0].astype(np.float64) - base
        fd.append(float(np.sum(adj.astype(np.float64) * d)))
    r.close()
    gc = g[3 * m.V: 6 * m.V][coords]
    e = rel_l2(gc, np.array(fd))
    say("aligned colour gradient vs render differences", e)
    assert e < 1e-4


def test_multiview_step_store_equals_recompute():
    """Fused steps over 4 views, hits stored vs recomputed: the same fitted colours."""
    m = S.with_random_colors(S.icosphere(3), 3)
    t = S.with_random_colors(S.radial_displace(S.icosphere(3), 0.08, seed=9), 4)
    cams = S.cameras_ring(4, 96, 96)
    r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=4)
    r.commit([t])
    tg = r.render(cams)
    r.close()
    out = []
    for store in (False, True):
        r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=4)
        r.commit([m])
        r.set_interior_samples(True, store)
        r.opt_init(drh.OPT_ADAM, 0.0, 1e-2)   # colours only: no edge-order sensitivity
        dt = r.device_array(np.concatenate([x.reshape(-1) for x in tg]))
        cs = drh.cameras_array(cams)
        for _ in range(4):
            r.opt_step_dev(cs, len(cams), dt.ptr, 1 << 12)
        out.append(r.get_params()[1])
        r.close()
    say("colour diff", np.abs(out[0] - out[1]).max())
    assert np.abs(out[0] - out[1]).max() < 1e-5
