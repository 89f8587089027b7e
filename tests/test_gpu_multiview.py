"""LAMBERT_VERTEX_COLOR (Lambert with the vertex colour as albedo) and the multi-view workload
as specified: the displaced 199,808-triangle sphere seen by 16 ring cameras at 512x512,
Lambert shading, gradients of vertex positions and colours, stored versus recomputed hits,
and the store / recompute choice made from an HBM budget.

Bars: images within 1e-5 of the oracle (hit ids bit-exact on the small scene), gradients
within 1e-3 relative L2 of the oracle on the same samples, stored and recomputed hits within
1e-5 of each other (float summation order only), and the colour gradient with render-aligned
interior samples within 1e-4 of finite differences of this library's own renders (the image
is linear in the colours)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu

IMG_TOL, GRAD_TOL = 1e-5, 1e-3
MODE = drh.LAMBERT_VERTEX_COLOR


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _scene():
    """An icosphere with random vertex colours and a cube that shadows part of its lit side."""
    sphere = S.with_random_colors(S.icosphere(3), 8)
    blocker = S.cube(0.25, 1)
    blocker.pos = (blocker.pos + np.array([0.8, 0.8, 0.8], np.float32)).astype(np.float32)
    blocker = S.with_random_colors(blocker, 9)
    return [sphere, blocker]


def _moved(meshes, seed):
    out = []
    for k, m in enumerate(meshes):
        mm = m.copy()
        if k == 0:
            mm = S.with_random_colors(S.radial_displace(m, 0.05, seed=seed, smooth=False), seed + 1)
        out.append(mm)
    return out


def test_lambert_vertex_colour_matches_oracle():
    lam = S.lambert_config2()
    meshes = _scene()
    cam = S.camera_look((0.3, -0.4, 2.9), 128, 128)
    ref = O.Scene(meshes, O.TRACER_BVH)
    img_o = ref.render(MODE, 4, [cam], lam=lam)[0]
    tgt = O.Scene(_moved(meshes, 30), O.TRACER_BVH).render(MODE, 4, [cam], lam=lam)[0]
    _, adj = O.loss(img_o, tgt)
    N = 1 << 16
    r = drh.DiffRender(mode=MODE, spp=4, light=lam)
    try:
        r.commit(meshes)
        prim, geom, *_ = r.hits(cam)
        prim_o, geom_o, *_ = ref.hits(cam)
        assert np.array_equal(prim, prim_o) and np.array_equal(geom, geom_o)
        img = r.render([cam])[0]
        assert np.abs(img - img_o).max() <= IMG_TOL
        V = meshes[0].V
        for aligned in (False, True):
            go = ref.d_render(MODE, 4, [cam], [adj], N, lam=lam, aligned=aligned)
            r.set_interior_samples(int(aligned), 0)
            g = r.d_render([cam], [adj], N)
            e, ep, ec = rel_l2(g, go), rel_l2(g[:3 * V], go[:3 * V]), rel_l2(g[3 * V:6 * V], go[3 * V:6 * V])
            say("LAMBERT_VCOL aligned", aligned, "rel L2", e, "positions", ep, "colours", ec)
            assert np.abs(go[:3 * V]).sum() > 0 and np.abs(go[3 * V:6 * V]).sum() > 0
            assert e < GRAD_TOL and ep < GRAD_TOL and ec < GRAD_TOL
    finally:
        r.close()


def test_lambert_vertex_colour_gradient_is_the_derivative_of_the_render():
    lam = S.lambert_config2()
    meshes = _scene()
    cam = S.camera_look((0.3, -0.4, 2.9), 64, 64)
    rng = np.random.default_rng(2)
    adj = rng.uniform(-1, 1, size=(64, 64, 3)).astype(np.float32)
    r = drh.DiffRender(mode=MODE, spp=4, light=lam)
    try:
        r.commit(meshes)
        r.set_interior_samples(1, 0)
        g = r.d_render([cam], [adj], 0)
        base = r.render([cam])[0].astype(np.float64)
        V = meshes[0].V
        coords = rng.choice(3 * V, 40, replace=False)
        fd = []
        for k in coords:
            mm = [m.copy() for m in meshes]
            mm[0].col = meshes[0].col.copy()
            mm[0].col.reshape(-1)[k] += 1.0
            r.commit(mm)
            d = r.render([cam])[0].astype(np.float64) - base
            fd.append(float(np.sum(adj.astype(np.float64) * d)))
        e = rel_l2(g[3 * V:6 * V][coords], np.array(fd))
        say("aligned LAMBERT_VCOL colour gradient vs render differences", e)
        assert np.abs(fd).max() > 0 and e < 1e-4
    finally:
        r.close()


def test_stored_hits_equal_recomputed_hits_with_shadows():
    lam = S.lambert_config2()
    meshes = _scene()
    cams = S.cameras_ring(3, 96, 96)
    t = drh.DiffRender(mode=MODE, spp=4, light=lam)
    t.commit(_moved(meshes, 40))
    targets = t.render(cams)
    t.close()
    out = []
    for store in (drh.HITS_RECOMPUTE, drh.HITS_STORE):
        r = drh.DiffRender(mode=MODE, spp=4, light=lam)
        r.commit(meshes)
        r.set_interior_samples(1, store)
        out.append(r.d_render_and_compare(cams, targets, 1 << 14))
        r.close()
    (l0, g0, i0), (l1, g1, i1) = out
    assert l0 == l1 and all(np.array_equal(a, b) for a, b in zip(i0, i1))
    e = rel_l2(g1, g0)
    say("LAMBERT_VCOL store vs recompute", e)
    assert e < 1e-5


def test_automatic_hit_policy_follows_the_budget():
    """DRH_HITS_AUTO stores the hits when 16 B x samples fit the limit and recomputes (holding
    no buffer) when they do not; either way the gradient is the same up to summation order."""
    lam = S.lambert_config2()
    meshes = _scene()
    cams = S.cameras_ring(2, 64, 64)
    t = drh.DiffRender(mode=MODE, spp=4, light=lam)
    t.commit(_moved(meshes, 50))
    targets = t.render(cams)
    t.close()
    need = 16 * 2 * 64 * 64 * 4
    res = {}
    for name, limit in (("fits", need), ("short", need - 16)):
        r = drh.DiffRender(mode=MODE, spp=4, light=lam)
        r.commit(meshes)
        r.set_interior_samples(1, drh.HITS_AUTO)
        r.set_hit_buffer_limit(limit)
        res[name] = r.d_render_and_compare(cams, targets, 1 << 13)[1]
        res[name + "_bytes"] = r.hit_buffer_bytes()
        r.close()
    assert res["fits_bytes"] == need and res["short_bytes"] == 0
    assert rel_l2(res["short"], res["fits"]) < 1e-5


def _config4(spp=1):
    base = S.uv_sphere(224, 448)
    m = S.with_random_colors(S.radial_displace(base, 0.02, seed=6, smooth=False), 6)
    t = S.with_random_colors(S.radial_displace(base, 0.02, seed=16, smooth=False), 16)
    return m, t, S.cameras_ring(16, 512, 512)


def test_config4_all_views_match_oracle():
    """199,808 triangles, all 16 ring views at 512x512, Lambert with vertex-colour albedo, spp 1
    (the reference's default), 2^16 edge samples per view: images and the position + colour
    gradient against the oracle on the same samples."""
    lam = S.lambert_config2()
    m, t, cams = _config4()
    N = 1 << 16
    r = drh.DiffRender(mode=MODE, spp=1, light=lam)
    try:
        r.commit([t])
        targets = r.render(cams)
        r.commit([m])
        imgs = r.render(cams)
        say("config 4:", m.F, "triangles,", len(cams), "views: oracle render")
        ref = O.Scene([m], O.TRACER_BVH)
        imgs_o = ref.render(MODE, 1, cams, lam=lam)
        worst = max(float(np.abs(a - b).max()) for a, b in zip(imgs, imgs_o))
        say("config 4 image max |diff|", worst)
        assert worst <= IMG_TOL
        adjs = [O.loss(a, b)[1] for a, b in zip(imgs_o, targets)]
        say("oracle d_render")
        go = ref.d_render(MODE, 1, cams, adjs, N, lam=lam)
        g = r.d_render(cams, adjs, N)
        V = m.V
        e, ep, ec = rel_l2(g, go), rel_l2(g[:3 * V], go[:3 * V]), rel_l2(g[3 * V:6 * V], go[3 * V:6 * V])
        say("config 4 rel L2", e, "positions", ep, "colours", ec)
        assert np.abs(go[:3 * V]).sum() > 0 and np.abs(go[3 * V:6 * V]).sum() > 0
        assert e < GRAD_TOL and ep < GRAD_TOL and ec < GRAD_TOL
    finally:
        r.close()


def test_config4_fused_steps_store_equals_recompute():
    """Three fused Adam steps on all 16 views fitting the colours, with deterministic sums:
    stored hits carry exactly the terms the interior pass would trace again (the closest hit
    and the light terms, bit for bit), so gradient and fitted colours are bit-identical."""
    lam = S.lambert_config2()
    m, t, cams = _config4()
    r = drh.DiffRender(mode=MODE, spp=1, light=lam)
    r.commit([t])
    tg = r.render(cams)
    r.close()
    out, grads = [], []
    for store in (drh.HITS_RECOMPUTE, drh.HITS_STORE):
        r = drh.DiffRender(mode=MODE, spp=1, light=lam)
        try:
            r.commit([m])
            r.set_interior_samples(1, store)
            r.set_deterministic(1)
            grads.append(r.d_render_and_compare(cams, tg, 1 << 16, images=False)[1])
            r.opt_init(drh.OPT_ADAM, 0.0, 1e-2)
            dt = r.device_array(np.concatenate([x.reshape(-1) for x in tg]))
            cs = drh.cameras_array(cams)
            for _ in range(3):
                r.opt_step_dev(cs, len(cams), dt.ptr, 1 << 16)
            out.append(r.get_params()[1])
            say("store", store, "hit buffer bytes", r.hit_buffer_bytes())
            assert r.hit_buffer_bytes() == (16 * 16 * 512 * 512 if store else 0)
        finally:
            r.close()
    say("config 4 store vs recompute gradient diff", np.abs(grads[1] - grads[0]).max(),
        "colour diff", np.abs(out[0] - out[1]).max())
    assert np.abs(grads[0]).sum() > 0 and np.abs(out[0] - m.col).max() > 0
    assert np.array_equal(grads[1], grads[0]) and np.array_equal(out[0], out[1])
