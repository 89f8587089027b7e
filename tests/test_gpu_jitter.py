"""Jittered pixel samples on the device (drh_set_pixel_jitter) against the oracle's pattern:
the render bit for bit (images within 1e-5, the parity bar), the aligned interior samples of
d_render_and_compare (gradient within 1e-3 relative L2, identical samples), and the fused
Adam step, whose t-th step renders with the pattern of seed + t (step 1's loss against the
oracle's loss of the start mesh rendered with seed + 1).  Config-2-like sizes: an icosphere,
Lambert shading, 4 spp."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _scene(mode):
    base = S.icosphere(3)
    m = S.with_albedo(base, 0.8) if mode == drh.LAMBERT else S.with_random_colors(base, 5)
    t = S.radial_displace(m, 0.05, seed=9, smooth=True)
    t.tc, t.tex, t.col = m.tc, m.tex, m.col
    cam = S.camera_look((0.3, 0.7, 2.8), 160, 120)
    return m, t, cam


@pytest.mark.parametrize("mode", [drh.VERTEX_COLOR, drh.LAMBERT])
def test_jittered_render_and_gradient_match_oracle(mode):
    lam = S.lambert_config2() if mode == drh.LAMBERT else None
    m, t, cam = _scene(mode)
    seed, N = 2024, 1 << 16
    sc = O.Scene([m], O.TRACER_BVH)
    sc.set_jitter(True, seed)
    st = O.Scene([t], O.TRACER_BVH)
    st.set_jitter(True, seed)
    img = sc.render(mode, 4, [cam], lam=lam)[0]
    tgt = st.render(mode, 4, [cam], lam=lam)[0]
    grid = O.Scene([m], O.TRACER_BVH).render(mode, 4, [cam], lam=lam)[0]
    assert np.abs(img - grid).max() > 1e-3        # the pattern is not the grid
    loss, adj = O.loss(img, tgt)
    go = sc.d_render(mode, 4, [cam], [adj], N, lam=lam, aligned=True)
    r = drh.DiffRender(mode=mode, spp=4, light=lam)
    try:
        r.commit([m])
        r.set_pixel_jitter(True, seed)
        r.set_interior_samples(True, drh.HITS_RECOMPUTE)
        dimg = r.render([cam])[0]
        print("mode", mode, "max |img - oracle|", float(np.abs(dimg - img).max()), flush=True)
        assert np.abs(dimg - img).max() <= 1e-5
        dl, g, imgs = r.d_render_and_compare([cam], [tgt], N)
        err = rel_l2(g, go)
        print("  gradient rel L2", err, "loss", dl, loss / (cam.width * cam.height), flush=True)
        assert err < 1e-3
        assert np.isclose(dl, loss / (cam.width * cam.height), rtol=1e-5)
        # stored hits give the same samples
        r.set_interior_samples(True, drh.HITS_STORE)
        _, gs, _ = r.d_render_and_compare([cam], [tgt], N)
        assert rel_l2(gs, go) < 1e-3
        # off again: the grid
        r.set_pixel_jitter(False)
        assert np.array_equal(r.render([cam])[0], r.render([cam])[0])
        assert np.abs(r.render([cam])[0] - grid).max() <= 1e-5
    finally:
        r.close()


def test_fused_step_draws_seed_plus_step():
    mode = drh.VERTEX_COLOR
    m, t, cam = _scene(mode)
    seed = 77
    st = O.Scene([t], O.TRACER_BVH)
    tgt = st.render(mode, 4, [cam])[0]
    sc = O.Scene([m], O.TRACER_BVH)
    expect = []
    for step in (1, 2):
        sc.set_jitter(True, seed + step)
        l, _ = O.loss(sc.render(mode, 4, [cam])[0], tgt)
        expect.append(l / (cam.width * cam.height))
    r = drh.DiffRender(mode=mode, spp=4)
    try:
        r.commit([m])
        r.set_pixel_jitter(True, seed)
        r.opt_init(drh.OPT_ADAM, 0.0, 0.0)     # zero rates: the mesh stays the start mesh
        dt = r.device_array(tgt.reshape(-1))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array([cam])
        got = []
        for _ in range(3):                      # step 1 eager, 2 captured, 3 replayed
            r.opt_step_dev(cs, 1, dt.ptr, cam.width * cam.height, dl.ptr)
            r.sync()
            got.append(float(dl.download(1)[0]))
    finally:
        r.close()
    print("fused step losses", got, "oracle seed+1, seed+2", expect, flush=True)
    assert np.isclose(got[0], expect[0], rtol=1e-5)
    assert np.isclose(got[1], expect[1], rtol=1e-5)
    assert got[2] != got[1]                     # the replay drew a fresh pattern
