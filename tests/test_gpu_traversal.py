"""The two LBVH walks (drh_set_traversal) give the same hits: a wave walking as one packet,
each lane walking alone, and the automatic choice between them render bit-identical images
and, with deterministic gradient sums, bit-identical gradients -- on scenes where the
automatic choice takes each walk (many triangles per pixel at 128 x 128, few at 512 x 512),
with shadow rays (Lambert), vertex colours and jittered samples."""
from __future__ import annotations

import numpy as np
import pytest

from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu


def _case(F, res, mode, seed):
    base = S.cube_sphere_for(F, 80000)
    t = S.radial_displace(base, 0.05, seed=seed, smooth=False)
    if mode == drh.VERTEX_COLOR:
        m, t = S.with_random_colors(base, seed + 1), S.with_random_colors(t, seed + 2)
    else:
        m, t = S.with_albedo(base, 0.8), S.with_albedo(t, 0.8)
    cam = S.camera_look((1.2, 0.8, 2.7), res, res)
    return m, t, cam


@pytest.mark.parametrize("F,res,mode,spp,jitter", [
    (50000, 128, drh.VERTEX_COLOR, 1, False),   # auto: lane walk for camera and edge rays
    (24000, 128, drh.LAMBERT, 4, True),         # lane walk, shadow rays, jittered 2x2 samples
    (20000, 512, drh.LAMBERT, 1, False),        # auto: wave walk
])
def test_traversal_modes_give_identical_results(F, res, mode, spp, jitter):
    lam = S.lambert_config2() if mode == drh.LAMBERT else None
    m, t, cam = _case(F, res, mode, 7)
    rt = drh.DiffRender(mode=mode, spp=spp, light=lam)
    rt.commit([t])
    tgt = rt.render([cam])
    rt.close()
    N = 4 * res * res
    out = {}
    for name, walk in (("wave", drh.TRAVERSAL_WAVE), ("lane", drh.TRAVERSAL_LANE), ("auto", drh.TRAVERSAL_AUTO)):
        r = drh.DiffRender(mode=mode, spp=spp, light=lam)
        try:
            r.commit([m])
            r.set_traversal(walk)
            r.set_deterministic(True)
            if jitter:
                r.set_pixel_jitter(True, 99)
            img = r.render([cam])[0]
            loss, g, _ = r.d_render_and_compare([cam], tgt, N)
            out[name] = (img, g, loss)
        finally:
            r.close()
    print("F", m.F, "res", res, "loss", out["wave"][2], flush=True)
    assert np.abs(out["wave"][1]).max() > 0
    for name in ("lane", "auto"):
        assert np.array_equal(out[name][0], out["wave"][0]), name
        assert np.array_equal(out[name][1], out["wave"][1]), name
        assert out[name][2] == out["wave"][2], name
