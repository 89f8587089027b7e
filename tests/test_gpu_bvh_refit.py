"""LBVH upkeep in the fused optimisation step: a full build every k steps and box refits of
the last topology in between.  Traversal is exact over any valid tree, so a refitted tree
must give the hits and the image of a freshly built one bit for bit, and fits with k = 1
(the reference's rebuild per iteration) and k = 8 must agree."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from diffrender_tutorials_amd import drh, scenes as S
from test_gpu_serving import SHORT_ITERS, LOSS_RTOL, assert_same_fit, _fit_alone, _job, _targets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _steps(r, cams, tg, n, N):
    dt = r.device_array(np.concatenate([t.reshape(-1) for t in tg]))
    dl = r.device_array(np.zeros(1, np.float32))
    cs = drh.cameras_array(cams)
    for _ in range(n):
        r.opt_step_dev(cs, len(cams), dt.ptr, N, dl.ptr)
    r.sync()
    return dt, dl


@pytest.mark.parametrize("graph", [0, 1])
def test_refitted_tree_renders_as_a_fresh_build(graph):
    lam = S.lambert_config2()
    m, t, cams = _job(3, 96, 13)
    tg = _targets(t, cams, 4, lam)
    r = drh.DiffRender(mode=drh.LAMBERT, spp=4, light=lam)
    try:
        r.commit([m])
        r.set_bvh_rebuild(8)
        r.opt_init(drh.OPT_ADAM, 3e-3, 0.0)
        drh.load().drh_opt_use_graph(r.h, graph)
        bufs = _steps(r, cams, tg, 11, 1 << 13)
        p11, _ = r.get_params()
        # step 12 (index 11) refits the tree over p11, renders, then moves the vertices on
        bufs2 = _steps(r, cams, tg, 1, 1 << 13)
        img_refit = r.render(cams)[0]
        hits_refit = r.hits(cams[0])
    finally:
        r.close()
    assert np.abs(p11 - m.pos).max() > 1e-3   # the geometry did move
    mm = m.copy()
    mm.pos = p11.astype(np.float32)
    f = drh.DiffRender(mode=drh.LAMBERT, spp=4, light=lam)
    try:
        f.commit([mm])
        img_fresh = f.render(cams)[0]
        hits_fresh = f.hits(cams[0])
    finally:
        f.close()
    say("max |refit - fresh|", float(np.abs(img_refit - img_fresh).max()))
    for a, b in zip(hits_refit, hits_fresh):
        assert np.array_equal(np.asarray(a), np.asarray(b))
    assert np.array_equal(img_refit, img_fresh)


def test_fits_with_and_without_refit_steps_agree():
    lam = S.lambert_config2()
    m, t, cams = _job(3, 96, 14)
    tg = _targets(t, cams, 4, lam)
    out = {}
    for every in (1, 8):
        r = drh.DiffRender(mode=drh.LAMBERT, spp=4, light=lam)
        try:
            r.commit([m])
            r.set_bvh_rebuild(every)
            r.opt_init(drh.OPT_ADAM, 1e-3, 1e-3)
            dt = r.device_array(tg[0].reshape(-1))
            dl = r.device_array(np.zeros(1, np.float32))
            cs = drh.cameras_array(cams)
            losses = []
            for _ in range(SHORT_ITERS):
                r.opt_step_dev(cs, 1, dt.ptr, 1 << 14, dl.ptr)
                losses.append(float(dl.download(1)[0]))
            out[every] = (r.get_params()[0], np.array(losses))
        finally:
            r.close()
    assert np.allclose(out[8][1], out[1][1], rtol=LOSS_RTOL), (out[8][1], out[1][1])
    assert_same_fit(out[8][0], out[1][0], m.pos)


def test_bvh_rebuild_rejects_non_positive_interval():
    r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=1)
    try:
        with pytest.raises(drh.DrhError):
            r.set_bvh_rebuild(0)
    finally:
        r.close()


def test_commit_refuses_more_triangles_than_the_walk_addresses():
    """The LBVH walks address nodes and leaf triangles by 32-bit byte offsets: a scene of more
    than 2^26 triangles is refused (DRH_E_INVALID) before any of its arrays is read, and the
    handle stays usable."""
    import ctypes
    m = S.cube()
    r = drh.DiffRender(mode=drh.VERTEX_COLOR, spp=1)
    try:
        d, keep = drh.scene_desc([m])
        keep[1][0] = (1 << 26) + 1   # the mesh's triangle count, as the descriptor sees it
        assert drh.load().drh_commit(r.h, ctypes.byref(d)) == -1
        say("refused:", drh.load().drh_last_error(r.h))
        r.commit([m])
        img = r.render([S.camera_look((2.0, 2.0, 2.0), 64, 64)])[0]
        assert np.isfinite(img).all() and img.max() > 0
    finally:
        r.close()
