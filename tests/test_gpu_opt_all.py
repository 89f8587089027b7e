"""The device-resident optimiser with every block moving (OptimizerParameters::OPT_ALL):
transform and texture rates next to the position rate, against the reference's own
OptSimple::Run on the same scenes (tests/golden/optall_*.npz, src/optimizer_simple.cpp:92-309).

Transform updates move rows 0-2 of the differentiable mesh's transform; the static mesh's
transform stays as committed.  With a position rate as well, the second step's position
gradient goes through the inverse of mesh 0's transform from the first commit, which the
reference keeps (Scene::transforms_inv is never cleared, common_src/scene_utils.cpp:110-122):
the two-step case with positions and transforms moving pins that."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import drh
from golden_cases import OPT_ALL_CASES, load, mesh_hash, opt_all_scene, sha

pytestmark = pytest.mark.gpu

# one step is bit-for-bit the reference's arithmetic up to float sum order; a few steps of
# the normalised (sign-like) update amplify ulp-level gradient differences a little
POS_TOL = {1: 1e-6, 2: 2e-6, 3: 5e-6}
TR_TOL = {1: 1e-6, 2: 2e-6, 3: 5e-6}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fit(name, graph=0):
    _, mode, spp, iters, (plr, tlr, trlr) = OPT_ALL_CASES[name]
    fx = load(name)
    meshes, targets, cam = opt_all_scene(name)
    for i, m in enumerate(meshes):
        assert mesh_hash(m) == str(fx[f"mesh{i}_sha"])
    tgt = O.Scene(targets).render(mode, spp, [cam])[0]
    assert sha(tgt) == str(fx["target_sha"])
    say("case", name, "iterations", iters, "rates", (plr, tlr, trlr))
    r = drh.DiffRender(mode=mode, spp=spp)
    try:
        r.commit(meshes)
        r.opt_init(drh.OPT_SIMPLE, plr, 0.1)
        r.opt_set_texture_lr(tlr)
        r.opt_set_transform_lr(trlr)
        drh.load().drh_opt_use_graph(r.h, graph)
        dt = r.device_array(tgt.reshape(-1))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array([cam])
        for _ in range(iters):
            r.opt_step_dev(cs, 1, dt.ptr, cam.width * cam.height, dl.ptr)
        r.sync()
        pos, _ = r.get_params()
        tr = r.get_transforms()
        tex = r.get_texels(meshes[0].tex.size) if tlr else None
    finally:
        r.close()
    return fx, meshes, iters, pos, tr, tex


@pytest.mark.parametrize("name", list(OPT_ALL_CASES))
def test_opt_all_matches_reference(name):
    fx, meshes, iters, pos, tr, tex = _fit(name)
    V0 = meshes[0].V
    dpos = float(np.abs(pos[:V0].reshape(-1) - fx["opt_pos"]).max())
    dtr = float(np.abs(tr.reshape(-1) - fx["opt_tr"]).max())
    say("max |pos - ref|", dpos, "max |transform - ref|", dtr)
    assert dpos <= POS_TOL[iters]
    assert dtr <= TR_TOL[iters]
    for i in range(1, len(meshes)):
        assert np.array_equal(tr[i], meshes[i].transform)       # not differentiable: untouched
        assert np.array_equal(pos[V0:V0 + meshes[i].V], meshes[i].pos)
    if tex is not None:
        dtex = float(np.abs(tex - fx["opt_tex"]).max())
        say("max |texture - ref|", dtex)
        assert dtex <= 1e-6 * iters
    trlr = OPT_ALL_CASES[name][4][2]
    if trlr:
        assert np.abs(tr[0] - meshes[0].transform).max() > 0.5 * trlr   # the transform did move


def test_transform_rate_zero_keeps_transforms():
    name = "optall_tfm_pos_sil_2"
    _, mode, spp, _, _ = OPT_ALL_CASES[name]
    meshes, targets, cam = opt_all_scene(name)
    tgt = O.Scene(targets).render(mode, spp, [cam])[0]
    r = drh.DiffRender(mode=mode, spp=spp)
    try:
        r.commit(meshes)
        r.opt_init(drh.OPT_SIMPLE, 0.02, 0.1)
        dt = r.device_array(tgt.reshape(-1))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array([cam])
        for _ in range(2):
            r.opt_step_dev(cs, 1, dt.ptr, cam.width * cam.height, dl.ptr)
        tr = r.get_transforms()
        with pytest.raises(drh.DrhError):
            r.opt_set_transform_lr(-1.0)
    finally:
        r.close()
    for i, m in enumerate(meshes):
        assert np.array_equal(tr[i], m.transform)


def test_transform_steps_replayed_from_a_graph_match_eager_steps():
    """Adam steps are captured once and replayed; the transform update is inside the graph."""
    name = "optall_tfm_sil_3"
    _, mode, spp, _, _ = OPT_ALL_CASES[name]
    meshes, targets, cam = opt_all_scene(name)
    tgt = O.Scene(targets).render(mode, spp, [cam])[0]
    out = []
    for graph in (0, 1):
        r = drh.DiffRender(mode=mode, spp=spp)
        try:
            r.commit(meshes)
            r.opt_init(drh.OPT_ADAM, 0.0, 0.0)
            r.opt_set_transform_lr(0.01)
            drh.load().drh_opt_use_graph(r.h, graph)
            dt = r.device_array(tgt.reshape(-1))
            dl = r.device_array(np.zeros(1, np.float32))
            cs = drh.cameras_array([cam])
            for _ in range(5):
                r.opt_step_dev(cs, 1, dt.ptr, cam.width * cam.height, dl.ptr)
            out.append(r.get_transforms())
        finally:
            r.close()
    say("max |graph - eager|", float(np.abs(out[0] - out[1]).max()))
    assert np.abs(out[0][0] - meshes[0].transform).max() > 0.01
    assert np.allclose(out[0], out[1], rtol=0, atol=1e-6)
