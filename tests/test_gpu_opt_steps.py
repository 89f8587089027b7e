"""drh_opt_steps (k fused steps captured as one graph) against the same steps one graph replay
at a time: with deterministic gradients the fitted parameters are bit-identical, across the
LBVH's build / refit pattern (rebuild every 3 steps, so the k-step graphs see every phase)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu


def _fit(stepper, mode):
    lam = S.lambert_config2() if mode == drh.LAMBERT_VERTEX_COLOR else None
    base = S.with_random_colors(S.icosphere(3), 2)
    target = S.with_random_colors(S.radial_displace(S.icosphere(3), 0.05, seed=3), 5)
    cam = S.camera_look((0.3, 0.5, 2.8), 96, 80)
    tgt = O.Scene([target], O.TRACER_BVH).render(mode, 1, [cam], lam=lam)[0]
    r = drh.DiffRender(mode=mode, spp=1, light=lam)
    try:
        r.commit([base])
        r.set_deterministic(True)
        r.set_bvh_rebuild(3)
        r.opt_init(drh.OPT_ADAM, 2e-3, 2e-3)
        dt = r.device_array(tgt.reshape(-1))
        dl = r.device_array(np.zeros(1, np.float32))
        cs = drh.cameras_array([cam])
        stepper(r, cs, dt, dl, cam.width * cam.height * 4)
        r.sync()
        pos, col = r.get_params()
        loss = float(dl.download(1)[0])
    finally:
        r.close()
    return pos, col, loss


@pytest.mark.parametrize("mode", [drh.VERTEX_COLOR, drh.LAMBERT_VERTEX_COLOR])
def test_k_step_graphs_equal_single_steps_bit_for_bit(mode):
    def single(r, cs, dt, dl, N):
        for _ in range(17):
            r.opt_step_dev(cs, 1, dt.ptr, N, dl.ptr)

    def batched(r, cs, dt, dl, N):
        for k in (1, 4, 4, 3, 4, 1):      # 17 steps: eager first, then graphs of 4, 3 and 1
            r.opt_steps_dev(cs, 1, dt.ptr, N, dl.ptr, k)

    a, b = _fit(single, mode), _fit(batched, mode)
    print("mode", mode, "loss", a[2], b[2], flush=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]
