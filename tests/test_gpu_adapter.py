"""The drop-in through the reference's own interface: oracle/_ref/adapter_check drives
libdrh.so with the DiffRenderHIP adapter (include/drh_idiffrender.h) as an IDiffRender
(src/virtual_drender.h:14-26), on Scene / DScene / CamInfo / Img objects of the reference's
own headers, and writes what it got.  Checked here against the oracle on the same scene:

* render images within 1e-5 of the oracle's (the parity bar of the C ABI tests);
* d_render into a DScene listing its meshes in reverse order, against the oracle's gradient
  (relative L2 < 1e-3, identical samples); a second d_render accumulates (2x);
* a DScene with only mesh 0 differentiable: the adapter commits again and the edge sampler
  draws mesh 0's edges only, as the oracle with mesh 1 fixed;
* d_render_and_compare: loss, images into outImages, and the gradient after the DScene is
  cleared (it was filled with 5.0 first);
* render / d_render of a scene other than the committed one are refused and leave their
  outputs untouched (src/drender.h:103-107, 148-152).

The binary is built in the container where the reference's headers are (make -C oracle
adapter); without it the test is skipped."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "..", "oracle", "_ref", "adapter_check")
GRAD_TOL = 1e-3
IMG_TOL = 1e-5


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        name = data[p:p + 32].split(b"\0")[0].decode()
        kind = chr(data[p + 32])
        n = int(np.frombuffer(data[p + 33:p + 41], np.int64)[0])
        p += 41
        out[name] = np.frombuffer(data[p:p + 4 * n], np.float32 if kind == "f" else np.int32).copy()
        p += 4 * n
    return out


def meshes_of(rec, tag, diff):
    ms = []
    for i in range(2):
        k = f"{tag}{i}_"
        V = rec[k + "pos"].size // 3
        ms.append(S.Mesh(pos=rec[k + "pos"].reshape(V, 3), idx=rec[k + "idx"].astype(np.uint32).reshape(-1, 3),
                         col=rec[k + "col"].reshape(V, 3), tc=rec[k + "tc"].reshape(V, 2),
                         nrm=rec[k + "nrm"].reshape(V, 3), tex=rec[k + "tex"].reshape(8, 8, 3),
                         transform=rec[k + "transform"], differentiable=bool(diff[i])))
    return ms


def mesh_order(rec, tag, mode):
    """DScene blocks (in the DScene's order) -> the flat gradient's mesh order."""
    ids, g = rec[tag + "_ids"], rec[tag + "_grad"]
    V = rec["s0_pos"].size // 3
    size = 3 * V + (3 * V if mode == drh.VERTEX_COLOR else 0) + \
        (8 * 8 * 3 if mode in (drh.TEXTURE_COLOR, drh.LAMBERT) else 0) + 12
    assert g.size == size * ids.size
    blocks = {int(m): g[k * size:(k + 1) * size] for k, m in enumerate(ids)}
    return np.concatenate([blocks[m] for m in sorted(blocks)])


@pytest.mark.skipif(not os.path.exists(EXE), reason="oracle/_ref/adapter_check is built only where the reference's headers are")
@pytest.mark.parametrize("mode", [drh.SILHOUETTE, drh.VERTEX_COLOR, drh.TEXTURE_COLOR, drh.LAMBERT])
def test_idiffrender_adapter_matches_oracle(tmp_path, mode):
    out = str(tmp_path / f"adapter_{mode}.bin")
    print("adapter_check mode", mode, flush=True)
    r = subprocess.run([EXE, str(mode), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    # the two refused calls printed the reference's message
    assert r.stdout.count("renderer was not prepared for this scene") == 2, r.stdout
    rec = read_records(out)
    W, H, views, N, m = (int(x) for x in rec["size"])
    assert m == mode
    cams = [S.Camera(rec[f"cam{i}_wv"], rec[f"cam{i}_proj"], W, H) for i in range(views)]
    spp = 4

    both = O.Scene(meshes_of(rec, "s", [1, 1]), O.TRACER_BVH)
    tgt = O.Scene(meshes_of(rec, "t", [1, 1]), O.TRACER_BVH).render(mode, spp, cams)
    img = both.render(mode, spp, cams)
    imgs = rec["images"].reshape(views, H, W, 3)
    assert np.abs(imgs - np.stack(img)).max() <= IMG_TOL
    assert np.abs(rec["target_images"].reshape(views, H, W, 3) - np.stack(tgt)).max() <= IMG_TOL
    loss = 0.0
    adj = []
    for v in range(views):
        l, a = O.loss(imgs[v], rec["target_images"].reshape(views, H, W, 3)[v])
        loss += l
        adj.append(a)

    gA = both.d_render(mode, spp, cams, adj, N)
    eA = rel_l2(mesh_order(rec, "dA1", mode), gA)
    print("  d_render (DScene {1, 0}) rel L2", eA, flush=True)
    assert eA < GRAD_TOL
    assert rel_l2(mesh_order(rec, "dA2", mode), 2 * mesh_order(rec, "dA1", mode)) < 1e-5

    only0 = O.Scene(meshes_of(rec, "s", [1, 0]), O.TRACER_BVH)
    gB = only0.d_render(mode, spp, cams, adj, N)
    eB = rel_l2(mesh_order(rec, "dB", mode), gB)
    print("  d_render (DScene {0}) rel L2", eB, flush=True)
    assert eB < GRAD_TOL

    eC = rel_l2(mesh_order(rec, "dC", mode), gA)
    print("  d_render_and_compare rel L2", eC, "loss", float(rec["compare_loss"][0]), loss / (views * W * H), flush=True)
    assert eC < GRAD_TOL
    assert np.isclose(float(rec["compare_loss"][0]), loss / (views * W * H), rtol=1e-5)
    assert np.abs(rec["compare_images"].reshape(views, H, W, 3) - imgs).max() <= IMG_TOL
    assert list(rec["refused_untouched"]) == [1, 1]
