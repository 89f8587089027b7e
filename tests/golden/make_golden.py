"""Write the golden vectors under tests/golden/ from the reference itself.

Runs only in a container where /root/reference exists and ``make -C oracle ref`` has
built oracle/_ref/ref_harness (the reference compiled from its own sources).  The
fixtures hold inputs and outputs only (no reference code).  Large buffers are stored
as a SHA-256 of the full float32/uint32 bytes plus a strided sample of values.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import dataclasses
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))   # tests/: the golden scenes

import refio  # noqa: E402
from diffrender_tutorials_amd import scenes as S  # noqa: E402
import oracle as O  # noqa: E402  (only to build targets / adjoints; each is re-checked against the reference)

SAMPLE_STRIDE = 97


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def mesh_hash(m) -> str:
    h = hashlib.sha256()
    for a in (m.pos, m.idx, m.col, m.tc, m.nrm, m.tex, m.transform):
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def big(d, name, a):
    a = np.ascontiguousarray(a).reshape(-1)
    d[name + "_sha"] = sha(a)
    d[name + "_sample"] = a[::SAMPLE_STRIDE].copy()
    d[name + "_size"] = a.size


def cam_fields(d, prefix, cam):
    d[prefix + "world_view"] = np.asarray(cam.world_view, np.float32)
    d[prefix + "proj"] = np.asarray(cam.proj, np.float32)
    d[prefix + "wh"] = np.array([cam.width, cam.height], np.int32)


def case(name, meshes, targets, cam, mode, spp, edge_samples, lam_ok=True, opt_iters=0):
    out = {"mode": mode, "spp": spp, "edge_samples": edge_samples}
    for i, m in enumerate(meshes):
        out[f"mesh{i}_sha"] = mesh_hash(m)
    cam_fields(out, "cam_", cam)
    r = refio.run(meshes, [cam], mode, spp, refio.OP_RENDER | refio.OP_HITS | refio.OP_EDGES)
    big(out, "image", r["images"])
    big(out, "hit_prim", r["hit_prim"])
    big(out, "hit_u", r["hit_u"])
    big(out, "hit_v", r["hit_v"])
    big(out, "hit_t", r["hit_t"])
    out["proj_inv"] = r["proj_inv"]
    out["wv_inv"] = r["wv_inv"]
    out["wvp"] = r["wvp"]
    big(out, "edge_v0", r["edge_v0"])
    big(out, "edge_v1", r["edge_v1"])
    big(out, "edge_pmf", r["edge_pmf"])
    big(out, "edge_cdf", r["edge_cdf"])
    big(out, "proj", r["proj"])
    out["n_edges"] = r["edge_v0"].size
    # target rendered by the reference, adjoint from the reference's loss
    rt = refio.run(targets, [cam], mode, spp, refio.OP_RENDER)
    img = r["images"].reshape(cam.height, cam.width, 3)
    tgt = rt["images"].reshape(cam.height, cam.width, 3)
    big(out, "target", tgt)
    if edge_samples:
        rc = refio.run(meshes, [cam], mode, spp, refio.OP_COMPARE, edge_samples=edge_samples, images=[tgt])
        out["compare_loss"] = rc["loss"]
        out["compare_grad"] = rc["compare_grad"]
        big(out, "compare_images", rc["compare_images"])
        # d_render with the adjoint 2*(I - T)
        _, adj = O.loss(img, tgt)
        rd = refio.run(meshes, [cam], mode, spp, refio.OP_DRENDER, edge_samples=edge_samples, images=[adj])
        out["grad"] = rd["grad"]
        big(out, "adjoint", adj)
    if opt_iters:
        ro = refio.run(meshes, [cam], mode, spp, refio.OP_OPTIMIZE, images=[tgt], opt_iters=opt_iters)
        out["opt_iters"] = opt_iters
        out["opt_pos"] = ro["opt_pos"]
        out["opt_col"] = ro["opt_col"]
        out["opt_err"] = ro["opt_err"]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, {k: (v.shape if isinstance(v, np.ndarray) else v) for k, v in out.items() if not k.endswith("_sample")})


def sequences():
    c = S.config1()
    r = refio.run(c["meshes"], c["cams"], 2, 16, refio.OP_SEQ, edge_samples=1 << 16)
    out = {"qmc": r["qmc"].reshape(11, 31), "hamm16": r["hamm"]}
    big(out, "prng65536", r["prng"])
    out["prng65536_head"] = r["prng"][:2048]
    r2 = refio.run(c["meshes"], c["cams"], 2, 4, refio.OP_SEQ, edge_samples=1 << 20)
    big(out, "prng1m", r2["prng"])
    out["hamm4"] = r2["hamm"]
    r3 = refio.run(c["meshes"], c["cams"], 2, 1, refio.OP_SEQ, edge_samples=1000)
    out["prng1000"] = r3["prng"]
    out["hamm1"] = r3["hamm"]
    np.savez_compressed(os.path.join(HERE, "sequences.npz"), **out)
    print("wrote sequences")


def adam_vectors():
    """Standard Adam golden vectors from torch.optim.Adam (run here only)."""
    import torch
    rng = np.random.default_rng(11)
    p0 = rng.standard_normal(64).astype(np.float32)
    grads = rng.standard_normal((5, 64)).astype(np.float32)
    p = torch.nn.Parameter(torch.tensor(p0.copy()))
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    traj = []
    for g in grads:
        opt.zero_grad()
        p.grad = torch.tensor(g)
        opt.step()
        traj.append(p.detach().numpy().copy())
    np.savez_compressed(os.path.join(HERE, "adam_torch.npz"), p0=p0, grads=grads, traj=np.array(traj))
    print("wrote adam_torch")


def _vsgf_mesh():
    """A small closed mesh with normals and texture coordinates for the VSGF fixtures."""
    m = S.radial_displace(S.icosphere(1), 0.05, seed=4)
    p = m.pos.astype(np.float64)
    tc = np.stack([np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi) + 0.5,
                   np.arccos(np.clip(p[:, 1] / np.linalg.norm(p, axis=1), -1, 1)) / np.pi], 1)
    m.tc = tc.astype(np.float32)
    m.nrm = S.vertex_normals(m.pos, m.idx)
    return m


def _fixture_image(H, W, seed):
    """Float RGB with the tone map's edge cases: 0, tiny, exactly 1, above 1, random."""
    rng = np.random.default_rng(seed)
    img = rng.uniform(0.0, 1.2, (H, W, 3)).astype(np.float32)
    img[0, 0] = (0.0, 1e-7, 1.0)
    img[0, 1] = (0.5, 2.0, 1e6)
    return img


def meshio_files():
    """The reference's VSGF and image writers and its VSGF reader (common_src/cmesh.cpp:139-205,
    common_src/Image2d.cpp:486-735) on small inputs: meshio_files.npz holds the inputs, the
    bytes the reference wrote and what the reference read back from this package's writer."""
    import tempfile
    from diffrender_tutorials_amd import meshio
    m = _vsgf_mesh()
    V = m.V
    pos4 = np.concatenate([m.pos, np.ones((V, 1), np.float32)], 1)
    nrm4 = np.concatenate([m.nrm, np.zeros((V, 1), np.float32)], 1)
    out = {"pos4": pos4, "nrm4": nrm4, "tc": m.tc, "idx": m.idx}
    with tempfile.TemporaryDirectory() as td:
        f = os.path.join(td, "ref.vsgf")
        refio.save_vsgf(f, pos4, nrm4, m.tc, m.idx)
        out["ref_vsgf"] = np.frombuffer(open(f, "rb").read(), np.uint8).copy()
        for tag, mm in (("ours", m), ("ours_nonrm", dataclasses.replace(m, nrm=None))):
            f = os.path.join(td, tag + ".vsgf")
            meshio.save_vsgf(f, mm)
            out[tag + "_vsgf_sha"] = hashlib.sha256(open(f, "rb").read()).hexdigest()
            back = refio.load_vsgf(f)
            for k, v in back.items():
                out[tag + "_ref_" + k] = v
        for tag, (H, W, seed) in (("img8", (5, 8, 11)), ("img6", (3, 6, 12))):
            img = _fixture_image(H, W, seed)
            out[tag] = img
            for ext, gamma in ((".bmp", 2.2), (".ppm", 2.2), (".image3f", 2.2), (".ppm", 1.0), (".png", 2.2)):
                f = os.path.join(td, tag + ext)
                refio.save_image(f, img, gamma)
                key = f"{tag}_ref{ext.replace('.', '_')}" + ("" if gamma == 2.2 else "_g1")
                out[key] = np.frombuffer(open(f, "rb").read(), np.uint8).copy()
    np.savez_compressed(os.path.join(HERE, "meshio_files.npz"), **out)
    print("wrote meshio_files")


def opt_all_case(name):
    """OptSimple (GD_Adam) with OPT_ALL flags: positions, transforms, colours and textures move
    with the case's rates (src/optimizer_simple.cpp:92-186, 272-309)."""
    import golden_cases as GC
    _, mode, spp, iters, lrs = GC.OPT_ALL_CASES[name]
    meshes, targets, cam = GC.opt_all_scene(name)
    out = {"mode": mode, "spp": spp, "opt_iters": iters, "lrs": np.array(lrs, np.float32)}
    for i, m in enumerate(meshes):
        out[f"mesh{i}_sha"] = mesh_hash(m)
    cam_fields(out, "cam_", cam)
    rt = refio.run(targets, [cam], mode, spp, refio.OP_RENDER)
    tgt = rt["images"].reshape(cam.height, cam.width, 3)
    out["target_sha"] = sha(tgt)
    ro = refio.run(meshes, [cam], mode, spp, refio.OP_OPTIMIZE, images=[tgt], opt_iters=iters, opt_lrs=lrs)
    for k in ("opt_pos", "opt_col", "opt_err", "opt_tr", "opt_tex"):
        out[k] = ro[k]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    print("wrote", name, {k: v.shape for k, v in out.items() if isinstance(v, np.ndarray)})


def transformr_vectors():
    """TransformR::to_mat (common_src/scene.h:99-104) of the reference on a few parameter sets:
    identity, the initial transform of the reference's restricted-transform test
    (src/tests.cpp:350-361), quarter turns and random sets."""
    rng = np.random.default_rng(17)
    x = [[0, 0, 0, 0, 0, 0, 1], [0, 0, 0, 0.0, 0.2, 0.2, 1], [0.1, -0.2, 0.3, 1.5707964, 0, 0, 1],
         [0, 0, 0, 0, 1.5707964, 0, 2], [0, 0, 0, 0, 0, 1.5707964, 0.5]]
    x += [list(rng.uniform(-1, 1, 6)) + [float(rng.uniform(0.5, 2))] for _ in range(6)]
    x = np.array(x, np.float32)
    r = refio.run([], [], 1, 1, 0, transformr=x)
    np.savez_compressed(os.path.join(HERE, "transformr.npz"), params=x, to_mat=r["transformr"].reshape(-1, 16))
    print("wrote transformr")


def main(only=None):
    """Write every fixture, or only those named in `only`."""
    if not os.path.exists(refio.HARNESS):
        raise SystemExit("oracle/_ref/ref_harness is missing: run `make -C oracle ref` (needs /root/reference)")
    want = (lambda n: True) if not only else (lambda n: n in only)
    if want("sequences"):
        sequences()
    if want("adam_torch"):
        adam_vectors()
    if want("meshio_files"):
        meshio_files()
    if want("transformr"):
        transformr_vectors()
    import golden_cases as GC
    for name in GC.OPT_ALL_CASES:
        if want(name):
            opt_all_case(name)
    c1 = S.config1()
    # one optimiser step: later steps amplify ulp-level position differences through the
    # sign-like normalised update, so the step itself is what is pinned
    if want("cfg1_cube_vcol"):
        case("cfg1_cube_vcol", c1["meshes"], c1["target"], c1["cams"][0], 2, 1, 1 << 16, opt_iters=1)
    if not only:
        _rest(c1)
    elif "tfm_lambert_2mesh" in only:
        import golden_cases as GC
        ms, ts, cam_t = GC.scene("tfm_lambert_2mesh")
        case("tfm_lambert_2mesh", ms, ts, cam_t, 4, 4, 1 << 14)


def _rest(c1):
    cube_t = S.perturb(c1["meshes"][0], 0.05, seed=2)
    cam_rot = S.camera_look((2.0, 1.5, 2.5), 256, 256)
    case("cube_rot_vcol", c1["meshes"], [cube_t], cam_rot, 2, 4, 1 << 16)
    case("cube_rot_sil", c1["meshes"], [cube_t], cam_rot, 1, 1, 1 << 16)
    ico = S.with_albedo(S.icosphere(3), 0.8)
    ico_t = S.with_albedo(S.radial_displace(S.icosphere(3), 0.1, seed=3), 0.8)
    cam_i = S.camera_look((0.5, 1.0, 2.8), 128, 128)
    case("ico3_lambert", [ico], [ico_t], cam_i, 4, 4, 1 << 14)
    ico_tx = S.with_uv_texture(S.icosphere(3), 8, 8, 5)
    ico_txt = S.with_uv_texture(S.radial_displace(S.icosphere(3), 0.1, seed=3), 8, 8, 6)
    case("ico3_texture", [ico_tx], [ico_txt], cam_i, 3, 4, 1 << 14)
    ico_v = S.with_random_colors(S.icosphere(2), 7)
    ico_vt = S.with_random_colors(S.radial_displace(S.icosphere(2), 0.1, seed=8), 9)
    box = S.with_random_colors(S.cube(0.3, 1), 10)
    box.pos = (box.pos + np.array([0.9, 0.5, 0.6], np.float32)).astype(np.float32)
    box_t = box.copy()
    box_t.pos = (box.pos + np.array([0.05, -0.03, 0.0], np.float32)).astype(np.float32)
    case("ico2_vcol_2mesh", [ico_v, box], [ico_vt, box_t], S.camera_look((0.3, 0.4, 3.0), 96, 96), 2, 1, 1 << 13)
    import golden_cases as GC
    ms, ts, cam_t = GC.scene("tfm_lambert_2mesh")
    case("tfm_lambert_2mesh", ms, ts, cam_t, 4, 4, 1 << 14)
    # config 2 geometry at full resolution: hits, image and the edge sampler (no gradient)
    c2 = S.config2()
    case("cfg2_ico5_lambert", c2["meshes"], c2["target"], c2["cams"][0], 4, 4, 0)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
