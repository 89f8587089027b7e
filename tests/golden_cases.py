"""Scenes of the golden fixtures (tests/golden/*.npz), rebuilt from the scene generators;
each fixture stores the SHA-256 of its meshes so a generator change is caught."""
from __future__ import annotations

import hashlib
import os

import numpy as np

from diffrender_tutorials_amd import scenes as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def mesh_hash(m) -> str:
    h = hashlib.sha256()
    for a in (m.pos, m.idx, m.col, m.tc, m.nrm, m.tex, m.transform):
        if a is not None:
            h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def _box():
    box = S.with_random_colors(S.cube(0.3, 1), 10)
    box.pos = (box.pos + np.array([0.9, 0.5, 0.6], np.float32)).astype(np.float32)
    box_t = box.copy()
    box_t.pos = (box.pos + np.array([0.05, -0.03, 0.0], np.float32)).astype(np.float32)
    return box, box_t


def tfm_scene():
    """A differentiable sphere under a rotated, translated transform beside a static box; the
    target moves the sphere's transform (transform optimisation, silhouette)."""
    ball = S.icosphere(2, 0.35)
    ball.transform = S.mat_mul(S.translate((0.15, -0.1, 0.05)), S.rotate_y(20.0))
    box = S.cube(0.25, 3)
    box.transform = S.translate((-0.55, 0.3, -0.3))
    box.differentiable = False
    ball_t = ball.copy()
    ball_t.transform = S.mat_mul(S.translate((0.22, -0.04, 0.05)), S.rotate_y(20.0))
    return [ball, box], [ball_t, box], S.camera_look((0.4, 0.5, 3.0), 128, 128)


# optimiser runs with every block moving (OptimizerParameters::OPT_ALL): name ->
# (scene, shading mode, spp, iterations, (position_lr, textures_lr, transforms_lr))
OPT_ALL_CASES = {
    "optall_tfm_sil_1": ("tfm", 1, 4, 1, (0.0, 0.0, 0.1)),
    "optall_tfm_sil_3": ("tfm", 1, 4, 3, (0.0, 0.0, 0.1)),
    "optall_tfm_pos_sil_2": ("tfm", 1, 4, 2, (0.02, 0.0, 0.05)),
    "optall_tex_2": ("ico3_texture", 3, 4, 2, (0.0, 0.05, 0.0)),
}


def opt_all_scene(name):
    key = OPT_ALL_CASES[name][0]
    return tfm_scene() if key == "tfm" else scene(key)


def scene(name):
    """(meshes, target_meshes, camera) of a golden case."""
    if name == "cfg1_cube_vcol":
        c = S.config1()
        return c["meshes"], c["target"], c["cams"][0]
    if name in ("cube_rot_vcol", "cube_rot_sil"):
        c = S.config1()
        return c["meshes"], [S.perturb(c["meshes"][0], 0.05, seed=2)], S.camera_look((2.0, 1.5, 2.5), 256, 256)
    cam_i = S.camera_look((0.5, 1.0, 2.8), 128, 128)
    if name == "ico3_lambert":
        return ([S.with_albedo(S.icosphere(3), 0.8)],
                [S.with_albedo(S.radial_displace(S.icosphere(3), 0.1, seed=3), 0.8)], cam_i)
    if name == "ico3_texture":
        return ([S.with_uv_texture(S.icosphere(3), 8, 8, 5)],
                [S.with_uv_texture(S.radial_displace(S.icosphere(3), 0.1, seed=3), 8, 8, 6)], cam_i)
    if name == "ico2_vcol_2mesh":
        box, box_t = _box()
        return ([S.with_random_colors(S.icosphere(2), 7), box],
                [S.with_random_colors(S.radial_displace(S.icosphere(2), 0.1, seed=8), 9), box_t],
                S.camera_look((0.3, 0.4, 3.0), 96, 96))
    if name == "tfm_lambert_2mesh":
        # both meshes under non-identity transforms, both differentiable: the tracer's doubled
        # instance transform and mesh 0's inverse transform in every mesh's edge gradient
        (ball, box), (ball_t, _), cam = tfm_scene()
        ball, box, ball_t = S.with_albedo(ball, 0.8), S.with_albedo(box, 0.6), S.with_albedo(ball_t, 0.8)
        box.differentiable = True
        return [ball, box], [ball_t, box], cam
    if name == "cfg2_ico5_lambert":
        c = S.config2()
        return c["meshes"], c["target"], c["cams"][0]
    raise KeyError(name)


SMALL_CASES = ["cfg1_cube_vcol", "cube_rot_vcol", "cube_rot_sil", "ico3_lambert", "ico3_texture", "ico2_vcol_2mesh",
               "tfm_lambert_2mesh"]
ALL_CASES = SMALL_CASES + ["cfg2_ico5_lambert"]


def check_sample(fx, name, arr):
    """Strided sample of a hashed buffer (for a readable diff when a hash differs)."""
    a = np.ascontiguousarray(arr).reshape(-1)
    return np.array_equal(a[::97], fx[name + "_sample"])


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
