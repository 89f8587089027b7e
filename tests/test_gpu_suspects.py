"""Clamped-determinant ("suspect") triangles in the binned form the LBVH tracer uses for large
scenes, against the brute-force tracer, which tests every triangle with the reference's
Moller-Trumbore arithmetic.

The subdivided cubes on the sphere below (49,152 and 78,732 triangles) are large enough for
the view and light suspect lists to be binned per cell, and their symmetric grids hold many
triangles seen exactly or almost edge-on: the brute-force tracer accepts a few dozen camera
rays per view through a clamped determinant (|det| < 1e-6), which the LBVH walk alone would
miss.  Bars: hits bit-exact, images bit-exact, gradients equal up to float summation order."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from diffrender_tutorials_amd import drh, scenes as S

pytestmark = pytest.mark.gpu

f32 = np.float32
EYES = ((0.0, 0.0, 3.0), (2.0, 2.0, 2.0), (0.0, 2.1, 2.1))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _mv4(B, x, y, z, w):
    return [((x * B[i] + y * B[4 + i]) + z * B[8 + i]) + w * B[12 + i] for i in range(4)]


def _normalize(v):
    return v * (f32(1) / np.sqrt(_dot(v, v)).astype(f32))[..., None]


def clamped_hits(m, cam, prim):
    """Pixels whose hit (pixel-centre camera ray) has |det| < 1e-6: the camera ray in float32
    as the tracer forms it (CastSingleRay), det = e1 . (d x e2) as Moller-Trumbore computes it."""
    P = np.asarray(m.pos, f32).reshape(-1, 3)
    I = np.asarray(m.idx).reshape(-1, 3)
    W, H = cam.width, cam.height
    Wi = np.linalg.inv(np.asarray(cam.world_view, np.float64).reshape(4, 4).T).astype(f32).T.reshape(-1)
    Pi = np.linalg.inv(np.asarray(cam.proj, np.float64).reshape(4, 4).T).astype(f32).T.reshape(-1)
    ys, xs = np.divmod(np.arange(W * H), W)
    sel = prim != 0xFFFFFFFF
    x = (xs[sel] + f32(0.5)).astype(f32)
    y = (ys[sel] + f32(0.5)).astype(f32)
    p = _mv4(Pi, f32(2) * (x / f32(W)) - f32(1), f32(-2) * (y / f32(H)) + f32(1), f32(0), f32(1))
    dr = _normalize(np.stack([p[0] / p[3], p[1] / p[3], p[2] / p[3]], -1).astype(f32))
    a = _mv4(Wi, f32(0), f32(0), f32(0), f32(1))
    b = _mv4(Wi, f32(100) * dr[:, 0], f32(100) * dr[:, 1], f32(100) * dr[:, 2], f32(1))
    d = _normalize(np.stack([b[0] - a[0], b[1] - a[1], b[2] - a[2]], -1).astype(f32))
    tri = I[prim[sel]]
    A = P[tri[:, 0]]
    det = _dot(P[tri[:, 1]] - A, _cross(d, P[tri[:, 2]] - A))
    return int((np.abs(det) < f32(1e-6)).sum())


def _hits(m, cam, tracer):
    r = drh.DiffRender(mode=drh.SILHOUETTE, spp=1, tracer=tracer)
    try:
        r.commit([m])
        return r.hits(cam)
    finally:
        r.close()


@pytest.mark.parametrize("n", [64, 81])
def test_binned_view_suspects_give_the_brute_force_hits(n):
    m = S.cube_sphere(n)
    clamped = 0
    for eye in EYES:
        cam = S.camera_look(eye, 512, 512)
        say("cube_sphere", n, m.F, "triangles, eye", eye)
        hb = _hits(m, cam, drh.TRACER_BF)
        hv = _hits(m, cam, drh.TRACER_BVH)
        for a, b in zip(hb, hv):   # primitive, mesh, t, u, v: bit for bit
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
        clamped += clamped_hits(m, cam, hb[0])
    say("   pixels hit through a clamped determinant:", clamped)
    assert clamped > 0   # the scenes exercise the suspect path


@pytest.mark.parametrize("light", ["reference", "config2"])
def test_binned_light_suspects_give_the_brute_force_shadows(light):
    m = S.with_albedo(S.cube_sphere(64), 0.8)
    lam = None if light == "reference" else S.lambert_config2()
    for eye in EYES:
        cam = S.camera_look(eye, 384, 384)
        imgs = []
        for tracer in (drh.TRACER_BF, drh.TRACER_BVH):
            r = drh.DiffRender(mode=drh.LAMBERT, spp=1, tracer=tracer, light=lam)
            try:
                r.commit([m])
                imgs.append(r.render([cam])[0])
            finally:
                r.close()
        say("light", light, "eye", eye, "differing pixels", int((imgs[0] != imgs[1]).any(-1).sum()))
        assert np.array_equal(imgs[0], imgs[1])


def test_binned_suspects_edge_gradient_matches_brute_force():
    base = S.cube_sphere(64)
    m = S.with_albedo(base, 0.8)
    target = S.with_albedo(S.radial_displace(base, 0.05, seed=3, smooth=False), 0.8)
    cam = S.camera_look((2.0, 2.0, 2.0), 256, 256)
    out = []
    for tracer in (drh.TRACER_BF, drh.TRACER_BVH):
        r = drh.DiffRender(mode=drh.LAMBERT, spp=1, tracer=tracer)
        try:
            r.commit([target])
            tg = r.render([cam])
            r.commit([m])
            out.append(r.d_render_and_compare([cam], tg, 1 << 17))
        finally:
            r.close()
    (lb, gb, ib), (lv, gv, iv) = out
    e = float(np.linalg.norm(gb - gv) / np.linalg.norm(gb))
    say("loss", lb, lv, "gradient rel L2 BVH vs BF", e)
    assert np.array_equal(ib[0], iv[0]) and lb == lv
    assert np.abs(gb).sum() > 0 and e < 1e-5
