"""Pins of the oracle's FP64 Lambert pixel-area functional (lambert_functional / fd_lambert),
the finite-difference reference of the LAMBERT / LAMBERT_VERTEX_COLOR position gradient:

* with the light's colour zero the shade is ambient x albedo everywhere on the mesh, so the
  functional is ambient x albedo x the silhouette functional (exact covered areas), which a
  wrong visibility rule, quadrature weight or pixel mapping breaks;
* the Gauss-Legendre order has converged (order 6 against order 10);
* the functional equals sum_p adj_p . image_p of the oracle's own render with a 16 x 16 grid of
  samples per pixel to the grid's accuracy, in both modes (a wrong normal blend, albedo or
  shadow rule moves it far more);
* its central differences are smooth in h (h = 1e-4 against h = 2e-4)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import scenes as S

LAM = {"light_dir": (np.array([-1.0, -1.0, -1.0]) / np.sqrt(3.0)).tolist(), "light_color": [1.0, 0.9, 0.8],
       "ambient": [0.1, 0.1, 0.1]}


def smooth_adjoint(W, H, seed=0):
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    a = np.stack([1.0 + 0.5 * np.sin(xx * (5.0 + k) / W) * np.cos(yy * (9.0 - k) / H) for k in range(3)], axis=2)
    return a.astype(np.float32)


def sphere(mode, subdiv=2):
    m = S.icosphere(subdiv, 1.0)
    return S.with_random_colors(m, 4) if mode == 8 else S.with_albedo(m, 0.8)


CAM = S.camera_look((0.5, 0.6, 2.9), 48, 40)


def test_ambient_only_is_albedo_times_silhouette():
    m = sphere(4)
    sc = O.Scene([m], O.TRACER_BVH)
    adj = smooth_adjoint(CAM.width, CAM.height)
    dark = {"light_dir": LAM["light_dir"], "light_color": [0.0, 0.0, 0.0], "ambient": [0.25, 0.25, 0.25]}
    F = sc.lambert_functional(4, CAM, adj, lam=dark)
    sil = sc.silhouette_functional(CAM, adj)
    assert abs(F - 0.25 * float(np.float32(0.8)) * sil) <= 1e-9 * abs(sil)   # the texel is a float32 0.8


@pytest.mark.parametrize("mode", [4, 8])
def test_quadrature_order_has_converged(mode):
    sc = O.Scene([sphere(mode)], O.TRACER_BVH)
    adj = smooth_adjoint(CAM.width, CAM.height)
    f6 = sc.lambert_functional(mode, CAM, adj, lam=LAM, order=6)
    f10 = sc.lambert_functional(mode, CAM, adj, lam=LAM, order=10)
    assert abs(f6 - f10) <= 2e-5 * abs(f10)


@pytest.mark.parametrize("mode", [4, 8])
def test_functional_is_the_pixel_integral_of_the_render(mode):
    m = sphere(mode)
    sc = O.Scene([m], O.TRACER_BVH)
    adj = smooth_adjoint(CAM.width, CAM.height)
    F = sc.lambert_functional(mode, CAM, adj, lam=LAM)
    img = sc.render(mode, 16 * 16, [CAM], lam=LAM)[0]
    R = float(np.sum(img.astype(np.float64) * adj))
    # the same with a shadow rule or normal blend changed: the light off everywhere
    dark = dict(LAM, light_color=[0.0, 0.0, 0.0])
    Fd = sc.lambert_functional(mode, CAM, adj, lam=dark)
    print("functional", F, "render", R, "unlit", Fd)
    assert abs(F - R) <= 2e-3 * abs(R)
    assert abs(Fd - R) > 20 * abs(F - R)


@pytest.mark.parametrize("mode", [4, 8])
def test_central_differences_are_smooth_in_h(mode):
    m = sphere(mode, 3)
    sc = O.Scene([m], O.TRACER_BVH)
    adj = smooth_adjoint(CAM.width, CAM.height)
    rng = np.random.default_rng(2)
    coords = rng.choice(3 * m.V, 24, replace=False)
    f1 = sc.fd_lambert(mode, CAM, adj, coords, 1e-4, lam=LAM)
    f2 = sc.fd_lambert(mode, CAM, adj, coords, 2e-4, lam=LAM)
    assert np.abs(f1).max() > 0
    assert np.linalg.norm(f1 - f2) <= 2e-3 * np.linalg.norm(f1)
