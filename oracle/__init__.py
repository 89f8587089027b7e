"""The oracle: a CPU restatement of the reference render / d_render path.

TEST INFRASTRUCTURE ONLY.  This package is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it, and
only to check or to time the CPU baseline; the shipped package
``diffrender_tutorials_amd`` never imports it.

``liboracle.so`` is built from ``cpu_ref.cpp`` by ``make -C oracle`` (done by
``__graft_entry__.build()``).  Each C entry point cites the reference function it
restates; see ``cpu_ref.h``.

Parity status of each function (how it is pinned):
  * orc_hits / orc_render (BF tracer, SIL / VCOL / TEXTURE / LAMBERT-reference):
    bit-exact against golden vectors from the reference compiled from its own
    sources (oracle/_ref), tests/golden/.
  * orc_d_render (interior + edge sampling, REF_PRNG14): golden gradients from the
    reference; tolerance stated in the tests (summation order differs).
  * orc_qmc_table / orc_prng_table / orc_hammersley / orc_edge_sampler / camera
    matrices: bit-exact against golden vectors.
  * LAMBERT with a non-reference light (config-2 light): the reference hard-codes
    its light, so only the light constants differ from the pinned reference preset;
    checked by the closed-form Lambert test in tests/test_oracle.py.
  * CPU BVH tracer: pinned against the BF tracer (same hits).
  * orc_optsimple_step: pinned by the reference OptSimple run (golden vectors).
  * orc_adam_step (standard Adam): pinned against torch.optim.Adam golden vectors.
  * orc_silhouette_functional / orc_fd_silhouette (FP64 box-filtered silhouette and its
    central differences): pinned by the closed-form area of a cube's front face and by
    analytic derivatives of that area (tests/test_oracle.py).
  * lambert_functional / fd_lambert (FP64 pixel-area functional of the Lambert image and its
    central differences): pinned by the silhouette functional (ambient-only light: the
    functional is ambient x albedo x covered area), by Gauss-Legendre order convergence and by
    the oracle's own render at 256 samples per pixel (tests/test_oracle_lambert_fd.py).
  * mt_barycentrics / barycentric_grad (the MT barycentrics of a ray and their derivatives
    with respect to the triangle's vertices, the reference's BarU_grad / BarV_grad): pinned
    against FP64 central differences of mt_barycentrics (tests/test_oracle.py).
  * restricted_matrix / restricted_jacobian (TransformR::to_mat and the Jacobian the
    reference takes with Enzyme, which this image lacks, so the reference cannot be run
    there): the matrix is pinned against LiteMath's product of its factors at special
    angles and the Jacobian against FP64 central differences of the matrix
    (tests/test_oracle.py).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

SILHOUETTE, VERTEX_COLOR, TEXTURE_COLOR, LAMBERT = 1, 2, 3, 4
LAMBERT_VERTEX_COLOR = 8   # LAMBERT with the vertex colour as albedo (no counterpart in the reference)
RNG_REF_PRNG14, RNG_QMC = 0, 1
TRACER_BF, TRACER_BVH = 0, 1


class _Cam(ctypes.Structure):
    _fields_ = [("world_view", ctypes.c_float * 16), ("proj", ctypes.c_float * 16),
                ("width", ctypes.c_float), ("height", ctypes.c_float)]


class _Lam(ctypes.Structure):
    _fields_ = [("light_dir", ctypes.c_float * 3), ("light_color", ctypes.c_float * 3),
                ("ambient", ctypes.c_float * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C oracle`")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_scene_create.restype = ctypes.c_void_p
        _lib.orc_loss.restype = ctypes.c_float
        _lib.orc_qmc_rnd.restype = ctypes.c_float
        _lib.orc_silhouette_functional.restype = ctypes.c_double
        _lib.orc_lambert_functional.restype = ctypes.c_double
    return _lib


def _p(a, t=ctypes.c_float):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(t))


def cam_struct(cam) -> _Cam:
    c = _Cam()
    c.world_view[:] = [float(x) for x in np.asarray(cam.world_view, np.float32)]
    c.proj[:] = [float(x) for x in np.asarray(cam.proj, np.float32)]
    c.width = float(cam.width)
    c.height = float(cam.height)
    return c


def lam_struct(lam):
    if lam is None or lam.get("light_dir") is None:
        return None
    s = _Lam()
    s.light_dir[:] = [float(x) for x in lam["light_dir"]]
    s.light_color[:] = [float(x) for x in lam["light_color"]]
    s.ambient[:] = [float(x) for x in lam["ambient"]]
    return ctypes.byref(s)


class Scene:
    """Flattened scene (prepare_for_render, common_src/scene_utils.cpp:98-173)."""

    def __init__(self, meshes, tracer=TRACER_BF):
        L = lib()
        nm = len(meshes)
        self.meshes = meshes
        vc = np.array([m.V for m in meshes], np.int32)
        fc = np.array([m.F for m in meshes], np.int32)
        tr = np.concatenate([np.asarray(m.transform, np.float32).reshape(16) for m in meshes])
        pos = np.ascontiguousarray(np.concatenate([m.pos for m in meshes]).astype(np.float32))
        has = lambda a: all(getattr(m, a) is not None for m in meshes)
        col = np.ascontiguousarray(np.concatenate([m.col for m in meshes]).astype(np.float32)) if has("col") else None
        tc = np.ascontiguousarray(np.concatenate([m.tc for m in meshes]).astype(np.float32)) if has("tc") else None
        nrm = np.ascontiguousarray(np.concatenate([m.nrm for m in meshes]).astype(np.float32)) if has("nrm") else None
        idx = np.ascontiguousarray(np.concatenate([m.idx for m in meshes]).astype(np.uint32))
        tw = np.array([0 if m.tex is None else m.tex.shape[1] for m in meshes], np.int32)
        th = np.array([0 if m.tex is None else m.tex.shape[0] for m in meshes], np.int32)
        texs = [m.tex.reshape(-1) for m in meshes if m.tex is not None]
        tex = np.ascontiguousarray(np.concatenate(texs).astype(np.float32)) if texs else None
        diff = np.array([1 if m.differentiable else 0 for m in meshes], np.int32)
        self._keep = (vc, fc, tr, pos, col, tc, nrm, idx, tw, th, tex, diff)
        self.h = L.orc_scene_create(nm, _p(vc, ctypes.c_int), _p(fc, ctypes.c_int), _p(tr), _p(pos), _p(col),
                                    _p(tc), _p(nrm), _p(idx, ctypes.c_uint32), _p(tw, ctypes.c_int),
                                    _p(th, ctypes.c_int), _p(tex), _p(diff, ctypes.c_int))
        self.h = ctypes.c_void_p(self.h)
        self.set_tracer(tracer)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.orc_scene_destroy(self.h)
            self.h = None

    def set_tracer(self, kind):
        lib().orc_scene_set_tracer(self.h, kind)

    def set_jitter(self, on, seed=0):
        """Jittered pixel samples of `seed` (the device's drh_set_pixel_jitter pattern)."""
        lib().orc_set_jitter(self.h, 1 if on else 0, ctypes.c_uint32(seed & 0xFFFFFFFF))

    def grad_size(self, mode):
        return lib().orc_grad_size(self.h, mode)

    def hits(self, cam):
        n = cam.width * cam.height
        prim = np.empty(n, np.uint32); geom = np.empty(n, np.uint32)
        t = np.empty(n, np.float32); u = np.empty(n, np.float32); v = np.empty(n, np.float32)
        c = cam_struct(cam)
        lib().orc_hits(self.h, ctypes.byref(c), _p(prim, ctypes.c_uint32), _p(geom, ctypes.c_uint32), _p(t), _p(u), _p(v))
        return prim, geom, t, u, v

    def render(self, mode, spp, cams, lam=None):
        cs = (_Cam * len(cams))(*[cam_struct(c) for c in cams])
        out = np.zeros(sum(c.width * c.height * 3 for c in cams), np.float32)
        lib().orc_render(self.h, mode, spp, lam_struct(lam), cs, len(cams), _p(out))
        res, off = [], 0
        for c in cams:
            res.append(out[off:off + c.width * c.height * 3].reshape(c.height, c.width, 3))
            off += c.width * c.height * 3
        return res

    def d_render(self, mode, spp, cams, adjoints, edge_samples, rng_mode=RNG_REF_PRNG14, lam=None, grad=None,
                 aligned=False):
        """aligned: interior samples at the render's grid offsets instead of Hammersley points."""
        cs = (_Cam * len(cams))(*[cam_struct(c) for c in cams])
        adj = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1) for a in adjoints]))
        if grad is None:
            grad = np.zeros(self.grad_size(mode), np.float32)
        lib().orc_d_render_ex(self.h, mode, spp, lam_struct(lam), cs, len(cams), _p(adj), edge_samples, rng_mode,
                              1 if aligned else 0, _p(grad))
        return grad

    def silhouette_functional(self, cam, adjoint):
        """FP64 box-filtered silhouette functional sum_p (sum of adjoint channels)_p * covered
        area of pixel p (closed convex mesh; front-facing triangles)."""
        c = cam_struct(cam)
        adj = np.ascontiguousarray(np.asarray(adjoint, np.float32).reshape(-1))
        return lib().orc_silhouette_functional(self.h, ctypes.byref(c), _p(adj))

    def fd_silhouette(self, cam, adjoint, coords, h=1e-4):
        """FP64 central differences of silhouette_functional for coordinates 3*vertex+axis."""
        c = cam_struct(cam)
        adj = np.ascontiguousarray(np.asarray(adjoint, np.float32).reshape(-1))
        co = np.ascontiguousarray(np.asarray(coords, np.int32))
        out = np.empty(len(co), np.float64)
        lib().orc_fd_silhouette(self.h, ctypes.byref(c), _p(adj), len(co), _p(co, ctypes.c_int), ctypes.c_double(h),
                                _p(out, ctypes.c_double))
        return out

    def lambert_functional(self, mode, cam, adjoint, lam=None, order=6):
        """FP64 pixel-area functional sum_p adj_p . integral of the LAMBERT / LAMBERT_VERTEX_COLOR
        shade over pixel p (closed convex mesh; front-facing triangles)."""
        c = cam_struct(cam)
        adj = np.ascontiguousarray(np.asarray(adjoint, np.float32).reshape(-1))
        return lib().orc_lambert_functional(self.h, mode, lam_struct(lam), ctypes.byref(c), _p(adj), order)

    def fd_lambert(self, mode, cam, adjoint, coords, h=1e-4, lam=None, order=6):
        """FP64 central differences of lambert_functional for coordinates 3*vertex+axis."""
        c = cam_struct(cam)
        adj = np.ascontiguousarray(np.asarray(adjoint, np.float32).reshape(-1))
        co = np.ascontiguousarray(np.asarray(coords, np.int32))
        out = np.empty(len(co), np.float64)
        lib().orc_fd_lambert(self.h, mode, lam_struct(lam), ctypes.byref(c), _p(adj), len(co), _p(co, ctypes.c_int),
                             ctypes.c_double(h), order, _p(out, ctypes.c_double))
        return out

    def edge_sampler(self, cam):
        c = cam_struct(cam)
        n = lib().orc_edge_sampler(self.h, ctypes.byref(c), None, None, None, None, None)
        v0 = np.empty(n, np.int32); v1 = np.empty(n, np.int32)
        pmf = np.empty(n, np.float32); cdf = np.empty(n + 1, np.float32)
        proj = np.empty(3 * sum(m.V for m in self.meshes), np.float32)
        lib().orc_edge_sampler(self.h, ctypes.byref(c), _p(v0, ctypes.c_int), _p(v1, ctypes.c_int), _p(pmf), _p(cdf), _p(proj))
        return v0, v1, pmf, cdf, proj.reshape(-1, 3)


def pixel_sample(on, seed, view, x, y, si, sq):
    """Offsets (xo, yo) in the pixel of sample si (orc_pixel_sample)."""
    xo, yo = ctypes.c_float(), ctypes.c_float()
    lib().orc_pixel_sample(1 if on else 0, ctypes.c_uint32(seed & 0xFFFFFFFF), view, x, y, si, sq,
                           ctypes.byref(xo), ctypes.byref(yo))
    return xo.value, yo.value


def loss(image, target):
    img = np.ascontiguousarray(image, np.float32).reshape(-1)
    tgt = np.ascontiguousarray(target, np.float32).reshape(-1)
    adj = np.empty_like(img)
    l = lib().orc_loss(_p(img), _p(tgt), img.size // 3, _p(adj))
    return float(l), adj.reshape(np.asarray(image).shape)


def qmc_table():
    t = np.empty(11 * 31, np.uint32)
    lib().orc_qmc_table(_p(t, ctypes.c_uint32))
    return t.reshape(11, 31)


def qmc_rnd(pos, dim, table):
    tab = np.ascontiguousarray(table, np.uint32).reshape(-1)
    return lib().orc_qmc_rnd(ctypes.c_uint32(pos), dim, _p(tab, ctypes.c_uint32))


def prng_table(n):
    out = np.empty(2 * n, np.float32)
    lib().orc_prng_table(n, _p(out))
    return out.reshape(n, 2)


def hammersley(n):
    out = np.empty(2 * n, np.float32)
    lib().orc_hammersley(n, _p(out))
    return out.reshape(n, 2)


def inverse4x4(m):
    a = np.ascontiguousarray(m, np.float32).reshape(16)
    out = np.empty(16, np.float32)
    lib().orc_inverse4x4(_p(a), _p(out))
    return out


def mat_mul(a, b):
    a = np.ascontiguousarray(a, np.float32).reshape(16)
    b = np.ascontiguousarray(b, np.float32).reshape(16)
    out = np.empty(16, np.float32)
    lib().orc_mat_mul(_p(a), _p(b), _p(out))
    return out


def optsimple_step(grad, m, G, block_start, block_lr):
    grad = np.ascontiguousarray(grad, np.float32).copy()
    bs = np.ascontiguousarray(block_start, np.int32)
    bl = np.ascontiguousarray(block_lr, np.float32)
    lib().orc_optsimple_step(grad.size, _p(grad), _p(m), _p(G), bs.size, _p(bs, ctypes.c_int), _p(bl))
    return grad


def adam_step(param, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    lib().orc_adam_step(param.size, _p(param), _p(np.ascontiguousarray(grad, np.float32)), _p(m), _p(v),
                        step, ctypes.c_float(lr), ctypes.c_float(b1), ctypes.c_float(b2), ctypes.c_float(eps))


def restricted_matrix(x):
    """TransformR::to_mat (common_src/scene.h:99-104) in FP64: translate(t) * rotZ(rz) *
    rotY(ry) * rotX(rx) * scale(s) with LiteMath's rotation conventions (LiteMath.h:1491-1519);
    returns the 4x4 matrix (row-major)."""
    x = np.asarray(x, np.float64)
    cx, sx, cy, sy, cz, sz = np.cos(x[3]), np.sin(x[3]), np.cos(x[4]), np.sin(x[4]), np.cos(x[5]), np.sin(x[5])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    M = np.eye(4)
    M[:3, :3] = Rz @ Ry @ Rx * x[6]
    M[:3, 3] = x[:3]
    return M


def restricted_jacobian(x):
    """d(rows 0-2 of the matrix, row-major 12)/dx, 7x12 in FP64: the Jacobian
    dmat_dtransforms_jac (src/drender_autodiff.cpp:78-107) computes with Enzyme, written out:
    rows 0-2 of column 3 are t, columns 0-2 are s * Rz Ry Rx."""
    x = np.asarray(x, np.float64)
    cx, sx, cy, sy, cz, sz = np.cos(x[3]), np.sin(x[3]), np.cos(x[4]), np.sin(x[4]), np.cos(x[5]), np.sin(x[5])
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    dRx = np.array([[0, 0, 0], [0, -sx, -cx], [0, cx, -sx]])
    dRy = np.array([[-sy, 0, cy], [0, 0, 0], [-cy, 0, -sy]])
    dRz = np.array([[-sz, -cz, 0], [cz, -sz, 0], [0, 0, 0]])
    J = np.zeros((7, 3, 4))
    for r in range(3):
        J[r, r, 3] = 1.0
    J[3, :, :3] = x[6] * (Rz @ Ry @ dRx)
    J[4, :, :3] = x[6] * (Rz @ dRy @ Rx)
    J[5, :, :3] = x[6] * (dRz @ Ry @ Rx)
    J[6, :, :3] = Rz @ Ry @ Rx
    return J.reshape(7, 12)


def mt_barycentrics(o, d, A, B, C):
    """Moller-Trumbore barycentrics in FP64 (raytrace_src/raytrace3d_bf.cpp:123-150 without the
    determinant clamp): u = weight of C, v = weight of B."""
    o, d, A, B, C = (np.asarray(x, np.float64) for x in (o, d, A, B, C))
    e1, e2 = B - A, C - A
    pvec = np.cross(d, e2)
    det = e1 @ pvec
    tvec = o - A
    return float(np.cross(tvec, e1) @ d / det), float(tvec @ pvec / det)


def barycentric_grad(o, d, A, B, C):
    """(du, dv): each 3x3, row k the derivative with respect to vertex k (A, B, C), closed form
    (the quantities BarU_grad / BarV_grad compute, src/functions.cpp:226-470)."""
    o, d, A, B, C = (np.asarray(x, np.float64) for x in (o, d, A, B, C))
    e1, e2 = B - A, C - A
    tvec, pvec, e1xd = o - A, np.cross(d, e2), np.cross(e1, d)
    det = e1 @ pvec
    u, v = mt_barycentrics(o, d, A, B, C)
    du = np.zeros((3, 3))
    dv = np.zeros((3, 3))
    du[1] = (np.cross(d, tvec) - u * pvec) / det
    du[2] = -u * e1xd / det
    du[0] = -(e1xd / det + du[1] + du[2])
    dv[1] = -v * pvec / det
    dv[2] = (np.cross(tvec, d) - v * e1xd) / det
    dv[0] = -(pvec / det + dv[1] + dv[2])
    return du, dv
