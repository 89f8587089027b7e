"""Python binding of libdrh.so (ctypes over the C ABI in include/drh.h).

Mirrors the reference's plugin interface ``IDiffRender`` (src/virtual_drender.h:14-26):
``MakeDifferentialRenderer(settings)`` returns a renderer with ``commit``, ``render``,
``d_render`` and ``d_render_and_compare``; scenes are lists of
:class:`diffrender_tutorials_amd.scenes.Mesh`, cameras :class:`scenes.Camera`, images
``(H, W, 3)`` float32 arrays, gradients the flat ``DScene`` vector.

The shared library is the product path.  There is no fallback: a missing library or
a missing GPU raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DRH_LIB: load another build of the library (e.g. to compare builds side by side)
LIB_PATH = os.environ.get("DRH_LIB") or os.path.join(HERE, "libdrh.so")

SILHOUETTE, VERTEX_COLOR, TEXTURE_COLOR, LAMBERT = 1, 2, 3, 4
LAMBERT_VERTEX_COLOR = 8   # LAMBERT with the vertex colour as albedo (no counterpart in the reference)
RNG_REF_PRNG14, RNG_QMC = 0, 1
TRACER_BVH, TRACER_BF = 0, 1
EDGE_ORDER_AUTO, EDGE_ORDER_SAMPLE, EDGE_ORDER_BINNED = 0, 1, 2
EDGE_GRAD_REFERENCE_DIAGONAL, EDGE_GRAD_FULL_JACOBIAN = 0, 1
OPT_SIMPLE, OPT_ADAM = 0, 1
HITS_RECOMPUTE, HITS_STORE, HITS_AUTO = 0, 1, 2
SIDE_TEST_REFERENCE, SIDE_TEST_WATERTIGHT = 0, 1
TRAVERSAL_AUTO, TRAVERSAL_WAVE, TRAVERSAL_LANE = 0, 1, 2
JOB_DETERMINISTIC, JOB_ALIGNED_INTERIOR, JOB_FULL_JACOBIAN = 1, 2, 4

STATUS = {0: "DRH_OK", -1: "DRH_E_INVALID", -2: "DRH_E_NOT_COMMITTED", -3: "DRH_E_HIP", -4: "DRH_E_OOM",
          -5: "DRH_E_UNSUPPORTED", -6: "DRH_E_DEVICE"}

# every symbol include/drh.h declares
EXPORTS = ["drh_create", "drh_destroy", "drh_last_error", "drh_commit", "drh_grad_size", "drh_render",
           "drh_d_render", "drh_d_render_and_compare", "drh_render_dev", "drh_d_render_dev",
           "drh_d_render_and_compare_dev", "drh_opt_init", "drh_opt_step", "drh_get_params",
           "drh_opt_set_texture_lr", "drh_opt_set_transform_lr", "drh_get_transforms", "drh_get_restricted", "drh_get_texels", "drh_set_device_sharing", "drh_set_interior_samples", "drh_set_pixel_jitter", "drh_set_edge_side_test", "drh_opt_steps", "drh_set_traversal",
           "drh_stream",
           "drh_sync", "drh_profile_enable", "drh_profile_read", "drh_debug_hits", "drh_debug_bvh",
           "drh_debug_edges", "drh_malloc", "drh_free", "drh_memcpy_htod", "drh_memcpy_dtoh",
           "drh_debug_copy", "drh_debug_cdf", "drh_opt_use_graph",
           "drh_server_create", "drh_server_submit", "drh_server_wait", "drh_server_stats", "drh_server_destroy",
           "drh_server_job_bytes", "drh_set_edge_order", "drh_set_bvh_rebuild", "drh_set_edge_grad_mode",
           "drh_set_deterministic", "drh_set_interior_pos_grad", "drh_build_edge_sampler_noalloc",
           "drh_set_hit_buffer_limit", "drh_get_hit_buffer_bytes", "drh_count_rays", "drh_read_ray_counts"]


class DrhError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Settings(ctypes.Structure):
    """DiffRenderSettings (src/virtual_drender.h:8-12) plus this path's options."""
    _fields_ = [("mode", ctypes.c_int), ("spp", ctypes.c_int), ("rng_mode", ctypes.c_int), ("tracer", ctypes.c_int),
                ("device", ctypes.c_int), ("light_custom", ctypes.c_int), ("light_dir", ctypes.c_float * 3),
                ("light_color", ctypes.c_float * 3), ("ambient", ctypes.c_float * 3)]


class CameraC(ctypes.Structure):
    _fields_ = [("world_view", ctypes.c_float * 16), ("proj", ctypes.c_float * 16),
                ("width", ctypes.c_float), ("height", ctypes.c_float)]


class SceneDesc(ctypes.Structure):
    _fields_ = [("n_meshes", ctypes.c_int), ("vcount", ctypes.POINTER(ctypes.c_int)),
                ("fcount", ctypes.POINTER(ctypes.c_int)), ("transforms", ctypes.POINTER(ctypes.c_float)),
                ("pos", ctypes.POINTER(ctypes.c_float)), ("col", ctypes.POINTER(ctypes.c_float)),
                ("tc", ctypes.POINTER(ctypes.c_float)), ("nrm", ctypes.POINTER(ctypes.c_float)),
                ("idx", ctypes.POINTER(ctypes.c_uint32)), ("tex_w", ctypes.POINTER(ctypes.c_int)),
                ("tex_h", ctypes.POINTER(ctypes.c_int)), ("texels", ctypes.POINTER(ctypes.c_float)),
                ("differentiable", ctypes.POINTER(ctypes.c_int)), ("restricted", ctypes.POINTER(ctypes.c_int)),
                ("restricted_params", ctypes.POINTER(ctypes.c_float))]


_lib = None


def load():
    """Load libdrh.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m diffrender_tutorials_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.drh_last_error.restype = ctypes.c_char_p
        L.drh_stream.restype = ctypes.c_void_p
        L.drh_create.argtypes = [ctypes.POINTER(Settings), ctypes.POINTER(ctypes.c_void_p)]
        for n in EXPORTS:
            f = getattr(L, n)
            if n not in ("drh_last_error", "drh_stream", "drh_destroy", "drh_server_destroy", "drh_server_job_bytes"):
                f.restype = ctypes.c_int
        L.drh_destroy.restype = None
        L.drh_server_destroy.restype = None
        L.drh_server_create.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.drh_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_int)]
        L.drh_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.POINTER(JobResult)]
        L.drh_server_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServerStats)]
        L.drh_server_destroy.argtypes = [ctypes.c_void_p]
        L.drh_server_job_bytes.restype = ctypes.c_size_t
        L.drh_server_job_bytes.argtypes = [ctypes.POINTER(JobDesc)]
        L.drh_malloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.drh_free.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.drh_memcpy_htod.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.drh_memcpy_dtoh.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.drh_debug_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_longlong]
        L.drh_set_hit_buffer_limit.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
        L.drh_get_hit_buffer_bytes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        L.drh_build_edge_sampler_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                      ctypes.c_void_p]
        _lib = L
    return _lib


def _fp(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _ip(a):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def camera_struct(cam) -> CameraC:
    c = CameraC()
    c.world_view[:] = [float(x) for x in np.asarray(cam.world_view, np.float32).reshape(16)]
    c.proj[:] = [float(x) for x in np.asarray(cam.proj, np.float32).reshape(16)]
    c.width = float(cam.width)
    c.height = float(cam.height)
    return c


def cameras_array(cams):
    return (CameraC * len(cams))(*[camera_struct(c) for c in cams])


class DiffRender:
    """IDiffRender over libdrh (one handle = one HIP stream + its device buffers)."""

    def __init__(self, mode=VERTEX_COLOR, spp=1, rng_mode=RNG_REF_PRNG14, tracer=TRACER_BVH, device=0, light=None):
        L = load()
        s = Settings()
        s.mode, s.spp, s.rng_mode, s.tracer, s.device = mode, spp, rng_mode, tracer, device
        if light is not None and light.get("light_dir") is not None:
            s.light_custom = 1
            s.light_dir[:] = [float(x) for x in light["light_dir"]]
            s.light_color[:] = [float(x) for x in light["light_color"]]
            s.ambient[:] = [float(x) for x in light["ambient"]]
        h = ctypes.c_void_p()
        st = L.drh_create(ctypes.byref(s), ctypes.byref(h))
        if st != 0:
            raise DrhError(st, "drh_create failed (no HIP device?)")
        self.h = h
        self.mode = mode
        self.spp = spp
        self._keep = None
        self.V = 0

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.drh_destroy(self.h)
            self.h = None

    __del__ = close

    def _check(self, st):
        if st != 0:
            raise DrhError(st, load().drh_last_error(self.h).decode())

    # -- IDiffRender -------------------------------------------------------
    def commit(self, meshes):
        """IDiffRender::commit (src/drender.h:91-96)."""
        L = load()
        d, keep = scene_desc(meshes)
        self._check(L.drh_commit(self.h, ctypes.byref(d)))
        self.V = int(sum(m.V for m in meshes))
        self.n_meshes = len(meshes)
        self.has_col = keep[5] is not None

    def grad_size(self):
        return load().drh_grad_size(self.h)

    def render(self, cams):
        """IDiffRender::render (src/drender.h:98-141); returns one (H, W, 3) image per camera."""
        cs = cameras_array(cams)
        out = np.empty(sum(c.width * c.height * 3 for c in cams), np.float32)
        self._check(load().drh_render(self.h, cs, len(cams), _fp(out)))
        return _split(out, cams)

    def d_render(self, cams, adjoints, edge_samples, grad=None):
        """IDiffRender::d_render (src/drender.h:144-167): accumulates into grad if given."""
        cs = cameras_array(cams)
        adj = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1) for a in adjoints]))
        acc = grad is not None
        if grad is None:
            grad = np.zeros(self.grad_size(), np.float32)
        self._check(load().drh_d_render(self.h, cs, len(cams), _fp(adj), int(edge_samples), _fp(grad), int(acc)))
        return grad

    def d_render_and_compare(self, cams, targets, edge_samples, images=True):
        """IDiffRender::d_render_and_compare (src/drender.h:169-191): returns
        (loss, grad, images); images is None when not requested (outImages = nullptr)."""
        cs = cameras_array(cams)
        tgt = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float32).reshape(-1) for a in targets]))
        grad = np.zeros(self.grad_size(), np.float32)
        imgs = np.empty_like(tgt) if images else None
        loss = ctypes.c_float()
        self._check(load().drh_d_render_and_compare(self.h, cs, len(cams), _fp(tgt), int(edge_samples), _fp(grad),
                                                    _fp(imgs) if images else None, ctypes.byref(loss)))
        return float(loss.value), grad, (_split(imgs, cams) if images else None)

    # -- device-resident optimisation --------------------------------------
    def opt_init(self, kind=OPT_SIMPLE, pos_lr=0.1, col_lr=0.1):
        self._check(load().drh_opt_init(self.h, kind, ctypes.c_float(pos_lr), ctypes.c_float(col_lr)))

    def opt_step_dev(self, cams_c, views, d_targets_ptr, edge_samples, d_loss_ptr=None):
        self._check(load().drh_opt_step(self.h, cams_c, views, ctypes.c_void_p(d_targets_ptr), int(edge_samples),
                                        ctypes.c_void_p(d_loss_ptr) if d_loss_ptr else None))

    def opt_steps_dev(self, cams_c, views, d_targets_ptr, edge_samples, d_loss_ptr, steps):
        """`steps` fused steps, captured together as one graph (drh_opt_steps)."""
        self._check(load().drh_opt_steps(self.h, cams_c, views, ctypes.c_void_p(d_targets_ptr), int(edge_samples),
                                         ctypes.c_void_p(d_loss_ptr) if d_loss_ptr else None, int(steps)))

    def set_interior_samples(self, aligned, store=False):
        """aligned: render-grid interior samples; store: HITS_RECOMPUTE / HITS_STORE / HITS_AUTO."""
        self._check(load().drh_set_interior_samples(self.h, int(aligned), int(store)))

    def set_edge_side_test(self, test):
        """SIDE_TEST_REFERENCE (Moller-Trumbore, the default) or SIDE_TEST_WATERTIGHT for the edge
        samples' side rays."""
        self._check(load().drh_set_edge_side_test(self.h, int(test)))

    def set_traversal(self, mode):
        """TRAVERSAL_AUTO (default), TRAVERSAL_WAVE (a wave walks the LBVH as one packet) or
        TRAVERSAL_LANE (each ray walks alone); the hits are the same in every mode."""
        self._check(load().drh_set_traversal(self.h, int(mode)))

    def set_pixel_jitter(self, on, seed=0):
        """Stratified jitter of the pixel samples with pattern `seed` (the fused step's t-th step
        uses seed + t); off: the reference's cell centres."""
        self._check(load().drh_set_pixel_jitter(self.h, int(bool(on)), ctypes.c_uint(seed & 0xFFFFFFFF)))

    def set_hit_buffer_limit(self, nbytes):
        self._check(load().drh_set_hit_buffer_limit(self.h, int(nbytes)))

    def hit_buffer_bytes(self):
        n = ctypes.c_size_t()
        self._check(load().drh_get_hit_buffer_bytes(self.h, ctypes.byref(n)))
        return n.value

    def set_device_sharing(self, handles):
        """How many handles share the device (1: this handle overlaps its own passes on a second
        stream; more: it leaves that to the other handles)."""
        self._check(load().drh_set_device_sharing(self.h, int(handles)))

    def set_edge_order(self, order):
        """EDGE_ORDER_AUTO / _SAMPLE (rnd0 slot order) / _BINNED (kept samples by 8x8 pixel bin)."""
        self._check(load().drh_set_edge_order(self.h, int(order)))

    def set_edge_grad_mode(self, mode):
        """EDGE_GRAD_REFERENCE_DIAGONAL (the reference's edge_grad) or EDGE_GRAD_FULL_JACOBIAN."""
        self._check(load().drh_set_edge_grad_mode(self.h, int(mode)))

    def set_bvh_rebuild(self, every):
        """Full LBVH build every `every` fused steps (1: each step), box refits between."""
        self._check(load().drh_set_bvh_rebuild(self.h, int(every)))

    def opt_set_texture_lr(self, lr):
        self._check(load().drh_opt_set_texture_lr(self.h, ctypes.c_float(lr)))

    def set_interior_pos_grad(self, on):
        """VERTEX_COLOR: add the interior term's vertex position gradient (1) or not (0, default)."""
        self._check(load().drh_set_interior_pos_grad(self.h, int(on)))

    def set_deterministic(self, on):
        """Fixed-point gradient sums, bit-reproducible run to run (1) or float atomics (0, default)."""
        self._check(load().drh_set_deterministic(self.h, int(on)))

    def opt_set_transform_lr(self, lr):
        """OptimizerParameters::transforms_lr with _OPT_TFM: 0 keeps the transforms fixed."""
        self._check(load().drh_opt_set_transform_lr(self.h, ctypes.c_float(lr)))

    def get_transforms(self):
        """The current transforms, (n_meshes, 16) float32, column-major as Mesh.transform."""
        out = np.empty(16 * self.n_meshes, np.float32)
        self._check(load().drh_get_transforms(self.h, _fp(out)))
        return out.reshape(-1, 16)

    def get_restricted(self):
        """The current TransformR parameters, (n_meshes, 7): translate, rotate, scale."""
        out = np.empty(7 * self.n_meshes, np.float32)
        self._check(load().drh_get_restricted(self.h, _fp(out)))
        return out.reshape(-1, 7)

    def get_texels(self, n):
        out = np.empty(n, np.float32)
        self._check(load().drh_get_texels(self.h, _fp(out)))
        return out

    def get_params(self):
        pos = np.empty(3 * self.V, np.float32)
        col = np.empty(3 * self.V, np.float32)
        self._check(load().drh_get_params(self.h, _fp(pos), _fp(col)))
        return pos.reshape(-1, 3), col.reshape(-1, 3)

    # -- diagnostics ---------------------------------------------------------
    def hits(self, cam):
        n = cam.width * cam.height
        prim = np.empty(n, np.uint32); geom = np.empty(n, np.uint32)
        t = np.empty(n, np.float32); u = np.empty(n, np.float32); v = np.empty(n, np.float32)
        c = camera_struct(cam)
        u32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self._check(load().drh_debug_hits(self.h, ctypes.byref(c), u32(prim), u32(geom), _fp(t), _fp(u), _fp(v)))
        return prim, geom, t, u, v

    def bvh(self, F):
        keys = np.empty(F, np.uint32); order = np.empty(F, np.uint32); b = np.empty(6, np.float32)
        u32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self._check(load().drh_debug_bvh(self.h, u32(keys), u32(order), _fp(b)))
        return keys, order, b

    def edges(self):
        n = ctypes.c_int()
        self._check(load().drh_debug_edges(self.h, ctypes.byref(n), None, None, None, None))
        v0 = np.empty(n.value, np.int32); v1 = np.empty(n.value, np.int32)
        pmf = np.empty(n.value, np.float32); cdf = np.empty(n.value + 1, np.float32)
        self._check(load().drh_debug_edges(self.h, ctypes.byref(n), _ip(v0), _ip(v1), _fp(pmf), _fp(cdf)))
        return v0, v1, pmf, cdf

    def count_rays(self, on=True):
        """Run the counting copies of the traversal kernels (eagerly) from now on (zeroes the counts)."""
        self._check(load().drh_count_rays(self.h, int(on)))

    def ray_counts(self):
        """Rays cast since counting was turned on: dict of render camera / shadow, interior
        closest-hit / any-hit, edge side / shadow rays."""
        c = (ctypes.c_ulonglong * 6)()
        self._check(load().drh_read_ray_counts(self.h, c))
        keys = ("render_camera", "render_shadow", "interior_closest", "interior_any", "edge_side", "edge_shadow")
        return dict(zip(keys, [int(x) for x in c]))

    def profile(self, on=True):
        self._check(load().drh_profile_enable(self.h, int(on)))

    def profile_read(self, name):
        ms = ctypes.c_double(); cnt = ctypes.c_int()
        self._check(load().drh_profile_read(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    # -- device memory plumbing ------------------------------------------------
    def device_array(self, a):
        """Upload a float32 numpy array; returns a DeviceBuffer (freed with the buffer)."""
        a = np.ascontiguousarray(a, np.float32)
        b = DeviceBuffer(self, a.nbytes)
        b.upload(a)
        return b

    def sync(self):
        self._check(load().drh_sync(self.h))

    def stream(self):
        return load().drh_stream(self.h)


class DeviceBuffer:
    """HBM allocation on a renderer's device (drh_malloc / drh_free)."""

    def __init__(self, r: DiffRender, nbytes: int):
        self.r = r
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        r._check(load().drh_malloc(r.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        self.r._check(load().drh_memcpy_htod(self.r.h, self.ptr, a.ctypes.data, a.nbytes))

    def download(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        self.r._check(load().drh_memcpy_dtoh(self.r.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None and getattr(self.r, "h", None) is not None:
            load().drh_free(self.r.h, self.ptr)
        self.ptr = None

    __del__ = free


def scene_desc(meshes):
    """drh_scene_desc over concatenated mesh arrays; returns (desc, arrays kept alive)."""
    n = len(meshes)
    vc = np.array([m.V for m in meshes], np.int32)
    fc = np.array([m.F for m in meshes], np.int32)
    tr = np.ascontiguousarray(np.concatenate([np.asarray(m.transform, np.float32).reshape(16) for m in meshes]))
    pos = np.ascontiguousarray(np.concatenate([m.pos for m in meshes]).astype(np.float32))
    has = lambda a: all(getattr(m, a) is not None for m in meshes)
    cat = lambda a: np.ascontiguousarray(np.concatenate([getattr(m, a) for m in meshes]).astype(np.float32)) if has(a) else None
    col, tc, nrm = cat("col"), cat("tc"), cat("nrm")
    idx = np.ascontiguousarray(np.concatenate([m.idx for m in meshes]).astype(np.uint32))
    tw = np.array([0 if m.tex is None else m.tex.shape[1] for m in meshes], np.int32)
    th = np.array([0 if m.tex is None else m.tex.shape[0] for m in meshes], np.int32)
    texs = [np.asarray(m.tex, np.float32).reshape(-1) for m in meshes if m.tex is not None]
    tex = np.ascontiguousarray(np.concatenate(texs)) if texs else None
    df = np.array([1 if m.differentiable else 0 for m in meshes], np.int32)
    rs = getattr  # RESTRICTED_TRANSFORM meshes carry 7 parameters in Mesh.restricted
    rf = np.array([0 if rs(m, "restricted", None) is None else 1 for m in meshes], np.int32)
    rp = np.ascontiguousarray(np.concatenate([np.zeros(7, np.float32) if rs(m, "restricted", None) is None
                                              else np.asarray(m.restricted, np.float32).reshape(7) for m in meshes]))
    if not rf.any():
        rf, rp = None, None
    d = SceneDesc(n, _ip(vc), _ip(fc), _fp(tr), _fp(pos), _fp(col), _fp(tc), _fp(nrm),
                  idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), _ip(tw), _ip(th), _fp(tex), _ip(df),
                  _ip(rf), _fp(rp))
    return d, (vc, fc, tr, pos, col, tc, nrm, idx, tw, th, tex, df, rf, rp)


class JobDesc(ctypes.Structure):
    _fields_ = [("settings", Settings), ("scene", SceneDesc), ("cams", ctypes.POINTER(CameraC)), ("views", ctypes.c_int),
                ("targets", ctypes.POINTER(ctypes.c_float)), ("edge_samples", ctypes.c_int),
                ("iterations", ctypes.c_int), ("lr_pos", ctypes.c_float), ("lr_col", ctypes.c_float),
                ("flags", ctypes.c_int)]


class JobResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("iterations", ctypes.c_int), ("final_loss", ctypes.c_float),
                ("queue_ms", ctypes.c_double), ("run_ms", ctypes.c_double), ("pos", ctypes.POINTER(ctypes.c_float)),
                ("col", ctypes.POINTER(ctypes.c_float)), ("latencies_ms", ctypes.POINTER(ctypes.c_float)),
                ("latencies_cap", ctypes.c_int), ("hbm_bytes", ctypes.c_size_t), ("hit_bytes", ctypes.c_size_t)]


class ServerStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_longlong), ("jobs_done", ctypes.c_int), ("jobs_failed", ctypes.c_int),
                ("jobs_queued", ctypes.c_int), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("peak_active", ctypes.c_int), ("mean_active", ctypes.c_double), ("hbm_used", ctypes.c_size_t),
                ("elapsed_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("step_s", ctypes.c_double),
                ("finish_s", ctypes.c_double)]


def settings_struct(mode, spp, rng_mode=RNG_REF_PRNG14, tracer=TRACER_BVH, device=0, light=None):
    s = Settings()
    s.mode, s.spp, s.rng_mode, s.tracer, s.device = mode, spp, rng_mode, tracer, device
    if light is not None and light.get("light_dir") is not None:
        s.light_custom = 1
        s.light_dir[:] = [float(x) for x in light["light_dir"]]
        s.light_color[:] = [float(x) for x in light["light_color"]]
        s.ambient[:] = [float(x) for x in light["ambient"]]
    return s


class Server:
    """Continuous-batching server (drh_server_*): submit mesh-fitting jobs, wait for results."""

    def __init__(self, device=0, hbm_budget=64 << 30, max_active=4):
        L = load()
        h = ctypes.c_void_p()
        st = L.drh_server_create(device, int(hbm_budget), max_active, ctypes.byref(h))
        if st != 0:
            raise DrhError(st, "drh_server_create failed")
        self.h = h
        self.V = {}

    def submit(self, meshes, cams, targets, iterations, edge_samples, mode=LAMBERT, spp=1, light=None,
               lr_pos=1e-3, lr_col=1e-3, rng_mode=RNG_REF_PRNG14, flags=0):
        d, keep = scene_desc(meshes)
        cs = cameras_array(cams)
        tg = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).reshape(-1) for t in targets]))
        j = JobDesc(settings_struct(mode, spp, rng_mode, light=light), d, cs, len(cams), _fp(tg), int(edge_samples),
                    int(iterations), float(lr_pos), float(lr_col), int(flags))
        self.last_job_bytes = load().drh_server_job_bytes(ctypes.byref(j))
        jid = ctypes.c_int()
        st = load().drh_server_submit(self.h, ctypes.byref(j), ctypes.byref(jid))
        if st != 0:
            raise DrhError(st, "drh_server_submit failed")
        self.V[jid.value] = (int(sum(m.V for m in meshes)), int(iterations))
        return jid.value

    def wait(self, job_id, timeout=600.0):
        V, iters = self.V[job_id]
        pos = np.empty(3 * V, np.float32)
        col = np.empty(3 * V, np.float32)
        lat = np.empty(iters, np.float32)
        r = JobResult(0, 0, 0.0, 0.0, 0.0, _fp(pos), _fp(col), _fp(lat), iters, 0, 0)
        st = load().drh_server_wait(self.h, job_id, float(timeout), ctypes.byref(r))
        return {"status": st, "iterations": r.iterations, "final_loss": r.final_loss, "queue_ms": r.queue_ms,
                "run_ms": r.run_ms, "pos": pos.reshape(-1, 3), "col": col.reshape(-1, 3),
                "latencies_ms": lat[:r.iterations], "hbm_bytes": r.hbm_bytes, "hit_bytes": r.hit_bytes}

    def stats(self):
        s = ServerStats()
        load().drh_server_stats(self.h, ctypes.byref(s))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.drh_server_destroy(self.h)
            self.h = None

    __del__ = close


def _split(flat, cams):
    res, off = [], 0
    for c in cams:
        n = c.width * c.height * 3
        res.append(flat[off:off + n].reshape(c.height, c.width, 3))
        off += n
    return res


def MakeDifferentialRenderer(mode=VERTEX_COLOR, spp=1, **kw) -> DiffRender:
    """MakeDifferentialRenderer (src/drender.cpp:89-147)."""
    return DiffRender(mode=mode, spp=spp, **kw)
