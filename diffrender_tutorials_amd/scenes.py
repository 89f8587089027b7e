"""Synthetic scenes and cameras for the benchmark configurations.

Host-side data generation only (numpy); nothing here is on the GPU path.

Meshes are returned as :class:`Mesh` (float32 positions / colours / tc / normals,
uint32 triangle indices, optional RGB float texture), the same per-mesh data the
reference's ``TriangleMesh`` holds (common_src/scene.h:40-70).  Cameras are
:class:`Camera` with the two matrices the reference's ``CamInfo`` carries
(common_src/camera.h:6-44) in float4x4 memory order (column-major).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Mesh:
    pos: np.ndarray                      # (V, 3) float32
    idx: np.ndarray                      # (F, 3) uint32
    col: np.ndarray | None = None        # (V, 3) float32
    tc: np.ndarray | None = None         # (V, 2) float32
    nrm: np.ndarray | None = None        # (V, 3) float32
    tex: np.ndarray | None = None        # (h, w, 3) float32
    transform: np.ndarray = dataclasses.field(
        default_factory=lambda: np.eye(4, dtype=np.float32).reshape(-1))  # 16 floats, column-major
    differentiable: bool = True
    # RESTRICTED_TRANSFORM mesh (TransformR, common_src/scene.h:91-116): 7 float32 parameters
    # translate xyz, rotate xyz (radians), scale; the transform is their matrix and the
    # gradient holds their 7 derivatives.  None: a plain 4x4 transform.
    restricted: np.ndarray | None = None

    @property
    def V(self) -> int:
        return int(self.pos.shape[0])

    @property
    def F(self) -> int:
        return int(self.idx.shape[0])

    def copy(self) -> "Mesh":
        return Mesh(**{f.name: (None if getattr(self, f.name) is None else
                                (getattr(self, f.name).copy() if isinstance(getattr(self, f.name), np.ndarray)
                                 else getattr(self, f.name)))
                       for f in dataclasses.fields(self)})


@dataclasses.dataclass
class Camera:
    world_view: np.ndarray   # 16 float32, column-major (float4x4 memory order)
    proj: np.ndarray         # 16 float32
    width: int
    height: int


# ---------------------------------------------------------------------------
# LiteMath-style matrices (common_src/LiteMath.h:1474,1679-1727), float32
# ---------------------------------------------------------------------------
f32 = np.float32


def perspective(fovy_deg: float, aspect: float, znear: float, zfar: float) -> np.ndarray:
    """perspectiveMatrix (common_src/LiteMath.h:1706-1727), column-major 16 floats."""
    ymax = f32(f32(znear) * f32(math.tan(float(f32(f32(fovy_deg) * f32(3.14159265358979323846)) / f32(360.0)))))
    xmax = f32(ymax * f32(aspect))
    left, right, bottom, top = -xmax, xmax, -ymax, ymax
    temp = f32(f32(2.0) * f32(znear))
    temp2 = f32(right - left)
    temp3 = f32(top - bottom)
    temp4 = f32(f32(zfar) - f32(znear))
    m = np.zeros(16, dtype=np.float32)
    m[0] = temp / temp2
    m[5] = temp / temp3
    m[8] = f32(right + left) / temp2
    m[9] = f32(top + bottom) / temp3
    m[10] = f32(-f32(zfar) - f32(znear)) / temp4
    m[11] = -1.0
    m[14] = f32(-temp * f32(zfar)) / temp4
    return m


def translate(t) -> np.ndarray:
    m = np.eye(4, dtype=np.float32).reshape(-1).copy()
    m[12:15] = np.asarray(t, dtype=np.float32)
    return m


def rotate_y(deg: float) -> np.ndarray:
    """rotate4x4Y (common_src/LiteMath.h:1501-1509), column-major float32."""
    phi = np.float32(np.deg2rad(deg))
    c, s = np.cos(phi), np.sin(phi)
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[0, 2], m[2, 0], m[2, 2] = c, -s, s, c     # rows of m are the columns
    return m.reshape(-1)


def mat_mul(a, b) -> np.ndarray:
    """Column-major a * b (float32)."""
    a4 = np.asarray(a, np.float32).reshape(4, 4)
    b4 = np.asarray(b, np.float32).reshape(4, 4)
    return (b4 @ a4).astype(np.float32).reshape(-1)


def _normalize(v):
    v = np.asarray(v, dtype=np.float32)
    n = f32(math.sqrt(float(f32(v[0] * v[0] + v[1] * v[1]) + f32(v[2] * v[2]))))
    return v * (f32(1.0) / n)


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]],
                    dtype=np.float32)


def look_at(eye, center, up) -> np.ndarray:
    """lookAt (common_src/LiteMath.h:1679-1704), column-major 16 floats."""
    eye = np.asarray(eye, dtype=np.float32)
    center = np.asarray(center, dtype=np.float32)
    z = _normalize(eye - center)
    y = np.asarray(up, dtype=np.float32)
    x = _cross(y, z)
    y = _cross(z, x)
    x = _normalize(x)
    y = _normalize(y)
    m = np.zeros(16, dtype=np.float32)
    m[0:4] = [x[0], y[0], z[0], 0]
    m[4:8] = [x[1], y[1], z[1], 0]
    m[8:12] = [x[2], y[2], z[2], 0]
    m[12] = -x[0] * eye[0] - x[1] * eye[1] - x[2] * eye[2]
    m[13] = -y[0] * eye[0] - y[1] * eye[1] - y[2] * eye[2]
    m[14] = -z[0] * eye[0] - z[1] * eye[1] - z[2] * eye[2]
    m[15] = 1
    return m


def camera_front(width: int, height: int, distance: float = 3.0, fov: float = 45.0) -> Camera:
    """Pinhole camera at `distance` on +z looking at the origin (the reference tests'
    camera: translate4x4(0,0,-3) and perspectiveMatrix(45, w/h, 0.1, 100), src/tests.cpp:39-44)."""
    return Camera(translate([0, 0, -distance]), perspective(fov, width / height, 0.1, 100.0), width, height)


def camera_look(eye, width: int, height: int, fov: float = 45.0, center=(0, 0, 0), up=(0, 1, 0)) -> Camera:
    return Camera(look_at(eye, center, up), perspective(fov, width / height, 0.1, 100.0), width, height)


def cameras_ring(n: int, width: int, height: int, radius: float = 3.0, elev_deg: float = 20.0) -> list[Camera]:
    """n cameras evenly on a ring around the y axis at the given elevation."""
    cams = []
    e = math.radians(elev_deg)
    for i in range(n):
        a = 2.0 * math.pi * i / n
        eye = (radius * math.cos(e) * math.sin(a), radius * math.sin(e), radius * math.cos(e) * math.cos(a))
        cams.append(camera_look(eye, width, height))
    return cams


# ---------------------------------------------------------------------------
# Meshes
# ---------------------------------------------------------------------------
def cube(half: float = 0.5, seed: int = 1) -> Mesh:
    """Unit cube, 8 vertices / 12 triangles, per-vertex colours ~U(0,1)."""
    h = half
    pos = np.array([[-h, -h, -h], [h, -h, -h], [h, h, -h], [-h, h, -h],
                    [-h, -h, h], [h, -h, h], [h, h, h], [-h, h, h]], dtype=np.float32)
    idx = np.array([[4, 5, 6], [4, 6, 7],   # +z
                    [1, 0, 3], [1, 3, 2],   # -z
                    [5, 1, 2], [5, 2, 6],   # +x
                    [0, 4, 7], [0, 7, 3],   # -x
                    [7, 6, 2], [7, 2, 3],   # +y
                    [0, 1, 5], [0, 5, 4]],  # -y
                   dtype=np.uint32)
    rng = np.random.default_rng(seed)
    col = rng.uniform(0, 1, size=(8, 3)).astype(np.float32)
    nrm = pos / np.linalg.norm(pos, axis=1, keepdims=True)
    return Mesh(pos=pos, idx=idx, col=col, nrm=nrm.astype(np.float32))


def icosphere(subdiv: int, radius: float = 1.0) -> Mesh:
    """Icosphere: icosahedron with `subdiv` 1->4 midpoint subdivisions projected to the
    sphere.  subdiv=5 gives 10242 vertices / 20480 triangles."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache: dict = {}
        nf = []

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    p = np.array(verts, dtype=np.float64)
    pos = (p * radius).astype(np.float32)
    nrm = p.astype(np.float32)
    return Mesh(pos=pos, idx=np.array(faces, dtype=np.uint32), nrm=nrm)


def cube_sphere(n: int, radius: float = 1.0) -> Mesh:
    """A cube whose faces are n x n grids of squares (two triangles each), every vertex
    pushed onto the sphere: F = 12 n^2 triangles, V = 6 n^2 + 2 vertices (n = 1 is the
    12-triangle cube, n = 81 gives 78,732 triangles, n = 129 gives 199,692).  Outward winding,
    normals = positions / radius."""
    n = int(n)
    if n < 1:
        raise ValueError("n >= 1")
    g = np.arange(n + 1)
    I, J = np.meshgrid(g, g, indexing="ij")
    pts, quads = [], []
    for axis in range(3):
        u, v = [a for a in range(3) if a != axis]
        for side in (0, n):
            P = np.zeros((n + 1, n + 1, 3), np.int64)
            P[..., axis] = side
            P[..., u] = I
            P[..., v] = J
            base = sum(len(p) for p in pts)
            pts.append(P.reshape(-1, 3))
            k = np.arange((n + 1) * (n + 1)).reshape(n + 1, n + 1) + base
            a, b, c, d = k[:-1, :-1], k[1:, :-1], k[1:, 1:], k[:-1, 1:]
            quads.append(np.stack([a, b, c, d], -1).reshape(-1, 4))
    P = np.concatenate(pts)
    key = (P[:, 0] * (n + 1) + P[:, 1]) * (n + 1) + P[:, 2]
    uniq, inv = np.unique(key, return_inverse=True)
    Q = inv[np.concatenate(quads)]
    tris = np.concatenate([Q[:, [0, 1, 2]], Q[:, [0, 2, 3]]])
    grid = np.stack([uniq // ((n + 1) * (n + 1)), (uniq // (n + 1)) % (n + 1), uniq % (n + 1)], -1)
    p = grid.astype(np.float64) * (2.0 / n) - 1.0
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    e1, e2 = p[tris[:, 1]] - p[tris[:, 0]], p[tris[:, 2]] - p[tris[:, 0]]
    inward = np.sum(np.cross(e1, e2) * (p[tris[:, 0]] + p[tris[:, 1]] + p[tris[:, 2]]), axis=1) < 0
    tris[inward] = tris[inward][:, [0, 2, 1]]
    return Mesh(pos=(p * radius).astype(np.float32), idx=tris.astype(np.uint32), nrm=p.astype(np.float32))


def cube_sphere_for(F: int, max_triangles: int) -> Mesh:
    """The cube_sphere whose 12 n^2 triangles come closest to F, at most max_triangles."""
    n = max(1, int(round(math.sqrt(F / 12.0))))
    while n > 1 and 12 * n * n > max_triangles:
        n -= 1
    return cube_sphere(n)


def one_ring(mesh: Mesh) -> list[np.ndarray]:
    nb = [set() for _ in range(mesh.V)]
    for a, b, c in mesh.idx:
        nb[a].update((b, c)); nb[b].update((a, c)); nb[c].update((a, b))
    return [np.array(sorted(s), dtype=np.int64) for s in nb]


def vertex_normals(pos: np.ndarray, idx: np.ndarray) -> np.ndarray:
    p = pos.astype(np.float64)
    fn = np.cross(p[idx[:, 1]] - p[idx[:, 0]], p[idx[:, 2]] - p[idx[:, 0]])
    n = np.zeros_like(p)
    for k in range(3):
        np.add.at(n, idx[:, k], fn)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-20)
    # orient outward for closed star-shaped meshes around the origin
    flip = np.sum(n * p, axis=1) < 0
    n[flip] *= -1
    return n.astype(np.float32)


def radial_displace(mesh: Mesh, sigma: float, seed: int, smooth: bool = True) -> Mesh:
    """Move every vertex along its radius by sigma*N(0,1), optionally averaged over the
    one-ring (the icosphere-5 fitting target); normals recomputed from the new geometry."""
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(mesh.V)
    if smooth:
        rings = one_ring(mesh)
        d = np.array([(d[i] + d[r].sum()) / (1 + len(r)) for i, r in enumerate(rings)])
    p = mesh.pos.astype(np.float64)
    r = np.linalg.norm(p, axis=1, keepdims=True)
    newp = p + (p / np.maximum(r, 1e-20)) * (sigma * d)[:, None]
    out = mesh.copy()
    out.pos = newp.astype(np.float32)
    out.nrm = vertex_normals(out.pos, out.idx)
    return out


def perturb(mesh: Mesh, sigma: float, seed: int) -> Mesh:
    """Vertices + N(0, sigma^2) (the unit-cube fitting target)."""
    rng = np.random.default_rng(seed)
    out = mesh.copy()
    out.pos = (mesh.pos + rng.normal(0.0, sigma, size=mesh.pos.shape)).astype(np.float32)
    return out


def with_albedo(mesh: Mesh, albedo: float = 0.8) -> Mesh:
    """Constant albedo as a 1x1 texture (tc = 0), the form LAMBERT shading reads."""
    out = mesh.copy()
    out.tc = np.zeros((mesh.V, 2), dtype=np.float32)
    out.tex = np.full((1, 1, 3), albedo, dtype=np.float32)
    return out


def with_uv_texture(mesh: Mesh, w: int, h: int, seed: int) -> Mesh:
    """Spherical (u, v) texture coordinates and a random RGB texture."""
    out = mesh.copy()
    p = mesh.pos.astype(np.float64)
    r = np.maximum(np.linalg.norm(p, axis=1), 1e-20)
    u = 0.5 + np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi)
    v = 0.5 - np.arcsin(np.clip(p[:, 1] / r, -1, 1)) / np.pi
    out.tc = np.stack([u, v], axis=1).astype(np.float32)
    out.tex = np.random.default_rng(seed).uniform(0, 1, size=(h, w, 3)).astype(np.float32)
    return out


def with_random_colors(mesh: Mesh, seed: int) -> Mesh:
    out = mesh.copy()
    out.col = np.random.default_rng(seed).uniform(0, 1, size=(mesh.V, 3)).astype(np.float32)
    return out


# ---------------------------------------------------------------------------
# Benchmark configurations
# ---------------------------------------------------------------------------
def config1():
    """Cube, vertex colours, 256x256, spp 1, 2^16 edge samples; target perturbed N(0, 0.05^2)."""
    m = cube(0.5, seed=1)
    target = perturb(m, 0.05, seed=2)
    cam = camera_front(256, 256)
    return dict(meshes=[m], target=[target], cams=[cam], mode=2, spp=1, edge_samples=1 << 16)


def config2(subdiv: int = 5, res: int = 1024, edge_samples: int = 1 << 20):
    """Icosphere-5 (10242 V / 20480 F), Lambert with constant albedo 0.8, 1024x1024, spp 4,
    2^20 edge samples; target displaced radially by 0.1*N(0,1) smoothed over one-ring."""
    m = with_albedo(icosphere(subdiv, 1.0), 0.8)
    target = with_albedo(radial_displace(icosphere(subdiv, 1.0), 0.1, seed=3), 0.8)
    cam = camera_front(res, res)
    return dict(meshes=[m], target=[target], cams=[cam], mode=4, spp=4, edge_samples=edge_samples)


LAMBERT_REFERENCE = dict(light_dir=None, light_color=None, ambient=None)


def lambert_config2():
    """Config-2 light: one directional light toward (1,1,1)/sqrt3, white, ambient 0.1."""
    d = -_normalize(np.array([1.0, 1.0, 1.0], dtype=np.float32))
    return dict(light_dir=d, light_color=np.ones(3, np.float32), ambient=np.full(3, 0.1, np.float32))


def config3_jobs(n: int = 64, seed: int = 5):
    """Config 3's uneven batch: per job a triangle count F = round(exp(U(ln 12, ln 80000)))
    (seed 5) met by the closest subdivided cube on the sphere (cube_sphere, 12 n^2 triangles,
    at most 80,000: from the 12-triangle cube to 78,732), resolution uniform in {128, 256,
    512}, vertex-colour or Lambert shading at random, edge samples = 4 x pixels; the target is
    the mesh with radial noise 0.05.  Returns dicts with meshes, target, cams, mode (2 VCOL /
    4 LAMBERT), spp, edge_samples and the drawn triangle count F."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        F = int(round(math.exp(rng.uniform(math.log(12), math.log(80000)))))
        res = int(rng.choice([128, 256, 512]))
        mode = 2 if rng.uniform() < 0.5 else 4
        az = rng.uniform(0, 2 * math.pi)
        jobs.append(config3_job(k, F, res, mode, az))
    return jobs


def config3_job(k: int, F: int, res: int, mode: int, az: float) -> dict:
    """One config-3 job: the cube_sphere closest to F triangles (at most 80,000), a target
    with radial noise 0.05 (seed 200 + k), vertex colours (mode 2, seeds 300 + k / 400 + k) or
    albedo 0.8 (mode 4), one camera at azimuth `az` and distance 3, res x res, spp 1, edge
    samples = 4 x pixels."""
    base = cube_sphere_for(F, 80000)
    target = radial_displace(base, 0.05, seed=200 + k, smooth=False)
    if mode == 2:
        m, t = with_random_colors(base, 300 + k), with_random_colors(target, 400 + k)
    else:
        m, t = with_albedo(base, 0.8), with_albedo(target, 0.8)
    eye = (3.0 * math.sin(az), 0.8, 3.0 * math.cos(az))
    cam = camera_look(eye, res, res)
    return dict(meshes=[m], target=[t], cams=[cam], mode=mode, spp=1, edge_samples=4 * res * res, F=F)


def config5_jobs(n: int = 32, seed: int = 7):
    """Config 5's serving stream: per job a triangle count log-uniform in [12, 200k] (the
    closest cube_sphere, at most 200,000 triangles: from the 12-triangle cube to 199,692),
    resolution uniform in {128, 256, 512, 1024}, 1-8 views on a ring, 50-500 Adam iterations,
    Lambert shading, radial-noise-0.05 targets, N = 4 x pixels of one view.  `n` jobs with
    seed 7; arrival times are drawn by the caller."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n):
        F = int(round(math.exp(rng.uniform(math.log(12), math.log(200000)))))
        res = int(rng.choice([128, 256, 512, 1024]))
        views = int(rng.integers(1, 9))
        iters = int(rng.integers(50, 501))
        jobs.append(config5_job(k, F, res, views, iters))
    return jobs


def config5_job(k: int, F: int, res: int, views: int, iters: int) -> dict:
    """One config-5 job: the cube_sphere closest to F triangles (at most 200,000) with albedo
    0.8, a target with radial noise 0.05 (seed 600 + k), `views` ring cameras at res x res,
    Lambert, spp 1, edge samples = 4 x pixels of one view, `iters` Adam iterations."""
    base = cube_sphere_for(F, 200000)
    target = radial_displace(base, 0.05, seed=600 + k, smooth=False)
    cams = cameras_ring(views, res, res)
    return dict(meshes=[with_albedo(base, 0.8)], target=[with_albedo(target, 0.8)], cams=cams, mode=4,
                spp=1, edge_samples=4 * res * res, iterations=iters, F=F)


def uv_sphere(n_lat: int, n_lon: int, radius: float = 1.0) -> Mesh:
    """Latitude-longitude sphere: two poles, n_lat - 1 rings of n_lon vertices, outward
    winding; F = 2 * n_lon * (n_lat - 1) (n_lat 224, n_lon 448: 199,808 triangles, 99,906
    vertices -- config 4's ~200k-triangle sphere)."""
    th = np.linspace(0.0, math.pi, n_lat + 1)[1:-1]
    ph = np.linspace(0.0, 2 * math.pi, n_lon, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pos = np.concatenate([[[0.0, 1.0, 0.0]], ring, [[0.0, -1.0, 0.0]]]) * radius
    nr = n_lat - 1
    S_ = pos.shape[0] - 1
    ring_id = lambda r, k: 1 + r * n_lon + (k % n_lon)
    tris = []
    for k in range(n_lon):
        tris.append((0, ring_id(0, k + 1), ring_id(0, k)))
        tris.append((S_, ring_id(nr - 1, k), ring_id(nr - 1, k + 1)))
    r = np.arange(nr - 1)[:, None]
    k = np.arange(n_lon)[None, :]
    a, b = 1 + r * n_lon + k, 1 + r * n_lon + (k + 1) % n_lon
    c, d = a + n_lon, b + n_lon
    quads = np.stack([np.stack([a, b, c], -1), np.stack([b, d, c], -1)], -2).reshape(-1, 3)
    idx = np.concatenate([np.array(tris), quads]).astype(np.uint32)
    pos = pos.astype(np.float32)
    return Mesh(pos=pos, idx=idx, nrm=vertex_normals(pos, idx))

