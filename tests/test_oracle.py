"""The oracle (oracle/cpu_ref.cpp) pinned against the reference's own outputs (golden
vectors written from the reference compiled from its sources) and against closed forms.
CPU only."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from diffrender_tutorials_amd import scenes as S
from golden_cases import SMALL_CASES, load, mesh_hash, rel_l2, scene, sha


def test_sequences_match_reference():
    fx = load("sequences")
    assert np.array_equal(O.qmc_table(), fx["qmc"])
    t = O.prng_table(1 << 16).reshape(-1)
    assert np.array_equal(t[:2048], fx["prng65536_head"])
    assert sha(t) == str(fx["prng65536_sha"])
    assert sha(O.prng_table(1 << 20).reshape(-1)) == str(fx["prng1m_sha"])
    assert np.array_equal(O.prng_table(1000).reshape(-1), fx["prng1000"])
    for n, key in ((1, "hamm1"), (4, "hamm4"), (16, "hamm16")):
        assert np.array_equal(O.hammersley(n).reshape(-1), fx[key])


def test_qmc_closed_forms():
    tab = O.qmc_table()
    # dimension 0 uses p(x) = x: its generator matrix is the identity, i.e. the
    # van der Corput radical inverse in base 2
    assert all(int(tab[0, b]) == 1 << (30 - b) for b in range(31))
    for pos in (1, 2, 3, 5, 1000, 123456):
        rev = int(f"{pos:031b}"[::-1], 2)
        assert O.qmc_rnd(pos, 0, tab) == np.float32(np.float32(rev + 1) * np.float32(2.0 ** -31))
    # dimensions 0 and 1 (polynomials x and x+1, the two the edge sampler draws) form a
    # (0,2)-sequence in base 2: 2^k points starting at a multiple of 2^k put exactly one
    # point in every elementary box [i/2^a, (i+1)/2^a) x [j/2^b, (j+1)/2^b), a + b = k
    def res(i, d):
        r = 0
        for b in range(31):
            if (i >> b) & 1:
                r ^= int(tab[d, b])
        return r

    for k in (4, 6, 8):
        for start in (0, 1 << k, 7 << k):
            pts = [(res(i, 0), res(i, 1)) for i in range(start, start + (1 << k))]
            for a in range(k + 1):
                assert len({(x >> (31 - a), y >> (31 - (k - a))) for x, y in pts}) == 1 << k


def _prng_stream(seed, draws):
    """RandomGenInit(seed) then `draws` rndFloat values, from the recurrence
    (common_src/qmc.h:81-87, common_src/qmc.cpp:250-268) in Python integers."""
    M = 0xFFFFFFFF
    st = [(seed * (seed * seed * 15731 + 74323) + 871483) & M, (seed * (seed * seed * 13734 + 37828) + 234234) & M]

    def nxt():
        v = (st[0] * 17 + st[1] * 13123) & M
        st[0] = ((v << 13) & M) ^ v
        st[1] ^= (v << 7) & M
        return v

    for _ in range(seed % 7):
        nxt()
    out = []
    for _ in range(draws):
        v = nxt()
        tmp = (v * ((v * v * 15731 + 74323) & M) + 871483) & M
        out.append(np.float32(np.float32(tmp) * np.float32(1.0 / 4294967296.0)))
    return out


def test_prng_closed_form():
    # 14 streams seeded 7777 + t*t + 1, thread t owning a contiguous block of samples
    n = 1000
    q, r = divmod(n, 14)
    tab = O.prng_table(n)
    for t in (0, 1, 5, 13):
        start = t * q + min(t, r)
        out = _prng_stream(7777 + t * t + 1, 6)
        assert list(tab[start:start + 3].reshape(-1)) == out, t


def test_camera_matrices_match_reference():
    for name in SMALL_CASES:
        fx = load(name)
        _, _, cam = scene(name)
        assert np.array_equal(O.inverse4x4(cam.proj), fx["proj_inv"])
        assert np.array_equal(O.inverse4x4(cam.world_view), fx["wv_inv"])
        assert np.array_equal(O.mat_mul(cam.proj, cam.world_view), fx["wvp"])
        # inverse is an inverse
        m = np.asarray(cam.world_view, np.float64).reshape(4, 4).T
        inv = np.asarray(fx["wv_inv"], np.float64).reshape(4, 4).T
        assert np.allclose(m @ inv, np.eye(4), atol=1e-5)


@pytest.mark.parametrize("name", SMALL_CASES)
def test_render_hits_edges_match_reference(name):
    fx = load(name)
    meshes, _, cam = scene(name)
    for i, m in enumerate(meshes):
        assert mesh_hash(m) == str(fx[f"mesh{i}_sha"])
    for tracer in (O.TRACER_BF, O.TRACER_BVH):
        sc = O.Scene(meshes, tracer)
        img = sc.render(int(fx["mode"]), int(fx["spp"]), [cam])[0]
        assert sha(img) == str(fx["image_sha"]), (name, tracer)
        prim, geom, t, u, v = sc.hits(cam)
        assert sha(prim) == str(fx["hit_prim_sha"])
        assert sha(u) == str(fx["hit_u_sha"]) and sha(v) == str(fx["hit_v_sha"]) and sha(t) == str(fx["hit_t_sha"])
    v0, v1, pmf, cdf, proj = sc.edge_sampler(cam)
    assert sha(v0) == str(fx["edge_v0_sha"]) and sha(v1) == str(fx["edge_v1_sha"])
    assert sha(pmf) == str(fx["edge_pmf_sha"]) and sha(cdf) == str(fx["edge_cdf_sha"])
    assert sha(proj) == str(fx["proj_sha"])


@pytest.mark.parametrize("name", SMALL_CASES)
def test_gradients_match_reference(name):
    fx = load(name)
    meshes, targets, cam = scene(name)
    mode, spp, N = int(fx["mode"]), int(fx["spp"]), int(fx["edge_samples"])
    sc = O.Scene(meshes, O.TRACER_BVH)
    img = sc.render(mode, spp, [cam])[0]
    tgt = O.Scene(targets, O.TRACER_BVH).render(mode, spp, [cam])[0]
    assert sha(tgt) == str(fx["target_sha"])
    loss, adj = O.loss(img, tgt)
    assert sha(adj) == str(fx["adjoint_sha"])
    g = sc.d_render(mode, spp, [cam], [adj], N)
    # identical samples; only the order of float accumulation differs from the reference
    assert rel_l2(g, fx["grad"]) < 1e-5, rel_l2(g, fx["grad"])
    assert np.isclose(loss / (cam.width * cam.height), float(fx["compare_loss"][0]), rtol=1e-6)
    assert rel_l2(g, fx["compare_grad"]) < 1e-5


def test_optsimple_run_matches_reference():
    """One OptSimple iteration (GD_Adam, positions + colours), composed from oracle pieces:
    d_render_and_compare, OptStep, OptUpdateScene (src/optimizer_simple.cpp:272-309)."""
    fx = load("cfg1_cube_vcol")
    meshes, targets, cam = scene("cfg1_cube_vcol")
    m = meshes[0].copy()
    tgt = O.Scene(targets).render(2, 1, [cam])[0]
    nv = m.V
    G = 6 * nv + 12
    mom = np.zeros(G, np.float32)
    gsq = np.zeros(G, np.float32)
    for it in range(int(fx["opt_iters"])):
        sc = O.Scene([m], O.TRACER_BVH)
        img = sc.render(2, 1, [cam])[0]
        _, adj = O.loss(img, tgt)
        g = sc.d_render(2, 1, [cam], [adj], cam.width * cam.height)
        step = O.optsimple_step(g, mom, gsq, [0, 3 * nv, 6 * nv], [0.1, 0.1, 0.1])
        m.pos = (m.pos - step[:3 * nv].reshape(-1, 3)).astype(np.float32)
        m.col = (m.col - step[3 * nv:6 * nv].reshape(-1, 3)).astype(np.float32)
    assert np.allclose(m.pos.reshape(-1), fx["opt_pos"], rtol=0, atol=1e-7)
    assert np.allclose(m.col.reshape(-1), fx["opt_col"], rtol=0, atol=1e-7)


def _rows_to_transform(T, rows):
    t = np.asarray(T, np.float32).reshape(4, 4).copy()     # t[c, r] = T(r, c)
    for r in range(3):
        for c in range(4):
            t[c, r] = rows[4 * r + c]
    return t.reshape(-1)


@pytest.mark.parametrize("name", ["optall_tfm_sil_1", "optall_tfm_sil_3", "optall_tex_2"])
def test_optsimple_opt_all_matches_reference(name):
    """OptSimple with OPT_ALL flags (transforms and textures move too), composed from oracle
    pieces, against the reference's OptSimple::Run: OptStep scales each block by its rate and
    OptUpdateScene subtracts the d_tr block from rows 0-2 of the transform
    (src/optimizer_simple.cpp:92-186, 188-236).  Position rates are 0 here, so the stale
    inverse transform the reference keeps for d_pos does not enter."""
    from golden_cases import OPT_ALL_CASES, opt_all_scene
    fx = load(name)
    _, mode, spp, iters, (plr, tlr, trlr) = OPT_ALL_CASES[name]
    meshes, targets, cam = opt_all_scene(name)
    for i, m in enumerate(meshes):
        assert mesh_hash(m) == str(fx[f"mesh{i}_sha"])
    tgt = O.Scene(targets).render(mode, spp, [cam])[0]
    assert sha(tgt) == str(fx["target_sha"])
    ms = [m.copy() for m in meshes]
    m0 = ms[0]
    nv = m0.V
    texo = 3 * nv
    ntex = 0 if m0.tex is None or mode not in (3, 4) else m0.tex.size
    tro = texo + ntex
    G = tro + 12
    mom = np.zeros(G, np.float32)
    gsq = np.zeros(G, np.float32)
    starts = [0] + ([texo] if ntex else []) + [tro]
    rates = [plr] + ([tlr] if ntex else []) + [trlr]
    for it in range(iters):
        sc = O.Scene(ms, O.TRACER_BVH)
        img = sc.render(mode, spp, [cam])[0]
        _, adj = O.loss(img, tgt)
        g = sc.d_render(mode, spp, [cam], [adj], cam.width * cam.height)
        assert g.size == G
        step = O.optsimple_step(g, mom, gsq, starts, rates)
        m0.pos = (m0.pos - step[:3 * nv].reshape(-1, 3)).astype(np.float32)
        if ntex:
            m0.tex = (m0.tex.reshape(-1) - step[texo:tro]).astype(np.float32).reshape(m0.tex.shape)
        rows = m0.transform.reshape(4, 4).T[:3].reshape(-1) - step[tro:tro + 12]
        m0.transform = _rows_to_transform(m0.transform, rows)
    assert np.allclose(m0.pos.reshape(-1), fx["opt_pos"], rtol=0, atol=1e-7)
    assert np.allclose(m0.transform, fx["opt_tr"][:16], rtol=0, atol=1e-6)
    for i in range(1, len(ms)):
        assert np.array_equal(ms[i].transform, fx["opt_tr"][16 * i:16 * i + 16])
    if ntex:
        assert np.allclose(m0.tex.reshape(-1), fx["opt_tex"], rtol=0, atol=1e-6)
    if trlr:
        assert np.abs(fx["opt_tr"][:16] - meshes[0].transform).max() > 0.05   # the transform did move


def test_restricted_matrix_matches_reference_to_mat():
    """TransformR::to_mat of the reference (float) on golden parameter sets."""
    fx = load("transformr")
    for x, m in zip(fx["params"], fx["to_mat"]):
        M = O.restricted_matrix(x)
        assert np.abs(M - m.reshape(4, 4).T).max() <= 3e-7 * (1 + np.abs(x).max()), x


def test_restricted_jacobian_matches_central_differences():
    """The closed-form Jacobian against FP64 central differences of the matrix."""
    rng = np.random.default_rng(3)
    xs = [np.array([0, 0, 0, 0, 0.2, 0.2, 1.0])] + [np.r_[rng.uniform(-2, 2, 6), rng.uniform(0.3, 3)] for _ in range(8)]
    h = 1e-6
    for x in xs:
        J = O.restricted_jacobian(x)
        for k in range(7):
            e = np.zeros(7)
            e[k] = h
            fd = (O.restricted_matrix(x + e)[:3] - O.restricted_matrix(x - e)[:3]).reshape(-1) / (2 * h)
            assert np.abs(J[k] - fd).max() < 1e-8, (x, k)


def test_barycentric_grad_matches_central_differences():
    """The closed-form derivatives of the MT barycentrics against FP64 central differences."""
    rng = np.random.default_rng(11)
    h = 1e-7
    for _ in range(20):
        A, B, C = rng.normal(size=(3, 3))
        o = rng.normal(size=3) * 0.2 + np.array([0.0, 0.0, 4.0])
        P = (A + B + C) / 3 + rng.normal(size=3) * 0.2
        d = (P - o) / np.linalg.norm(P - o)
        du, dv = O.barycentric_grad(o, d, A, B, C)
        V = [A, B, C]
        for k in range(3):
            for c in range(3):
                p, m = [x.copy() for x in V], [x.copy() for x in V]
                p[k][c] += h
                m[k][c] -= h
                up, vp = O.mt_barycentrics(o, d, *p)
                um, vm = O.mt_barycentrics(o, d, *m)
                assert abs(du[k, c] - (up - um) / (2 * h)) < 1e-6 * (1 + abs(du[k, c]))
                assert abs(dv[k, c] - (vp - vm) / (2 * h)) < 1e-6 * (1 + abs(dv[k, c]))


def test_adam_matches_torch_vectors():
    fx = load("adam_torch")
    p = fx["p0"].copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    for k, g in enumerate(fx["grads"]):
        O.adam_step(p, g, m, v, k + 1, 1e-3)
        assert np.allclose(p, fx["traj"][k], rtol=0, atol=2e-7)


def test_moller_trumbore_against_exact_geometry():
    """Hit / miss of a single triangle against float64 ray-triangle geometry, away from
    the triangle's edges (where float rounding decides)."""
    tri = S.Mesh(pos=np.array([[-0.6, -0.4, 0.1], [0.7, -0.5, -0.2], [0.1, 0.6, 0.3]], np.float32),
                 idx=np.array([[0, 1, 2]], np.uint32), col=np.eye(3, dtype=np.float32))
    cam = S.camera_look((0.3, 0.2, 3.0), 64, 64)
    prim, _, t, u, v = O.Scene([tri]).hits(cam)
    wv = np.asarray(cam.world_view, np.float64).reshape(4, 4).T
    pr = np.asarray(cam.proj, np.float64).reshape(4, 4).T
    pinv, wvinv = np.linalg.inv(pr), np.linalg.inv(wv)
    A, B, C = tri.pos.astype(np.float64)
    checked = 0
    for y in range(64):
        for x in range(64):
            p = pinv @ np.array([2 * (x + 0.5) / 64 - 1, -2 * (y + 0.5) / 64 + 1, 0, 1])
            d = p[:3] / p[3]
            o = (wvinv @ np.array([0, 0, 0, 1.0]))[:3]
            d = (wvinv @ np.append(100 * d / np.linalg.norm(d), 1))[:3] - o
            d /= np.linalg.norm(d)
            M = np.column_stack([-d, B - A, C - A])
            tt, bv, bu = np.linalg.solve(M, o - A)
            margin = min(bu, bv, 1 - bu - bv)
            if abs(margin) < 1e-4:
                continue
            checked += 1
            hit = margin > 0 and tt > 0
            i = y * 64 + x
            assert (prim[i] != 0xFFFFFFFF) == hit
            if hit:
                assert abs(t[i] - tt) < 1e-5 and abs(u[i] - bu) < 1e-5 and abs(v[i] - bv) < 1e-5
    assert checked > 3000


def test_lambert_closed_form_config_light():
    """A single large square facing the camera, albedo a, the config-2 light: every covered
    pixel is (Lc * max(0, n.(-L)) + ambient) * a (no occluder, so no shadow)."""
    quad = S.Mesh(pos=np.array([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], np.float32),
                  idx=np.array([[0, 1, 2], [0, 2, 3]], np.uint32),
                  nrm=np.tile(np.array([[0, 0, 1]], np.float32), (4, 1)))
    quad = S.with_albedo(quad, 0.8)
    lam = S.lambert_config2()
    cam = S.camera_front(32, 32)
    img = O.Scene([quad]).render(O.LAMBERT, 1, [cam], lam=lam)[0]
    n = np.array([0, 0, 1.0])
    expect = (lam["light_color"] * max(0.0, float(n @ -lam["light_dir"])) + lam["ambient"]) * 0.8
    covered = img.sum(-1) > 0
    assert covered.mean() > 0.3
    assert np.allclose(img[covered], expect, atol=1e-6)
    # and the reference light puts the square in shade of its hard-coded direction
    img2 = O.Scene([quad]).render(O.LAMBERT, 1, [cam])[0]
    L = np.array([-1, -1, 0]) / np.sqrt(2)
    expect2 = (np.array([1, 0.5, 0.2]) * max(0.0, float(n @ -L)) + 0.1) * 0.8
    assert np.allclose(img2[covered], expect2, atol=1e-6)


def test_edge_sampler_closed_form():
    """pmf = normalised float3 edge length (x, y in pixels, z world); cdf its running sum."""
    meshes, _, cam = scene("cube_rot_vcol")
    v0, v1, pmf, cdf, proj = O.Scene(meshes).edge_sampler(cam)
    assert len(v0) == 18 and np.all(v0 < v1)
    keys = v0.astype(np.int64) * 1000 + v1
    assert np.all(np.diff(keys) > 0)
    L = np.linalg.norm(proj[v1].astype(np.float64) - proj[v0], axis=1)
    assert np.allclose(pmf, L / L.sum(), rtol=1e-6)
    assert np.allclose(cdf[1:], np.cumsum(L) / L.sum(), rtol=1e-6) and cdf[0] == 0


def test_config2_bvh_equals_reference_bf():
    """Config 2 at full size (1024^2, 20480 triangles): the oracle's BVH tracer, with the
    clamped-determinant suspect triangles, reproduces the reference BF hit map and image."""
    fx = load("cfg2_ico5_lambert")
    meshes, _, cam = scene("cfg2_ico5_lambert")
    assert mesh_hash(meshes[0]) == str(fx["mesh0_sha"])
    sc = O.Scene(meshes, O.TRACER_BVH)
    prim, _, t, u, v = sc.hits(cam)
    assert sha(prim) == str(fx["hit_prim_sha"]) and sha(t) == str(fx["hit_t_sha"])
    assert sha(u) == str(fx["hit_u_sha"]) and sha(v) == str(fx["hit_v_sha"])
    img = sc.render(O.LAMBERT, 4, [cam])[0]
    assert sha(img) == str(fx["image_sha"])
    v0, v1, pmf, cdf, proj = sc.edge_sampler(cam)
    assert sha(cdf) == str(fx["edge_cdf_sha"]) and sha(pmf) == str(fx["edge_pmf_sha"])
    assert sha(v0) == str(fx["edge_v0_sha"]) and int(fx["n_edges"]) == 30720


# ---------------------------------------------------------------------------
# FP64 finite differences of the box-filtered silhouette (the fin_diff.cpp check)
# ---------------------------------------------------------------------------
def _cube_front_setup(W=128):
    from diffrender_tutorials_amd import scenes as S
    m = S.cube(0.5, seed=1)
    cam = S.camera_front(W, W)
    f = 1.0 / np.tan(np.radians(22.5))
    half = 0.5 * f / 2.5 * (W / 2)          # half side of the front face (z = 0.5) in pixels
    return m, cam, half


def test_box_filtered_silhouette_is_the_projected_area():
    """Adjoint 1 per pixel: the functional is the covered area, for a cube seen head-on the
    front face's square (side faces are back-facing): closed form (2*half)^2."""
    m, cam, half = _cube_front_setup()
    ref = O.Scene([m], O.TRACER_BF)
    adj = np.full((128, 128, 3), 1.0 / 3.0, np.float32)
    L = ref.silhouette_functional(cam, adj)
    assert abs(L - (2 * half) ** 2) < 1e-6 * L


def test_silhouette_central_differences_match_the_area_derivative():
    """Moving a front corner by dx changes the area by -side/2 * dsx/dx (the two edges at the
    corner); toward the camera both screen coordinates grow: side/2 * (|dsx/dz| + |dsy/dz|)."""
    m, cam, half = _cube_front_setup()
    ref = O.Scene([m], O.TRACER_BF)
    adj = np.full((128, 128, 3), 1.0 / 3.0, np.float32)
    k = [i for i in range(8) if abs(m.pos[i, 2] - 0.5) < 1e-6 and m.pos[i, 0] < 0 and m.pos[i, 1] < 0][0]
    fd = ref.fd_silhouette(cam, adj, [3 * k, 3 * k + 1, 3 * k + 2], 1e-4)
    f = 1.0 / np.tan(np.radians(22.5))
    dsx_dx = f / 2.5 * 64.0
    dsx_dz = 0.5 * f * 64.0 / 2.5 ** 2
    side = 2 * half
    assert np.allclose(fd, [-side / 2 * dsx_dx, -side / 2 * dsx_dx, side / 2 * 2 * dsx_dz], rtol=1e-6), fd


def test_oracle_edge_gradient_matches_fp64_central_differences():
    """The reference's edge-sampled SILHOUETTE gradient (the oracle restates it) against FP64
    central differences (h = 1e-4) of the box-filtered image: every vertex coordinate of the
    cube, a non-uniform adjoint.  Monte Carlo error with 2^20 samples is well under 2%."""
    m, cam, _ = _cube_front_setup()
    ref = O.Scene([m], O.TRACER_BF)
    yy, xx = np.mgrid[0:128, 0:128].astype(np.float32)
    a = (1.0 + 0.5 * np.sin(xx / 9.0) * np.cos(yy / 13.0)).astype(np.float32)
    adj = np.repeat(a[..., None] / 3.0, 3, axis=2)
    coords = np.arange(24)
    fd = ref.fd_silhouette(cam, adj, coords, 1e-4)
    g = ref.d_render(O.SILHOUETTE, 1, [cam], [adj], 1 << 20)[:24]
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert err < 0.02, (err, g, fd)
