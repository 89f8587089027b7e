// oracle/ref_harness.cpp -- drives the reference implementation, compiled from its
// own sources under /root/reference (see oracle/Makefile), to write golden vectors.
//
// TEST INFRASTRUCTURE ONLY.  This file is the harness's own code; it includes the
// reference's headers and links the reference's objects, and calls the reference's
// public entry points (MakeDifferentialRenderer, IDiffRender::render / d_render /
// d_render_and_compare, IRayTracer::CastSingleRay, build_edge_sampler, qmc::, prng::,
// OptSimple via CreateSimpleOptimizer).  Nothing here re-implements reference logic
// except the edge-set construction needed to call build_edge_sampler, which repeats
// the calls DiffRender::edge_derivatives makes (src/drender.h:243-302).
//
// Usage: ref_harness <input.bin> <output.bin>
// Input (little endian):
//   int32 magic(0x44524831), mode, spp, n_meshes, n_views, edge_samples, ops, opt_iters
//   if ops & OP_OPT_ALL: float32 position_lr, textures_lr, transforms_lr (flags OPT_ALL)
//   if ops & OP_TRANSFORMR: int32 n, float32 params[7n] (TransformR::to_mat of each)
//   per mesh: int32 nv, nf, has_col, has_tc, has_nrm, tex_w, tex_h, diff;
//             float32 transform[16]; pos[3nv]; [col[3nv]]; [tc[2nv]]; [nrm[3nv]]; uint32 idx[3nf];
//             float32 tex[tex_w*tex_h*3]
//   per view: float32 world_view[16], proj[16], width, height
//   if ops & (OP_DRENDER|OP_COMPARE|OP_OPTIMIZE): per view float32 image[H*W*3] (adjoint or target)
// Output: records {int32 tag, int32 nbytes, bytes}.  OP_OPTIMIZE also records the wall time of
// the optimisation loop (float64 seconds) and the OpenMP thread count.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "drender.h"
#include "optimizer.h"
#include "qmc.h"

using namespace diff_render;

enum {
  OP_RENDER = 1, OP_HITS = 2, OP_DRENDER = 4, OP_COMPARE = 8, OP_SEQ = 16, OP_EDGES = 32, OP_OPTIMIZE = 64,
  OP_OPT_ALL = 128, OP_TRANSFORMR = 256,
};
enum {
  T_IMAGES = 1, T_HIT_PRIM = 2, T_HIT_GEOM = 3, T_HIT_T = 4, T_HIT_U = 5, T_HIT_V = 6, T_GRAD = 7, T_LOSS = 8,
  T_QMC = 9, T_PRNG = 10, T_HAMM = 11, T_EDGE_V0 = 12, T_EDGE_V1 = 13, T_EDGE_PMF = 14, T_EDGE_CDF = 15,
  T_COMPARE_GRAD = 16, T_COMPARE_IMAGES = 17, T_PROJ = 18, T_OPT_POS = 19, T_OPT_COL = 20, T_OPT_ERR = 21,
  T_PROJ_INV = 22, T_WV_INV = 23, T_WVP = 24, T_OPT_TR = 25, T_OPT_TEX = 26, T_TRANSFORMR = 27,
  T_OPT_SECONDS = 28, T_OMP_THREADS = 29,
};

static FILE* g_in;
static FILE* g_out;
template <class T> static T rd() { T v; if (fread(&v, sizeof v, 1, g_in) != 1) { fprintf(stderr, "short read\n"); exit(2); } return v; }
template <class T> static void rdv(std::vector<T>& v, size_t n) { v.resize(n); if (n && fread(v.data(), sizeof(T), n, g_in) != n) { fprintf(stderr, "short read\n"); exit(2); } }
static void put(int tag, const void* p, size_t nbytes) {
  int32_t h[2] = {tag, (int32_t)nbytes};
  fwrite(h, sizeof h, 1, g_out);
  if (nbytes) fwrite(p, 1, nbytes, g_out);
}

int main(int argc, char** argv) {
  if (argc != 3) { fprintf(stderr, "usage: %s in out\n", argv[0]); return 1; }
  g_in = fopen(argv[1], "rb");
  g_out = fopen(argv[2], "wb");
  if (!g_in || !g_out) { perror("open"); return 1; }
  if (rd<int32_t>() != 0x44524831) { fprintf(stderr, "bad magic\n"); return 1; }
  const int mode = rd<int32_t>(), spp = rd<int32_t>(), nm = rd<int32_t>(), nviews = rd<int32_t>();
  const int edge_samples = rd<int32_t>(), ops = rd<int32_t>(), opt_iters = rd<int32_t>();
  float lr_pos = 0.1f, lr_tex = 0.1f, lr_tfm = 0.1f;
  if (ops & OP_OPT_ALL) { lr_pos = rd<float>(); lr_tex = rd<float>(); lr_tfm = rd<float>(); }
  if (ops & OP_TRANSFORMR) {
    // TransformR::to_mat of each 7-parameter set (translate, rotate, scale), float4x4 memory
    const int n = rd<int32_t>();
    std::vector<float> out;
    for (int k = 0; k < n; k++) {
      std::vector<float> x;
      rdv(x, 7);
      const TransformR t(float3(x[0], x[1], x[2]), float3(x[3], x[4], x[5]), x[6]);
      const float4x4 m = t.to_mat();
      out.insert(out.end(), (const float*)&m, (const float*)&m + 16);
    }
    put(T_TRANSFORMR, out.data(), out.size() * 4);
  }

  Scene scene;
  std::vector<int> diff_ids;
  for (int i = 0; i < nm; i++) {
    const int nv = rd<int32_t>(), nf = rd<int32_t>(), hc = rd<int32_t>(), ht = rd<int32_t>(), hn = rd<int32_t>();
    const int tw = rd<int32_t>(), th = rd<int32_t>(), df = rd<int32_t>();
    std::vector<float> T, P, C, TC, N, TX;
    std::vector<uint32_t> I;
    rdv(T, 16); rdv(P, 3 * nv);
    if (hc) rdv(C, 3 * nv);
    if (ht) rdv(TC, 2 * nv);
    if (hn) rdv(N, 3 * nv);
    rdv(I, 3 * nf);
    rdv(TX, (size_t)tw * th * 3);
    TriangleMesh m;
    for (int k = 0; k < nv; k++) m.vertices.push_back(float3(P[3 * k], P[3 * k + 1], P[3 * k + 2]));
    if (hc) for (int k = 0; k < nv; k++) m.colors.push_back(float3(C[3 * k], C[3 * k + 1], C[3 * k + 2]));
    if (ht) for (int k = 0; k < nv; k++) m.tc.push_back(float2(TC[2 * k], TC[2 * k + 1]));
    if (hn) for (int k = 0; k < nv; k++) m.normals.push_back(float3(N[3 * k], N[3 * k + 1], N[3 * k + 2]));
    for (auto x : I) m.indices.push_back(x);
    if (tw > 0) {
      CPUTexture tex;
      tex.w = tw; tex.h = th; tex.channels = 3;
      tex.m_data = TX;
      m.textures.push_back(tex);
    }
    float4x4 tr;
    memcpy((void*)&tr, T.data(), 16 * sizeof(float));
    scene.add_mesh(m, {tr});
    if (df) diff_ids.push_back(i);
  }
  std::vector<CamInfo> cams(nviews);
  for (int v = 0; v < nviews; v++) {
    std::vector<float> wv, pr;
    rdv(wv, 16); rdv(pr, 16);
    cams[v].width = rd<float>();
    cams[v].height = rd<float>();
    memcpy((void*)&cams[v].mWorldView, wv.data(), 64);
    memcpy((void*)&cams[v].mProj, pr.data(), 64);
    cams[v].commit();
  }
  std::vector<Img> given;
  if (ops & (OP_DRENDER | OP_COMPARE | OP_OPTIMIZE)) {
    for (int v = 0; v < nviews; v++) {
      const int W = (int)cams[v].width, H = (int)cams[v].height;
      std::vector<float> buf;
      rdv(buf, (size_t)W * H * 3);
      Img im(W, H);
      memcpy((void*)im.data(), buf.data(), buf.size() * sizeof(float));
      given.push_back(im);
    }
  }

  DiffRenderSettings settings;
  settings.mode = (SHADING_MODEL)mode;
  settings.spp = spp;
  auto R = MakeDifferentialRenderer(settings);

  if (ops & OP_RENDER) {
    std::vector<Img> imgs;
    for (int v = 0; v < nviews; v++) imgs.emplace_back((int)cams[v].width, (int)cams[v].height);
    R->commit(scene);
    R->render(scene, cams.data(), imgs.data(), nviews);
    std::vector<float> out;
    for (auto& im : imgs) out.insert(out.end(), (float*)im.data(), (float*)im.data() + 3 * im.width() * im.height());
    put(T_IMAGES, out.data(), out.size() * 4);
  }
  if (ops & OP_HITS) {
    scene.prepare_for_render();
    auto tracer = MakeRayTracer3D("");
    tracer->Init(&scene);
    std::vector<uint32_t> prim, geom;
    std::vector<float> t, u, vv;
    for (int v = 0; v < nviews; v++) {
      tracer->SetCamera(cams[v]);
      for (int y = 0; y < (int)cams[v].height; y++)
        for (int x = 0; x < (int)cams[v].width; x++) {
          SurfaceInfo h = tracer->CastSingleRay(x + 0.5f, y + 0.5f);
          prim.push_back(h.primId); geom.push_back(h.geomId); t.push_back(h.t); u.push_back(h.u); vv.push_back(h.v);
        }
    }
    put(T_HIT_PRIM, prim.data(), prim.size() * 4);
    put(T_HIT_GEOM, geom.data(), geom.size() * 4);
    put(T_HIT_T, t.data(), t.size() * 4);
    put(T_HIT_U, u.data(), u.size() * 4);
    put(T_HIT_V, vv.data(), vv.size() * 4);
    // camera matrices as BruteForce3D::SetCamera computes them
    std::vector<float> pinv, wvinv, wvp;
    for (int v = 0; v < nviews; v++) {
      float4x4 a = LiteMath::inverse4x4(cams[v].mProj), b = LiteMath::inverse4x4(cams[v].mWorldView);
      pinv.insert(pinv.end(), (float*)&a, (float*)&a + 16);
      wvinv.insert(wvinv.end(), (float*)&b, (float*)&b + 16);
      wvp.insert(wvp.end(), cams[v].mWVP, cams[v].mWVP + 16);
    }
    put(T_PROJ_INV, pinv.data(), pinv.size() * 4);
    put(T_WV_INV, wvinv.data(), wvinv.size() * 4);
    put(T_WVP, wvp.data(), wvp.size() * 4);
  }
  if (ops & OP_DRENDER) {
    DScene d(scene, settings.mode, diff_ids);
    R->commit(scene);
    R->d_render(scene, cams.data(), given.data(), nviews, edge_samples, d);
    std::vector<float> out;
    for (auto& dm : d.get_dmeshes()) out.insert(out.end(), dm.full_data(), dm.full_data() + dm.full_size());
    put(T_GRAD, out.data(), out.size() * 4);
  }
  if (ops & OP_COMPARE) {
    DScene d(scene, settings.mode, diff_ids);
    std::vector<Img> imgs(nviews);
    const float loss = R->d_render_and_compare(scene, cams.data(), given.data(), nviews, edge_samples, d, imgs.data());
    std::vector<float> out;
    for (auto& dm : d.get_dmeshes()) out.insert(out.end(), dm.full_data(), dm.full_data() + dm.full_size());
    put(T_COMPARE_GRAD, out.data(), out.size() * 4);
    put(T_LOSS, &loss, 4);
    std::vector<float> io;
    for (auto& im : imgs) io.insert(io.end(), (float*)im.data(), (float*)im.data() + 3 * im.width() * im.height());
    put(T_COMPARE_IMAGES, io.data(), io.size() * 4);
  }
  if (ops & OP_SEQ) {
    unsigned int table[qmc::QRNG_DIMENSIONS][qmc::QRNG_RESOLUTION];
    qmc::init(table);
    put(T_QMC, table, sizeof table);
    // the edge-sample loop's stream assignment, reproduced with the same pragma
    // (src/drender.h:306-331)
    const int MT = 14;
    prng::RandomGen gens[MT];
    for (int i = 0; i < MT; i++) gens[i] = prng::RandomGenInit(7777 + i * i + 1);
    std::vector<float> rn(2 * (size_t)edge_samples);
#pragma omp parallel for num_threads(MT)
    for (int i = 0; i < edge_samples; i++) {
      auto& gen = gens[omp_get_thread_num()];
      const float r0 = prng::rndFloat(&gen);
      const float r1 = prng::rndFloat(&gen);
      rn[2 * i] = r0; rn[2 * i + 1] = r1;
    }
    put(T_PRNG, rn.data(), rn.size() * 4);
    std::vector<float> hm(2 * spp);
    qmc::planeHammersley(hm.data(), spp);
    put(T_HAMM, hm.data(), hm.size() * 4);
  }
  if (ops & OP_EDGES) {
    // the calls DiffRender::edge_derivatives makes for view 0 (src/drender.h:243-302)
    scene.prepare_for_render();
    Scene s2d_diff, s2d_full;
    std::set<Edge> es;
    for (int i = 0; i < (int)scene.get_meshes().size(); i++) {
      for (int j = 0; j < (int)scene.get_transform(i).size(); j++) {
        TriangleMesh m = scene.get_mesh(i);
        transform(m, scene.get_transform(i)[j]);
        for (auto& v : m.vertices) { auto c = v; VertexShader(cams[0], c.x, c.y, c.z, v.M); }
        s2d_full.add_mesh(m);
        bool is_diff = false;
        for (int id : diff_ids) is_diff |= (id == i);
        if (is_diff) {
          s2d_diff.add_mesh(m);
          for (size_t k = 0; k < scene.get_mesh(i).indices.size(); k += 3) {
            auto A = scene.get_index(i, j, k), B = scene.get_index(i, j, k + 1), C = scene.get_index(i, j, k + 2);
            auto An = scene.get_vertex_n(i, k), Bn = scene.get_vertex_n(i, k + 1), Cn = scene.get_vertex_n(i, k + 2);
            es.insert(Edge(A, B, An, Bn, i, j));
            es.insert(Edge(B, C, Bn, Cn, i, j));
            es.insert(Edge(C, A, Cn, An, i, j));
          }
        }
      }
    }
    std::vector<int> map2d(scene.get_instance_id_mapping().size(), 0);
    s2d_diff.set_instance_id_mapping(map2d);
    s2d_diff.prepare_for_render();
    s2d_full.set_instance_id_mapping(map2d);
    s2d_full.prepare_for_render();
    std::vector<Edge> edges(es.begin(), es.end());
    Sampler smp = build_edge_sampler(s2d_diff, edges);
    std::vector<int32_t> v0, v1;
    for (auto& e : edges) { v0.push_back(e.v0); v1.push_back(e.v1); }
    put(T_EDGE_V0, v0.data(), v0.size() * 4);
    put(T_EDGE_V1, v1.data(), v1.size() * 4);
    put(T_EDGE_PMF, smp.pmf.data(), smp.pmf.size() * 4);
    put(T_EDGE_CDF, smp.cdf.data(), smp.cdf.size() * 4);
    std::vector<float> proj;
    TriangleMesh pm;
    s2d_full.get_prepared_mesh(pm);
    for (auto& v : pm.vertices) { proj.push_back(v.x); proj.push_back(v.y); proj.push_back(v.z); }
    put(T_PROJ, proj.data(), proj.size() * 4);
  }
  if (ops & OP_OPTIMIZE) {
    // OptSimple, GD_Adam with positions and colours only (no transform / texture updates)
    OptimizerParameters p(OptimizerParameters::GD_Adam);
    p.verbose = false;
    p.flags = OptimizerParameters::_OPT_POS | OptimizerParameters::_OPT_COL;
    if (ops & OP_OPT_ALL) {
      // every block: positions, transforms, colours, textures with the given rates
      p.flags = OptimizerParameters::OPT_ALL;
      p.position_lr = lr_pos; p.textures_lr = lr_tex; p.transforms_lr = lr_tfm;
    }
    p.differentiable_mesh_ids = diff_ids;
    IOptimizer* opt = CreateSimpleOptimizer();
    opt->Init(scene, R, cams.data(), given.data(), nviews, p);
    float err = 0.0f;
    // wall time of the reference's whole optimisation loop (commit + d_render_and_compare +
    // OptStep per iteration), for the CPU baseline
    const auto t0 = std::chrono::steady_clock::now();
    Scene res = opt->Run(opt_iters, err);
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    put(T_OPT_SECONDS, &secs, sizeof secs);
    const int threads = omp_get_max_threads();
    put(T_OMP_THREADS, &threads, sizeof threads);
    std::vector<float> pos, col;
    for (auto& v : res.get_mesh(0).vertices) { pos.push_back(v.x); pos.push_back(v.y); pos.push_back(v.z); }
    for (auto& c : res.get_mesh(0).colors) { col.push_back(c.x); col.push_back(c.y); col.push_back(c.z); }
    put(T_OPT_POS, pos.data(), pos.size() * 4);
    put(T_OPT_COL, col.data(), col.size() * 4);
    put(T_OPT_ERR, &err, 4);
    if (ops & OP_OPT_ALL) {
      std::vector<float> trs, tex;
      for (int i = 0; i < nm; i++) {
        const float4x4 t = res.get_transform(i)[0];
        const float* f = (const float*)&t;
        trs.insert(trs.end(), f, f + 16);
      }
      if (!res.get_mesh(0).textures.empty()) tex = res.get_mesh(0).textures[0].m_data;
      put(T_OPT_TR, trs.data(), trs.size() * 4);
      put(T_OPT_TEX, tex.data(), tex.size() * 4);
    }
    delete opt;
  }
  fclose(g_out);
  fclose(g_in);
  return 0;
}
