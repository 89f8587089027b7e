// oracle/adapter_check.cpp -- test harness (not the product): drives libdrh.so through the
// reference's own plugin interface, IDiffRender (src/virtual_drender.h:14-26), with the
// DiffRenderHIP adapter of include/drh_idiffrender.h, on a scene built from the reference's
// own types (Scene, TriangleMesh, CamInfo, DScene, Img).  Built by `make -C oracle adapter`
// against the reference's headers (and its scene_utils.cpp, which DScene's constructor calls),
// into _ref/adapter_check.  tests/test_gpu_adapter.py runs it on the GPU and checks what it
// writes against the oracle.
//
//   adapter_check <mode 1..4> <output file>
//
// Output: records of [name: 32 bytes, NUL-padded][kind: 'f' float32 | 'i' int32][count: int64]
// [count values], in this order: the two scenes' mesh arrays and transforms, the cameras, the
// images, the gradients of three DScenes and the checks' flags.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "drh_idiffrender.h"

using namespace diff_render;

namespace {

FILE* g_out = nullptr;

void put(const char* name, char kind, const void* data, int64_t count) {
  char nm[32] = {0};
  std::strncpy(nm, name, 31);
  std::fwrite(nm, 1, 32, g_out);
  std::fwrite(&kind, 1, 1, g_out);
  std::fwrite(&count, 8, 1, g_out);
  std::fwrite(data, 4, (size_t)count, g_out);
}
void putf(const std::string& name, const std::vector<float>& v) { put(name.c_str(), 'f', v.data(), (int64_t)v.size()); }
void puti(const std::string& name, const std::vector<int>& v) { put(name.c_str(), 'i', v.data(), (int64_t)v.size()); }

uint32_t g_rng = 12345u;
float rnd01() {
  g_rng = g_rng * 1664525u + 1013904223u;
  return (float)(g_rng >> 8) * (1.0f / 16777216.0f);
}

// A box of half-size s with per-vertex colours, texture coordinates, normals and an 8x8 RGB
// texture; `bulge` pushes the +x face out (the target's difference).
TriangleMesh box(float s, float bulge) {
  TriangleMesh m;
  for (int k = 0; k < 8; k++) {
    const float x = (k & 1) ? s + bulge : -s, y = (k & 2) ? s : -s, z = (k & 4) ? s : -s;
    m.vertices.push_back(float3(x, y, z));
    m.colors.push_back(float3(0.2f + 0.7f * rnd01(), 0.2f + 0.7f * rnd01(), 0.2f + 0.7f * rnd01()));
    m.tc.push_back(float2((k & 1) ? 0.9f : 0.1f, (k & 2) ? 0.85f : 0.15f));
    const float l = std::sqrt(x * x + y * y + z * z);
    m.normals.push_back(float3(x / l, y / l, z / l));
  }
  const unsigned f[12][3] = {{0, 2, 1}, {1, 2, 3}, {4, 5, 6}, {5, 7, 6}, {0, 1, 4}, {1, 5, 4},
                             {2, 6, 3}, {3, 6, 7}, {0, 4, 2}, {2, 4, 6}, {1, 3, 5}, {3, 7, 5}};
  for (auto& t : f) m.indices.insert(m.indices.end(), t, t + 3);
  CPUTexture tex;
  tex.w = 8; tex.h = 8; tex.channels = 3;
  for (int i = 0; i < 8 * 8 * 3; i++) tex.m_data.push_back(0.15f + 0.8f * rnd01());
  m.textures.push_back(tex);
  return m;
}

void put_scene(const std::string& tag, const Scene& sc) {
  for (int i = 0; i < (int)sc.get_meshes().size(); i++) {
    const TriangleMesh& m = sc.get_mesh(i);
    const std::string p = tag + std::to_string(i) + "_";
    putf(p + "pos", std::vector<float>((const float*)m.vertices.data(), (const float*)m.vertices.data() + 3 * m.vertices.size()));
    putf(p + "col", std::vector<float>((const float*)m.colors.data(), (const float*)m.colors.data() + 3 * m.colors.size()));
    putf(p + "tc", std::vector<float>((const float*)m.tc.data(), (const float*)m.tc.data() + 2 * m.tc.size()));
    putf(p + "nrm", std::vector<float>((const float*)m.normals.data(), (const float*)m.normals.data() + 3 * m.normals.size()));
    puti(p + "idx", std::vector<int>(m.indices.begin(), m.indices.end()));
    putf(p + "tex", m.textures[0].m_data);
    const float4x4& T = sc.get_transform(i)[0];
    putf(p + "transform", std::vector<float>((const float*)&T, (const float*)&T + 16));
  }
}

std::vector<float> flat(const Img* im, int n) {
  std::vector<float> v;
  for (int i = 0; i < n; i++) v.insert(v.end(), (const float*)im[i].data(), (const float*)im[i].data() + 3 * im[i].width() * im[i].height());
  return v;
}

// DScene blocks in the DScene's own order, with the mesh id of each
void put_dscene(const std::string& tag, DScene& d) {
  std::vector<int> ids;
  std::vector<float> g;
  for (auto& dm : d.get_dmeshes()) {
    ids.push_back(dm.get_mesh_id());
    g.insert(g.end(), dm.full_data(), dm.full_data() + dm.full_size());
  }
  puti(tag + "_ids", ids);
  putf(tag + "_grad", g);
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) { std::fprintf(stderr, "usage: adapter_check <mode 1..4> <out>\n"); return 2; }
  const int mode = std::atoi(argv[1]);
  if (mode < 1 || mode > 4) { std::fprintf(stderr, "mode must be 1..4\n"); return 2; }
  g_out = std::fopen(argv[2], "wb");
  if (!g_out) { std::perror(argv[2]); return 2; }

  const float4x4 T0 = LiteMath::translate4x4(float3(-0.35f, 0.05f, 0.0f)) * LiteMath::rotate4x4Y(0.5f) * LiteMath::rotate4x4X(0.3f);
  const float4x4 T1 = LiteMath::translate4x4(float3(0.45f, -0.1f, -0.3f)) * LiteMath::rotate4x4Y(-0.4f);
  Scene scene, target;
  g_rng = 7u;
  scene.add_mesh(box(0.35f, 0.0f), {T0});
  scene.add_mesh(box(0.25f, 0.0f), {T1});
  g_rng = 7u;
  target.add_mesh(box(0.35f, 0.08f), {T0});
  target.add_mesh(box(0.25f, 0.0f), {T1});
  put_scene("s", scene);
  put_scene("t", target);

  const int W = 96, H = 64, views = 2, N = 20000;
  CamInfo cams[views] = {CamInfo(float3(0.3f, 0.4f, 2.6f), float3(0, 0, 0), float3(0, 1, 0), (float)W, (float)H),
                         CamInfo(float3(-1.2f, 0.6f, 2.2f), float3(0, 0, 0), float3(0, 1, 0), (float)W, (float)H)};
  for (int i = 0; i < views; i++) {
    putf("cam" + std::to_string(i) + "_wv", std::vector<float>((const float*)&cams[i].mWorldView, (const float*)&cams[i].mWorldView + 16));
    putf("cam" + std::to_string(i) + "_proj", std::vector<float>((const float*)&cams[i].mProj, (const float*)&cams[i].mProj + 16));
  }
  puti("size", {W, H, views, N, mode});

  DiffRenderSettings settings;
  settings.mode = (SHADING_MODEL)mode;
  settings.spp = 4;
  std::shared_ptr<IDiffRender> r = MakeDifferentialRendererHIP(settings);
  if (!r || !static_cast<DiffRenderHIP*>(r.get())->handle()) { std::fclose(g_out); return 3; }

  Img tgt[views], img[views], adj[views];
  r->commit(target);
  r->render(target, cams, tgt, views);
  r->commit(scene);
  r->render(scene, cams, img, views);
  putf("target_images", flat(tgt, views));
  putf("images", flat(img, views));
  for (int i = 0; i < views; i++) {
    adj[i].resize(W, H);
    for (int p = 0; p < W * H; p++) adj[i].data()[p] = 2.0f * (img[i].data()[p] - tgt[i].data()[p]);
  }

  // both meshes differentiable, listed in reverse order; d_render accumulates
  DScene dA(scene, settings.mode, {1, 0});
  r->d_render(scene, cams, adj, views, N, dA);
  put_dscene("dA1", dA);
  r->d_render(scene, cams, adj, views, N, dA);
  put_dscene("dA2", dA);
  // only mesh 0 differentiable: the adapter commits again with mesh 1 fixed
  DScene dB(scene, settings.mode, {0});
  r->d_render(scene, cams, adj, views, N, dB);
  put_dscene("dB", dB);
  // d_render_and_compare clears the DScene and writes the images
  DScene dC(scene, settings.mode, {0, 1});
  for (auto& dm : dC.get_dmeshes()) for (int k = 0; k < dm.full_size(); k++) dm.full_data()[k] = 5.0f;
  Img out[views];
  const float loss = r->d_render_and_compare(scene, cams, tgt, views, N, dC, out);
  put_dscene("dC", dC);
  putf("compare_images", flat(out, views));
  putf("compare_loss", {loss});

  // a scene that is not the committed one: refused, outputs untouched
  Scene other = scene;
  Img probe[1] = {Img(W, H, float3(-7.0f, -7.0f, -7.0f))};
  r->render(other, cams, probe, 1);
  bool untouched = true;
  for (int p = 0; p < W * H; p++) untouched &= probe[0].data()[p].x == -7.0f;
  DScene dD(scene, settings.mode, {0, 1});
  r->d_render(other, cams, adj, views, N, dD);
  bool zero = true;
  for (auto& dm : dD.get_dmeshes()) for (int k = 0; k < dm.full_size(); k++) zero &= dm.full_data()[k] == 0.0f;
  puti("refused_untouched", {untouched ? 1 : 0, zero ? 1 : 0});
  std::fclose(g_out);
  std::printf("adapter_check mode %d done\n", mode);
  return 0;
}
