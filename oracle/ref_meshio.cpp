// oracle/ref_meshio.cpp -- drives the reference's mesh and image file code, compiled from
// its own sources under /root/reference (common_src/cmesh.cpp, common_src/Image2d.cpp; see
// oracle/Makefile), to write golden files for the host-side file IO.
//
// TEST INFRASTRUCTURE ONLY.  Calls mesh_utils::LoadMeshFromVSGF / SaveMeshToVSGF and
// LiteImage::SaveImage<float3>; re-implements none of them.
//
// Usage:
//   ref_meshio load  <in.vsgf> <out.bin>   out: int32 V, I; float32 pos4[4V], nrm4[4V],
//                                          tang4[4V], tc[2V]; uint32 idx[I], mat[I/3]
//   ref_meshio save  <in.bin> <out.vsgf>   in: int32 V, I; float32 pos4[4V], nrm4[4V], tc[2V];
//                                          uint32 idx[I]   (materials all zero)
//   ref_meshio image <in.bin> <out.ext> <gamma>
//                                          in: int32 W, H; float32 rgb[3WH], row 0 at the top
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "LiteMath.h"
#include "Image2d.h"
#include "cmesh.h"

static std::vector<char> slurp(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
  std::vector<char> b;
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
  fclose(f);
  return b;
}

struct Reader {
  const std::vector<char>& b;
  size_t off = 0;
  void get(void* dst, size_t n) {
    if (off + n > b.size()) { fprintf(stderr, "short input\n"); exit(2); }
    memcpy(dst, b.data() + off, n);
    off += n;
  }
  int32_t i32() { int32_t v; get(&v, 4); return v; }
};

int main(int argc, char** argv) {
  if (argc < 4) { fprintf(stderr, "usage: ref_meshio load|save|image in out [gamma]\n"); return 1; }
  const std::string op = argv[1];
  if (op == "load") {
    mesh_utils::SimpleMesh m = mesh_utils::LoadMeshFromVSGF(argv[2]);
    FILE* f = fopen(argv[3], "wb");
    int32_t hdr[2] = {(int32_t)m.VerticesNum(), (int32_t)m.IndicesNum()};
    fwrite(hdr, sizeof hdr, 1, f);
    for (auto& p : m.vPos4f) fwrite(&p, sizeof(float), 4, f);
    for (auto& p : m.vNorm4f) fwrite(&p, sizeof(float), 4, f);
    for (auto& p : m.vTang4f) fwrite(&p, sizeof(float), 4, f);
    for (auto& p : m.vTexCoord2f) fwrite(&p, sizeof(float), 2, f);
    fwrite(m.indices.data(), sizeof(unsigned), m.indices.size(), f);
    fwrite(m.matIndices.data(), sizeof(unsigned), m.matIndices.size(), f);
    fclose(f);
    return 0;
  }
  if (op == "save") {
    std::vector<char> b = slurp(argv[2]);
    Reader r{b};
    const int V = r.i32(), I = r.i32();
    mesh_utils::SimpleMesh m(V, I);
    for (int i = 0; i < V; i++) { float v[4]; r.get(v, 16); m.vPos4f[i] = LiteMath::float4(v[0], v[1], v[2], v[3]); }
    for (int i = 0; i < V; i++) { float v[4]; r.get(v, 16); m.vNorm4f[i] = LiteMath::float4(v[0], v[1], v[2], v[3]); }
    for (int i = 0; i < V; i++) { float v[2]; r.get(v, 8); m.vTexCoord2f[i] = LiteMath::float2(v[0], v[1]); }
    r.get(m.indices.data(), sizeof(unsigned) * I);
    for (auto& x : m.matIndices) x = 0;
    mesh_utils::SaveMeshToVSGF(argv[3], m);
    return 0;
  }
  if (op == "image" && argc == 5) {
    std::vector<char> b = slurp(argv[2]);
    Reader r{b};
    const int W = r.i32(), H = r.i32();
    LiteImage::Image2D<LiteMath::float3> img(W, H);
    for (int i = 0; i < W * H; i++) { float v[3]; r.get(v, 12); img.data()[i] = LiteMath::float3(v[0], v[1], v[2]); }
    return LiteImage::SaveImage(argv[3], img, (float)atof(argv[4])) ? 0 : 3;
  }
  fprintf(stderr, "bad arguments\n");
  return 1;
}
