/*
 * drh_idiffrender.h -- the reference's renderer plugin interface over libdrh.so.
 *
 * DiffRenderHIP implements diff_render::IDiffRender (src/virtual_drender.h:14-26) with the C
 * ABI of drh.h, and MakeDifferentialRendererHIP mirrors the factory MakeDifferentialRenderer
 * (src/drender.cpp:89-147).  Header-only C++14; it needs the reference's own headers on the
 * include path (src/ and common_src/) and links with -ldrh.  A maintainer adds it to the
 * reference tree and returns MakeDifferentialRendererHIP(settings) from
 * MakeDifferentialRenderer for the four shading models this path covers.
 *
 * What it does with the reference's types:
 *   commit    flattens Scene (meshes, per-mesh transform or TransformR, colours, texture
 *             coordinates, normals, texture 0) into a drh_scene_desc and uploads it; the
 *             meshes marked differentiable are those of the last DScene seen (all meshes
 *             before any), and d_render commits again when a DScene names another set,
 *             because the edge sampler draws only the differentiable meshes' edges
 *             (src/drender.h:244-280);
 *   render    Img = LiteImage::Image2D<float3> is W x H x 3 float32, copied in and out;
 *   d_render  adds into DScene's DMesh blocks, matched by mesh id (the DScene may list its
 *             meshes in any order); refuses (prints, returns) a scene that was not the last
 *             committed, as the reference does (src/drender.h:148-152);
 *   d_render_and_compare  commits, renders into outImages (resized to the cameras), clears
 *             the DScene and returns the mean squared error (src/drender.h:169-191).
 * Errors print "[DiffRenderHIP::<call>]: <drh_last_error>" and return, as the reference's
 * DiffRender prints and returns; d_render_and_compare then returns NaN.
 *
 * Not on this path (the call prints and does nothing): PHONG / GGX / PATH_TEST, meshes with
 * several BASE_TRANSFORM instances, more than one texture or other than 3 channels on a
 * differentiable mesh in TEXTURE_COLOR / LAMBERT.
 */
#pragma once
#include <cmath>
#include <cstring>
#include <iostream>
#include <limits>
#include <memory>
#include <vector>

#include "virtual_drender.h"
#include "drh.h"

namespace diff_render {

struct DiffRenderHIP : IDiffRender {
  DiffRenderHIP() = default;
  DiffRenderHIP(const DiffRenderHIP&) = delete;
  DiffRenderHIP& operator=(const DiffRenderHIP&) = delete;
  ~DiffRenderHIP() override {
    if (h_) drh_destroy(h_);
  }

  void init(const DiffRenderSettings& settings) override {
    mode = settings.mode;
    const int m = (int)settings.mode;
    if (m < DRH_SILHOUETTE || m > DRH_LAMBERT) {
      std::cout << "[DiffRenderHIP::init]: shading model " << m << " is not on this path" << std::endl;
      return;
    }
    drh_settings ds;
    std::memset(&ds, 0, sizeof(ds));
    ds.mode = m;                          // SILHOUETTE = 1 ... LAMBERT = 4, the same numbering
    ds.spp = settings.spp;
    ds.rng_mode = DRH_RNG_REF_PRNG14;     // the reference's 14 OpenMP PRNG streams
    ds.tracer = DRH_TRACER_BVH;
    ds.device = 0;
    if (h_) drh_destroy(h_);
    h_ = nullptr;
    const drh_status s = drh_create(&ds, &h_);
    if (s != DRH_OK) {
      std::cout << "[DiffRenderHIP::init]: drh_create failed with status " << (int)s
                << (s == DRH_E_HIP ? " (no usable HIP device)" : "") << std::endl;
      h_ = nullptr;
    }
    last_ = nullptr;
  }

  void commit(const Scene& scene) override {
    std::vector<int> diff(scene.get_meshes().size(), 1);
    if (diff_.size() == diff.size()) diff = diff_;
    upload(scene, diff, "commit");
  }

  void render(const Scene& scene, const CamInfo* cams, Img* images, int views) override {
    if (!ready(scene, "render")) return;
    const std::vector<drh_camera> c = cameras(cams, views);
    std::vector<float> buf(pixels(cams, views) * 3);
    if (report(drh_render(h_, c.data(), views, buf.data()), "render")) return;
    size_t off = 0;
    for (int i = 0; i < views; i++) {
      images[i].resize((unsigned)cams[i].width, (unsigned)cams[i].height);
      const size_t n = 3 * (size_t)images[i].width() * images[i].height();
      std::memcpy((float*)images[i].data(), buf.data() + off, n * sizeof(float));
      off += n;
    }
  }

  void d_render(const Scene& scene, const CamInfo* cams, const Img* adjoints, int views,
                const int edge_samples_in_total, DScene& d_scene, Img* = nullptr, int = 0) override {
    if (!ready(scene, "d_render")) return;
    if (!match(scene, d_scene, "d_render")) return;
    const std::vector<drh_camera> c = cameras(cams, views);
    std::vector<float> adj;
    adj.reserve(pixels(cams, views) * 3);
    for (int i = 0; i < views; i++) {
      if (adjoints[i].width() != (unsigned)cams[i].width || adjoints[i].height() != (unsigned)cams[i].height) {
        std::cout << "[DiffRenderHIP::d_render]: adjoint " << i << " is not the camera's size" << std::endl;
        return;
      }
      const float* a = (const float*)adjoints[i].data();
      adj.insert(adj.end(), a, a + 3 * (size_t)adjoints[i].width() * adjoints[i].height());
    }
    std::vector<float> g = gather(d_scene);
    if (report(drh_d_render(h_, c.data(), views, adj.data(), edge_samples_in_total, g.data(), 1), "d_render")) return;
    scatter(g, d_scene);
  }

  float d_render_and_compare(const Scene& scene, const CamInfo* cams, const Img* targets, int views,
                             const int edge_samples_in_total, DScene& d_scene, Img* out_images = nullptr) override {
    const float nan = std::numeric_limits<float>::quiet_NaN();
    if (!h_) { std::cout << "[DiffRenderHIP::d_render_and_compare]: not initialised" << std::endl; return nan; }
    if (!upload(scene, flags_of(scene, d_scene), "d_render_and_compare")) return nan;
    if (!match(scene, d_scene, "d_render_and_compare")) return nan;
    const std::vector<drh_camera> c = cameras(cams, views);
    std::vector<float> tgt;
    tgt.reserve(pixels(cams, views) * 3);
    for (int i = 0; i < views; i++) {
      if (targets[i].width() != (unsigned)cams[i].width || targets[i].height() != (unsigned)cams[i].height) {
        std::cout << "[DiffRenderHIP::d_render_and_compare]: target " << i << " is not the camera's size" << std::endl;
        return nan;
      }
      const float* t = (const float*)targets[i].data();
      tgt.insert(tgt.end(), t, t + 3 * (size_t)targets[i].width() * targets[i].height());
    }
    std::vector<float> imgs(tgt.size());
    std::vector<float> g(d_scene.full_size(), 0.0f);
    float loss = nan;
    if (report(drh_d_render_and_compare(h_, c.data(), views, tgt.data(), edge_samples_in_total, g.data(),
                                        out_images ? imgs.data() : nullptr, &loss), "d_render_and_compare"))
      return nan;
    d_scene.clear();
    scatter(g, d_scene);
    if (out_images) {
      size_t off = 0;
      for (int i = 0; i < views; i++) {
        out_images[i].resize((unsigned)cams[i].width, (unsigned)cams[i].height);
        const size_t n = 3 * (size_t)out_images[i].width() * out_images[i].height();
        std::memcpy((float*)out_images[i].data(), imgs.data() + off, n * sizeof(float));
        off += n;
      }
    }
    return loss;
  }

  // The handle, for the library's own settings (drh_set_deterministic and the like).
  drh_handle handle() const { return h_; }

 private:
  drh_handle h_ = nullptr;
  const Scene* last_ = nullptr;    // the scene of the last successful commit
  std::vector<int> diff_;          // its differentiable-mesh flags

  bool report(drh_status s, const char* call) const {
    if (s == DRH_OK) return false;
    std::cout << "[DiffRenderHIP::" << call << "]: " << drh_last_error(h_) << std::endl;
    return true;
  }

  bool ready(const Scene& scene, const char* call) const {
    if (!h_) { std::cout << "[DiffRenderHIP::" << call << "]: not initialised" << std::endl; return false; }
    if (&scene != last_) {
      std::cout << "[DiffRenderHIP::" << call << "]: error, renderer was not prepared for this scene!" << std::endl;
      return false;
    }
    return true;
  }

  static std::vector<int> flags_of(const Scene& scene, DScene& d) {
    std::vector<int> f(scene.get_meshes().size(), 0);
    for (auto& dm : d.get_dmeshes())
      if (dm.get_mesh_id() >= 0 && dm.get_mesh_id() < (int)f.size()) f[dm.get_mesh_id()] = 1;
    return f;
  }

  // Commit again when the DScene differentiates another set of meshes than the last commit.
  bool match(const Scene& scene, DScene& d, const char* call) {
    const std::vector<int> f = flags_of(scene, d);
    if (f != diff_ && !upload(scene, f, call)) return false;
    int want = 0;
    for (auto& dm : d.get_dmeshes()) want += dm.full_size();
    if (drh_grad_size(h_) != want) {
      std::cout << "[DiffRenderHIP::" << call << "]: DScene holds " << want << " floats, the committed scene's"
                << " gradient " << drh_grad_size(h_) << " (a DScene of another scene or shading model?)" << std::endl;
      return false;
    }
    return true;
  }

  // The flat gradient lists the differentiable meshes in mesh order; DScene in its own order.
  static std::vector<DMesh*> by_mesh(DScene& d) {
    std::vector<DMesh*> v;
    for (auto& dm : d.get_dmeshes()) v.push_back(&dm);
    for (size_t i = 1; i < v.size(); i++)
      for (size_t j = i; j > 0 && v[j - 1]->get_mesh_id() > v[j]->get_mesh_id(); j--) std::swap(v[j - 1], v[j]);
    return v;
  }
  static std::vector<float> gather(DScene& d) {
    std::vector<float> g;
    for (DMesh* dm : by_mesh(d)) g.insert(g.end(), dm->full_data(), dm->full_data() + dm->full_size());
    return g;
  }
  static void scatter(const std::vector<float>& g, DScene& d) {
    size_t off = 0;
    for (DMesh* dm : by_mesh(d)) {
      std::memcpy(dm->full_data(), g.data() + off, sizeof(float) * dm->full_size());
      off += dm->full_size();
    }
  }

  static std::vector<drh_camera> cameras(const CamInfo* cams, int views) {
    std::vector<drh_camera> c(views > 0 ? views : 0);
    for (int i = 0; i < views; i++) {
      std::memcpy(c[i].world_view, (const float*)&cams[i].mWorldView, 16 * sizeof(float));  // column-major
      std::memcpy(c[i].proj, (const float*)&cams[i].mProj, 16 * sizeof(float));
      c[i].width = cams[i].width;
      c[i].height = cams[i].height;
    }
    return c;
  }
  static size_t pixels(const CamInfo* cams, int views) {
    size_t n = 0;
    for (int i = 0; i < views; i++) n += (size_t)cams[i].width * (size_t)cams[i].height;
    return n;
  }

  // Scene -> drh_scene_desc (the flattening of Scene::prepare_for_render,
  // common_src/scene_utils.cpp:98-173, is done by drh_commit on the device).
  bool upload(const Scene& scene, const std::vector<int>& diff, const char* call) {
    if (!h_) { std::cout << "[DiffRenderHIP::" << call << "]: not initialised" << std::endl; return false; }
    const int nm = (int)scene.get_meshes().size();
    const bool textured = mode == SHADING_MODEL::TEXTURE_COLOR || mode == SHADING_MODEL::LAMBERT;
    std::vector<int> vc, fc, tw, th, rflag;
    std::vector<float> tr, pos, col, tc, nrm, tex, rpar;
    std::vector<uint32_t> idx;
    bool any_col = false, any_tc = false, any_nrm = false;
    for (int i = 0; i < nm; i++) {
      const TriangleMesh& m = scene.get_mesh(i);
      any_col |= !m.colors.empty();
      any_tc |= !m.tc.empty();
      any_nrm |= !m.normals.empty();
    }
    for (int i = 0; i < nm; i++) {
      const TriangleMesh& m = scene.get_mesh(i);
      const int V = m.vertex_count();
      vc.push_back(V);
      fc.push_back(m.face_count());
      const bool restricted = scene.get_instancing_type(i) == Scene::RESTRICTED_TRANSFORM;
      const size_t instances = restricted ? scene.get_restricted_transform(i).size() : scene.get_transform(i).size();
      if (instances != 1) {
        std::cout << "[DiffRenderHIP::" << call << "]: mesh " << i << " has " << instances
                  << " instances; one per mesh is on this path" << std::endl;
        return false;
      }
      const float4x4 T = restricted ? float4x4() : scene.get_transform(i)[0];
      tr.insert(tr.end(), (const float*)&T, (const float*)&T + 16);
      rflag.push_back(restricted ? 1 : 0);
      for (int k = 0; k < 7; k++) rpar.push_back(restricted ? scene.get_restricted_transform(i)[0].M[k] : 0.0f);
      for (const float3& v : m.vertices) { pos.push_back(v.x); pos.push_back(v.y); pos.push_back(v.z); }
      // absent attributes take prepare_for_render's defaults (scene_utils.cpp:128-133)
      for (int k = 0; k < V; k++) {
        const float3 c = m.colors.empty() ? float3(0, 0, 0) : m.colors[k];
        const float2 t = m.tc.empty() ? float2(0, 0) : m.tc[k];
        const float3 n = m.normals.empty() ? float3(1, 0, 0) : m.normals[k];
        col.push_back(c.x); col.push_back(c.y); col.push_back(c.z);
        tc.push_back(t.x); tc.push_back(t.y);
        nrm.push_back(n.x); nrm.push_back(n.y); nrm.push_back(n.z);
      }
      idx.insert(idx.end(), m.indices.begin(), m.indices.end());
      if (textured && diff[i] && (m.textures.size() > 1 || (m.textures.size() == 1 && m.textures[0].channels != 3))) {
        std::cout << "[DiffRenderHIP::" << call << "]: mesh " << i
                  << " needs exactly one RGB texture for a texture gradient on this path" << std::endl;
        return false;
      }
      if (!m.textures.empty() && m.textures[0].channels == 3) {
        tw.push_back(m.textures[0].w);
        th.push_back(m.textures[0].h);
        tex.insert(tex.end(), m.textures[0].m_data.begin(), m.textures[0].m_data.end());
      } else {
        tw.push_back(0);
        th.push_back(0);
      }
    }
    drh_scene_desc d;
    std::memset(&d, 0, sizeof(d));
    d.n_meshes = nm;
    d.vcount = vc.data();
    d.fcount = fc.data();
    d.transforms = tr.data();
    d.pos = pos.data();
    d.col = any_col ? col.data() : nullptr;
    d.tc = any_tc ? tc.data() : nullptr;
    d.nrm = any_nrm ? nrm.data() : nullptr;
    d.idx = idx.data();
    d.tex_w = tw.data();
    d.tex_h = th.data();
    d.texels = tex.empty() ? nullptr : tex.data();
    d.differentiable = diff.data();
    d.restricted = rflag.data();
    d.restricted_params = rpar.data();
    last_ = nullptr;
    if (report(drh_commit(h_, &d), call)) return false;
    last_ = &scene;
    diff_ = diff;
    return true;
  }
};

// MakeDifferentialRenderer (src/drender.cpp:89-147) for the shading models on this path;
// nullptr (after a message) for PHONG, GGX and PATH_TEST.
inline std::shared_ptr<IDiffRender> MakeDifferentialRendererHIP(const DiffRenderSettings& settings) {
  const int m = (int)settings.mode;
  if (m < (int)SHADING_MODEL::SILHOUETTE || m > (int)SHADING_MODEL::LAMBERT) {
    std::cout << "[MakeDifferentialRendererHIP]: shading model " << m << " is not on this path" << std::endl;
    return nullptr;
  }
  auto impl = std::make_shared<DiffRenderHIP>();
  impl->init(settings);
  return impl;
}

}  // namespace diff_render
