// cpt_rt.hpp -- header-only C++ facade over cpt.h with the shape of CudaPT's
// RTScene / RTRenderer (include/raytracer.h:228-323), for callers that drive
// the reference's classes: the same method names, argument meaning and bool
// results, with the scene given as arrays instead of an assimp import.
//
//   cpt::RTScene scene;
//   scene.SetMesh(pos, nrm, nv, idx, mat, nt);      // RTScene::AddModel's arrays
//   scene.SetMaterials(mats, nmat);
//   cpt::RTRenderer r;
//   r.SetRendererMode(cpt::RTRenderer::BDPT_STREAM);
//   r.Init(1280, 720);
//   r.Render(camPos, camDir, camUp, fov, scene);    // one accumulation frame
//   const float* rgb = r.GetResult();                // W*H*3 running mean
#ifndef CPT_RT_HPP
#define CPT_RT_HPP

#include <string>
#include <vector>

#include "cpt.h"

namespace cpt {

inline std::string LastError() { return cpt_last_error(); }

class RTScene {
 public:
  RTScene() { cpt_scene_create(&h_); }
  ~RTScene() {
    if (h_) cpt_scene_destroy(h_);
  }
  RTScene(const RTScene&) = delete;
  RTScene& operator=(const RTScene&) = delete;

  // RTScene::AddModel (src/raytracer.cpp:238-282) minus the file import:
  // positions / normals (3 floats per vertex), triangles (3 vertex indices) and
  // one material index per triangle; then the SAH BVH of
  // src/raytracer.cpp:267-279 (max depth 32).
  bool SetMesh(const float* pos3, const float* nrm3, uint32_t nv, const uint32_t* idx3, const uint32_t* mat,
               uint32_t nt, uint32_t max_depth = 32, const float* uv2 = nullptr) {
    return h_ && cpt_scene_set_mesh(h_, pos3, nrm3, uv2, nv, idx3, mat, nt) == CPT_OK &&
           cpt_bvh_build_sah(h_, max_depth) == CPT_OK;
  }
  // RTScene::m_pMaterials.
  bool SetMaterials(const cpt_material* m, uint32_t n) { return h_ && cpt_scene_set_materials(h_, m, n) == CPT_OK; }
  // RTScene::m_pTextures after LoadTexture (src/raytracer.cpp:161-179): a
  // width x height RGBA float texture; id is what materials' texture ids name.
  bool AddTexture(uint32_t width, uint32_t height, const float* rgba, uint32_t& id) {
    return h_ && cpt_scene_add_texture(h_, width, height, rgba, &id) == CPT_OK;
  }
  // RTScene::UpdateMaterialsDirtyFlag (src/raytracer.cpp:284-300): pushes edited
  // materials; renderers drawing the scene restart accumulation.
  bool UpdateMaterialsDirtyFlag() { return h_ && cpt_scene_update_materials(h_) == CPT_OK; }
  // RTScene::GetCompactBVH: node bounds (6 floats), (offset-or-triStart, triN)
  // pairs and the leaf-order triangle ids.
  bool GetCompactBVH(std::vector<float>& bounds, std::vector<uint32_t>& off, std::vector<uint32_t>& reorder,
                     uint32_t nt) const {
    uint32_t n = 0;
    if (!h_ || cpt_bvh_export(h_, nullptr, nullptr, nullptr, &n) != CPT_OK) return false;
    bounds.resize((size_t)n * 6);
    off.resize((size_t)n * 2);
    reorder.resize(nt);
    return cpt_bvh_export(h_, bounds.data(), off.data(), reorder.data(), &n) == CPT_OK;
  }
  cpt_scene* handle() const { return h_; }

 private:
  cpt_scene* h_ = nullptr;
};

class RTRenderer {
 public:
  // RTRenderer::RENDERER_MODE entries this library implements.
  enum RENDERER_MODE {
    BDPT_STREAM = CPT_MODE_BDPT_STREAM,
    BDPT_STREAM_ADAP_BILAT_ATOMIC = CPT_MODE_BDPT_ADAPTIVE_PLL,
    BDPT_STREAM_ADAP_BILAT_BINARY = CPT_MODE_BDPT_ADAPTIVE_BINARY,
    BDPT_STREAM_ADAP_EMP = CPT_MODE_BDPT_ADAPTIVE_EMP,
    PT_STREAM = CPT_MODE_PT_STREAM,               // DEFINE_RT(PTSC, cudaRTPTStream)
    PT_STREAM_REGEN = CPT_MODE_PT_STREAM_REGEN,   // DEFINE_RT(PTSCRegen, cudaRTPTStreamRegen), commented out there
    PT_DISNEY_MIS = CPT_MODE_PT_DISNEY_MIS,       // no reference entry: Disney BRDF path tracer with MIS
    BDPT_REGEN = CPT_MODE_BDPT_REGEN              // DEFINE_RT(BDPTRegen, cudaRTBDPTRegen)
  };
  RTRenderer() = default;
  ~RTRenderer() { release(); }
  RTRenderer(const RTRenderer&) = delete;
  RTRenderer& operator=(const RTRenderer&) = delete;

  RENDERER_MODE GetRendererMode() const { return mode_; }
  // RTRenderer::SetRendererMode (src/raytracer.cpp:593-608): the next Init or
  // Render runs the new mode from a fresh accumulation.
  void SetRendererMode(RENDERER_MODE mode) {
    if (mode != mode_) {
      mode_ = mode;
      release();
    }
  }
  // RTRenderer::Init (src/raytracer.cpp:574-591).
  bool Init(unsigned int width, unsigned int height) {
    release();
    if (cpt_renderer_create(&h_, (int)mode_, width, height) != CPT_OK) return false;
    w_ = width;
    h_px_ = height;
    result_.assign((size_t)width * height * 3, 0.f);
    return true;
  }
  // RTRenderer::Render (src/raytracer.cpp:610-625): one frame of the mode's
  // Render; the running mean is copied to GetResult().
  bool Render(const float camPos[3], const float camDir[3], const float camUp[3], float fov, RTScene& scene) {
    if (!h_ && (w_ == 0 || !Init(w_, h_px_))) return false;
    return cpt_render_iteration(h_, scene.handle(), camPos, camDir, camUp, fov, result_.data()) == CPT_OK;
  }
  const float* GetResult() const { return result_.empty() ? nullptr : result_.data(); }

  // <ns>::GetAttribsN / GetAttribute (src/raytracer.cpp:33-58).
  uint32_t GetAttribsN() const {
    uint32_t n = 0;
    if (h_) cpt_renderer_attrib_count(h_, &n);
    return n;
  }
  bool GetAttribute(uint32_t i, std::string& name, double& value) const {
    char buf[64];
    int type = 0;
    if (!h_ || cpt_renderer_get_attrib(h_, i, buf, sizeof buf, &type, &value) != CPT_OK) return false;
    name = buf;
    return true;
  }
  bool SetAttribute(const std::string& name, double value) {
    for (uint32_t i = 0; i < GetAttribsN(); i++) {
      std::string n;
      double v;
      if (GetAttribute(i, n, v) && n == name) return cpt_renderer_set_attrib(h_, i, value) == CPT_OK;
    }
    return false;
  }
  cpt_renderer* handle() const { return h_; }

 private:
  void release() {
    if (h_) cpt_renderer_destroy(h_);
    h_ = nullptr;
  }
  RENDERER_MODE mode_ = BDPT_STREAM;
  cpt_renderer* h_ = nullptr;
  unsigned int w_ = 0, h_px_ = 0;
  std::vector<float> result_;
};

}  // namespace cpt

#endif  // CPT_RT_HPP
