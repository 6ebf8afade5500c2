// C ABI of libcpt_hip.so (include/cpt.h): scene store, BVH build + upload,
// batched traversal, and the wavefront BDPT renderer driver.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpt.h"
#include "cpt_internal.h"
#include "kernels.h"
#include "mlp.h"

using namespace cpt;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) return fail(CPT_E_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
    if (e == hipSuccess) n = count;
    return e;
  }
  hipError_t upload(const T* src, size_t count, hipStream_t st) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    e = hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
  }
};



}  // namespace

struct cpt_scene {
  std::vector<float> pos, nrm, uv;
  std::vector<float4> tex;              // RTScene::m_pTextures' texels (RGBA), concatenated
  std::vector<uint4> tex_desc;          // first texel, width, height
  std::vector<float4> env;              // environment radiance, env_w x env_h (empty: none)
  u32 env_w = 0, env_h = 0;
  std::vector<u32> idx, mat;
  std::vector<cpt_material> mats;
  u32 nv = 0, nt = 0;
  BvhHost bvh;
  bool has_bvh = false;
  bool uploaded = false;
  // leaves of the 4-wide BVH hold at most leaf_max triangles: more leaf levels
  // cost fewer triangle tests than they add node visits (2: 16 % more rays/s
  // than 4 on the 262k-triangle sphere grid, 1 is level with 2)
  u32 mesh_version = 0, mat_version = 0, leaf_max = 2;
  hipStream_t stream = nullptr;
  DevArray<Node4> d_nodes4;
  DevArray<Node4Q> d_nodes4q;
  DevArray<float4> d_tris, d_tris_bin, d_vpos, d_vnrm, d_bnodes, d_tex, d_surf, d_env;
  DevArray<uint4> d_tex_desc;
  DevArray<uint4> d_tri_verts;
  DevArray<uint2> d_boff;
  DevArray<DevMaterial> d_mats;
  DevArray<u32> d_light;
  DevArray<float4> d_rays, d_hits;
  DevArray<DevStats> d_stats;
  u32 n_light = 0;
  DevScene dev() const {
    DevScene s;
    s.nodes4 = d_nodes4.p;
    s.nodes4q = d_nodes4q.p;
    s.tris = d_tris.p;
    s.tri_verts = d_tri_verts.p;
    s.vpos = d_vpos.p;
    s.vnrm = d_vnrm.p;
    s.mats = d_mats.p;
    s.surf = d_surf.p;
    s.light_tris = d_light.p;
    s.n_light_tris = n_light;
    s.n_tris = nt;
    s.bnodes = d_bnodes.p;
    s.boff = d_boff.p;
    s.tris_bin = d_tris_bin.p;
    s.tex = d_tex.p;
    s.tex_desc = d_tex_desc.p;
    s.n_tex = (u32)tex_desc.size();
    s.env = d_env.p;
    s.env_w = env.empty() ? 0u : env_w;
    s.env_h = env.empty() ? 0u : env_h;
    s.qmag = bvh.qmag;
    return s;
  }
  ~cpt_scene() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int ensure_stream(hipStream_t& st) {
  if (!st) HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  return CPT_OK;
}

// updateLightTriCudaMem (src/cudaRTBDPTStream.cu:57-77) over leaf order.
int upload_materials(cpt_scene* s) {
  std::vector<DevMaterial> dm(s->mats.size());
  static_assert(sizeof(DevMaterial) == sizeof(cpt_material), "material layouts agree");
  if (!s->mats.empty()) std::memcpy(dm.data(), s->mats.data(), sizeof(cpt_material) * s->mats.size());
  HIP_TRY(s->d_mats.upload(dm.data(), dm.size(), s->stream));
  // per-triangle shading records in the device tree's leaf order
  std::vector<float4> surf((size_t)s->nt * 8);
  auto bits = [](u32 x) { float f; std::memcpy(&f, &x, 4); return f; };
  for (u32 i = 0; i < s->nt; i++) {
    const u32 o = s->bvh.dreorder[i];
    const u32 v[3] = {s->idx[o * 3], s->idx[o * 3 + 1], s->idx[o * 3 + 2]};
    const cpt_material& m = s->mats[s->mat[o]];
    float4* r = &surf[(size_t)i * 8];
    for (int k = 0; k < 3; k++)
      r[k] = make_float4(s->nrm[v[k] * 3], s->nrm[v[k] * 3 + 1], s->nrm[v[k] * 3 + 2], s->uv[v[k] * 2 + 1]);
    r[3] = make_float4(m.diffuse[0], m.diffuse[1], m.diffuse[2], m.transparency);
    r[4] = make_float4(m.emissive[0], m.emissive[1], m.emissive[2], m.specular);
    r[5] = make_float4(m.metallic, m.roughness, bits((u32)m.diffuse_tex_id), bits((u32)m.emissive_tex_id));
    r[6] = make_float4(s->uv[v[0] * 2], s->uv[v[1] * 2], s->uv[v[2] * 2], 0.f);
    r[7] = make_float4(bits(v[0]), bits(v[1]), bits(v[2]), bits(s->mat[o]));
  }
  HIP_TRY(s->d_surf.upload(surf.data(), surf.size(), s->stream));
  // emissive triangles listed in the reference's leaf order (the order light
  // subpaths pick them in), each as its index in the device tree's order
  std::vector<u32> dev_index(s->nt);
  for (u32 i = 0; i < s->nt; i++) dev_index[s->bvh.dreorder[i]] = i;
  std::vector<u32> light;
  for (u32 i = 0; i < s->nt; i++) {
    u32 o = s->bvh.reorder[i];
    u32 m = s->mat[o];
    const float* e = s->mats[m].emissive;
    if (std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]) > 0.f) light.push_back(dev_index[o]);
  }
  s->n_light = (u32)light.size();
  HIP_TRY(s->d_light.upload(light.data(), light.size(), s->stream));
  return CPT_OK;
}

// cudaBilateralFilter::updateGaussian (src/cudaBilateralFilter.cu:5-24): the
// spatial table of 2 * radius + 1 taps, sigma = the colour sigma as the
// reference's callers pass it, zero beyond.
constexpr int kFilterTable = 2 * kMaxFilterRadius + 2;
void filter_table(float sigma_color, u32 radius, float g[kFilterTable]) {
  for (int i = 0; i < kFilterTable; i++) g[i] = 0.f;
  for (u32 i = 0; i < 2 * radius + 1; i++) {
    int x = (int)i - (int)radius;
    g[i] = std::exp(-(float)(x * x) / (2 * sigma_color * sigma_color));
  }
}

FilterArgs filter_args(u32 width, u32 height, float sc, float sp, float sn, float sd, u32 radius, u32 nd,
                       bool use_diff) {
  FilterArgs f;
  f.width = width; f.height = height; f.radius = radius; f.nd = nd;
  f.kc = 2.f * sc * sc; f.kp = 2.f * sp * sp;
  f.kn = 2.f * sn * sn; f.kd = 2.f * sd * sd;
  f.use_diff = use_diff;
  f.kz = 0.f;
  f.cam[0] = f.cam[1] = f.cam[2] = 0.f;
  return f;
}

// Staging of cpt_filter_cross_bilateral_device_noalloc: the spatial table the
// filter kernel reads, kept in a file-scope device array instead of a
// per-call allocation.  Every device the library runs on holds its own copy
// (256 bytes) until the process ends.  In host code `scratch` names a host-side
// stand-in, never the device copy: kernels and copies take the address that
// hipGetSymbolAddress returns for the current device, cached per device.
__device__ __align__(256) float scratch[kFilterTable];
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kScratchDevices] = {};

int filter_scratch(float** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return fail(CPT_E_LIMIT, "filter scratch: device %d beyond the %d the library tracks", dev, kScratchDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(CPT_E_HIP, "filter scratch: hipGetSymbolAddress returned no address on device %d", dev);
    g_scratch_ptr[dev] = (float*)p;
  }
  *out = g_scratch_ptr[dev];
  return CPT_OK;
}

}  // namespace

extern "C" {

const char* cpt_last_error(void) { return g_err.c_str(); }

int cpt_device_count(int* n) {
  if (!n) return fail(CPT_E_ARG, "cpt_device_count: null output");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *n = e == hipSuccess ? c : 0;
  return CPT_OK;
}

int cpt_device_alloc(void** ptr, uint64_t bytes) {
  if (!ptr) return fail(CPT_E_ARG, "cpt_device_alloc: null output");
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return fail(CPT_E_NOMEM, "cpt_device_alloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
  return CPT_OK;
}
int cpt_device_free(void* ptr) {
  if (ptr) HIP_TRY(hipFree(ptr));
  return CPT_OK;
}
int cpt_memcpy_htod(void* dst, const void* src, uint64_t bytes) {
  if (bytes && (!dst || !src)) return fail(CPT_E_ARG, "cpt_memcpy_htod: null pointer");
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return CPT_OK;
}
int cpt_memcpy_dtoh(void* dst, const void* src, uint64_t bytes) {
  if (bytes && (!dst || !src)) return fail(CPT_E_ARG, "cpt_memcpy_dtoh: null pointer");
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return CPT_OK;
}
int cpt_memcpy_dtod(void* dst, const void* src, uint64_t bytes) {
  if (bytes && (!dst || !src)) return fail(CPT_E_ARG, "cpt_memcpy_dtod: null pointer");
  if (bytes) HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return CPT_OK;
}
int cpt_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return CPT_OK;
}

int cpt_scene_create(cpt_scene** out) {
  if (!out) return fail(CPT_E_ARG, "cpt_scene_create: null output");
  *out = new cpt_scene();
  return CPT_OK;
}

int cpt_scene_set_mesh(cpt_scene* s, const float* pos3, const float* nrm3, const float* uv2, uint32_t nv,
                       const uint32_t* idx3, const uint32_t* mat, uint32_t nt) {
  if (!s || (!pos3 && nv) || (!idx3 && nt)) return fail(CPT_E_ARG, "cpt_scene_set_mesh: null argument");
  for (size_t i = 0; i < (size_t)nt * 3; i++)
    if (idx3[i] >= nv) return fail(CPT_E_ARG, "cpt_scene_set_mesh: index %u out of range (nv=%u)", idx3[i], nv);
  s->pos.assign(pos3, pos3 + (size_t)nv * 3);
  if (nrm3) s->nrm.assign(nrm3, nrm3 + (size_t)nv * 3); else s->nrm.assign((size_t)nv * 3, 0.f);
  if (uv2) s->uv.assign(uv2, uv2 + (size_t)nv * 2); else s->uv.assign((size_t)nv * 2, 0.f);
  s->idx.assign(idx3, idx3 + (size_t)nt * 3);
  if (mat) s->mat.assign(mat, mat + nt); else s->mat.assign(nt, 0u);
  s->nv = nv;
  s->nt = nt;
  s->has_bvh = false;
  s->uploaded = false;
  s->mesh_version++;
  return CPT_OK;
}

int cpt_scene_add_texture(cpt_scene* s, uint32_t width, uint32_t height, const float* rgba, uint32_t* id) {
  if (!s || !rgba) return fail(CPT_E_ARG, "cpt_scene_add_texture: null argument");
  if (width == 0 || height == 0 || width > 65536 || height > 65536)
    return fail(CPT_E_ARG, "cpt_scene_add_texture: size %ux%u not in 1..65536", width, height);
  const size_t n = (size_t)width * height;
  if (s->tex.size() + n > 0xFFFFFFFFull) return fail(CPT_E_LIMIT, "cpt_scene_add_texture: more than 2^32 texels");
  const u32 first = (u32)s->tex.size();
  s->tex.resize(s->tex.size() + n);
  std::memcpy(&s->tex[first], rgba, n * sizeof(float4));
  s->tex_desc.push_back(make_uint4(first, width, height, 0u));
  if (id) *id = (u32)s->tex_desc.size() - 1;
  s->uploaded = false;
  s->mat_version++;                             // shading changes: renderers start over
  return CPT_OK;
}

int cpt_scene_set_environment(cpt_scene* s, uint32_t width, uint32_t height, const float* rgb) {
  if (!s) return fail(CPT_E_ARG, "cpt_scene_set_environment: null scene");
  if (width == 0 || height == 0) {
    s->env.clear();
    s->env_w = s->env_h = 0;
  } else {
    if (!rgb) return fail(CPT_E_ARG, "cpt_scene_set_environment: null texels");
    if ((uint64_t)width * height > (1ull << 26))
      return fail(CPT_E_LIMIT, "cpt_scene_set_environment: %ux%u texels exceed 2^26", width, height);
    s->env.resize((size_t)width * height);
    for (size_t i = 0; i < s->env.size(); i++) s->env[i] = make_float4(rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2], 0.f);
    s->env_w = width;
    s->env_h = height;
  }
  s->uploaded = false;
  s->mesh_version++;                    // renderers drawing this scene restart
  return CPT_OK;
}

int cpt_scene_clear_textures(cpt_scene* s) {
  if (!s) return fail(CPT_E_ARG, "cpt_scene_clear_textures: null scene");
  s->tex.clear();
  s->tex_desc.clear();
  s->uploaded = false;
  s->mat_version++;
  return CPT_OK;
}

int cpt_scene_set_materials(cpt_scene* s, const cpt_material* m, uint32_t n) {
  if (!s || (!m && n)) return fail(CPT_E_ARG, "cpt_scene_set_materials: null argument");
  s->mats.assign(m, m + n);
  return CPT_OK;
}

int cpt_scene_update_materials(cpt_scene* s) {
  if (!s) return fail(CPT_E_ARG, "cpt_scene_update_materials: null scene");
  if (!s->uploaded) return fail(CPT_E_STATE, "cpt_scene_update_materials: scene not uploaded");
  for (u32 m : s->mat)
    if (m >= s->mats.size()) return fail(CPT_E_ARG, "cpt_scene_update_materials: material %u missing", m);
  int rc = upload_materials(s);
  if (rc) return rc;
  s->mat_version++;
  return CPT_OK;
}

int cpt_bvh_build_sah(cpt_scene* s, uint32_t max_depth) {
  if (!s) return fail(CPT_E_ARG, "cpt_bvh_build_sah: null scene");
  if (s->nt == 0) return fail(CPT_E_STATE, "cpt_bvh_build_sah: empty mesh");
  if (max_depth == 0 || max_depth > 64) return fail(CPT_E_ARG, "cpt_bvh_build_sah: max_depth %u not in 1..64", max_depth);
  if (!build_bvh_sah(s->pos.data(), s->idx.data(), s->nt, max_depth, s->bvh) ||
      !build_bvh_device(s->pos.data(), s->idx.data(), s->nt, s->bvh))
    return fail(CPT_E_STATE, "cpt_bvh_build_sah: build failed");
  collapse_bvh4(s->pos.data(), s->idx.data(), s->bvh, s->leaf_max);
  if (!quantize_bvh4(s->bvh)) return fail(CPT_E_STATE, "cpt_bvh_build_sah: a node cannot be quantized to 64 bytes");
  if (s->bvh.depth4 > 31) return fail(CPT_E_LIMIT, "cpt_bvh_build_sah: 4-wide depth %u exceeds 31", s->bvh.depth4);
  s->has_bvh = true;
  s->uploaded = false;
  s->mesh_version++;             // device node arrays move: renderers and captured graphs start over
  return CPT_OK;
}

int cpt_bvh_export(cpt_scene* s, float* bounds6, uint32_t* off2, uint32_t* reorder, uint32_t* node_n) {
  if (!s) return fail(CPT_E_ARG, "cpt_bvh_export: null scene");
  if (!s->has_bvh) return fail(CPT_E_STATE, "cpt_bvh_export: no BVH");
  const auto& nodes = s->bvh.nodes;
  if (node_n) *node_n = (u32)nodes.size();
  for (size_t i = 0; i < nodes.size(); i++) {
    if (bounds6)
      for (int k = 0; k < 3; k++) { bounds6[i * 6 + k] = nodes[i].mn[k]; bounds6[i * 6 + 3 + k] = nodes[i].mx[k]; }
    if (off2) { off2[i * 2] = nodes[i].a; off2[i * 2 + 1] = nodes[i].n; }
  }
  if (reorder) std::memcpy(reorder, s->bvh.reorder.data(), sizeof(u32) * s->nt);
  return CPT_OK;
}

int cpt_bvh4_set_leaf_size(cpt_scene* s, uint32_t max_tris) {
  if (!s) return fail(CPT_E_ARG, "cpt_bvh4_set_leaf_size: null scene");
  if (max_tris < 1 || max_tris > 64) return fail(CPT_E_ARG, "cpt_bvh4_set_leaf_size: %u not in 1..64", max_tris);
  s->leaf_max = max_tris;
  if (!s->has_bvh) return CPT_OK;
  collapse_bvh4(s->pos.data(), s->idx.data(), s->bvh, s->leaf_max);
  if (!quantize_bvh4(s->bvh)) return fail(CPT_E_STATE, "cpt_bvh4_set_leaf_size: a node cannot be quantized to 64 bytes");
  if (s->bvh.depth4 > 31) return fail(CPT_E_LIMIT, "cpt_bvh4_set_leaf_size: 4-wide depth %u exceeds 31", s->bvh.depth4);
  s->uploaded = false;
  s->mesh_version++;
  return CPT_OK;
}

int cpt_bvh4_info(cpt_scene* s, uint32_t* node_n, uint32_t* depth, uint32_t* leaf_n) {
  if (!s) return fail(CPT_E_ARG, "cpt_bvh4_info: null scene");
  if (!s->has_bvh) return fail(CPT_E_STATE, "cpt_bvh4_info: no BVH");
  if (node_n) *node_n = (u32)s->bvh.nodes4.size();
  if (depth) *depth = s->bvh.depth4;
  if (leaf_n) *leaf_n = s->bvh.leaves4;
  return CPT_OK;
}

int cpt_scene_upload(cpt_scene* s) {
  if (!s) return fail(CPT_E_ARG, "cpt_scene_upload: null scene");
  if (!s->has_bvh) return fail(CPT_E_STATE, "cpt_scene_upload: build the BVH first");
  for (u32 m : s->mat)
    if (m >= s->mats.size()) return fail(CPT_E_ARG, "cpt_scene_upload: material %u missing", m);
  int rc = ensure_stream(s->stream);
  if (rc) return rc;
  const u32 nt = s->nt;
  // triangle intersection data in a tree's leaf order (src/cudaRTCommon.cu:503-514,
  // host FP32 edges): the device tree's for the 4-wide layouts, the reference
  // tree's for its binary layout
  auto tri_data = [&](const std::vector<u32>& order) {
    std::vector<float4> out((size_t)nt * 3);
    for (u32 i = 0; i < nt; i++) {
      u32 o = order[i];
      const float* a = &s->pos[(size_t)s->idx[o * 3] * 3];
      const float* b = &s->pos[(size_t)s->idx[o * 3 + 1] * 3];
      const float* c = &s->pos[(size_t)s->idx[o * 3 + 2] * 3];
      float ow;
      std::memcpy(&ow, &o, 4);
      out[i * 3] = make_float4(a[0], a[1], a[2], ow);
      out[i * 3 + 1] = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.f);
      out[i * 3 + 2] = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.f);
    }
    return out;
  };
  std::vector<float4> tris = tri_data(s->bvh.dreorder), tris_bin = tri_data(s->bvh.reorder);
  std::vector<uint4> tv(nt);
  for (u32 i = 0; i < nt; i++) {
    u32 o = s->bvh.dreorder[i];
    tv[i] = make_uint4(s->idx[o * 3], s->idx[o * 3 + 1], s->idx[o * 3 + 2], s->mat[o]);
  }
  std::vector<float4> vp(s->nv), vn(s->nv);
  for (u32 i = 0; i < s->nv; i++) {
    vp[i] = make_float4(s->pos[i * 3], s->pos[i * 3 + 1], s->pos[i * 3 + 2], s->uv[i * 2]);
    vn[i] = make_float4(s->nrm[i * 3], s->nrm[i * 3 + 1], s->nrm[i * 3 + 2], s->uv[i * 2 + 1]);
  }
  const auto& bn = s->bvh.nodes;
  std::vector<float4> bnodes(bn.size() * 2);
  std::vector<uint2> boff(bn.size());
  for (size_t i = 0; i < bn.size(); i++) {
    bnodes[i * 2] = make_float4(bn[i].mn[0], bn[i].mn[1], bn[i].mn[2], 0.f);
    bnodes[i * 2 + 1] = make_float4(bn[i].mx[0], bn[i].mx[1], bn[i].mx[2], 0.f);
    boff[i] = make_uint2(bn[i].a, bn[i].n);
  }
  HIP_TRY(s->d_nodes4.upload(s->bvh.nodes4.data(), s->bvh.nodes4.size(), s->stream));
  HIP_TRY(s->d_nodes4q.upload(s->bvh.nodes4q.data(), s->bvh.nodes4q.size(), s->stream));
  HIP_TRY(s->d_tris.upload(tris.data(), tris.size(), s->stream));
  HIP_TRY(s->d_tris_bin.upload(tris_bin.data(), tris_bin.size(), s->stream));
  HIP_TRY(s->d_tri_verts.upload(tv.data(), tv.size(), s->stream));
  HIP_TRY(s->d_vpos.upload(vp.data(), vp.size(), s->stream));
  HIP_TRY(s->d_vnrm.upload(vn.data(), vn.size(), s->stream));
  HIP_TRY(s->d_bnodes.upload(bnodes.data(), bnodes.size(), s->stream));
  HIP_TRY(s->d_tex.upload(s->tex.data(), s->tex.size(), s->stream));
  HIP_TRY(s->d_tex_desc.upload(s->tex_desc.data(), s->tex_desc.size(), s->stream));
  HIP_TRY(s->d_env.upload(s->env.data(), s->env.size(), s->stream));
  HIP_TRY(s->d_boff.upload(boff.data(), boff.size(), s->stream));
  rc = upload_materials(s);
  if (rc) return rc;
  if (!s->d_stats.p) {
    HIP_TRY(s->d_stats.alloc(1));
    HIP_TRY(hipMemsetAsync(s->d_stats.p, 0, sizeof(DevStats), s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
  }
  s->uploaded = true;
  s->mat_version++;
  return CPT_OK;
}

int cpt_scene_destroy(cpt_scene* s) {
  if (!s) return fail(CPT_E_ARG, "cpt_scene_destroy: null scene");
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  delete s;
  return CPT_OK;
}

int cpt_trace_closest_device(cpt_scene* s, const cpt_ray* d_rays, uint32_t n, uint32_t cull, cpt_hit* d_hits,
                             void* stream, uint32_t layout) {
  if (!s || (n && (!d_rays || !d_hits))) return fail(CPT_E_ARG, "cpt_trace_closest_device: null argument");
  if (!s->uploaded) return fail(CPT_E_STATE, "cpt_trace_closest_device: scene not uploaded");
  if (layout > 2) return fail(CPT_E_ARG, "cpt_trace_closest_device: layout %u not 0, 1 or 2", layout);
  hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  HIP_TRY(launch_trace(s->dev(), (const float4*)d_rays, n, nullptr, cull, (float4*)d_hits, 1, 0, layout,
                       s->d_stats.p, st));
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  return CPT_OK;
}

int cpt_trace_occluded_device(cpt_scene* s, const cpt_ray* d_rays, uint32_t n, uint32_t cull, uint32_t* d_occluded,
                              void* stream) {
  if (!s || (n && (!d_rays || !d_occluded))) return fail(CPT_E_ARG, "cpt_trace_occluded_device: null argument");
  if (!s->uploaded) return fail(CPT_E_STATE, "cpt_trace_occluded_device: scene not uploaded");
  hipStream_t st = stream ? (hipStream_t)stream : s->stream;
  HIP_TRY(launch_trace(s->dev(), (const float4*)d_rays, n, nullptr, cull, (float4*)d_occluded, 0, 1, 0, s->d_stats.p,
                       st));
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  return CPT_OK;
}

int cpt_trace_closest(cpt_scene* s, const cpt_ray* rays, uint32_t n, uint32_t cull, cpt_hit* hits) {
  if (!s || (n && (!rays || !hits))) return fail(CPT_E_ARG, "cpt_trace_closest: null argument");
  if (!s->uploaded) return fail(CPT_E_STATE, "cpt_trace_closest: scene not uploaded");
  if (n == 0) return CPT_OK;
  HIP_TRY(s->d_rays.alloc((size_t)n * 2));
  HIP_TRY(s->d_hits.alloc(n));
  HIP_TRY(hipMemcpyAsync(s->d_rays.p, rays, sizeof(cpt_ray) * n, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(launch_trace(s->dev(), s->d_rays.p, n, nullptr, cull, s->d_hits.p, 1, 0, 0, s->d_stats.p, s->stream));
  HIP_TRY(hipMemcpyAsync(hits, s->d_hits.p, sizeof(cpt_hit) * n, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return CPT_OK;
}

int cpt_trace_cost(cpt_scene* s, const cpt_ray* rays, uint32_t n, uint32_t cull, uint32_t layout, uint32_t* cost2) {
  if (!s || (n && (!rays || !cost2))) return fail(CPT_E_ARG, "cpt_trace_cost: null argument");
  if (!s->uploaded) return fail(CPT_E_STATE, "cpt_trace_cost: scene not uploaded");
  if (layout > 2) return fail(CPT_E_ARG, "cpt_trace_cost: layout %u not 0, 1 or 2", layout);
  if (n == 0) return CPT_OK;
  HIP_TRY(s->d_rays.alloc((size_t)n * 2));
  HIP_TRY(s->d_hits.alloc(n));                  // 16 bytes per ray; the costs take 8
  HIP_TRY(hipMemcpyAsync(s->d_rays.p, rays, sizeof(cpt_ray) * n, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(launch_trace(s->dev(), s->d_rays.p, n, nullptr, cull, s->d_hits.p, 0, 0, layout, s->d_stats.p, s->stream,
                       1));
  HIP_TRY(hipMemcpyAsync(cost2, s->d_hits.p, sizeof(uint32_t) * 2 * n, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return CPT_OK;
}

int cpt_trace_depth(cpt_scene* s, const cpt_ray* rays, uint32_t n, uint32_t spec_depth, uint32_t* out3) {
  if (!s || (n && (!rays || !out3))) return fail(CPT_E_ARG, "cpt_trace_depth: null argument");
  if (!s->uploaded) return fail(CPT_E_STATE, "cpt_trace_depth: scene not uploaded");
  if (n == 0) return CPT_OK;
  HIP_TRY(s->d_rays.alloc((size_t)n * 2));
  HIP_TRY(s->d_hits.alloc(n));                  // 16 bytes per ray; the results take 12
  HIP_TRY(hipMemcpyAsync(s->d_rays.p, rays, sizeof(cpt_ray) * n, hipMemcpyHostToDevice, s->stream));
  HIP_TRY(launch_bvh_depth(s->dev(), s->d_rays.p, n, spec_depth, (uint32_t*)s->d_hits.p, s->stream));
  HIP_TRY(hipMemcpyAsync(out3, s->d_hits.p, sizeof(uint32_t) * 3 * n, hipMemcpyDeviceToHost, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  return CPT_OK;
}

int cpt_filter_cross_bilateral_device(const float* d_color, const float* d_pos, const float* d_norm,
                                      const float* d_diff, uint32_t width, uint32_t height, float sigma_color,
                                      float sigma_pos, float sigma_norm, float sigma_diff, uint32_t radius,
                                      const uint32_t* d_sample_n, uint32_t nd, float* d_filtered, float* d_err,
                                      void* stream) {
  if (!d_color || !d_pos || !d_norm || !d_sample_n || !d_filtered || !d_err)
    return fail(CPT_E_ARG, "cpt_filter_cross_bilateral_device: null argument");
  if (width == 0 || height == 0) return fail(CPT_E_ARG, "cpt_filter_cross_bilateral_device: empty image");
  if (radius > (u32)kMaxFilterRadius) return fail(CPT_E_LIMIT, "cpt_filter_cross_bilateral_device: radius %u > 31", radius);
  hipStream_t st = (hipStream_t)stream;
  float g[kFilterTable];
  filter_table(sigma_color, radius, g);
  float* d_g = nullptr;
  HIP_TRY(hipMalloc((void**)&d_g, sizeof g));
  HIP_TRY(hipMemcpyAsync(d_g, g, sizeof g, hipMemcpyHostToDevice, st));
  const FilterArgs f = filter_args(width, height, sigma_color, sigma_pos, sigma_norm, sigma_diff, radius, nd,
                                   d_diff != nullptr);
  hipError_t e = launch_filter_error(f, d_color, d_pos, d_norm, d_diff, d_g, d_sample_n, d_filtered, d_err, st);
  hipError_t e2 = hipStreamSynchronize(st);
  (void)hipFree(d_g);
  HIP_TRY(e);
  HIP_TRY(e2);
  return CPT_OK;
}

int cpt_filter_cross_bilateral_device_noalloc(const float* d_color, const float* d_pos, const float* d_norm,
                                              const float* d_diff, uint32_t width, uint32_t height,
                                              float sigma_color, float sigma_pos, float sigma_norm, float sigma_diff,
                                              uint32_t radius, const uint32_t* d_sample_n, uint32_t nd,
                                              float* d_filtered, float* d_err, void* stream) {
  if (!d_color || !d_pos || !d_norm || !d_sample_n || !d_filtered || !d_err)
    return fail(CPT_E_ARG, "cpt_filter_cross_bilateral_device_noalloc: null argument");
  if (width == 0 || height == 0) return fail(CPT_E_ARG, "cpt_filter_cross_bilateral_device_noalloc: empty image");
  if (radius > (u32)kMaxFilterRadius)
    return fail(CPT_E_LIMIT, "cpt_filter_cross_bilateral_device_noalloc: radius %u > %d: the %d-entry scratch table "
                "holds 2 * radius + 1 taps", radius, kMaxFilterRadius, kFilterTable);
  float* d_g = nullptr;
  int rc = filter_scratch(&d_g);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  float g[kFilterTable];
  filter_table(sigma_color, radius, g);
  // every entry is rewritten, so nothing of an earlier call's table survives
  HIP_TRY(hipMemcpyAsync(d_g, g, sizeof g, hipMemcpyHostToDevice, st));
  const FilterArgs f = filter_args(width, height, sigma_color, sigma_pos, sigma_norm, sigma_diff, radius, nd,
                                   d_diff != nullptr);
  hipError_t e = launch_filter_error(f, d_color, d_pos, d_norm, d_diff, d_g, d_sample_n, d_filtered, d_err, st);
  hipError_t e2 = hipStreamSynchronize(st);       // g is on this stack frame until the copy has run
  HIP_TRY(e);
  HIP_TRY(e2);
  return CPT_OK;
}

int cpt_trace_counters(cpt_scene* s, uint64_t* nodes, uint64_t* tris, uint64_t* rays, int reset) {
  if (!s) return fail(CPT_E_ARG, "cpt_trace_counters: null scene");
  if (!s->d_stats.p) return fail(CPT_E_STATE, "cpt_trace_counters: scene not uploaded");
  DevStats h;
  HIP_TRY(hipStreamSynchronize(s->stream));
  HIP_TRY(hipMemcpy(&h, s->d_stats.p, sizeof h, hipMemcpyDeviceToHost));
  if (nodes) *nodes = h.ext_nodes;
  if (tris) *tris = h.ext_tris;
  if (rays) *rays = h.ext_rays;
  if (reset) HIP_TRY(hipMemset(s->d_stats.p, 0, sizeof(DevStats)));
  return CPT_OK;
}

}  // extern "C"

// ============================================================== renderer

namespace {

bool is_adaptive(int mode) { return mode >= CPT_MODE_BDPT_ADAPTIVE_PLL && mode <= CPT_MODE_BDPT_ADAPTIVE_EMP; }
bool is_pt(int mode) { return mode == CPT_MODE_PT_STREAM || mode == CPT_MODE_PT_STREAM_REGEN; }
bool is_disney(int mode) { return mode == CPT_MODE_PT_DISNEY_MIS; }
bool is_bdpt_regen(int mode) { return mode == CPT_MODE_BDPT_REGEN; }
bool uses_filter(int mode) { return mode == CPT_MODE_BDPT_ADAPTIVE_PLL || mode == CPT_MODE_BDPT_ADAPTIVE_BINARY; }

struct Attrib {
  const char* name;
  int type;  // 0 uint, 1 int, 2 bool, 3 float, 4 enum
  double value;
};

struct PathSet {
  DevArray<float4> ray_o, ray_d, mul, smp, acc;
  DevArray<u32> pix, sid, depth;
  DevArray<uint64_t> rng;
  hipError_t alloc(size_t n) {
    hipError_t e;
    if ((e = ray_o.alloc(n)) || (e = ray_d.alloc(n)) || (e = mul.alloc(n)) || (e = smp.alloc(n)) ||
        (e = acc.alloc(n)) || (e = pix.alloc(n)) || (e = sid.alloc(n)) || (e = depth.alloc(n)) ||
        (e = rng.alloc(n)))
      return e;
    return hipSuccess;
  }
  PathBuf view() {
    PathBuf b;
    b.ray_o = ray_o.p; b.ray_d = ray_d.p; b.pix = pix.p; b.sid = sid.p; b.depth = depth.p;
    b.mul = mul.p; b.smp = smp.p; b.acc = acc.p; b.rng = rng.p;
    return b;
  }
};

// Attribute names: the reference's own list for the adaptive renderer
// (src/cudaRTBDPTStreamAdapBilatAtomic.cu:53-71, same order), then this
// build's estimator parameters.
const char* const kEyeDepth = "Eye Depth";
const char* const kLightDepth = "Light Depth";
const char* const kLightPaths = "Light Paths";
const char* const kSeed = "Seed";
const char* const kLiteralShadow = "Literal Shadow Endpoint";
const char* const kWarmup = "Warm-up Frames";
// the adaptive files' output selection (AdapBilatAtomic.cu:30-31): which image a Render call returns
const char* const kDebugMode = "Debug Mode";
// Filter and allocation extensions, each off at 0 (the reference's behaviour):
// the spatial sigma of the cross-bilateral filter (0 = the colour sigma, as the
// reference passes it); a relative-depth term (z_q - z_p) / z_p with this sigma,
// z the first-hit distance from the camera; the fewest paths a pixel gets (the
// warm-up frames are raised to it); the most paths a pixel gets (a pixel that has
// them is given no more, and a pass's proportional count stops at the cap)
const char* const kFilterSpatial = "Filter Spatial EuD";
const char* const kFilterDepth = "Filter Depth EuD";
const char* const kMinPaths = "Min Pixel Paths";
const char* const kMaxPaths = "Max Pixel Paths";
const char* const kInFlight = "Frames In Flight";
// adaptive mode: frames' worth of samples (k * W * H) one adaptive pass
// allocates from one error map and traces as one wavefront (1 = the
// reference's one frame per allocation)
const char* const kPassFrames = "Adaptive Pass Frames";
// BVH node format of the traversal kernels: 0 = 128-byte FP32 4-wide nodes,
// 1 = the same nodes quantized to 64 bytes
const char* const kNodeFormat = "Node Format";
// traversal scheduling: 0 = one ray per thread; 1..64 = a persistent grid
// whose lanes take the next ray from a global queue when theirs is finished,
// a wave refilling once that many of its lanes are idle
const char* const kRefill = "Traversal Refill";
// the same for the extension rays of the second and later bounces, which are
// less coherent than the first bounce's (one pixel's jittered rays per wave):
// refilling at 48 idle lanes there is 4 % faster on the sphere grid
const char* const kRefillSecondary = "Secondary Refill";
// the same for shadow rays
const char* const kRefillShadow = "Shadow Refill";
// workgroups per compute unit of the refilling grid
const char* const kRefillGroups = "Refill Workgroups";
// path capacity below which rays go one thread per ray whatever the refill
// settings: a grid that can hold the whole launch gains nothing from refilling
// (one frame of 800 x 600 per call: shadow 10 % faster one thread per ray)
const char* const kRefillMinPaths = "Refill Min Paths";
// allocation signal from the learned model's output y (log per-sample
// variance): 0 = nd / (nd + N) * exp(y) / N, the filter-residual estimate's
// shape; 1 = exp(y) / N, the expected squared error (samples settle where N
// grows with the per-sample standard deviation); 2 = exp(y) / N^2, the error
// one more sample removes (sigma^2 / N - sigma^2 / (N + 1) ~ sigma^2 / N^2);
// 3 = sqrt(exp(y) / N * e_plain), the geometric mean with the mode's own error
// map (filter residual or empirical variance), a hedge where either misjudges
const char* const kLearnedSignal = "Learned Signal";
// 1: the plain stream renderer traces each batch with one megakernel (one thread
// carries its sample through every bounce, the reference's regeneration
// baseline) instead of the wavefront passes; same image, bit for bit.  Its
// time is reported as extension time.
const char* const kMegakernel = "Megakernel";
const char* const kTailBounce = "Tail Bounce";
const char* const kLdsStack = "Traversal LDS Stack";
const char* const kLdsStackShadow = "Shadow LDS Stack";
constexpr u32 kMaxInFlight = 90;   // 30 three-frame light-cache groups + partial ones at either end

}  // namespace

// ---- learned per-pixel error model ---------------------------------------

struct cpt_mlp {
  hipStream_t stream = nullptr;
  DevArray<float> w, m, v, grad, loss, b1, b2, w3, b3, part, loss_part;
  DevArray<uint16_t> w1f, w2tf, w2f;
  int step = 0;
  ~cpt_mlp() {
    if (stream) (void)hipStreamDestroy(stream);
  }
  MlpState state() {
    MlpState s;
    s.w = w.p; s.m = m.p; s.v = v.p; s.grad = grad.p; s.loss = loss.p;
    s.dev.w1f = w1f.p; s.dev.w2tf = w2tf.p; s.dev.w2f = w2f.p;
    s.dev.b1 = b1.p; s.dev.b2 = b2.p; s.dev.w3 = w3.p; s.dev.b3 = b3.p;
    return s;
  }
};

struct cpt_renderer {
  int mode = 0;
  u32 W = 0, H = 0;
  std::vector<Attrib> attribs;
  hipStream_t stream = nullptr;
  PathSet A, B;
  DevArray<float4> hits;
  DevArray<float4> sh_o, sh_d, sh_c;
  DevArray<LightVert> lv;
  DevArray<FrameCounters> fc;
  DevArray<DevStats> stats;
  DevArray<float> result;
  DevArray<u32> sample_n;
  bool allocated = false;
  u32 frame = 0;
  bool started = false;
  bool have_cam = false;
  float cam[9] = {0};
  const cpt_scene* scene = nullptr;
  u32 scene_mesh_version = 0, scene_mat_version = 0;
  bool profiling = false;
  uint64_t extend_launches = 0;
  struct Ev { hipEvent_t a, b; int kind; };
  std::vector<Ev> events;          // recorded pairs awaiting readback
  std::vector<hipEvent_t> pool;    // reusable events
  double ms[11] = {0};  // extend, shadow, shade, light, total, filter, error sum, generation, accumulation, model,
                        // tail
  // shadow rays of bounce b traced on `side` while bounce b + 1's extension
  // rays run on `stream` (a captured batch iteration takes the side stream
  // into its graph through the same events)
  bool overlap_shadow = true;
  hipStream_t side = nullptr;
  hipEvent_t ev_shaded = nullptr, ev_shadow = nullptr;
  // the light pass of a plain batch also runs on `side`, beside the first
  // extension (which reads no light vertex); the first shading waits for it
  hipEvent_t ev_light = nullptr;
  bool light_pending = false;
  ~cpt_renderer() {
    for (auto& e : events) { pool.push_back(e.a); pool.push_back(e.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
    if (ev_shaded) (void)hipEventDestroy(ev_shaded);
    if (ev_shadow) (void)hipEventDestroy(ev_shadow);
    if (ev_light) (void)hipEventDestroy(ev_light);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }
  hipError_t take(hipEvent_t* e) {
    if (!pool.empty()) { *e = pool.back(); pool.pop_back(); return hipSuccess; }
    return hipEventCreate(e);
  }
  double attr(const char* name) const {
    for (const auto& a : attribs)
      if (std::strcmp(a.name, name) == 0) return a.value;
    return 0.0;
  }
  void set_attr(const char* name, double v) {
    for (auto& a : attribs)
      if (std::strcmp(a.name, name) == 0) a.value = v;
  }
  // adaptive mode
  DevArray<float> gpos, gnrm, gdif, filtered, err, gauss, err_sum;
  DevArray<u32> counts, offsets, block_tot, total, big_list, big_n;
  DevArray<u32> paths;                  // eye paths traced per pixel since the last restart
  DevArray<float> sq, maxe;                  // second moment (W*H*3), max of the error map
  DevArray<u32> flags, rank, kept;           // round-based allocation
  DevArray<RoundState> rounds;
  u32 last_rounds = 0;
  DevArray<double> block_sums;
  DevArray<float4> rec_acc, rec_pos, rec_nrm, rec_dif;
  DevArray<u32> term_bounce, tend;           // path regeneration: where each path ended, per-frame last bounce
  DevArray<uint64_t> term_rng;               // and the RNG state it ended in
  float gauss_sigma = -1.f;
  u32 gauss_radius = 0;
  u32 persist_chunk = 0, persist_grid = 0;   // persistent extend grid (0: one thread per ray)
  int cus = 0;                               // compute units of the device
  DevArray<u32> queue;                       // its global work-queue counter
  cpt_mlp* mlp = nullptr;              // learned error model (adaptive mode)
  DevArray<float> mlp_x, mlp_y, mlp_e, sq_f, sq_e;
  u32 cap = 0;
  u32 slot_margin = 1024;              // adaptive path slots beyond W*H*pass frames (rounding of the counts)
  long long slot_group[kLightSlots];   // three-frame group whose light cache each slot holds (-1: none)
  void forget_light_caches() { for (auto& g : slot_group) g = -1; }
  SampleRec rec() {
    SampleRec r;
    r.acc = rec_acc.p; r.gpos = rec_pos.p; r.gnrm = rec_nrm.p; r.gdif = rec_dif.p;
    return r;
  }
};

namespace {

int renderer_alloc(cpt_renderer* r) {
  size_t n = (size_t)r->W * r->H;
  // adaptive frames draw floor(W*H*e_i/sum e) samples: at most W*H plus float
  // rounding of the ratios, bounded by the margin (and clamped on the device);
  // plain frames trace "Frames In Flight" frames as one wavefront
  size_t cap = is_adaptive(r->mode) ? n * (size_t)r->attr(kPassFrames) + r->slot_margin
                                    : n * (size_t)r->attr(kInFlight);
  if (cap > 0xFFFFFFFFull / 2) return fail(CPT_E_LIMIT, "render: %zu path slots exceed 2^31", cap);
  r->cap = (u32)cap;
  r->forget_light_caches();
  HIP_TRY(r->A.alloc(cap));
  HIP_TRY(r->B.alloc(cap));
  HIP_TRY(r->hits.alloc(cap));
  HIP_TRY(r->sh_o.alloc(cap));
  HIP_TRY(r->sh_d.alloc(cap));
  HIP_TRY(r->sh_c.alloc(cap));
  if (is_adaptive(r->mode)) {
    HIP_TRY(r->gpos.alloc(n * 3));
    HIP_TRY(r->sq.alloc(n * 3));
    HIP_TRY(r->flags.alloc(n));
    HIP_TRY(r->rank.alloc(n));
    HIP_TRY(r->rounds.alloc(1));
    HIP_TRY(r->maxe.alloc(1));
    HIP_TRY(r->kept.alloc(1));
    HIP_TRY(r->gnrm.alloc(n * 3));
    HIP_TRY(r->gdif.alloc(n * 3));
    HIP_TRY(r->filtered.alloc(n * 3));
    HIP_TRY(r->err.alloc(n));
    HIP_TRY(r->gauss.alloc(64));
    HIP_TRY(r->err_sum.alloc(1));
    HIP_TRY(r->counts.alloc(n));
    HIP_TRY(r->paths.alloc(n));
    HIP_TRY(r->offsets.alloc(n));
    HIP_TRY(r->block_tot.alloc(n / 1024 + 2));
    HIP_TRY(r->total.alloc(1));
    HIP_TRY(r->block_sums.alloc(n / 256 + 2));
    HIP_TRY(r->rec_pos.alloc(cap));
    HIP_TRY(r->rec_nrm.alloc(cap));
    HIP_TRY(r->rec_dif.alloc(cap));
    HIP_TRY(r->big_list.alloc(n));
    HIP_TRY(r->big_n.alloc(1));
    for (auto* a : {&r->gpos, &r->gnrm, &r->gdif, &r->filtered, &r->sq})
      HIP_TRY(hipMemsetAsync(a->p, 0, sizeof(float) * n * 3, r->stream));
    HIP_TRY(hipMemsetAsync(r->err.p, 0, sizeof(float) * n, r->stream));
    HIP_TRY(hipMemsetAsync(r->counts.p, 0, sizeof(u32) * n, r->stream));
    HIP_TRY(hipMemsetAsync(r->paths.p, 0, sizeof(u32) * n, r->stream));
  }
  HIP_TRY(r->rec_acc.alloc(cap));
  if (r->mode == CPT_MODE_PT_STREAM_REGEN) {
    HIP_TRY(r->term_bounce.alloc(cap));
    HIP_TRY(r->term_rng.alloc(cap));
    HIP_TRY(r->tend.alloc(kMaxInFlight));
  }
  HIP_TRY(r->queue.alloc(2));
  HIP_TRY(r->fc.alloc(1));
  HIP_TRY(r->stats.alloc(1));
  HIP_TRY(r->result.alloc(n * 3));
  HIP_TRY(r->sample_n.alloc(n));
  HIP_TRY(hipMemsetAsync(r->stats.p, 0, sizeof(DevStats), r->stream));
  HIP_TRY(hipMemsetAsync(r->result.p, 0, sizeof(float) * n * 3, r->stream));
  HIP_TRY(hipMemsetAsync(r->sample_n.p, 0, sizeof(u32) * n, r->stream));
  HIP_TRY(hipMemsetAsync(r->fc.p, 0, sizeof(FrameCounters), r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  r->allocated = true;
  return CPT_OK;
}

int lv_alloc(cpt_renderer* r) {
  size_t need = (size_t)r->attr(kLightPaths) * ((size_t)r->attr(kLightDepth) + 1) * kLightSlots;
  if (is_bdpt_regen(r->mode)) need = kRegenLightVerts;     // one fixed 128 x 5 cache
  if (need > r->lv.n) r->forget_light_caches();
  HIP_TRY(r->lv.alloc(need));
  return CPT_OK;
}

// Light caches for frames [f0, f0 + nf): one per three-frame group, in slot
// group % kLightSlots, rebuilt only when the slot holds another group.  A
// cache depends only on (seed, group), so a rebuild reproduces it exactly.
LightJobs light_jobs(cpt_renderer* r, u32 f0, u32 nf) {
  LightJobs j;
  std::memset(&j, 0, sizeof j);
  for (u32 g = f0 / 3; g <= (f0 + nf - 1) / 3; g++) {
    u32 slot = g % kLightSlots;
    if (r->slot_group[slot] != (long long)g) {
      j.slot[j.n] = slot;
      j.frame[j.n] = 3 * g;
      j.n++;
      r->slot_group[slot] = g;
    }
  }
  return j;
}

hipError_t timed_begin(cpt_renderer* r, int kind, hipEvent_t* a, hipStream_t st = nullptr) {
  if (!r->profiling) return hipSuccess;
  hipError_t e = r->take(a);
  if (e) return e;
  (void)kind;
  return hipEventRecord(*a, st ? st : r->stream);
}
hipError_t timed_end(cpt_renderer* r, int kind, hipEvent_t a, hipStream_t st = nullptr) {
  if (!r->profiling) return hipSuccess;
  hipEvent_t b;
  hipError_t e = r->take(&b);
  if (e) return e;
  e = hipEventRecord(b, st ? st : r->stream);
  r->events.push_back({a, b, kind});
  return e;
}

// Camera basis of src/cudaRTBDPTStream.cu:823-824 (host float Vec3 math).
void camera_basis(const float* d, const float* up, float* right, float* up2) {
  float rx = d[1] * up[2] - d[2] * up[1], ry = d[2] * up[0] - d[0] * up[2], rz = d[0] * up[1] - d[1] * up[0];
  float rl = std::sqrt(rx * rx + ry * ry + rz * rz);
  right[0] = rx / rl; right[1] = ry / rl; right[2] = rz / rl;
  float ux = right[1] * d[2] - right[2] * d[1], uy = right[2] * d[0] - right[0] * d[2], uz = right[0] * d[1] - right[1] * d[0];
  float ul = std::sqrt(ux * ux + uy * uy + uz * uz);
  up2[0] = ux / ul; up2[1] = uy / ul; up2[2] = uz / ul;
}

// Gaussian table of cudaBilateralFilter::updateGaussian (cudaBilateralFilter.cu:5-24),
// spatial sigma = colour sigma as the reference passes it (AdapBilatAtomic.cu:1327).
int update_gauss(cpt_renderer* r, float sigma, u32 radius) {
  if (sigma == r->gauss_sigma && radius == r->gauss_radius) return CPT_OK;
  float g[64] = {0};
  for (u32 i = 0; i < 2 * radius + 1; i++) {
    int x = (int)i - (int)radius;
    g[i] = std::exp(-(float)(x * x) / (2 * sigma * sigma));
  }
  HIP_TRY(hipMemcpyAsync(r->gauss.p, g, sizeof g, hipMemcpyHostToDevice, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  r->gauss_sigma = sigma;
  r->gauss_radius = radius;
  return CPT_OK;
}

// gen_eye: the batch starts from fresh eye paths in r->A.  With refilling
// waves at bounce 0, quantized nodes and one view nothing writes them: the first extension and
// shading compute each slot's eye path where k_gen_eye would have stored it
// (3 GB of path state at 1280x720 and 64 frames).  Otherwise k_gen_eye writes
// them here: a multi-view batch finds each pixel's view by a search of
// dependent loads, which a refilling wave would wait for (4% slower over 32
// views).
int bounce_loop(cpt_renderer* r, const DevScene& sc, const FrameArgs& fa, u32 max_n, const Accum& ac,
                bool gen_eye = false) {
  hipEvent_t ev = nullptr;
  PathSet* cur = &r->A;
  PathSet* nxt = &r->B;
  ShadowBuf sb;
  sb.ray_o = r->sh_o.p; sb.ray_d = r->sh_d.p; sb.contrib = r->sh_c.p;
  const bool big = (double)max_n >= r->attr(kRefillMinPaths);
  const uint32_t refill_sh = big ? (uint32_t)std::min(64.0, std::max(0.0, r->attr(kRefillShadow))) : 0u;
  // Shadow rays of bounce b go to the side stream once bounce b is shaded and
  // overlap bounce b + 1's extension rays (which read only ray origins and
  // directions; the shadow kernel writes only accumulators); shading of bounce
  // b + 1, which reads those accumulators and rewrites the shadow queue, waits
  // for it.  The two refilling grids take their rays from separate counters.
  const bool ov = r->overlap_shadow;
  if (ov) {
    if (!r->side) HIP_TRY(hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
    if (!r->ev_shaded) HIP_TRY(hipEventCreateWithFlags(&r->ev_shaded, hipEventDisableTiming));
    if (!r->ev_shadow) HIP_TRY(hipEventCreateWithFlags(&r->ev_shadow, hipEventDisableTiming));
  }
  hipStream_t sst = ov ? r->side : r->stream;
  const bool eye0 =
      gen_eye && big && r->attr(kRefill) >= 1.0 && r->attr(kNodeFormat) != 0 && fa.n_views == 0;
  if (gen_eye && !eye0) HIP_TRY(launch_gen_eye(fa, r->A.view(), r->stream));
  bool shadow_pending = false;
  // "Tail Bounce" t > 0: bounces t.. run as one k_tail launch (plain BDPT's
  // frames and the adaptive frames after the warm-up, whose accumulation goes
  // through the sample records; the first-hit G-buffer is written at bounce 0,
  // before any tail)
  const double tail_attr = r->attr(kTailBounce);
  const u32 tail = tail_attr >= 1.0 && tail_attr < (double)fa.eye_depth && ac.apply == 2 &&
                           !(ac.variant & (kVariantPT | kVariantDisney))
                       ? (u32)tail_attr
                       : 0u;
  for (u32 b = 0; b < fa.eye_depth; b++) {
    if (tail && b == tail) {
      if (shadow_pending) {                      // the tail shades: bounce b - 1's occlusion results first
        HIP_TRY(hipStreamWaitEvent(r->stream, r->ev_shadow, 0));
        shadow_pending = false;
      }
      if (r->light_pending) {
        HIP_TRY(hipStreamWaitEvent(r->stream, r->ev_light, 0));
        r->light_pending = false;
      }
      const uint32_t grid =
          (uint32_t)(r->cus > 0 ? r->cus : 256) * (uint32_t)std::min(32.0, std::max(1.0, r->attr(kRefillGroups)));
      HIP_TRY(timed_begin(r, 10, &ev));
      HIP_TRY(launch_tail(sc, fa, cur->view(), r->fc.p, b, grid, r->lv.p, ac, (int)r->attr(kLdsStack),
                          r->attr(kNodeFormat) != 0, r->stats.p, r->stream));
      HIP_TRY(timed_end(r, 10, ev));
      break;
    }
    r->extend_launches++;
    HIP_TRY(timed_begin(r, 0, &ev));
    const int quant = r->attr(kNodeFormat) != 0;
    const uint32_t refill =
        big ? (uint32_t)std::min(64.0, std::max(0.0, r->attr(b == 0 ? kRefill : kRefillSecondary))) : 0u;
    const uint32_t dyn_grid =
        (uint32_t)(r->cus > 0 ? r->cus : 256) * (uint32_t)std::min(32.0, std::max(1.0, r->attr(kRefillGroups)));
    if (refill)
      HIP_TRY(launch_trace_dyn(sc, cur->view().ray_o, cur->view().ray_d, r->fc.p->active + b, r->hits.p, nullptr,
                               PathBuf{}, r->stats.p, (int)r->attr(kLdsStack), quant, 0, dyn_grid, refill, r->queue.p,
                               r->stream, eye0 && b == 0 ? &fa : nullptr));
    else if (r->persist_chunk)
      HIP_TRY(launch_extend_persist(sc, cur->view(), r->fc.p, (int)b, r->hits.p, r->stats.p, (int)r->attr(kLdsStack),
                                    quant, r->persist_chunk, r->persist_grid, r->queue.p, r->stream));
    else
      HIP_TRY(launch_extend(sc, cur->view(), r->fc.p, (int)b, max_n, r->hits.p, r->stats.p, (int)r->attr(kLdsStack),
                            quant, r->stream));
    HIP_TRY(timed_end(r, 0, ev));
    if (shadow_pending) {
      HIP_TRY(hipStreamWaitEvent(r->stream, r->ev_shadow, 0));
      shadow_pending = false;
    }
    if (r->light_pending) {
      HIP_TRY(hipStreamWaitEvent(r->stream, r->ev_light, 0));
      r->light_pending = false;
    }
    HIP_TRY(timed_begin(r, 2, &ev));
    HIP_TRY(launch_shade(sc, fa, cur->view(), nxt->view(), r->hits.p, r->fc.p, (int)b, max_n, sb, r->lv.p, ac,
                         r->stream, eye0 && b == 0));
    HIP_TRY(timed_end(r, 2, ev));
    // no emissive triangle: no light vertex to connect to (and the path
    // tracers connect nothing), so no shading emits an occlusion ray
    const bool may_emit = sc.n_light_tris > 0 && !(ac.variant & kVariantPT);
    if (b + 1 < fa.eye_depth && may_emit) {
      if (ov) {
        HIP_TRY(hipEventRecord(r->ev_shaded, r->stream));
        HIP_TRY(hipStreamWaitEvent(sst, r->ev_shaded, 0));
      }
      HIP_TRY(timed_begin(r, 1, &ev, sst));
      if (refill_sh)
        HIP_TRY(launch_trace_dyn(sc, sb.ray_o, sb.ray_d, r->fc.p->shadow + b, nullptr, sb.contrib, nxt->view(),
                                 r->stats.p, (int)r->attr(kLdsStackShadow), quant, 1, dyn_grid, refill_sh,
                                 r->queue.p + (ov ? 1 : 0), sst));
      else
        HIP_TRY(launch_shadow(sc, sb, nxt->view(), r->fc.p, (int)b, max_n, (int)r->attr(kLdsStackShadow), quant,
                              r->stats.p, sst));
      HIP_TRY(timed_end(r, 1, ev, sst));
      if (ov) {
        HIP_TRY(hipEventRecord(r->ev_shadow, sst));
        shadow_pending = true;
      }
    }
    std::swap(cur, nxt);
  }
  if (shadow_pending) HIP_TRY(hipStreamWaitEvent(r->stream, r->ev_shadow, 0));
  if (r->light_pending) {                        // no bounce shaded: join the light pass here
    HIP_TRY(hipStreamWaitEvent(r->stream, r->ev_light, 0));
    r->light_pending = false;
  }
  return CPT_OK;
}

// After a failed enqueue: waits for whatever the side stream still runs (the
// light pass, a bounce's shadow rays), so no kernel writes the light caches or
// accumulators after the call has returned, and drops the pending light join.
void join_side(cpt_renderer* r) {
  if (r->side) (void)hipStreamSynchronize(r->side);
  (void)hipStreamSynchronize(r->stream);
  r->light_pending = false;
}

// One frame of cudaRTBDPTStreamAdapBilatAtomic::Render (AdapBilatAtomic.cu:1338-1538).
// (off, cam: the first pixel and the camera of one view of a multi-view batch)
MlpFeatureArgs feature_args(const cpt_renderer* r, size_t off = 0, const float* cam = nullptr) {
  MlpFeatureArgs a;
  const size_t o3 = 3 * off;
  a.result = r->result.p + o3; a.filtered = r->filtered.p + o3; a.err = r->err.p + off; a.sq = r->sq.p + o3;
  a.sq_f = r->sq_f.p + o3;
  a.gpos = r->gpos.p + o3; a.gnrm = r->gnrm.p + o3; a.gdif = r->gdif.p + o3;
  a.sample_n = r->sample_n.p + off;
  if (!cam) cam = r->cam;
  a.cam[0] = cam[0]; a.cam[1] = cam[1]; a.cam[2] = cam[2];
  return a;
}

// Cross-bilateral filter of the running mean and the weighted squared residual
// (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1409-1430) into r->filtered / r->err.
// One W x H image starting at pixel `off` of the renderer's buffers, seen
// from cam (a view of a multi-view batch; the renderer's own image: 0, W, H).
int run_filter_at(cpt_renderer* r, size_t off, u32 W, u32 H, const float* cam, const float* src, float* dst,
                  float* err) {
  float sc_ = (float)r->attr("Filter Color EuD"), sp_ = (float)r->attr("Filter Pos EuD");
  float sn_ = (float)r->attr("Filter Norm EuD"), sd_ = (float)r->attr("Filter Diff EuD");
  u32 radius = (u32)r->attr("Filter Radius");
  if (radius > (u32)kMaxFilterRadius) return fail(CPT_E_LIMIT, "render: filter radius %u exceeds 31", radius);
  // the spatial table takes the colour sigma as the reference's callers pass it
  // (AdapBilatAtomic.cu:1327) unless a spatial sigma of its own is set
  const float ss = (float)r->attr(kFilterSpatial);
  int rc = update_gauss(r, ss > 0.f ? ss : sc_, radius);
  if (rc) return rc;
  FilterArgs f = filter_args(W, H, sc_, sp_, sn_, sd_, radius, (u32)r->attr("Desired Max Sampling"),
                             r->attr("Filter Diffuse Flag") != 0);
  const float sz = (float)r->attr(kFilterDepth);
  if (sz > 0.f) {
    f.kz = 2.f * sz * sz;
    f.cam[0] = cam[0]; f.cam[1] = cam[1]; f.cam[2] = cam[2];
  }
  const size_t o3 = 3 * off;
  HIP_TRY(launch_filter_error(f, src + o3, r->gpos.p + o3, r->gnrm.p + o3, r->gdif.p + o3, r->gauss.p,
                              r->sample_n.p + off, dst + o3, err + off, r->stream));
  return CPT_OK;
}
int run_filter(cpt_renderer* r, const float* src, float* dst, float* err) {
  return run_filter_at(r, 0, r->W, r->H, r->cam, src, dst, err);
}
int run_filter(cpt_renderer* r) { return run_filter(r, r->result.p, r->filtered.p, r->err.p); }

// Inputs of the learned model: the second moment filtered like the image (its
// residual is discarded), then the per-pixel features.
int model_buffers(cpt_renderer* r) {
  const size_t n = (size_t)r->W * r->H;
  HIP_TRY(r->sq_f.alloc(n * 3));
  HIP_TRY(r->sq_e.alloc(n));
  if (r->mlp) {
    HIP_TRY(r->mlp_x.alloc(n * kMlpIn));
    HIP_TRY(r->mlp_y.alloc(n));
    HIP_TRY(r->mlp_e.alloc(n));
  }
  return CPT_OK;
}
int model_features(cpt_renderer* r, float* d_x) {
  const size_t n = (size_t)r->W * r->H;
  int rc = model_buffers(r);
  if (rc) return rc;
  rc = run_filter(r, r->sq.p, r->sq_f.p, r->sq_e.p);
  if (rc) return rc;
  HIP_TRY(launch_mlp_features(feature_args(r), (u32)n, d_x, r->stream));
  return CPT_OK;
}

// Round-based sample allocation into r->counts / r->offsets / r->total:
// window bisection (AdapBilatBinary) or window = max error (AdapEmp).  Rounds
// run on the device in batches; the host reads the done flag between batches.
// After kMaxRounds rounds the remaining slots are filled as the reference's
// fill kernel does.
constexpr u32 kMaxRounds = 1024;
constexpr u32 kRoundBatch = 8;
int allocate_rounds(cpt_renderer* r, const FrameArgs& fa, const float* err, u32 gen, u32 nd, u32 maxp) {
  const u32 n = r->W * r->H;
  const bool bisect = r->mode == CPT_MODE_BDPT_ADAPTIVE_BINARY;
  const float desired = (float)n / (float)(nd ? nd : 1u);
  const float min_prob = (float)r->attr("Min Trace Probability");
  HIP_TRY(launch_error_max(err, n, r->maxe.p, r->stream));
  HIP_TRY(launch_round_init(r->rounds.p, r->maxe.p, gen, bisect ? 1u : 0u, desired, min_prob, r->counts.p, n,
                            r->stream));
  for (u32 done_rounds = 0; done_rounds < kMaxRounds; done_rounds += kRoundBatch) {
    for (u32 j = 0; j < kRoundBatch; j++)
      HIP_TRY(launch_round(fa, err, n, r->rounds.p, r->flags.p, r->rank.p, r->block_tot.p, r->kept.p, r->counts.p,
                           r->stream));
    RoundState h;
    HIP_TRY(hipMemcpyAsync(&h, r->rounds.p, sizeof h, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    r->last_rounds = h.rounds;
    if (h.done) break;
  }
  HIP_TRY(launch_round_fill(r->rounds.p, n, r->counts.p, r->stream));
  HIP_TRY(launch_cap_counts(r->counts.p, r->paths.p, maxp, n, r->stream));
  HIP_TRY(launch_scan_counts(r->counts.p, n, r->offsets.p, r->block_tot.p, r->total.p, r->stream));
  return CPT_OK;
}

// Plain 1-spp warm-up frames before adaptive sampling starts (AdapBilatAtomic.cu:1372): at least
// "Warm-up Frames", raised to "Min Pixel Paths" so every pixel has that many paths.
u32 warmup_frames(const cpt_renderer* r) {
  return (u32)std::max(r->attr(kWarmup), r->attr(kMinPaths));
}

// One view of a multi-view adaptive batch: its first pixel in the
// concatenated buffers, its size and its camera position.
struct ViewSpan {
  size_t off;
  u32 W, H;
  const float* cam;
};
// Per-view allocation state of a multi-view adaptive batch: error sums and
// path totals (one per view), which become the views' queue bases.
struct ViewAlloc {
  const std::vector<ViewSpan>* views;
  float* sums;
  u32* totals;
};

int render_adaptive_frame(cpt_renderer* r, cpt_scene* s, const FrameArgs& fa, const ViewAlloc* va = nullptr) {
  DevScene sc = s->dev();
  const u32 n = r->W * r->H;
  const std::vector<ViewSpan> one{{0, r->W, r->H, r->cam}};
  const std::vector<ViewSpan>& views = va ? *va->views : one;
  hipEvent_t ev = nullptr;
  LightJobs jobs = light_jobs(r, fa.frame0, 1);
  if (jobs.n) {
    HIP_TRY(timed_begin(r, 3, &ev));
    HIP_TRY(launch_light_pass(sc, fa, jobs, r->lv.p, r->fc.p, r->stream));
    HIP_TRY(timed_end(r, 3, ev));
  }
  Accum ac;
  std::memset(&ac, 0, sizeof ac);
  ac.result = r->result.p; ac.sample_n = r->sample_n.p;
  ac.gpos = r->gpos.p; ac.gnrm = r->gnrm.p; ac.gdif = r->gdif.p;
  ac.sq = r->sq.p;
  ac.rec = r->rec();
  // AdapBilatAtomic's eye bounce adds the reflection offset; the bisection and
  // empirical files use the plain stream's bounce
  ac.variant = kVariantGbuf | (r->mode == CPT_MODE_BDPT_ADAPTIVE_PLL ? kVariantReflOffset : 0);
  const u32 maxp = (u32)r->attr(kMaxPaths);
  if (fa.frame0 < warmup_frames(r)) {              // plain 1 spp frames with G-buffer (:1372-1406)
    ac.apply = 1;
    HIP_TRY(launch_frame_begin(r->fc.p, n, r->stream));
    HIP_TRY(launch_gen_eye_rec(fa, r->A.view(), ac.rec, r->stream));
    int rc = bounce_loop(r, sc, fa, n, ac);
    if (rc) return rc;
    HIP_TRY(launch_add_paths(r->paths.p, nullptr, n, fa.frame0 == 0, r->stream));
  } else {
    int rc = CPT_OK;
    const u32 nd = (u32)r->attr("Desired Max Sampling");
    const float* alloc_err = r->err.p;
    if (r->mode == CPT_MODE_BDPT_ADAPTIVE_EMP) {    // AdapEmp.cu:1209
      HIP_TRY(timed_begin(r, 6, &ev));
      HIP_TRY(launch_emp_error(r->result.p, r->sq.p, r->sample_n.p, n, r->err.p, r->stream));
      HIP_TRY(timed_end(r, 6, ev));
    } else {
      HIP_TRY(timed_begin(r, 5, &ev));
      for (const ViewSpan& v : views) {
        rc = run_filter_at(r, v.off, v.W, v.H, v.cam, r->result.p, r->filtered.p, r->err.p);
        if (rc) return rc;
      }
      HIP_TRY(timed_end(r, 5, ev));
    }
    if (r->mlp) {                                  // learned estimate replaces the filter residual
      HIP_TRY(timed_begin(r, 9, &ev));
      rc = model_buffers(r);
      if (rc) return rc;
      for (const ViewSpan& v : views) {            // second moment filtered like the image, then the features
        rc = run_filter_at(r, v.off, v.W, v.H, v.cam, r->sq.p, r->sq_f.p, r->sq_e.p);
        if (rc) return rc;
        HIP_TRY(launch_mlp_features(feature_args(r, v.off, v.cam), v.W * v.H, r->mlp_x.p + v.off * kMlpIn,
                                    r->stream));
      }
      HIP_TRY(launch_mlp_infer(r->mlp->state(), r->mlp_x.p, n, r->mlp_y.p, r->stream));
      HIP_TRY(launch_mlp_err(r->mlp_y.p, r->sample_n.p, nd, (u32)r->attr(kLearnedSignal), n, r->err.p, r->mlp_e.p,
                             r->stream));
      HIP_TRY(timed_end(r, 9, ev));
      alloc_err = r->mlp_e.p;
    }
    // pixels that have their "Max Pixel Paths" get no weight in the allocation
    HIP_TRY(launch_cap_error(const_cast<float*>(alloc_err), r->paths.p, maxp, n, r->stream));
    if (r->mode == CPT_MODE_BDPT_ADAPTIVE_PLL && !va) {   // proportional counts (AdapBilatAtomic.cu:817-842)
      HIP_TRY(timed_begin(r, 6, &ev));
      HIP_TRY(launch_error_sum(alloc_err, n, r->block_sums.p, r->err_sum.p, r->stream));
      HIP_TRY(timed_end(r, 6, ev));
      HIP_TRY(timed_begin(r, 7, &ev));
      HIP_TRY(launch_alloc_scan(alloc_err, r->err_sum.p, n, n * fa.nframes, r->counts.p, r->offsets.p,
                                r->block_tot.p, r->total.p, r->cap, r->paths.p, maxp, r->stream));
    } else if (va) {                                 // the same per view, then one queue over all views
      HIP_TRY(timed_begin(r, 6, &ev));
      for (size_t k = 0; k < views.size(); k++)
        HIP_TRY(launch_error_sum(alloc_err + views[k].off, views[k].W * views[k].H, r->block_sums.p, va->sums + k,
                                 r->stream));
      HIP_TRY(timed_end(r, 6, ev));
      HIP_TRY(timed_begin(r, 7, &ev));
      for (size_t k = 0; k < views.size(); k++) {
        const ViewSpan& v = views[k];
        const u32 nv = v.W * v.H;
        HIP_TRY(launch_alloc_scan(alloc_err + v.off, va->sums + k, nv, nv * fa.nframes, r->counts.p + v.off,
                                  r->offsets.p + v.off, r->block_tot.p, va->totals + k, r->cap, r->paths.p + v.off,
                                  maxp, r->stream));
      }
      HIP_TRY(launch_view_bases(va->totals, (u32)views.size(), r->total.p, r->stream));
      HIP_TRY(launch_offset_views(r->offsets.p, n, fa, va->totals, r->stream));
    } else {                                         // rounds (AdapBilatBinary.cu:1348-1388, AdapEmp.cu:1232-1260)
      HIP_TRY(timed_begin(r, 7, &ev));
      rc = allocate_rounds(r, fa, alloc_err, n * fa.nframes, nd, maxp);
      if (rc) return rc;
    }
    HIP_TRY(launch_add_paths(r->paths.p, r->counts.p, n, 0, r->stream));
    HIP_TRY(launch_frame_begin_dev(r->fc.p, r->total.p, r->cap, r->stream));
    HIP_TRY(launch_gen_adaptive(fa, r->offsets.p, r->counts.p, r->total.p, r->cap, r->A.view(), ac.rec, r->stream));
    HIP_TRY(timed_end(r, 7, ev));
    ac.apply = 2;
    rc = bounce_loop(r, sc, fa, r->cap, ac);
    if (rc) return rc;
    HIP_TRY(timed_begin(r, 8, &ev));
    HIP_TRY(launch_accum_adaptive(r->offsets.p, r->counts.p, n, r->cap, ac.rec, r->result.p, r->sample_n.p, r->gpos.p,
                                  r->gnrm.p, r->gdif.p, r->sq.p, r->big_list.p, r->big_n.p, r->stream));
    HIP_TRY(timed_end(r, 8, ev));
  }
  HIP_TRY(launch_frame_end(r->fc.p, r->stats.p, fa.nframes, fa.eye_depth, r->stream));
  return CPT_OK;
}

// fa.nframes frames of cudaRTBDPTStream::Render traced as one wavefront
// (src/cudaRTBDPTStream.cu:874-946), then applied to the running mean in
// frame order: identical to fa.nframes sequential calls.
// cudaRTBDPTRegen::Render (src/cudaRTBDPTRegen.cu:728-792), frame by frame: the
// light cache when frame % 3 == 0, then one thread per pixel.
int render_bdpt_regen(cpt_renderer* r, cpt_scene* s, const FrameArgs& fa0) {
  DevScene sc = s->dev();
  FrameArgs fa = fa0;
  fa.nframes = 1;
  hipEvent_t ev = nullptr;
  for (u32 k = 0; k < fa0.nframes; k++) {
    fa.frame0 = fa0.frame0 + k;
    if (fa.frame0 % 3 == 0) {
      HIP_TRY(timed_begin(r, 3, &ev));
      HIP_TRY(launch_regen_light(sc, fa, r->lv.p, r->stream));
      HIP_TRY(timed_end(r, 3, ev));
    }
    r->extend_launches++;
    HIP_TRY(timed_begin(r, 0, &ev));
    HIP_TRY(launch_regen_eye(sc, fa, r->lv.p, r->result.p, r->sample_n.p, (int)r->attr(kLdsStack),
                             r->attr(kNodeFormat) != 0, r->stats.p, r->stream));
    HIP_TRY(timed_end(r, 0, ev));
  }
  return CPT_OK;
}

int render_batch(cpt_renderer* r, cpt_scene* s, const FrameArgs& fa) {
  if (is_adaptive(r->mode)) return render_adaptive_frame(r, s, fa);
  if (is_bdpt_regen(r->mode)) return render_bdpt_regen(r, s, fa);
  DevScene sc = s->dev();
  const u32 n = r->W * r->H * fa.nframes;
  hipEvent_t ev = nullptr;
  const bool pt = is_pt(r->mode), regen = r->mode == CPT_MODE_PT_STREAM_REGEN, disney = is_disney(r->mode);
  LightJobs jobs;
  std::memset(&jobs, 0, sizeof jobs);
  // the path tracers have no light pass (the Disney mode samples the emitters directly)
  if (!pt && !disney) jobs = light_jobs(r, fa.frame0, fa.nframes);
  const bool mega = !pt && r->attr(kMegakernel) != 0;
  if (jobs.n) {
    // The first extension reads no light vertex: with the side stream the
    // pass runs there, after everything queued so far (earlier shading reads
    // the cache slots it rebuilds), and the first shading waits for it.
    // It writes only the cache and FrameCounters::light_n.
    hipStream_t ls = r->stream;
    if (r->overlap_shadow && !mega) {
      if (!r->side) HIP_TRY(hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
      if (!r->ev_light) HIP_TRY(hipEventCreateWithFlags(&r->ev_light, hipEventDisableTiming));
      HIP_TRY(hipEventRecord(r->ev_light, r->stream));
      HIP_TRY(hipStreamWaitEvent(r->side, r->ev_light, 0));
      ls = r->side;
    }
    HIP_TRY(timed_begin(r, 3, &ev, ls));
    HIP_TRY(launch_light_pass(sc, fa, jobs, r->lv.p, r->fc.p, ls));
    HIP_TRY(timed_end(r, 3, ev, ls));
    if (ls != r->stream) {
      HIP_TRY(hipEventRecord(r->ev_light, ls));
      r->light_pending = true;
    }
  }
  HIP_TRY(launch_frame_begin(r->fc.p, n, r->stream));
  Accum ac;
  std::memset(&ac, 0, sizeof ac);
  ac.result = r->result.p;
  ac.sample_n = r->sample_n.p;
  ac.rec.acc = r->rec_acc.p;
  ac.apply = 2;
  ac.variant = pt ? kVariantPT | (regen ? kVariantPTRegen : 0) : disney ? kVariantDisney : 0;
  if (regen) {
    ac.term_bounce = r->term_bounce.p;
    ac.term_rng = r->term_rng.p;
    ac.tend = r->tend.p;
  }
  if (mega) {
    r->extend_launches++;
    HIP_TRY(timed_begin(r, 0, &ev));
    HIP_TRY(launch_mega(sc, fa, r->fc.p, r->lv.p, ac, (int)r->attr(kLdsStack), r->attr(kNodeFormat) != 0,
                        r->stats.p, r->stream));
    HIP_TRY(timed_end(r, 0, ev));
  } else {
    int rc = bounce_loop(r, sc, fa, n, ac, true);
    if (rc) return rc;
  }
  if (regen) {                                     // the restarts, as a second wavefront
    HIP_TRY(launch_frame_end(r->fc.p, r->stats.p, fa.nframes, fa.eye_depth, r->stream));
    HIP_TRY(timed_begin(r, 7, &ev));
    HIP_TRY(launch_regen(fa, r->A.view(), ac, r->fc.p, r->stream));
    HIP_TRY(timed_end(r, 7, ev));
    ac.term_bounce = nullptr;
    ac.term_rng = nullptr;
    int rc = bounce_loop(r, sc, fa, n, ac);
    if (rc) return rc;
  }
  HIP_TRY(timed_begin(r, 8, &ev));
  HIP_TRY(launch_apply_frames(fa, r->rec_acc.p, r->result.p, r->sample_n.p, pt ? 1 : 0, r->stream));
  HIP_TRY(timed_end(r, 8, ev));
  HIP_TRY(launch_frame_end(r->fc.p, r->stats.p, regen ? 0u : fa.nframes, fa.eye_depth, r->stream));
  return CPT_OK;
}

// The image a Render call hands back (src/cudaRTBDPTStreamAdapBilatAtomic.cu:1539-1552 and the same lines of the
// Binary and Emp files), chosen by the adaptive modes' "Debug Mode": 0 "None" the running mean; 1 "Traced" the
// paths traced per pixel in the last frame as grey (the reference adds 1 per traced path through its running
// mean, AdapBilatBinary.cu:1417-1421; here the running mean is left intact so the mode can be switched back);
// 2 "Prob" max(0, e) and 3 "Prob With Limit" max(Min Trace Probability, e) as grey, e the per-pixel error map
// (pt_applyPixelProbToResult_kernel, AdapBilatAtomic.cu:944-951); 4 "Filtered Result" the cross-bilateral image.
int copy_result(cpt_renderer* r, float* result) {
  const size_t n = (size_t)r->W * r->H;
  const u32 debug = is_adaptive(r->mode) ? (u32)r->attr(kDebugMode) : 0u;
  if (debug == 0) {
    HIP_TRY(hipMemcpyAsync(result, r->result.p, n * 12, hipMemcpyDeviceToHost, r->stream));
  } else if (debug == 4) {
    HIP_TRY(hipMemcpyAsync(result, r->filtered.p, n * 12, hipMemcpyDeviceToHost, r->stream));
  } else {
    std::vector<float> grey(n);
    HIP_TRY(hipMemcpyAsync(grey.data(), debug == 1 ? (const void*)r->counts.p : (const void*)r->err.p, n * 4,
                           hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    const float lo = debug == 3 ? (float)r->attr("Min Trace Probability") : 0.f;
    for (size_t i = 0; i < n; i++) {
      float v = debug == 1 ? (float)((const u32*)grey.data())[i] : std::fmax(lo, grey[i]);
      result[i * 3] = result[i * 3 + 1] = result[i * 3 + 2] = v;
    }
  }
  HIP_TRY(hipStreamSynchronize(r->stream));
  return CPT_OK;
}

int render_frames(cpt_renderer* r, cpt_scene* s, const float* pos, const float* dir, const float* up, float fov,
                  uint32_t nframes, float* result) {
  if (!r || !s || !pos || !dir || !up) return fail(CPT_E_ARG, "render: null argument");
  if (!s->has_bvh) return fail(CPT_E_STATE, "render: scene has no BVH");
  if (!s->uploaded) {                              // GetIsCudaDirty -> initAllSceneCudaMem (:830-838)
    int rc = cpt_scene_upload(s);
    if (rc) return rc;
  }
  u32 eye_depth = (u32)r->attr(kEyeDepth), light_depth = (u32)r->attr(kLightDepth);
  u32 light_paths = (u32)r->attr(kLightPaths);
  if (is_adaptive(r->mode)) {
    u32 radius = (u32)r->attr("Filter Radius");
    if (radius > (u32)kMaxFilterRadius) return fail(CPT_E_LIMIT, "render: filter radius %u exceeds 31", radius);
    if (r->attr(kWarmup) < 1) return fail(CPT_E_LIMIT, "render: warm-up frames must be >= 1");
    const double dm = r->attr(kDebugMode);
    if (dm != std::floor(dm) || dm > 4) return fail(CPT_E_ARG, "render: Debug Mode %g not in 0..4", dm);
    if (dm == 4 && !uses_filter(r->mode))
      return fail(CPT_E_ARG, "render: Debug Mode 4 (Filtered Result) needs a filtering mode; the empirical mode has none");
  }
  if (eye_depth < 1 || eye_depth > (u32)kMaxDepth) return fail(CPT_E_LIMIT, "render: eye depth %u not in 1..15", eye_depth);
  if (is_bdpt_regen(r->mode) && eye_depth > 8)
    return fail(CPT_E_LIMIT, "render: eye depth %u: the Path Regen renderer carries at most 8 levels", eye_depth);
  if (light_depth < 1 || light_depth > 64) return fail(CPT_E_LIMIT, "render: light depth %u not in 1..64", light_depth);
  if (light_paths < 1 || light_paths > (u32)kBlock) return fail(CPT_E_LIMIT, "render: light paths %u not in 1..256", light_paths);
  for (const char* a : {kLdsStack, kLdsStackShadow})
    if (!traversal_variant_ok((int)r->attr(a)))
      return fail(CPT_E_ARG, "render: %s %g is not a traversal variant (8, 16, 24, 806, 807, 808, 1606)", a, r->attr(a));
  const bool adaptive = is_adaptive(r->mode);
  u32 in_flight = (u32)r->attr(adaptive ? kPassFrames : kInFlight);
  if (in_flight < 1 || in_flight > kMaxInFlight)
    return fail(CPT_E_LIMIT, "render: %s %u not in 1..%u", adaptive ? kPassFrames : kInFlight, in_flight,
                kMaxInFlight);
  if (!r->allocated || (size_t)r->cap < (size_t)r->W * r->H * in_flight + (adaptive ? r->slot_margin : 0)) {
    int rc = renderer_alloc(r);
    if (rc) return rc;
  }
  int rc = lv_alloc(r);
  if (rc) return rc;
  if (r->mlp) {
    rc = model_buffers(r);
    if (rc) return rc;
  }
  float cam[9] = {pos[0], pos[1], pos[2], dir[0], dir[1], dir[2], up[0], up[1], up[2]};
  bool reset = !r->have_cam || std::memcmp(cam, r->cam, sizeof cam) != 0 || r->scene != s ||
               r->scene_mesh_version != s->mesh_version || r->scene_mat_version != s->mat_version;
  if (reset) r->forget_light_caches();
  std::memcpy(r->cam, cam, sizeof cam);
  r->have_cam = true;
  r->scene = s;
  r->scene_mesh_version = s->mesh_version;
  r->scene_mat_version = s->mat_version;
  FrameArgs fa;
  std::memset(&fa, 0, sizeof fa);
  fa.width = r->W;
  fa.height = r->H;
  for (int k = 0; k < 3; k++) { fa.cam_pos[k] = pos[k]; fa.cam_dir[k] = dir[k]; }
  camera_basis(dir, up, fa.cam_right, fa.cam_up);
  fa.fov = fov;
  fa.tan_half = std::tan(fov * 0.5f);             // the oracle's tanf of the same float
  fa.eye_depth = eye_depth;
  fa.light_depth = light_depth;
  fa.light_paths = light_paths;
  fa.flags = r->attr(kLiteralShadow) != 0 ? 1u : 0u;
  uint64_t seed = (uint64_t)r->attr(kSeed);
  fa.seed = seed;
  hipEvent_t ev_all = nullptr;
  HIP_TRY(timed_begin(r, 4, &ev_all));
  fa.lv_stride = light_paths * (light_depth + 1);
  u32 next = (reset || !r->started) ? 0u : r->frame + 1;   // frame counter (:826-828)
  for (u32 done = 0; done < nframes;) {
    u32 k = std::min(in_flight, nframes - done);
    if (adaptive && next < warmup_frames(r)) k = 1;        // warm-up frames one at a time
    fa.frame0 = next;
    fa.nframes = k;
    rc = render_batch(r, s, fa);
    if (rc) {
      join_side(r);
      return rc;
    }
    r->frame = next + k - 1;
    r->started = true;
    next += k;
    done += k;
  }
  HIP_TRY(timed_end(r, 4, ev_all));
  if (result) return copy_result(r, result);
  HIP_TRY(hipStreamSynchronize(r->stream));
  return CPT_OK;
}

// Pixels off .. off + n of one buffer (CPT_BUF_*) to the host.
int read_pixels(cpt_renderer* r, int which, size_t off, size_t n, void* out, const char* fn) {
  if (!r->allocated) return fail(CPT_E_STATE, "%s: nothing rendered yet", fn);
  const void* src = nullptr;
  size_t per = 4;                                  // bytes per pixel
  switch (which) {
    case CPT_BUF_RESULT: src = r->result.p + 3 * off; per = 12; break;
    case CPT_BUF_SAMPLE_N: src = r->sample_n.p + off; break;
    case CPT_BUF_GBUF_POS:
    case CPT_BUF_GBUF_NORM:
    case CPT_BUF_GBUF_DIFF:
    case CPT_BUF_FILTERED:
    case CPT_BUF_ERROR:
    case CPT_BUF_SAMPLE_COUNTS:
    case CPT_BUF_PIXEL_PATHS:
    case CPT_BUF_SECOND_MOMENT:
    case CPT_BUF_ALLOC_ERROR:
      if (!is_adaptive(r->mode)) return fail(CPT_E_ARG, "%s: buffer %d exists only in the adaptive modes", fn, which);
      if (which == CPT_BUF_FILTERED && !uses_filter(r->mode))
        return fail(CPT_E_ARG, "%s: the empirical-variance mode does not filter", fn);
      switch (which) {
        case CPT_BUF_GBUF_POS: src = r->gpos.p + 3 * off; per = 12; break;
        case CPT_BUF_GBUF_NORM: src = r->gnrm.p + 3 * off; per = 12; break;
        case CPT_BUF_GBUF_DIFF: src = r->gdif.p + 3 * off; per = 12; break;
        case CPT_BUF_FILTERED: src = r->filtered.p + 3 * off; per = 12; break;
        case CPT_BUF_SECOND_MOMENT: src = r->sq.p + 3 * off; per = 12; break;
        case CPT_BUF_ERROR: src = r->err.p + off; break;
        case CPT_BUF_SAMPLE_COUNTS: src = r->counts.p + off; break;
        case CPT_BUF_PIXEL_PATHS: src = r->paths.p + off; break;
        default: src = (r->mlp && r->mlp_e.p ? r->mlp_e.p : r->err.p) + off; break;
      }
      break;
    default:
      return fail(CPT_E_ARG, "%s: buffer %d not produced by this mode", fn, which);
  }
  HIP_TRY(hipMemcpyAsync(out, src, n * per, hipMemcpyDeviceToHost, r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return CPT_OK;
}

}  // namespace

extern "C" {

int cpt_renderer_create(cpt_renderer** out, int mode, uint32_t width, uint32_t height) {
  if (!out) return fail(CPT_E_ARG, "cpt_renderer_create: null output");
  if (mode < CPT_MODE_BDPT_STREAM || mode > CPT_MODE_BDPT_REGEN)
    return fail(CPT_E_ARG, "cpt_renderer_create: unknown mode %d", mode);
  if (width == 0 || height == 0 || (uint64_t)width * height > (1ull << 26))
    return fail(CPT_E_LIMIT, "cpt_renderer_create: %ux%u outside 1..2^26 pixels", width, height);
  cpt_renderer* r = new cpt_renderer();
  r->mode = mode;
  r->W = width;
  r->H = height;
  if (mode == CPT_MODE_BDPT_ADAPTIVE_EMP) {
    // src/cudaRTBDPTStreamAdapEmp.cu:24-35, indices 0-3
    r->attribs = {{"Adaptive Mode", 4, 0}, {"Desired Max Sampling", 0, 5},
                  {"Min Trace Probability", 3, 0}, {kDebugMode, 4, 0}};
  } else if (is_adaptive(mode)) {
    // src/cudaRTBDPTStreamAdapBilatAtomic.cu:26-71 and src/cudaRTBDPTStreamAdapBilatBinary.cu:26-68,
    // indices 0-14 as the reference numbers them
    r->attribs = {{"Adaptive Mode", 4, 0},          {"Desired Max Sampling", 0, 5},
                  {"Min Trace Probability", 3, 0},  {kDebugMode, 4, 0},
                  {"Filter Color EuD", 3, 50},      {"Filter Pos EuD", 3, 1},
                  {"Filter Norm EuD", 3, 0.25},     {"Filter Diff EuD", 3, 0.1},
                  {"Filter Radius", 0, 5},          {"Filter Diffuse Flag", 2, 0},
                  {"Avg Filter Time", 3, 0},        {"Avg MSE Time", 3, 0},
                  {"Avg Gen Time", 3, 0},           {"Avg Trace Time", 3, 0},
                  {"Avg Accum Time", 3, 0}};
  }
  // NORMALRAY_BOUND_MAX: 5 in the BDPT files, 10 in the path tracers (cudaRTPTStream.cu:7)
  // (cudaRTBDPTRegen: NORMALRAY_BOUND_MAX 5, src/cudaRTBDPTRegen.cu:4)
  r->attribs.push_back({kEyeDepth, 0, is_pt(mode) ? 10.0 : is_disney(mode) ? 8.0 : 5.0});
  r->attribs.push_back({kLightDepth, 0, 5});
  r->attribs.push_back({kLightPaths, 0, 128});
  r->attribs.push_back({kSeed, 0, (double)0x5EED0001u});
  r->attribs.push_back({kLiteralShadow, 2, 0});
  if (is_adaptive(mode)) {
    r->attribs.push_back({kWarmup, 0, 5});
    r->attribs.push_back({kMinPaths, 0, 0});
    r->attribs.push_back({kMaxPaths, 0, 0});
    if (mode != CPT_MODE_BDPT_ADAPTIVE_EMP) {
      r->attribs.push_back({kFilterSpatial, 3, 0});
      r->attribs.push_back({kFilterDepth, 3, 0});
    }
    r->attribs.push_back({kPassFrames, 0, 1});
    r->attribs.push_back({kLearnedSignal, 0, 0});
    // a pass traces one frame's paths, most of which end within two bounces:
    // from bounce 2 on one k_tail launch (90 ms against 102 ms for the 32-frame
    // 1920 x 1080 adaptive workload on the environment-lit heightfield)
    r->attribs.push_back({kTailBounce, 0, 2});
  } else {
    r->attribs.push_back({kInFlight, 0, is_bdpt_regen(mode) ? 1.0 : 24.0});
    if (mode == CPT_MODE_BDPT_STREAM) {
      r->attribs.push_back({kMegakernel, 0, 0});
      r->attribs.push_back({kTailBounce, 0, 0});
    }
  }
  // traversal kernel variant: 8 LDS stack entries, >= 6 waves per SIMD
  // (swept on the sphere grid: 1.58 vs 1.36 Grays/s for 16 entries)
  r->attribs.push_back({kLdsStack, 0, 806});
  r->attribs.push_back({kLdsStackShadow, 0, 806});
  r->attribs.push_back({kNodeFormat, 0, 1});
  r->attribs.push_back({kRefill, 0, 64});
  r->attribs.push_back({kRefillSecondary, 0, 48});
  r->attribs.push_back({kRefillShadow, 0, 16});
  r->attribs.push_back({kRefillMinPaths, 0, (double)(1u << 20)});
  r->attribs.push_back({kRefillGroups, 0, 8});
  int rc = ensure_stream(r->stream);
  if (rc) { delete r; return rc; }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&r->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    r->cus = 256;
  *out = r;
  return CPT_OK;
}

int cpt_render_iteration(cpt_renderer* r, cpt_scene* s, const float cam_pos[3], const float cam_dir[3],
                         const float cam_up[3], float fov, float* result) {
  return render_frames(r, s, cam_pos, cam_dir, cam_up, fov, 1, result);
}

int cpt_render_frames(cpt_renderer* r, cpt_scene* s, const float cam_pos[3], const float cam_dir[3],
                      const float cam_up[3], float fov, uint32_t n_frames, float* result) {
  if (n_frames == 0) return fail(CPT_E_ARG, "cpt_render_frames: n_frames = 0");
  return render_frames(r, s, cam_pos, cam_dir, cam_up, fov, n_frames, result);
}

int cpt_renderer_attrib_count(cpt_renderer* r, uint32_t* n) {
  if (!r || !n) return fail(CPT_E_ARG, "cpt_renderer_attrib_count: null argument");
  *n = (u32)r->attribs.size();
  return CPT_OK;
}

int cpt_renderer_get_attrib(cpt_renderer* r, uint32_t idx, char* name, uint32_t name_cap, int* type, double* value) {
  if (!r) return fail(CPT_E_ARG, "cpt_renderer_get_attrib: null renderer");
  if (idx >= r->attribs.size()) return fail(CPT_E_ARG, "cpt_renderer_get_attrib: index %u out of range", idx);
  const Attrib& a = r->attribs[idx];
  if (name && name_cap) { std::strncpy(name, a.name, name_cap - 1); name[name_cap - 1] = 0; }
  if (type) *type = a.type;
  if (value) *value = a.value;
  return CPT_OK;
}

int cpt_renderer_set_attrib(cpt_renderer* r, uint32_t idx, double value) {
  if (!r) return fail(CPT_E_ARG, "cpt_renderer_set_attrib: null renderer");
  if (idx >= r->attribs.size()) return fail(CPT_E_ARG, "cpt_renderer_set_attrib: index %u out of range", idx);
  if (!(value >= 0) || value > 1.8e19) return fail(CPT_E_ARG, "cpt_renderer_set_attrib: value out of range");
  r->attribs[idx].value = value;
  // parameters change the estimator: restart; the output selection only picks the image handed back
  if (std::strcmp(r->attribs[idx].name, kDebugMode) != 0) r->have_cam = false;
  return CPT_OK;
}

int cpt_renderer_reset(cpt_renderer* r) {
  if (!r) return fail(CPT_E_ARG, "cpt_renderer_reset: null renderer");
  r->have_cam = false;
  return CPT_OK;
}

int cpt_renderer_read_buffer(cpt_renderer* r, int which, void* out) {
  if (!r || !out) return fail(CPT_E_ARG, "cpt_renderer_read_buffer: null argument");
  if (which == CPT_BUF_BOUNCE_RAYS) {
    if (!r->allocated) return fail(CPT_E_STATE, "cpt_renderer_read_buffer: nothing rendered yet");
    FrameCounters h;
    HIP_TRY(hipMemcpyAsync(&h, r->fc.p, sizeof h, hipMemcpyDeviceToHost, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    std::memcpy(out, h.active, sizeof h.active);
    std::memcpy((uint32_t*)out + 16, h.shadow, sizeof h.shadow);
    return CPT_OK;
  }
  return read_pixels(r, which, 0, (size_t)r->W * r->H, out, "cpt_renderer_read_buffer");
}

int cpt_renderer_stats(cpt_renderer* r, cpt_stats* out, int reset) {
  if (!r || !out) return fail(CPT_E_ARG, "cpt_renderer_stats: null argument");
  std::memset(out, 0, sizeof *out);
  if (!r->allocated) return CPT_OK;
  HIP_TRY(hipStreamSynchronize(r->stream));
  DevStats h;
  HIP_TRY(hipMemcpy(&h, r->stats.p, sizeof h, hipMemcpyDeviceToHost));
  for (auto& e : r->events) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e.a, e.b));
    r->ms[e.kind] += ms;
    r->pool.push_back(e.a);
    r->pool.push_back(e.b);
  }
  r->events.clear();
  out->extension_rays = h.ext_rays;
  out->shadow_rays = h.shadow_rays;
  out->eye_paths = h.eye_paths;
  out->frames = h.frames;
  out->extend_nodes = h.ext_nodes;
  out->extend_tris = h.ext_tris;
  out->shadow_nodes = h.sh_nodes;
  out->shadow_tris = h.sh_tris;
  out->extend_spills = h.ext_spill;
  out->shadow_spills = h.sh_spill;
  out->extend_launches = r->extend_launches;
  out->extend_ms = r->ms[0];
  out->shadow_ms = r->ms[1];
  out->shade_ms = r->ms[2];
  out->light_ms = r->ms[3];
  out->total_ms = r->ms[4];
  out->filter_ms = r->ms[5];
  out->error_ms = r->ms[6];
  out->gen_ms = r->ms[7];
  out->accum_ms = r->ms[8];
  out->mlp_ms = r->ms[9];
  out->tail_rays = h.tail_rays;
  out->tail_nodes = h.tail_nodes;
  out->tail_tris = h.tail_tris;
  out->tail_ms = r->ms[10];
  if (reset) {
    HIP_TRY(hipMemset(r->stats.p, 0, sizeof(DevStats)));
    for (double& m : r->ms) m = 0;
    r->extend_launches = 0;
  }
  return CPT_OK;
}

int cpt_renderer_set_profiling(cpt_renderer* r, int on) {
  if (!r) return fail(CPT_E_ARG, "cpt_renderer_set_profiling: null renderer");
  r->profiling = on != 0;
  return CPT_OK;
}

int cpt_renderer_destroy(cpt_renderer* r) {
  if (!r) return fail(CPT_E_ARG, "cpt_renderer_destroy: null renderer");
  if (r->stream) (void)hipStreamSynchronize(r->stream);
  delete r;
  return CPT_OK;
}

}  // extern "C"

namespace {

// Host PCG32 (O'Neill 2014), the generator of the device kernels.
struct HostPcg {
  uint64_t state = 0, inc = 0;
  HostPcg(uint64_t initstate, uint64_t initseq) {
    inc = (initseq << 1u) | 1u;
    next();
    state += initstate;
    next();
  }
  uint32_t next() {
    uint64_t old = state;
    state = old * 6364136223846793005ULL + inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((-rot) & 31));
  }
  float uniform01() { return (float)((next() >> 8) + 1u) * (1.0f / 16777216.0f); }
};

int mlp_upload(cpt_mlp* m, const float* params) {
  hipStream_t st = m->stream;
  HIP_TRY(hipMemcpyAsync(m->w.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemsetAsync(m->m.p, 0, sizeof(float) * kMlpParams, st));
  HIP_TRY(hipMemsetAsync(m->v.p, 0, sizeof(float) * kMlpParams, st));
  HIP_TRY(hipMemsetAsync(m->grad.p, 0, sizeof(float) * kMlpParams, st));
  HIP_TRY(launch_mlp_pack(m->state(), st));
  HIP_TRY(hipStreamSynchronize(st));
  m->step = 0;
  return CPT_OK;
}

// One training step on `stream`: forward + backward into per-workgroup
// partials, then the fixed-order sum and (apply) Adam.
int mlp_step(cpt_mlp* m, const float* d_x, const float* d_target, uint32_t n, float lr, int apply, hipStream_t st) {
  const u32 nwg = mlp_train_groups(n);
  HIP_TRY(m->part.alloc((size_t)nwg * kMlpParams));
  HIP_TRY(m->loss_part.alloc(nwg));
  MlpState s = m->state();
  HIP_TRY(launch_mlp_train(s, d_x, d_target, n, m->part.p, m->loss_part.p, st));
  int step = apply ? m->step + 1 : (m->step > 0 ? m->step : 1);
  HIP_TRY(launch_mlp_adam(s, m->part.p, nwg, m->loss_part.p, lr, step, apply != 0, st));
  if (apply) m->step = step;
  return CPT_OK;
}
}  // namespace

extern "C" {

int cpt_mlp_create(cpt_mlp** out, uint64_t seed) {
  if (!out) return fail(CPT_E_ARG, "cpt_mlp_create: null output");
  *out = nullptr;
  auto* m = new (std::nothrow) cpt_mlp();
  if (!m) return fail(CPT_E_NOMEM, "cpt_mlp_create: out of host memory");
  auto bad = [&](int rc) { delete m; return rc; };
  if (int rc = ensure_stream(m->stream)) return bad(rc);
  for (auto* a : {&m->w, &m->m, &m->v, &m->grad})
    if (a->alloc(kMlpParams) != hipSuccess) return bad(fail(CPT_E_NOMEM, "cpt_mlp_create: device allocation"));
  if (m->loss.alloc(1) || m->b1.alloc(64) || m->b2.alloc(64) || m->w3.alloc(64) || m->b3.alloc(1) ||
      m->w1f.alloc(kMlpW1Frag) || m->w2tf.alloc(kMlpW2Frag) || m->w2f.alloc(kMlpW2Frag))
    return bad(fail(CPT_E_NOMEM, "cpt_mlp_create: device allocation"));
  // the K padding of layer 1 (features 9..15) stays zero
  if (hipMemset(m->w1f.p, 0, sizeof(uint16_t) * kMlpW1Frag) != hipSuccess)
    return bad(fail(CPT_E_HIP, "cpt_mlp_create: hipMemset"));
  std::vector<float> p(kMlpParams, 0.f);
  HostPcg rng(seed, 0);
  auto glorot = [&](int off, int fin, int fout) {
    float a = std::sqrt(6.0f / (float)(fin + fout));
    for (int i = 0; i < fin * fout; i++) p[off + i] = (2.0f * rng.uniform01() - 1.0f) * a;
  };
  glorot(kMlpOffW1, kMlpIn, kMlpHidden);
  glorot(kMlpOffW2, kMlpHidden, kMlpHidden);
  glorot(kMlpOffW3, kMlpHidden, 1);
  if (int rc = mlp_upload(m, p.data())) return bad(rc);
  *out = m;
  return CPT_OK;
}

int cpt_mlp_destroy(cpt_mlp* m) {
  if (!m) return fail(CPT_E_ARG, "cpt_mlp_destroy: null model");
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  delete m;
  return CPT_OK;
}

int cpt_mlp_set_params(cpt_mlp* m, const float* params, uint32_t n) {
  if (!m || !params) return fail(CPT_E_ARG, "cpt_mlp_set_params: null argument");
  if (n != (uint32_t)kMlpParams) return fail(CPT_E_ARG, "cpt_mlp_set_params: %u parameters, expected 4865", n);
  return mlp_upload(m, params);
}

int cpt_mlp_get_params(cpt_mlp* m, float* params, uint32_t n) {
  if (!m || !params) return fail(CPT_E_ARG, "cpt_mlp_get_params: null argument");
  if (n != (uint32_t)kMlpParams) return fail(CPT_E_ARG, "cpt_mlp_get_params: %u parameters, expected 4865", n);
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(params, m->w.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return CPT_OK;
}

int cpt_mlp_get_grad(cpt_mlp* m, float* grad, uint32_t n) {
  if (!m || !grad) return fail(CPT_E_ARG, "cpt_mlp_get_grad: null argument");
  if (n != (uint32_t)kMlpParams) return fail(CPT_E_ARG, "cpt_mlp_get_grad: %u parameters, expected 4865", n);
  HIP_TRY(hipStreamSynchronize(m->stream));
  HIP_TRY(hipMemcpy(grad, m->grad.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return CPT_OK;
}

int cpt_mlp_train_step(cpt_mlp* m, const float* d_x, const float* d_target, uint32_t n, float lr, int apply,
                       float* loss, void* stream) {
  if (!m || !d_x || !d_target) return fail(CPT_E_ARG, "cpt_mlp_train_step: null argument");
  if (n == 0) return fail(CPT_E_ARG, "cpt_mlp_train_step: empty batch");
  if (n > (1u << 24)) return fail(CPT_E_LIMIT, "cpt_mlp_train_step: batch %u exceeds 2^24", n);
  if (!(lr >= 0.f)) return fail(CPT_E_ARG, "cpt_mlp_train_step: learning rate must be >= 0");
  hipStream_t st = stream ? (hipStream_t)stream : m->stream;
  int rc = mlp_step(m, d_x, d_target, n, lr, apply, st);
  if (rc) return rc;
  if (loss) {
    HIP_TRY(hipStreamSynchronize(st));
    float l = 0.f;
    HIP_TRY(hipMemcpy(&l, m->loss.p, sizeof l, hipMemcpyDeviceToHost));
    *loss = l / (float)n;
  }
  return CPT_OK;
}

int cpt_mlp_train_steps(cpt_mlp* m, const float* d_x, const float* d_target, uint32_t n_samples, uint32_t batch,
                        uint32_t first_batch, uint32_t steps, float lr, float* loss, void* stream) {
  if (!m || !d_x || !d_target) return fail(CPT_E_ARG, "cpt_mlp_train_steps: null argument");
  if (batch == 0 || steps == 0 || n_samples < batch)
    return fail(CPT_E_ARG, "cpt_mlp_train_steps: need 0 < batch <= samples and steps > 0");
  if (batch > (1u << 24)) return fail(CPT_E_LIMIT, "cpt_mlp_train_steps: batch %u exceeds 2^24", batch);
  if (!(lr >= 0.f)) return fail(CPT_E_ARG, "cpt_mlp_train_steps: learning rate must be >= 0");
  hipStream_t st = stream ? (hipStream_t)stream : m->stream;
  const uint32_t nb = n_samples / batch;
  for (uint32_t k = 0; k < steps; k++) {
    const size_t off = (size_t)((first_batch + k) % nb) * batch;
    int rc = mlp_step(m, d_x + off * kMlpIn, d_target + off, batch, lr, 1, st);
    if (rc) return rc;
  }
  if (loss) {
    HIP_TRY(hipStreamSynchronize(st));
    float l = 0.f;
    HIP_TRY(hipMemcpy(&l, m->loss.p, sizeof l, hipMemcpyDeviceToHost));
    *loss = l / (float)batch;
  }
  return CPT_OK;
}

int cpt_mlp_infer(cpt_mlp* m, const float* d_x, uint32_t n, float* d_y, void* stream) {
  if (!m || ((!d_x || !d_y) && n)) return fail(CPT_E_ARG, "cpt_mlp_infer: null argument");
  hipStream_t st = stream ? (hipStream_t)stream : m->stream;
  HIP_TRY(launch_mlp_infer(m->state(), d_x, n, d_y, st));
  if (!stream) HIP_TRY(hipStreamSynchronize(st));
  return CPT_OK;
}

int cpt_renderer_mlp_features(cpt_renderer* r, float* d_x) {
  if (!r || !d_x) return fail(CPT_E_ARG, "cpt_renderer_mlp_features: null argument");
  if (!uses_filter(r->mode) || !r->allocated || !r->have_cam)
    return fail(CPT_E_STATE, "cpt_renderer_mlp_features: needs an adaptive renderer that has rendered");
  // the filter state the next adaptive frame allocates from
  int rc = run_filter(r);
  if (rc) return rc;
  rc = model_features(r, d_x);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(r->stream));
  return CPT_OK;
}

int cpt_renderer_mlp_target(cpt_renderer* r, const float* d_ref, float clamp, float* d_t) {
  if (!r || !d_ref || !d_t) return fail(CPT_E_ARG, "cpt_renderer_mlp_target: null argument");
  if (!r->allocated) return fail(CPT_E_STATE, "cpt_renderer_mlp_target: renderer has not rendered");
  if (!(clamp >= 0.f)) return fail(CPT_E_ARG, "cpt_renderer_mlp_target: clamp must be >= 0");
  HIP_TRY(launch_mlp_target(r->result.p, d_ref, r->sample_n.p, r->W * r->H, clamp > 0.f ? clamp : INFINITY, d_t,
                            r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return CPT_OK;
}

int cpt_renderer_set_mlp(cpt_renderer* r, cpt_mlp* m) {
  if (!r) return fail(CPT_E_ARG, "cpt_renderer_set_mlp: null renderer");
  if (m && !uses_filter(r->mode))
    return fail(CPT_E_STATE, "cpt_renderer_set_mlp: the model drives the filter-based adaptive renderers only");
  r->mlp = m;
  return CPT_OK;
}

}  // extern "C"

// ---- multi-view batch ---------------------------------------------------------

struct cpt_batch {
  cpt_renderer* r = nullptr;           // the renderer over the concatenated pixels (W = total, H = 1)
  std::vector<cpt_view> views;
  std::vector<ViewDesc> vd;
  DevArray<ViewDesc> d_views;
  DevArray<u32> frame_dev;
  u32 total = 0;
  cpt_batch_config cfg;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  u32 graph_k = 0;
  bool started = false;
  const cpt_scene* scene = nullptr;
  u32 mesh_v = 0, mat_v = 0;
  const void* buf_key = nullptr;       // path buffer the graph was captured on
  int cus = 0;
  double ms = 0.0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // per-view adaptive sampling (cpt_batch_set_adaptive): the views as spans
  // of the concatenated buffers, their error sums and path totals
  bool adaptive = false;
  std::vector<ViewSpan> spans;
  DevArray<float> vsum;
  DevArray<u32> vtot;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    graph_k = 0;
  }
  ~cpt_batch() {
    drop_graph();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (r) cpt_renderer_destroy(r);
  }
};

namespace {

constexpr u32 kMaxViews = 4096;
// extension-ray schedules the autotuner tries, each with every traversal
// variant: {persistent chunk, refill threshold, workgroups per CU}
struct Schedule {
  u32 chunk, refill, wgs;
};
const Schedule kSchedules[] = {{0, 0, 8},     {256, 0, 8},  {1024, 0, 8}, {4096, 0, 8}, {1024, 0, 4},
                               {1024, 0, 16}, {0, 64, 6},   {0, 64, 8},   {0, 48, 8},   {0, 32, 8}};
const u32 kVariants[] = {806u, 807u};

cpt_batch_config default_batch_config() {
  cpt_batch_config c;
  std::memset(&c, 0, sizeof c);
  c.frames_in_flight = 24;
  c.extend_chunk = 0;
  c.extend_workgroups = 8;
  c.refill = 64;
  c.node_format = 1;
  c.traversal_variant = 806;
  c.use_graph = 1;
  c.eye_depth = 5;
  c.light_depth = 5;
  c.light_paths = 128;
  c.seed = 0x5EED0001u;
  return c;
}

int batch_check_config(const cpt_batch_config& c) {
  if (c.frames_in_flight < 1 || c.frames_in_flight > kMaxInFlight)
    return fail(CPT_E_LIMIT, "batch: frames in flight %u not in 1..%u", c.frames_in_flight, kMaxInFlight);
  if ((c.extend_chunk || c.refill) && (c.extend_workgroups < 1 || c.extend_workgroups > 32))
    return fail(CPT_E_LIMIT, "batch: extend workgroups per CU %u not in 1..32", c.extend_workgroups);
  if (c.refill > 64) return fail(CPT_E_LIMIT, "batch: refill threshold %u not in 0..64", c.refill);
  if (c.eye_depth < 1 || c.eye_depth > (u32)kMaxDepth) return fail(CPT_E_LIMIT, "batch: eye depth not in 1..15");
  if (c.light_depth < 1 || c.light_depth > 64) return fail(CPT_E_LIMIT, "batch: light depth not in 1..64");
  if (c.light_paths < 1 || c.light_paths > (u32)kBlock) return fail(CPT_E_LIMIT, "batch: light paths not in 1..256");
  if (c.node_format > 1) return fail(CPT_E_ARG, "batch: node format %u not 0 or 1", c.node_format);
  if (!traversal_variant_ok((int)c.traversal_variant))
    return fail(CPT_E_ARG, "batch: traversal variant %u not 8, 16, 24, 806, 807, 808 or 1606", c.traversal_variant);
  return CPT_OK;
}

void batch_apply_config(cpt_batch* b) {
  cpt_renderer* r = b->r;
  r->set_attr(kEyeDepth, b->cfg.eye_depth);
  r->set_attr(kLightDepth, b->cfg.light_depth);
  r->set_attr(kLightPaths, b->cfg.light_paths);
  r->set_attr(kSeed, (double)b->cfg.seed);
  r->set_attr(kInFlight, b->cfg.frames_in_flight);
  r->set_attr(kLdsStack, b->cfg.traversal_variant);
  r->set_attr(kNodeFormat, b->cfg.node_format ? 1 : 0);
  r->set_attr(kRefill, b->cfg.refill);
  r->set_attr(kRefillSecondary, b->cfg.refill);
  r->set_attr(kRefillGroups, b->cfg.extend_workgroups);
  r->set_attr(kRefillMinPaths, 0);               // the batch's schedule applies at every size
  r->persist_chunk = b->cfg.refill ? 0u : b->cfg.extend_chunk;
  r->persist_grid = r->persist_chunk ? b->cfg.extend_workgroups * (u32)(b->cus > 0 ? b->cus : 256) : 0u;
  b->drop_graph();
  b->started = false;
}

// One iteration of k frames of every view: light caches of the frames'
// groups, eye paths, the bounce loop, the per-frame running-mean updates, and
// the device frame counter advance.  No host synchronisation: capturable.
int batch_enqueue(cpt_batch* b, cpt_scene* s, u32 k) {
  cpt_renderer* r = b->r;
  DevScene sc = s->dev();
  FrameArgs fa;
  std::memset(&fa, 0, sizeof fa);
  fa.width = b->total;
  fa.height = 1;
  fa.eye_depth = b->cfg.eye_depth;
  fa.light_depth = b->cfg.light_depth;
  fa.light_paths = b->cfg.light_paths;
  fa.seed = b->cfg.seed;
  fa.nframes = k;
  fa.lv_stride = b->cfg.light_paths * (b->cfg.light_depth + 1);
  fa.n_views = (u32)b->vd.size();
  fa.views = b->d_views.p;
  fa.frame_dev = b->frame_dev.p;
  LightJobs jobs;
  std::memset(&jobs, 0, sizeof jobs);
  jobs.dev_groups = (k + 2) / 3 + 1;                 // groups k frames can span
  const u32 n = b->total * k;
  HIP_TRY(launch_light_pass(sc, fa, jobs, r->lv.p, r->fc.p, r->stream));
  HIP_TRY(launch_frame_begin(r->fc.p, n, r->stream));
  Accum ac;
  std::memset(&ac, 0, sizeof ac);
  ac.result = r->result.p;
  ac.sample_n = r->sample_n.p;
  ac.rec.acc = r->rec_acc.p;
  ac.apply = 2;
  int rc = bounce_loop(r, sc, fa, n, ac, true);
  if (rc) return rc;
  HIP_TRY(launch_apply_frames(fa, r->rec_acc.p, r->result.p, r->sample_n.p, 0, r->stream));
  HIP_TRY(launch_frame_end(r->fc.p, r->stats.p, k, fa.eye_depth, r->stream));
  HIP_TRY(launch_frame_advance(b->frame_dev.p, k, r->stream));
  return CPT_OK;
}

int batch_prepare(cpt_batch* b, cpt_scene* s) {
  if (!s->has_bvh) return fail(CPT_E_STATE, "cpt_batch_render: scene has no BVH");
  if (!s->uploaded) {
    int rc = cpt_scene_upload(s);
    if (rc) return rc;
  }
  if (b->scene != s || b->mesh_v != s->mesh_version || b->mat_v != s->mat_version) {
    b->drop_graph();
    b->started = false;
    b->scene = s;
    b->mesh_v = s->mesh_version;
    b->mat_v = s->mat_version;
  }
  cpt_renderer* r = b->r;
  const size_t need = b->adaptive ? (size_t)b->total * (size_t)r->attr(kPassFrames) + r->slot_margin
                                  : (size_t)b->total * b->cfg.frames_in_flight;
  if (!r->allocated || (size_t)r->cap < need) {
    int rc = renderer_alloc(r);
    if (rc) return rc;
  }
  int rc = lv_alloc(r);
  if (rc) return rc;
  HIP_TRY(r->queue.alloc(2));
  if (b->buf_key != (const void*)r->A.ray_o.p) {   // buffers moved: a captured graph points at the old ones
    b->drop_graph();
    b->buf_key = (const void*)r->A.ray_o.p;
  }
  if (!b->started) {
    u32 zero = 0;
    HIP_TRY(hipMemcpyAsync(b->frame_dev.p, &zero, sizeof zero, hipMemcpyHostToDevice, r->stream));
    HIP_TRY(hipStreamSynchronize(r->stream));
    r->started = false;                  // adaptive frames count on the host from 0 again
    r->forget_light_caches();
    b->started = true;
  }
  return CPT_OK;
}

// Per-view adaptive frames (cpt_batch_set_adaptive): the frame loop of a
// one-view adaptive renderer over every view at once.  Warm-up frames trace
// one path per pixel of every view; then per pass each view allocates
// pass_frames x its pixel count paths over its own pixels (its own filter,
// model features and error sum), and all views' paths are traced as one
// wavefront.  Host-driven (no graph): a pass reads nothing back, but its
// shape changes after the warm-up.
int batch_render_adaptive(cpt_batch* b, cpt_scene* s, u32 frames) {
  cpt_renderer* r = b->r;
  const u32 radius = (u32)r->attr("Filter Radius");
  if (radius > (u32)kMaxFilterRadius) return fail(CPT_E_LIMIT, "cpt_batch_render: filter radius %u exceeds 31", radius);
  if (r->attr(kWarmup) < 1) return fail(CPT_E_LIMIT, "cpt_batch_render: warm-up frames must be >= 1");
  const u32 pass = (u32)r->attr(kPassFrames);
  if (pass < 1 || pass > kMaxInFlight)
    return fail(CPT_E_LIMIT, "cpt_batch_render: %s %u not in 1..%u", kPassFrames, pass, kMaxInFlight);
  int rc = batch_prepare(b, s);
  if (rc) return rc;
  if (r->mlp) {
    rc = model_buffers(r);
    if (rc) return rc;
  }
  FrameArgs fa;
  std::memset(&fa, 0, sizeof fa);
  fa.width = b->total;
  fa.height = 1;
  fa.eye_depth = b->cfg.eye_depth;
  fa.light_depth = b->cfg.light_depth;
  fa.light_paths = b->cfg.light_paths;
  fa.seed = b->cfg.seed;
  fa.lv_stride = b->cfg.light_paths * (b->cfg.light_depth + 1);
  fa.flags = r->attr(kLiteralShadow) != 0 ? 1u : 0u;
  fa.n_views = (u32)b->vd.size();
  fa.views = b->d_views.p;
  const ViewAlloc va{&b->spans, b->vsum.p, b->vtot.p};
  HIP_TRY(hipEventRecord(b->e0, r->stream));
  u32 next = r->started ? r->frame + 1 : 0u;
  for (u32 done = 0; done < frames;) {
    u32 k = std::min(pass, frames - done);
    if (next < warmup_frames(r)) k = 1;
    fa.frame0 = next;
    fa.nframes = k;
    rc = render_adaptive_frame(r, s, fa, &va);
    if (rc) {
      join_side(r);
      return rc;
    }
    r->frame = next + k - 1;
    r->started = true;
    next += k;
    done += k;
  }
  HIP_TRY(hipEventRecord(b->e1, r->stream));
  HIP_TRY(hipEventSynchronize(b->e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->ms += ms;
  return CPT_OK;
}

int batch_capture(cpt_batch* b, cpt_scene* s, u32 k) {
  b->drop_graph();
  cpt_renderer* r = b->r;
  HIP_TRY(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
  const uint64_t launches = r->extend_launches;       // counted per replay instead
  int rc = batch_enqueue(b, s, k);
  r->extend_launches = launches;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(r->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  HIP_TRY(e);
  b->graph = g;
  HIP_TRY(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_k = k;
  return CPT_OK;
}

int batch_render(cpt_batch* b, cpt_scene* s, u32 frames) {
  if (b->adaptive) return batch_render_adaptive(b, s, frames);
  int rc = batch_prepare(b, s);
  if (rc) return rc;
  cpt_renderer* r = b->r;
  const u32 K = b->cfg.frames_in_flight;
  HIP_TRY(hipEventRecord(b->e0, r->stream));
  for (u32 done = 0; done < frames;) {
    u32 k = std::min(K, frames - done);
    if (b->cfg.use_graph && k == K) {
      if (!b->exec || b->graph_k != K) {
        rc = batch_capture(b, s, K);
        if (rc) return rc;
      }
      HIP_TRY(hipGraphLaunch(b->exec, r->stream));
      r->extend_launches += b->cfg.eye_depth;
    } else {
      rc = batch_enqueue(b, s, k);
      if (rc) {
        join_side(r);
        return rc;
      }
    }
    done += k;
  }
  HIP_TRY(hipEventRecord(b->e1, r->stream));
  HIP_TRY(hipEventSynchronize(b->e1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, b->e0, b->e1));
  b->ms += ms;
  return CPT_OK;
}

int batch_candidate(const cpt_batch* b, u32 i, cpt_batch_config* out) {
  u32 idx = 0;
  for (const Schedule& sc : kSchedules)
    for (u32 v : kVariants) {
      if (idx++ == i) {
        *out = b->cfg;
        out->extend_chunk = sc.chunk;
        out->refill = sc.refill;
        out->extend_workgroups = sc.wgs;
        out->traversal_variant = v;
        return CPT_OK;
      }
    }
  return fail(CPT_E_ARG, "cpt_batch_candidate: index %u out of range", i);
}
u32 batch_candidate_count() { return (u32)(sizeof(kSchedules) / sizeof(kSchedules[0]) * 2); }

}  // namespace

extern "C" {

int cpt_batch_create(cpt_batch** out, const cpt_view* views, uint32_t n_views) {
  if (!out || !views) return fail(CPT_E_ARG, "cpt_batch_create: null argument");
  *out = nullptr;
  if (n_views == 0 || n_views > kMaxViews) return fail(CPT_E_LIMIT, "cpt_batch_create: %u views not in 1..4096", n_views);
  auto* b = new (std::nothrow) cpt_batch();
  if (!b) return fail(CPT_E_NOMEM, "cpt_batch_create: out of host memory");
  uint64_t total = 0;
  for (u32 v = 0; v < n_views; v++) {
    const cpt_view& V = views[v];
    if (V.width == 0 || V.height == 0) { delete b; return fail(CPT_E_ARG, "cpt_batch_create: view %u is empty", v); }
    ViewDesc d;
    std::memset(&d, 0, sizeof d);
    d.offset = (u32)total;
    d.width = V.width;
    d.height = V.height;
    for (int k = 0; k < 3; k++) { d.cam_pos[k] = V.cam_pos[k]; d.cam_dir[k] = V.cam_dir[k]; }
    camera_basis(V.cam_dir, V.cam_up, d.cam_right, d.cam_up);
    d.fov = V.fov;
    d.tan_half = std::tan(V.fov * 0.5f);
    b->vd.push_back(d);
    total += (uint64_t)V.width * V.height;
  }
  if (total > (1ull << 26)) { delete b; return fail(CPT_E_LIMIT, "cpt_batch_create: %llu pixels exceed 2^26", (unsigned long long)total); }
  b->views.assign(views, views + n_views);
  b->total = (u32)total;
  b->cfg = default_batch_config();
  int rc = cpt_renderer_create(&b->r, CPT_MODE_BDPT_STREAM, b->total, 1);
  if (rc) { delete b; return rc; }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    b->cus = 256;
  if (b->d_views.upload(b->vd.data(), b->vd.size(), b->r->stream) != hipSuccess || b->frame_dev.alloc(1) != hipSuccess ||
      hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
    delete b;
    return fail(CPT_E_NOMEM, "cpt_batch_create: device allocation");
  }
  batch_apply_config(b);
  *out = b;
  return CPT_OK;
}

int cpt_batch_get_config(cpt_batch* b, cpt_batch_config* cfg) {
  if (!b || !cfg) return fail(CPT_E_ARG, "cpt_batch_get_config: null argument");
  *cfg = b->cfg;
  return CPT_OK;
}

int cpt_batch_set_config(cpt_batch* b, const cpt_batch_config* cfg) {
  if (!b || !cfg) return fail(CPT_E_ARG, "cpt_batch_set_config: null argument");
  int rc = batch_check_config(*cfg);
  if (rc) return rc;
  b->cfg = *cfg;
  batch_apply_config(b);
  return CPT_OK;
}

int cpt_batch_render(cpt_batch* b, cpt_scene* s, uint32_t frames) {
  if (!b || !s) return fail(CPT_E_ARG, "cpt_batch_render: null argument");
  return batch_render(b, s, frames);
}

int cpt_batch_read_result(cpt_batch* b, uint32_t view, float* out) {
  if (!b || !out) return fail(CPT_E_ARG, "cpt_batch_read_result: null argument");
  if (view >= b->vd.size()) return fail(CPT_E_ARG, "cpt_batch_read_result: view %u out of range", view);
  if (!b->r->allocated) return fail(CPT_E_STATE, "cpt_batch_read_result: nothing rendered yet");
  const ViewDesc& d = b->vd[view];
  HIP_TRY(hipMemcpyAsync(out, b->r->result.p + (size_t)d.offset * 3, (size_t)d.width * d.height * 12,
                         hipMemcpyDeviceToHost, b->r->stream));
  HIP_TRY(hipStreamSynchronize(b->r->stream));
  return CPT_OK;
}

int cpt_batch_sample_counts(cpt_batch* b, uint32_t view, uint32_t* out) {
  if (!b || !out) return fail(CPT_E_ARG, "cpt_batch_sample_counts: null argument");
  if (view >= b->vd.size()) return fail(CPT_E_ARG, "cpt_batch_sample_counts: view %u out of range", view);
  if (!b->r->allocated) return fail(CPT_E_STATE, "cpt_batch_sample_counts: nothing rendered yet");
  const ViewDesc& d = b->vd[view];
  HIP_TRY(hipMemcpyAsync(out, b->r->sample_n.p + d.offset, (size_t)d.width * d.height * 4, hipMemcpyDeviceToHost,
                         b->r->stream));
  HIP_TRY(hipStreamSynchronize(b->r->stream));
  return CPT_OK;
}

int cpt_batch_read_buffer(cpt_batch* b, uint32_t view, int which, void* out) {
  if (!b || !out) return fail(CPT_E_ARG, "cpt_batch_read_buffer: null argument");
  if (view >= b->vd.size()) return fail(CPT_E_ARG, "cpt_batch_read_buffer: view %u out of range", view);
  const ViewDesc& d = b->vd[view];
  return read_pixels(b->r, which, d.offset, (size_t)d.width * d.height, out, "cpt_batch_read_buffer");
}

int cpt_batch_set_adaptive(cpt_batch* b, int mode, uint32_t pass_frames, cpt_mlp* model) {
  if (!b) return fail(CPT_E_ARG, "cpt_batch_set_adaptive: null batch");
  if (mode != CPT_MODE_BDPT_STREAM && mode != CPT_MODE_BDPT_ADAPTIVE_PLL)
    return fail(CPT_E_ARG, "cpt_batch_set_adaptive: mode %d is neither plain streaming BDPT (0) nor the proportional "
                           "adaptive mode (1)", mode);
  const bool on = mode == CPT_MODE_BDPT_ADAPTIVE_PLL;
  if (on && (pass_frames < 1 || pass_frames > kMaxInFlight))
    return fail(CPT_E_LIMIT, "cpt_batch_set_adaptive: pass frames %u not in 1..%u", pass_frames, kMaxInFlight);
  if (!on && model) return fail(CPT_E_ARG, "cpt_batch_set_adaptive: a model needs the adaptive mode");
  const u32 nv = (u32)b->vd.size();
  if (on && (b->vsum.alloc(nv) != hipSuccess || b->vtot.alloc(nv) != hipSuccess))
    return fail(CPT_E_NOMEM, "cpt_batch_set_adaptive: device allocation");
  cpt_renderer* r = nullptr;
  int rc = cpt_renderer_create(&r, mode, b->total, 1);
  if (rc) return rc;
  cpt_renderer_destroy(b->r);
  b->r = r;
  b->adaptive = on;
  b->spans.clear();
  for (const ViewDesc& d : b->vd) b->spans.push_back(ViewSpan{d.offset, d.width, d.height, d.cam_pos});
  r->slot_margin = 1024u * nv;              // each view's counts may round past its budget
  r->set_attr(kPassFrames, on ? pass_frames : 1u);
  r->mlp = model;
  batch_apply_config(b);
  return CPT_OK;
}

int cpt_batch_set_attrib(cpt_batch* b, const char* name, double value) {
  if (!b || !name) return fail(CPT_E_ARG, "cpt_batch_set_attrib: null argument");
  for (const char* own : {kEyeDepth, kLightDepth, kLightPaths, kSeed, kInFlight, kPassFrames, kLdsStack, kNodeFormat,
                          kRefill, kRefillSecondary, kRefillGroups, kRefillMinPaths})
    if (std::strcmp(name, own) == 0)
      return fail(CPT_E_ARG, "cpt_batch_set_attrib: \"%s\" belongs to the batch configuration", name);
  for (uint32_t i = 0; i < b->r->attribs.size(); i++)
    if (std::strcmp(b->r->attribs[i].name, name) == 0) {
      int rc = cpt_renderer_set_attrib(b->r, i, value);
      if (rc) return rc;
      b->started = false;                   // a changed estimator restarts accumulation
      return CPT_OK;
    }
  return fail(CPT_E_ARG, "cpt_batch_set_attrib: no attribute \"%s\" in this batch's mode", name);
}

int cpt_batch_stats(cpt_batch* b, cpt_stats* out, int reset) {
  if (!b || !out) return fail(CPT_E_ARG, "cpt_batch_stats: null argument");
  int rc = cpt_renderer_stats(b->r, out, reset);
  if (rc) return rc;
  out->total_ms = b->ms;
  if (reset) b->ms = 0.0;
  return CPT_OK;
}

int cpt_batch_reset(cpt_batch* b) {
  if (!b) return fail(CPT_E_ARG, "cpt_batch_reset: null batch");
  b->started = false;
  return CPT_OK;
}

int cpt_batch_candidates(cpt_batch* b, uint32_t* n) {
  if (!b || !n) return fail(CPT_E_ARG, "cpt_batch_candidates: null argument");
  *n = batch_candidate_count();
  return CPT_OK;
}

int cpt_batch_candidate(cpt_batch* b, uint32_t i, cpt_batch_config* cfg) {
  if (!b || !cfg) return fail(CPT_E_ARG, "cpt_batch_candidate: null argument");
  return batch_candidate(b, i, cfg);
}

int cpt_batch_autotune(cpt_batch* b, cpt_scene* s, uint32_t iterations, cpt_batch_config* best, double* ms,
                       uint32_t n_ms) {
  if (!b || !s) return fail(CPT_E_ARG, "cpt_batch_autotune: null argument");
  if (b->adaptive)
    return fail(CPT_E_STATE, "cpt_batch_autotune: tune the plain batch, then call cpt_batch_set_adaptive (it keeps the "
                             "schedule)");
  if (iterations < 1) return fail(CPT_E_ARG, "cpt_batch_autotune: iterations must be >= 1");
  const cpt_batch_config base = b->cfg;
  const u32 n = batch_candidate_count();
  double best_ms = 0.0;
  cpt_batch_config best_cfg = base;
  for (u32 i = 0; i < n; i++) {
    cpt_batch_config c;
    int rc = batch_candidate(b, i, &c);
    if (rc) return rc;
    b->cfg = c;
    batch_apply_config(b);
    const u32 K = c.frames_in_flight;
    rc = batch_render(b, s, K);                      // capture + warm
    if (rc) return rc;
    b->ms = 0.0;
    rc = batch_render(b, s, K * iterations);
    if (rc) return rc;
    const double per = b->ms / iterations;
    if (ms && i < n_ms) ms[i] = per;
    if (i == 0 || per < best_ms) { best_ms = per; best_cfg = c; }
  }
  b->ms = 0.0;
  b->cfg = best_cfg;
  batch_apply_config(b);
  HIP_TRY(hipMemset(b->r->stats.p, 0, sizeof(DevStats)));
  if (best) *best = best_cfg;
  return CPT_OK;
}

int cpt_batch_destroy(cpt_batch* b) {
  if (!b) return fail(CPT_E_ARG, "cpt_batch_destroy: null batch");
  if (b->r && b->r->stream) (void)hipStreamSynchronize(b->r->stream);
  delete b;
  return CPT_OK;
}

}  // extern "C"
