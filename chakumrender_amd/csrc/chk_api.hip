// C ABI of libchakum (include/chakum_c.h): context, scene upload, trace,
// BRDF and render entry points.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "chk_device.h"
#include "chk_internal.h"
#include "chk_light64.h"

namespace chk {

thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

namespace {

// Load a direction given in FP64 at precision T and normalise it there: the
// formulas assume unit vectors (materialTypes.cpp:8).
template <typename T>
__device__ inline v3<T> load_dir(const double *__restrict__ p) {
  return normalize(mk3<T>((T)p[0], (T)p[1], (T)p[2]));
}

// T = double: the reference's precision; T = float: the arithmetic of the
// fused render kernels, on the FP32 rounding of the inputs.
template <typename T>
__global__ void k_brdf_eval(const chk_material *__restrict__ mats, uint32_t n_mats,
                            const uint32_t *__restrict__ mat_idx, const double *__restrict__ n,
                            const double *__restrict__ wo_ref, const double *__restrict__ wo,
                            const double *__restrict__ wi, uint32_t k, double *__restrict__ f,
                            double *__restrict__ pdf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const uint32_t mi = mat_idx[i];
  if (mi >= n_mats) {
    f[i] = NAN;
    pdf[i] = NAN;
    return;
  }
  const MatT<T> m = load_mat<T>(mats + mi);
  const size_t o = 3 * (size_t)i;
  T fv, pv;
  brdf_pdf(m, load_dir<T>(n + o), load_dir<T>(wo_ref + o), load_dir<T>(wo + o),
           load_dir<T>(wi + o), fv, pv);
  f[i] = fv;
  pdf[i] = pv;
}

template <typename T>
__global__ void k_brdf_sample(const chk_material *__restrict__ mats, uint32_t n_mats,
                              const uint32_t *__restrict__ mat_idx, const double *__restrict__ n,
                              const double *__restrict__ wo_ref, const double *__restrict__ wo,
                              const double *__restrict__ u, uint32_t k, double *__restrict__ wi,
                              double *__restrict__ pdf) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const uint32_t mi = mat_idx[i];
  const size_t o = 3 * (size_t)i;
  if (mi >= n_mats) {
    wi[o] = wi[o + 1] = wi[o + 2] = NAN;
    pdf[i] = NAN;
    return;
  }
  const MatT<T> m = load_mat<T>(mats + mi);
  const v3<T> N = load_dir<T>(n + o), R = load_dir<T>(wo_ref + o), O = load_dir<T>(wo + o);
  const v3<T> w = brdf_sample(m, N, R, O, (T)u[o], (T)u[o + 1], (T)u[o + 2]);
  T fv, pv;
  brdf_pdf(m, N, R, O, w, fv, pv);
  wi[o] = w.x;
  wi[o + 1] = w.y;
  wi[o + 2] = w.z;
  pdf[i] = pv;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__global__ void k_fill_zero(uint4 *p, size_t n16, unsigned char *tail, size_t ntail) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n16; k += stride) p[k] = make_uint4(0u, 0u, 0u, 0u);
  if (i < ntail) tail[i] = 0;
}

}  // namespace
}  // namespace chk

using namespace chk;


namespace {
// The path tracer of chk_render and its batched and learned forms traces
// triangles and samples mesh lights only.
chk_status reject_extras(const chk_scene_s *sc, const char *who) {
  if (sc && sc->extras) {
    set_error(std::string(who) + ": the scene has spheres or triangle, sphere, point or "
              "environment sources; "
              "only the direct-lighting entry points and chk_render_shade render those");
    return CHK_EINVAL;
  }
  return CHK_OK;
}
chk_status reject_extras(const chk_request *reqs, uint32_t n, const char *who) {
  for (uint32_t i = 0; i < n; ++i) {
    const chk_status s = reject_extras(reqs[i].scene, who);
    if (s != CHK_OK) return s;
  }
  return CHK_OK;
}
}  // namespace

extern "C" {

const char *chk_status_string(chk_status s) {
  switch (s) {
    case CHK_OK: return "ok";
    case CHK_EINVAL: return "invalid argument";
    case CHK_ENOMEM: return "out of memory";
    case CHK_EDEVICE: return "device error";
    case CHK_ECAPACITY: return "capacity exceeded";
    case CHK_EINTERNAL: return "internal error";
  }
  return "unknown status";
}

const char *chk_last_error(void) { return g_err.c_str(); }

const char *chk_version(void) { return "chakum-mi355x 0.1 (gfx950)"; }

chk_status chk_device_alloc(int device, size_t bytes, void **dev) {
  if (!dev) {
    set_error("chk_device_alloc: null out");
    return CHK_EINVAL;
  }
  *dev = nullptr;
  CHK_HIP(hipSetDevice(device));
  if (bytes == 0) return CHK_OK;
  if (hipMalloc(dev, bytes) != hipSuccess) {
    *dev = nullptr;
    set_error("chk_device_alloc: cannot allocate " + std::to_string(bytes) + " bytes");
    return CHK_ENOMEM;
  }
  return CHK_OK;
}

chk_status chk_device_free(void *dev) {
  if (dev) CHK_HIP(hipFree(dev));
  return CHK_OK;
}

chk_status chk_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
  if (bytes == 0) return CHK_OK;
  if (!dst || !src || kind < 0 || kind > 2) {
    set_error("chk_memcpy: null pointer or bad kind");
    return CHK_EINVAL;
  }
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                    : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  CHK_HIP(hipMemcpyAsync(dst, src, bytes, k, (hipStream_t)stream));
  if (kind == 1) CHK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CHK_OK;
}

chk_status chk_memzero(void *dev, size_t bytes, void *stream) {
  if (bytes == 0) return CHK_OK;
  if (!dev) {
    set_error("chk_memzero: null pointer");
    return CHK_EINVAL;
  }
  if (((uintptr_t)dev & 15u) != 0) {
    set_error("chk_memzero: pointer must be 16-byte aligned");
    return CHK_EINVAL;
  }
  const size_t n16 = bytes / 16, nt = bytes % 16;
  const unsigned blocks = (unsigned)std::min<size_t>(std::max<size_t>((n16 + 255) / 256, 1), 8192);
  hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, (hipStream_t)stream, (uint4 *)dev,
                     n16, (unsigned char *)dev + 16 * n16, nt);
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

chk_status chk_stream_create(int device, void **stream) {
  if (!stream) {
    set_error("chk_stream_create: null out");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(device));
  hipStream_t s;
  CHK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = (void *)s;
  return CHK_OK;
}

chk_status chk_stream_destroy(void *stream) {
  if (stream) CHK_HIP(hipStreamDestroy((hipStream_t)stream));
  return CHK_OK;
}

chk_status chk_stream_sync(void *stream) {
  CHK_HIP(hipStreamSynchronize((hipStream_t)stream));
  return CHK_OK;
}

chk_status chk_create(int device, chk_ctx_t *out) {
  CHK_TRY_BEGIN
  if (!out) {
    set_error("chk_create: null out");
    return CHK_EINVAL;
  }
  *out = nullptr;
  int n = 0;
  CHK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("chk_create: no such device");
    return CHK_EINVAL;
  }
  CHK_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  CHK_HIP(hipGetDeviceProperties(&prop, device));
  chk_ctx_s *c = new chk_ctx_s();
  c->device = device;
  c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  // the device's copy of chk_render_noalloc's scratch array; a failed
  // lookup is reported by that entry point
  (void)noalloc_scratch_init(device);
  *out = c;
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_destroy(chk_ctx_t c) {
  if (!c) return CHK_OK;
  (void)hipSetDevice(c->device);
  if (c->ltc_dev) (void)hipFree(c->ltc_dev);
  if (c->batch_state) batch_state_free(c->batch_state);
  delete c;
  return CHK_OK;
}

chk_status chk_scene_create(chk_ctx_t ctx, const float *tris, uint32_t n_tris,
                            const uint32_t *tri_mat, const chk_material *mats, uint32_t n_mats,
                            const chk_mesh_light *lights, uint32_t n_lights, chk_scene_t *out) {
  chk_scene_desc D;
  std::memset(&D, 0, sizeof(D));
  D.tris = tris;
  D.n_tris = n_tris;
  D.tri_mat = tri_mat;
  D.mats = mats;
  D.n_mats = n_mats;
  D.mesh_lights = lights;
  D.n_mesh_lights = n_lights;
  return chk_scene_create_desc(ctx, &D, out);
}

chk_status chk_scene_create_desc(chk_ctx_t ctx, const chk_scene_desc *desc, chk_scene_t *out) {
  CHK_TRY_BEGIN
  if (!ctx || !out || !desc) {
    set_error("chk_scene_create: null argument");
    return CHK_EINVAL;
  }
  const chk_scene_desc &D = *desc;
  const float *tris = D.tris;
  const uint32_t *tri_mat = D.tri_mat;
  const chk_material *mats = D.mats;
  const uint32_t n_tris = D.n_tris, n_mats = D.n_mats;
  const uint32_t n_lights = D.n_mesh_lights;
  if ((n_tris && (!tris || !tri_mat)) || (n_mats && !mats) || (n_lights && !D.mesh_lights) ||
      (D.n_tri_lights && !D.tri_lights) || (D.n_spheres && !D.spheres) ||
      (D.n_sphere_sources && !D.sphere_sources) || (D.n_point_sources && !D.point_sources)) {
    set_error("chk_scene_create: null argument");
    return CHK_EINVAL;
  }
  *out = nullptr;
  for (uint32_t i = 0; i < n_tris; ++i)
    if (tri_mat[i] >= n_mats) {
      set_error("chk_scene_create: triangle " + std::to_string(i) + " has material index " +
                std::to_string(tri_mat[i]) + " >= n_mats");
      return CHK_EINVAL;
    }
  if (D.n_spheres > kMaxSpheres || D.n_tri_lights > kMaxSources ||
      D.n_sphere_sources > kMaxSources || D.n_point_sources > kMaxSources) {
    set_error("chk_scene_create: more than " + std::to_string(kMaxSpheres) + " spheres or " +
              std::to_string(kMaxSources) + " triangle, sphere or point sources");
    return CHK_ECAPACITY;
  }
  // light runs: the mesh lights, then one run per triangle light
  std::vector<chk_mesh_light> runs(D.mesh_lights, D.mesh_lights + n_lights);
  for (uint32_t j = 0; j < D.n_tri_lights; ++j) {
    const uint32_t o = D.tri_lights[j];
    if (o >= n_tris || mats[tri_mat[o]].light == CHK_LIGHT_NONE) {
      set_error("chk_scene_create: triangle light " + std::to_string(j) +
                " names no emitter triangle");
      return CHK_EINVAL;
    }
    runs.push_back(chk_mesh_light{o, 1u});
  }
  for (uint32_t j = 0; j < D.n_spheres; ++j) {
    const chk_sphere &q = D.spheres[j];
    if (!(q.radius > 0) || !std::isfinite(q.radius) || !std::isfinite(q.center[0]) ||
        !std::isfinite(q.center[1]) || !std::isfinite(q.center[2]) || q.mat >= n_mats) {
      set_error("chk_scene_create: sphere " + std::to_string(j) +
                " has a bad centre, radius or material");
      return CHK_EINVAL;
    }
  }
  for (uint32_t j = 0; j < D.n_sphere_sources; ++j) {
    const uint32_t s = D.sphere_sources[j];
    if (s >= D.n_spheres || mats[D.spheres[s].mat].light == CHK_LIGHT_NONE) {
      set_error("chk_scene_create: sphere source " + std::to_string(j) +
                " names no emitter sphere");
      return CHK_EINVAL;
    }
  }
  if (D.env) {
    const chk_env_source &E = *D.env;
    const uint64_t px = (uint64_t)E.width * E.height;
    if (!E.image || E.width == 0 || E.height == 0 || E.axis > 2 || px >= kMaxEnvPixels ||
        !std::isfinite(E.radiance[0]) || !std::isfinite(E.radiance[1]) ||
        !std::isfinite(E.radiance[2])) {
      set_error("chk_scene_create: the environment light needs an image of 1 .. 2^28 pixels, "
                "axis 0, 1 or 2 and a finite radiance");
      return CHK_EINVAL;
    }
    for (uint64_t i = 0; i < 3 * px; ++i)
      if (!std::isfinite(E.image[i])) {
        set_error("chk_scene_create: the environment image is not finite");
        return CHK_EINVAL;
      }
  }
  for (uint32_t j = 0; j < D.n_point_sources; ++j)
    for (int c = 0; c < 3; ++c)
      if (!std::isfinite(D.point_sources[j].pos[c]) || !std::isfinite(D.point_sources[j].power[c])) {
        set_error("chk_scene_create: point source " + std::to_string(j) + " is not finite");
        return CHK_EINVAL;
      }
  const chk_mesh_light *lights = runs.data();
  const uint32_t n_runs = (uint32_t)runs.size();
  for (uint32_t j = 0; j < n_lights; ++j) {
    if (lights[j].n_tris == 0 || (uint64_t)lights[j].first_tri + lights[j].n_tris > n_tris) {
      set_error("chk_scene_create: mesh light " + std::to_string(j) + " is empty or out of range");
      return CHK_EINVAL;
    }
  }
  CHK_HIP(hipSetDevice(ctx->device));
  auto t0 = std::chrono::steady_clock::now();
  Bvh4 bvh;
  std::string msg;
  if (!build_bvh4(tris, n_tris, bvh, &msg)) {
    set_error("chk_scene_create: " + msg);
    return CHK_ECAPACITY;
  }
  const int stack = pick_stack_size(bvh.stats.stack_bound);
  if (stack < 0) {
    set_error("chk_scene_create: BVH needs a traversal stack of " +
              std::to_string(bvh.stats.stack_bound) + " entries (> 96)");
    return CHK_ECAPACITY;
  }
  auto t1 = std::chrono::steady_clock::now();
  const uint32_t m = bvh.stats.n_tris;

  // per-leaf (material, light)
  std::vector<int32_t> tri_light(n_tris, -1);
  for (uint32_t j = 0; j < n_lights; ++j)
    for (uint32_t i = 0; i < lights[j].n_tris; ++i) tri_light[lights[j].first_tri + i] = (int32_t)j;
  std::vector<uint32_t> info(2 * (size_t)std::max(m, 1u), 0xFFFFFFFFu);
  for (uint32_t li = 0; li < m; ++li) {
    uint32_t o = bvh.leaf_to_orig[li];
    info[2 * li] = tri_mat[o];
    info[2 * li + 1] = tri_light[o] >= 0 ? (uint32_t)tri_light[o] : 0xFFFFFFFFu;
  }
  // light tables: exact area CDF per light (the reference quantises it to
  // 300 bins, lightSources.h:151, lightSources.cpp:784-807)
  std::vector<LightTri> ltris;
  std::vector<LightRec> lrecs;
  std::vector<double> larea64;
  std::vector<LightTri64> ltri64;
  uint32_t used_lights = 0;
  for (uint32_t j = 0; j < n_runs; ++j) {
    LightRec R;
    R.first = (uint32_t)ltris.size();
    R.count = 0;
    double area = 0;
    std::vector<double> areas;
    for (uint32_t i = 0; i < lights[j].n_tris; ++i) {
      const float *p = tris + 9 * (size_t)(lights[j].first_tri + i);
      double e1[3], e2[3];
      for (int k = 0; k < 3; ++k) {
        e1[k] = (double)p[3 + k] - p[k];
        e2[k] = (double)p[6 + k] - p[k];
      }
      double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
             cz = e1[0] * e2[1] - e1[1] * e2[0];
      double a = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
      areas.push_back(a);
      area += a;
    }
    if (!(area > 0)) {
      set_error(std::string("chk_scene_create: ") + (j < n_lights ? "mesh light " : "triangle light ") +
                std::to_string(j < n_lights ? j : j - n_lights) + " has zero area");
      return CHK_EINVAL;
    }
    double cdf = 0;
    for (uint32_t i = 0; i < lights[j].n_tris; ++i) {
      const uint32_t o = lights[j].first_tri + i;
      const float *p = tris + 9 * (size_t)o;
      const chk_material &mt = mats[tri_mat[o]];
      cdf += areas[i] / area;
      // the reference's bin: index = cdf * 300 clamped to 299; the last
      // triangle also takes every bin after it (lightSources.cpp:793-806)
      uint32_t bin_end = (uint32_t)(cdf * 300);
      if (bin_end >= 300 || i + 1 == lights[j].n_tris) bin_end = 299;
      LightTri T;
      T.a_cdf = make_float4(p[0], p[1], p[2], i + 1 == lights[j].n_tris ? 2.0f : (float)cdf);
      T.b_area = make_float4(p[3], p[4], p[5], (float)areas[i]);
      uint32_t leaf = bvh.orig_to_leaf[o] >= 0 ? (uint32_t)bvh.orig_to_leaf[o] : 0xFFFFFFFFu;
      float lf;
      std::memcpy(&lf, &leaf, 4);
      T.c_leaf = make_float4(p[6], p[7], p[8], lf);
      T.rad = make_float4(mt.radiance[0], mt.radiance[1], mt.radiance[2], (float)bin_end);
      ltris.push_back(T);
      ltri64.push_back(make_light64(p));
      ++R.count;
    }
    R.area = (float)area;
    R.pad = 0;
    lrecs.push_back(R);
    larea64.push_back(area);
    ++used_lights;
  }

  // one device blob
  const size_t b_nodes = align256(bvh.qnodes.size() * sizeof(Node4Q));
  const size_t b_tris = align256(std::max<size_t>(bvh.tris.size(), 1) * sizeof(TriRec));
  const size_t b_info = align256(info.size() * 4);
  const size_t b_mats = align256(std::max<uint32_t>(n_mats, 1) * sizeof(chk_material));
  const size_t b_lt = align256(std::max<size_t>(ltris.size(), 1) * sizeof(LightTri));
  const size_t b_lr = align256(std::max<size_t>(lrecs.size(), 1) * sizeof(LightRec));
  // per original triangle (leaf, material) for entry points that name a
  // triangle by its caller index
  std::vector<uint32_t> oinfo(2 * (size_t)std::max(n_tris, 1u), 0u);
  for (uint32_t o = 0; o < n_tris; ++o) {
    oinfo[2 * (size_t)o] = bvh.orig_to_leaf[o] >= 0 ? (uint32_t)bvh.orig_to_leaf[o] : 0xFFFFFFFFu;
    oinfo[2 * (size_t)o + 1] = tri_mat[o];
  }
  const size_t b_oi = align256(oinfo.size() * 4);
  const size_t b_la = align256(std::max<size_t>(larea64.size(), 1) * sizeof(double));
  const size_t b_l64 = align256(std::max<size_t>(ltri64.size(), 1) * sizeof(LightTri64));
  // spheres and the other sources
  std::vector<SphereRec> sph(std::max<uint32_t>(D.n_spheres, 1u));
  std::vector<uint32_t> smat(std::max<uint32_t>(D.n_spheres, 1u), 0u);
  for (uint32_t j = 0; j < D.n_spheres; ++j) {
    for (int c = 0; c < 3; ++c) sph[j].c[c] = D.spheres[j].center[c];
    sph[j].r = D.spheres[j].radius;
    smat[j] = D.spheres[j].mat;
  }
  const size_t b_sp = align256(sph.size() * sizeof(SphereRec));
  const size_t b_sm = align256(smat.size() * 4);
  const size_t b_ss = align256(std::max<uint32_t>(D.n_sphere_sources, 1u) * 4);
  const size_t b_pt = align256(std::max<uint32_t>(D.n_point_sources, 1u) * sizeof(chk_point_source));
  const size_t total = b_nodes + b_tris + b_info + b_mats + b_lt + b_lr + b_oi + b_la + b_l64 +
                       b_sp + b_sm + b_ss + b_pt;
  void *blob = nullptr;
  if (hipMalloc(&blob, total) != hipSuccess) {
    set_error("chk_scene_create: cannot allocate " + std::to_string(total) + " device bytes");
    return CHK_ENOMEM;
  }
  char *p = (char *)blob;
  chk_scene_s *sc = new chk_scene_s();
  sc->ctx = ctx;
  sc->blob = blob;
  sc->blob_bytes = total;
  sc->d.nodes = (const Node4Q *)p;
  hipError_t e = hipMemcpy(p, bvh.qnodes.data(), bvh.qnodes.size() * sizeof(Node4Q),
                           hipMemcpyHostToDevice);
  p += b_nodes;
  sc->d.tris = (const float4 *)p;
  if (e == hipSuccess && !bvh.tris.empty())
    e = hipMemcpy(p, bvh.tris.data(), bvh.tris.size() * sizeof(TriRec), hipMemcpyHostToDevice);
  p += b_tris;
  sc->d.leaf_info = (const uint2 *)p;
  if (e == hipSuccess) e = hipMemcpy(p, info.data(), info.size() * 4, hipMemcpyHostToDevice);
  p += b_info;
  sc->d.mats = (const chk_material *)p;
  if (e == hipSuccess && n_mats)
    e = hipMemcpy(p, mats, n_mats * sizeof(chk_material), hipMemcpyHostToDevice);
  p += b_mats;
  sc->d.light_tris = (const LightTri *)p;
  if (e == hipSuccess && !ltris.empty())
    e = hipMemcpy(p, ltris.data(), ltris.size() * sizeof(LightTri), hipMemcpyHostToDevice);
  p += b_lt;
  sc->d.lights = (const LightRec *)p;
  if (e == hipSuccess && !lrecs.empty())
    e = hipMemcpy(p, lrecs.data(), lrecs.size() * sizeof(LightRec), hipMemcpyHostToDevice);
  p += b_lr;
  sc->d.orig_info = (const uint2 *)p;
  if (e == hipSuccess) e = hipMemcpy(p, oinfo.data(), oinfo.size() * 4, hipMemcpyHostToDevice);
  p += b_oi;
  sc->d.light_area = (const double *)p;
  if (e == hipSuccess && !larea64.empty())
    e = hipMemcpy(p, larea64.data(), larea64.size() * sizeof(double), hipMemcpyHostToDevice);
  p += b_la;
  sc->d.light64 = (const LightTri64 *)p;
  if (e == hipSuccess && !ltri64.empty())
    e = hipMemcpy(p, ltri64.data(), ltri64.size() * sizeof(LightTri64), hipMemcpyHostToDevice);
  p += b_l64;
  sc->d.spheres = (const SphereRec *)p;
  if (e == hipSuccess) e = hipMemcpy(p, sph.data(), sph.size() * sizeof(SphereRec), hipMemcpyHostToDevice);
  p += b_sp;
  sc->d.sphere_mat = (const uint32_t *)p;
  if (e == hipSuccess) e = hipMemcpy(p, smat.data(), smat.size() * 4, hipMemcpyHostToDevice);
  p += b_sm;
  sc->d.sph_src = (const uint32_t *)p;
  if (e == hipSuccess && D.n_sphere_sources)
    e = hipMemcpy(p, D.sphere_sources, D.n_sphere_sources * 4, hipMemcpyHostToDevice);
  p += b_ss;
  sc->d.points = (const chk_point_source *)p;
  if (e == hipSuccess && D.n_point_sources)
    e = hipMemcpy(p, D.point_sources, D.n_point_sources * sizeof(chk_point_source),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("chk_scene_create: upload: ") + hipGetErrorString(e));
    (void)hipFree(blob);
    delete sc;
    return CHK_EDEVICE;
  }
  sc->d.n_lights = n_lights;
  sc->d.n_tri_lights = D.n_tri_lights;
  sc->d.n_spheres = D.n_spheres;
  sc->d.n_sph_src = D.n_sphere_sources;
  sc->d.n_points = D.n_point_sources;
  sc->extras = D.n_tri_lights || D.n_spheres || D.n_sphere_sources || D.n_point_sources || D.env;
  if (D.env) {
    std::string why;
    if (!build_env(sc, *D.env, why)) {
      set_error("chk_scene_create: " + why);
      (void)hipFree(blob);
      delete sc;
      return CHK_EDEVICE;
    }
  }
  (void)used_lights;
  sc->d.n_leaf_tris = m;
  sc->d.n_tris = n_tris;
  sc->n_mats = n_mats;
  for (uint32_t i = 0; i < n_mats; ++i) {
    const uint32_t b = mats[i].brdf;
    const bool ok = b == CHK_BRDF_CLASSIC_PHONG || b == CHK_BRDF_GGX || b == CHK_BRDF_BECKMANN ||
                    b == CHK_BRDF_PHONG;
    if (mats[i].light == CHK_LIGHT_NONE && !ok) {
      sc->bad_sampler_mat = (int)i;
      sc->bad_sampler_brdf = b;
      break;
    }
  }
  {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const TriRec &t : bvh.tris)
      for (int k = 0; k < 3; ++k) {
        lo[k] = std::min(lo[k], std::min(t.a[k], std::min(t.b[k], t.c[k])));
        hi[k] = std::max(hi[k], std::max(t.a[k], std::max(t.b[k], t.c[k])));
      }
    for (int k = 0; k < 3; ++k) {
      const bool ok = m > 0 && hi[k] > lo[k];
      sc->d.box_lo[k] = ok ? lo[k] : 0.0f;
      sc->d.box_inv[k] = ok ? 1.0f / (hi[k] - lo[k]) : 1.0f;
    }
  }
  sc->stack_size = stack;
  {
    static std::atomic<uint64_t> next_gen{1};
    sc->gen = next_gen.fetch_add(1);
  }
  chk_scene_info &I = sc->info;
  I.n_tris = n_tris;
  I.n_tris_bvh = m;
  I.n_nodes = bvh.stats.n_nodes;
  I.n_leaves = bvh.stats.n_leaves;
  I.max_depth = bvh.stats.max_depth;
  I.stack_bound = bvh.stats.stack_bound;
  I.stack_size = (uint32_t)stack;
  I.n_lights = n_lights;
  I.device_bytes = total;
  I.sah_cost = bvh.stats.sah_cost;
  I.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  *out = sc;
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_scene_destroy(chk_scene_t sc) {
  if (!sc) return CHK_OK;
  (void)hipSetDevice(sc->ctx->device);
  if (sc->work) (void)hipFree(sc->work);
  if (sc->work2) (void)hipFree(sc->work2);
  if (sc->lwork2) (void)hipFree(sc->lwork2);
  if (sc->st2) (void)hipStreamDestroy(sc->st2);
  if (sc->ev_fork) (void)hipEventDestroy(sc->ev_fork);
  if (sc->ev_join) (void)hipEventDestroy(sc->ev_join);
  if (sc->lwork) (void)hipFree(sc->lwork);
  if (sc->dwork) (void)hipFree(sc->dwork);
  if (sc->env_mem) (void)hipFree(sc->env_mem);
  if (sc->dummy_box) (void)hipFree(sc->dummy_box);
  if (sc->blob) (void)hipFree(sc->blob);
  delete sc;
  return CHK_OK;
}

chk_status chk_scene_set_trace_stack(chk_scene_t sc, uint32_t entries) {
  if (!sc || (entries != 8 && entries != 16 && entries != 32)) {
    set_error("chk_scene_set_trace_stack: null scene, or entries not 8, 16 or 32");
    return CHK_EINVAL;
  }
  sc->shallow_stack = (int)entries;
  return CHK_OK;
}

chk_status chk_scene_set_accel(chk_scene_t sc, uint32_t kind) {
  CHK_TRY_BEGIN
  if (!sc || (kind != CHK_ACCEL_BVH && kind != CHK_ACCEL_DUMMY)) {
    set_error("chk_scene_set_accel: null scene, or kind not CHK_ACCEL_BVH / CHK_ACCEL_DUMMY");
    return CHK_EINVAL;
  }
  if (kind == CHK_ACCEL_DUMMY && !sc->dummy_box && sc->d.n_tris) {
    CHK_HIP(hipSetDevice(sc->ctx->device));
    // dummyAccel.cpp:13-23: nBoxT = (int)sqrt(n) runs of nBoxT triangles
    const uint32_t runs = std::max<uint32_t>(1u, (uint32_t)std::sqrt((double)sc->d.n_tris));
    float4 *box = nullptr;
    if (hipMalloc(&box, (size_t)runs * 2 * sizeof(float4)) != hipSuccess) {
      set_error("chk_scene_set_accel: cannot allocate the run boxes");
      return CHK_ENOMEM;
    }
    const hipError_t e = launch_dummy_boxes(sc, runs, runs, box, nullptr);
    const hipError_t e2 = e == hipSuccess ? hipDeviceSynchronize() : e;
    if (e2 != hipSuccess) {
      (void)hipFree(box);
      set_error(std::string("chk_scene_set_accel: ") + hipGetErrorString(e2));
      return CHK_EDEVICE;
    }
    sc->dummy_box = box;
    sc->dummy_runs = runs;
  }
  sc->accel = kind;
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_scene_get_info(chk_scene_t sc, chk_scene_info *info) {
  if (!sc || !info) {
    set_error("chk_scene_get_info: null argument");
    return CHK_EINVAL;
  }
  *info = sc->info;
  return CHK_OK;
}

chk_status chk_trace_closest(chk_scene_t sc, const chk_ray *rays, uint32_t n, chk_hit *hits,
                             void *stream) {
  CHK_TRY_BEGIN
  if (!sc || (n && (!rays || !hits))) {
    set_error("chk_trace_closest: null argument");
    return CHK_EINVAL;
  }
  if (n == 0) return CHK_OK;
  if (n > kMaxTraceBatch) {
    set_error("chk_trace_closest: more than 2^31 rays in one call");
    return CHK_ECAPACITY;
  }
  CHK_HIP(hipSetDevice(sc->ctx->device));
  if (sc->accel == CHK_ACCEL_DUMMY) {
    CHK_HIP(launch_dummy_trace(sc, false, sc->dummy_box, sc->dummy_runs, (const float4 *)rays, n,
                               (float4 *)hits, nullptr, (hipStream_t)stream));
    return CHK_OK;
  }
  // no scratch: the stack depth bound of the tree is exact, so calls on
  // distinct streams share nothing but the read-only scene
  CHK_HIP(launch_closest(sc, TraceShape{sc->stack_size, 128}, (const float4 *)rays,
                         linear_range(n), (float4 *)hits, false, false, nullptr, SlotScenes{},
                         (hipStream_t)stream));
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_trace_any(chk_scene_t sc, const chk_ray *rays, uint32_t n, uint8_t *occluded,
                         void *stream) {
  CHK_TRY_BEGIN
  if (!sc || (n && (!rays || !occluded))) {
    set_error("chk_trace_any: null argument");
    return CHK_EINVAL;
  }
  if (n == 0) return CHK_OK;
  if (n > kMaxTraceBatch) {
    set_error("chk_trace_any: more than 2^31 rays in one call");
    return CHK_ECAPACITY;
  }
  CHK_HIP(hipSetDevice(sc->ctx->device));
  if (sc->accel == CHK_ACCEL_DUMMY) {
    CHK_HIP(launch_dummy_trace(sc, true, sc->dummy_box, sc->dummy_runs, (const float4 *)rays, n,
                               nullptr, occluded, (hipStream_t)stream));
    return CHK_OK;
  }
  CHK_HIP(launch_any(sc, TraceShape{sc->stack_size, 128}, (const float4 *)rays,
                     linear_range(n), occluded, nullptr, nullptr, false, false, nullptr,
                     SlotScenes{}, (hipStream_t)stream));
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_trace_spheres(const float *spheres, uint32_t n_spheres, const chk_ray *rays,
                             uint32_t n, chk_hit *hits, void *stream) {
  CHK_TRY_BEGIN
  if ((n && (!rays || !hits)) || (n_spheres && !spheres)) {
    set_error("chk_trace_spheres: null argument");
    return CHK_EINVAL;
  }
  if (n == 0) return CHK_OK;
  CHK_HIP(launch_spheres(spheres, n_spheres, (const float4 *)rays, n, (float4 *)hits,
                         (hipStream_t)stream));
  return CHK_OK;
  CHK_TRY_END
}

chk_status chk_brdf_eval(const chk_material *mats, uint32_t n_mats, const uint32_t *mat_idx,
                         const double *n, const double *wo_ref, const double *wo,
                         const double *wi, uint32_t k, uint32_t flags, double *f, double *pdf,
                         void *stream) {
  if (k == 0) return CHK_OK;
  if (!mats || !mat_idx || !n || !wo_ref || !wo || !wi || !f || !pdf || n_mats == 0) {
    set_error("chk_brdf_eval: null argument");
    return CHK_EINVAL;
  }
  if (flags & ~(uint32_t)CHK_BRDF_FP32) {
    set_error("chk_brdf_eval: unknown flags");
    return CHK_EINVAL;
  }
  const dim3 grid((k + 255) / 256), block(256);
  if (flags & CHK_BRDF_FP32)
    hipLaunchKernelGGL(k_brdf_eval<float>, grid, block, 0, (hipStream_t)stream, mats, n_mats,
                       mat_idx, n, wo_ref, wo, wi, k, f, pdf);
  else
    hipLaunchKernelGGL(k_brdf_eval<double>, grid, block, 0, (hipStream_t)stream, mats, n_mats,
                       mat_idx, n, wo_ref, wo, wi, k, f, pdf);
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

chk_status chk_brdf_sample(const chk_material *mats, uint32_t n_mats, const uint32_t *mat_idx,
                           const double *n, const double *wo_ref, const double *wo,
                           const double *u, uint32_t k, uint32_t flags, double *wi, double *pdf,
                           void *stream) {
  if (k == 0) return CHK_OK;
  if (!mats || !mat_idx || !n || !wo_ref || !wo || !u || !wi || !pdf || n_mats == 0) {
    set_error("chk_brdf_sample: null argument");
    return CHK_EINVAL;
  }
  if (flags & ~(uint32_t)CHK_BRDF_FP32) {
    set_error("chk_brdf_sample: unknown flags");
    return CHK_EINVAL;
  }
  const dim3 grid((k + 255) / 256), block(256);
  if (flags & CHK_BRDF_FP32)
    hipLaunchKernelGGL(k_brdf_sample<float>, grid, block, 0, (hipStream_t)stream, mats, n_mats,
                       mat_idx, n, wo_ref, wo, u, k, wi, pdf);
  else
    hipLaunchKernelGGL(k_brdf_sample<double>, grid, block, 0, (hipStream_t)stream, mats, n_mats,
                       mat_idx, n, wo_ref, wo, u, k, wi, pdf);
  CHK_HIP(hipGetLastError());
  return CHK_OK;
}

chk_status chk_render(chk_scene_t sc, const chk_camera *camera, const chk_render_params *params,
                      double *accum, double *sumsq, chk_render_stats *stats, void *stream) {
  CHK_TRY_BEGIN
  if (!sc) {
    set_error("chk_render: null scene");
    return CHK_EINVAL;
  }
  if (reject_extras(sc, "chk_render") != CHK_OK) return CHK_EINVAL;
  CHK_HIP(hipSetDevice(sc->ctx->device));
  return render_impl(sc, camera, params, accum, sumsq, stats, false, nullptr, nullptr,
                     (hipStream_t)stream);
  CHK_TRY_END
}

chk_status chk_render_count(chk_scene_t sc, const chk_camera *camera,
                            const chk_render_params *params, double *accum,
                            chk_render_stats *stats, void *stream) {
  CHK_TRY_BEGIN
  if (!sc || !stats) {
    set_error("chk_render_count: null scene or stats");
    return CHK_EINVAL;
  }
  if (reject_extras(sc, "chk_render_count") != CHK_OK) return CHK_EINVAL;
  CHK_HIP(hipSetDevice(sc->ctx->device));
  return render_impl(sc, camera, params, accum, nullptr, stats, true, nullptr, nullptr,
                     (hipStream_t)stream);
  CHK_TRY_END
}

chk_status chk_render_learned(chk_scene_t sc, const chk_camera *camera,
                              const chk_render_params *params, chk_proposal_t proposal,
                              const chk_train_params *train, double *accum, double *sumsq,
                              chk_render_stats *stats, void *stream) {
  CHK_TRY_BEGIN
  if (!sc || !proposal) {
    set_error("chk_render_learned: null scene or proposal");
    return CHK_EINVAL;
  }
  if (reject_extras(sc, "chk_render_learned") != CHK_OK) return CHK_EINVAL;
  CHK_HIP(hipSetDevice(sc->ctx->device));
  return render_impl(sc, camera, params, accum, sumsq, stats, false, proposal, train,
                     (hipStream_t)stream);
  CHK_TRY_END
}

chk_status chk_render_batch(chk_ctx_t ctx, const chk_request *reqs, uint32_t n,
                            const chk_batch_tuning *tuning, double *latency_ms,
                            chk_batch_stats *stats, void *stream) {
  CHK_TRY_BEGIN
  if (!ctx || (n && !reqs)) {
    set_error("chk_render_batch: null context or requests");
    return CHK_EINVAL;
  }
  if (reject_extras(reqs, n, "chk_render_batch") != CHK_OK) return CHK_EINVAL;
  CHK_HIP(hipSetDevice(ctx->device));
  return batch_impl(ctx, reqs, n, tuning, latency_ms, stats, (hipStream_t)stream);
  CHK_TRY_END
}

chk_status chk_batch_autotune(chk_ctx_t ctx, const chk_request *reqs, uint32_t n,
                              chk_batch_tuning *best, double *best_ms, void *stream) {
  CHK_TRY_BEGIN
  if (!ctx || !best || (n && !reqs)) {
    set_error("chk_batch_autotune: null argument");
    return CHK_EINVAL;
  }
  if (reject_extras(reqs, n, "chk_batch_autotune") != CHK_OK) return CHK_EINVAL;
  CHK_HIP(hipSetDevice(ctx->device));
  return batch_autotune(ctx, reqs, n, best, best_ms, (hipStream_t)stream);
  CHK_TRY_END
}

}  // extern "C"
