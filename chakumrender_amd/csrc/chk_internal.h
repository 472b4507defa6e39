// Internal structures shared by the host API and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "chakum_c.h"
#include "chk_bvh.h"

namespace chk {

void set_error(const std::string &msg);

#define CHK_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::chk::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));         \
      return CHK_EDEVICE;                                                          \
    }                                                                              \
  } while (0)

// Device-resident light data for next-event estimation.
struct LightTri {
  float4 a_cdf;    // A, cumulative area fraction within its light
  float4 b_area;   // B, triangle area
  float4 c_leaf;   // C, leaf index (raw bits)
  float4 rad;      // radiance (emitter material), w unused
};
struct LightRec {
  uint32_t first;  // into the LightTri table
  uint32_t count;
  float area;      // total area of the mesh light
  float pad;
};

struct DeviceScene {
  const Node4 *nodes = nullptr;
  const float4 *tris = nullptr;        // 3 float4 per leaf triangle
  const uint2 *leaf_info = nullptr;    // (material, light or 0xFFFFFFFF)
  const chk_material *mats = nullptr;
  const LightTri *light_tris = nullptr;
  const LightRec *lights = nullptr;
  uint32_t n_lights = 0;
  uint32_t n_leaf_tris = 0;
};

}  // namespace chk

struct chk_ctx_s {
  int device = 0;
  int n_cu = 256;
  double *ltc_dev = nullptr;   // packed FP64 table for the FP64 LTC pass (6 per record)
  float *ltc_tab = nullptr;    // packed FP32 table for the LTC kernel (see chk_ltc.hip)
  std::vector<double> ltc_host;
  void *ltc_scratch = nullptr; // FP64-pass point list + counter
  size_t ltc_scratch_bytes = 0;
  float ltc_kappa = 8.0f;  // FP64 re-evaluation threshold of chk_ltc_eval
};

struct chk_scene_s {
  chk_ctx_s *ctx = nullptr;
  chk::DeviceScene d;
  void *blob = nullptr;        // one device allocation holding all scene arrays
  size_t blob_bytes = 0;
  chk_scene_info info{};
  int stack_size = 32;
  // render scratch (grown on demand)
  void *work = nullptr;
  size_t work_bytes = 0;
};

namespace chk {

// Kernel launchers (chk_trace.hip).
// Closest hit over rays (chk_ray layout).  count_dev != nullptr: ray count
// read from device memory (at most n).  leaf_ids: report/skip leaf indices
// instead of original ones.  work_counter must be zero at launch.
// counters (dev, 3 x u64): [node visits, tri tests] accumulated when
// `count`, [2] = lanes whose stack overflowed (always).
hipError_t launch_closest(const chk_scene_s *s, const float4 *rays, uint32_t n,
                          const uint32_t *count_dev, float4 *hits, bool leaf_ids,
                          uint32_t *work_counter, bool count, unsigned long long *counters,
                          int grid, hipStream_t st);
// Any hit.  occluded (may be null) receives 0/1; with path_L, unoccluded
// rays add shadow_L[i].xyz to path_L[bits(shadow_L[i].w)].
hipError_t launch_any(const chk_scene_s *s, const float4 *rays, uint32_t n,
                      const uint32_t *count_dev, uint8_t *occluded, const float4 *shadow_L,
                      float4 *path_L, bool leaf_ids, uint32_t *work_counter, bool count,
                      unsigned long long *counters, int grid, hipStream_t st);
hipError_t launch_zero_u32(uint32_t *p, uint32_t n, hipStream_t st);

int pick_stack_size(uint32_t bound);
int trace_grid(const chk_scene_s *s);

}  // namespace chk
