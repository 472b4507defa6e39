// The environment light on the device: EnvSource::getRadiance
// (lightSources.h:57-101), invertedCDF (lightSources.cpp:662-678) and the
// importance sample of initSamples (lightSources.cpp:765-779), in FP64 with
// the reference's operation order (no contraction into FMAs).
#pragma once

#include "chk_internal.h"
#include "chk_ltc_dev.h"

namespace chk {

// Grid and bin counts of initSamples (lightSources.cpp:681-682,743-754).
constexpr uint32_t kEnvAreaRows = 2000, kEnvRows = 3000, kEnvCols = 4000, kEnvBins = 2000;

// A double converted to uint32_t as the reference's x86 build does it: NaN
// gives 0 (cvttsd2si's "integer indefinite", low word).
__device__ inline uint32_t to_u32(double v) { return v == v ? (uint32_t)v : 0u; }

// EnvSource::getThetaPhi + getRadiance: the pixel of unit direction d.
// Where a rounding excess makes acos or asin NaN the reference's conversion
// gives row or column 0; the row is clamped to the image (the reference reads
// one row past it at theta = pi).
__device__ inline uint32_t env_pixel(const EnvDev &E, double dx, double dy, double dz) {
#pragma clang fp contract(off)
  double x, y, z;
  if (E.axis == 0) {
    x = dy;
    y = dz;
    z = dx;
  } else if (E.axis == 1) {
    x = dx;
    y = dz;
    z = dy;
  } else {
    x = dx;
    y = dy;
    z = dz;
  }
  const double theta = acos(z);
  const double st = sqrt(1 - z * z);
  double phi;
  if (st < 1e-12) {
    phi = 0.0;
  } else {
    x /= st;
    y /= st;
    if (y >= 0 && x >= 0)
      phi = asin(y);
    else if (y >= 0 && x <= 0)
      phi = kPi64 - asin(y);
    else if (y <= 0 && x <= 0)
      phi = kPi64 + asin(-y);
    else
      phi = 2 * kPi64 - asin(-y);
  }
  uint32_t col = to_u32((phi / (2 * kPi64)) * E.w), row = to_u32((theta / kPi64) * E.h);
  col = col >= E.w ? E.w - 1 : col;
  row = row >= E.h ? E.h - 1 : row;
  return row * E.w + col;
}
// envMap[p].mult(radiance)
__device__ inline d3 env_pixel_radiance(const EnvDev &E, uint32_t p) {
#pragma clang fp contract(off)
  return dk(E.img[3 * (size_t)p] * E.rad[0], E.img[3 * (size_t)p + 1] * E.rad[1],
            E.img[3 * (size_t)p + 2] * E.rad[2]);
}
// to_greyScale (mathPrimitives.h:10)
__device__ inline double grey(d3 L) {
#pragma clang fp contract(off)
  return L.x * 0.3 + L.y * 0.59 + L.z * 0.11;
}
// The importance sample of (e1, e2): the marginal's inverse CDF picks the
// row, that row's conditional inverse CDF the column; the direction is the
// cell's corner (theta, phi) about world z.
__device__ inline d3 env_sample(const EnvDev &E, double e1, double e2) {
#pragma clang fp contract(off)
  const uint32_t ti = E.inv_m[(uint32_t)(e1 * kEnvBins)];
  const double theta = ti * (kPi64 / kEnvRows);
  const uint32_t pj = E.inv_c[(size_t)ti * kEnvBins + (uint32_t)(e2 * kEnvBins)];
  const double phi = pj * (2.0 * kPi64) / kEnvCols;
  double x = sin(theta), y = x;
  x *= cos(phi);
  y *= sin(phi);
  return dk(x, y, cos(theta));
}

}  // namespace chk
