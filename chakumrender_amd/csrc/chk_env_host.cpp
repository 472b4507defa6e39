// The environment light's tables (EnvSource::initSamples,
// lightSources.cpp:662-782), built once on the host at scene creation:
//   normalizationArea  sum of grey(getRadiance) sin(theta) over a 2000 x 4000
//                      (theta, phi) grid, times d_theta d_phi (calcArea)
//   marginal           per row of a 3000 x 4000 grid: row sum sin(theta)
//                      d_phi / area (calcMarginalDensity)
//   conditionals       grey sin(theta) / (area pMarginal) per cell
//                      (calcConditionalDensity)
//   inverse CDFs       2000 bins each (invertedCDF)
// Grid directions can fall exactly on pixel edges of the map (for a map 64
// pixels wide, every 62.5th of the 4000 columns), where the pixel a
// direction reads depends on the last bit of sin, cos, acos and asin.  The
// tables are therefore made with the host's libm, as the reference makes
// them, rows in parallel and each row's sum in the reference's order; the
// GPU reads them at render time.  The reference then
// pre-draws 200,000 samples from clock()-seeded tables; the estimators here
// draw from the inverse CDFs with Philox instead.
#include <cmath>
#include <cstdint>
#include <vector>

#include "chakum_types.h"

namespace chk {

namespace {

constexpr double kPi = 3.14159265359;  // mathPrimitives.h:4
constexpr uint32_t kAreaRows = 2000, kRows = 3000, kCols = 4000, kBins = 2000;

// A double converted to uint32_t as the reference's x86 build converts it:
// NaN gives 0.
inline uint32_t to_u32(double v) { return v == v ? (uint32_t)v : 0u; }

struct Env {
  const chk_env_source *s;
  // EnvSource::getThetaPhi + getRadiance + to_greyScale of unit direction
  // (dx, dy, dz); the row is clamped to the image (the reference reads one
  // row past it at theta = pi).
  double grey(double dx, double dy, double dz) const {
    double x, y, z;
    if (s->axis == 0) {
      x = dy; y = dz; z = dx;
    } else if (s->axis == 1) {
      x = dx; y = dz; z = dy;
    } else {
      x = dx; y = dy; z = dz;
    }
    const double theta = std::acos(z);
    const double st = std::sqrt(1 - z * z);
    double phi;
    if (st < 1e-12) {
      phi = 0.0;
    } else {
      x /= st;
      y /= st;
      if (y >= 0 && x >= 0)
        phi = std::asin(y);
      else if (y >= 0 && x <= 0)
        phi = kPi - std::asin(y);
      else if (y <= 0 && x <= 0)
        phi = kPi + std::asin(-y);
      else
        phi = 2 * kPi - std::asin(-y);
    }
    uint32_t col = to_u32((phi / (2 * kPi)) * s->width), row = to_u32((theta / kPi) * s->height);
    col = col >= s->width ? s->width - 1 : col;
    row = row >= s->height ? s->height - 1 : row;
    const double *p = s->image + 3 * ((size_t)row * s->width + col);
    const double L0 = p[0] * s->radiance[0], L1 = p[1] * s->radiance[1], L2 = p[2] * s->radiance[2];
    return L0 * 0.3 + L1 * 0.59 + L2 * 0.11;  // mathPrimitives.h:10
  }
  // the cells of row i of a rows x kCols grid (theta = i pi / rows)
  double row(uint32_t rows, uint32_t i, double *vals) const {
    const double dth = kPi / rows, dph = 2.0 * kPi / kCols;
    const double st = std::sin(dth * i), ct = std::cos(dth * i);
    double sum = 0;
    for (uint32_t j = 0; j < kCols; ++j) {
      const double cp = std::cos(dph * j), sp = std::sin(dph * j);
      const double g = grey(st * cp, st * sp, ct);
      if (vals) vals[j] = g;
      sum += g;
    }
    return sum;
  }
};

// invertedCDF (lightSources.cpp:662-678).
void inverted_cdf(const double *pdf, uint32_t length, double delta, uint32_t *inv, uint32_t nb) {
  double sum = 0;
  for (uint32_t j = 0; j < nb; ++j) inv[j] = 0;
  for (uint32_t i = 0; i < length; ++i) {
    sum += pdf[i] * delta;
    const uint32_t k = to_u32(sum * nb);
    inv[k >= nb ? nb - 1 : k] = i;
    for (uint32_t j = k - 1; j < nb && inv[j] == 0; j--) inv[j] = i;
  }
  for (uint32_t j = to_u32(sum * nb); j < nb; ++j) inv[j] = length - 1;
}

}  // namespace

// normalizationArea and the inverse CDFs (inv_m: 2000, inv_c: 3000 x 2000).
double env_tables(const chk_env_source &src, uint32_t *inv_m, uint32_t *inv_c) {
  const Env E{&src};
  std::vector<double> rs(kAreaRows);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < (int)kAreaRows; ++i) rs[i] = E.row(kAreaRows, (uint32_t)i, nullptr);
  const double dth_a = kPi / kAreaRows, dph = 2.0 * kPi / kCols;
  double area = 0;
  for (uint32_t i = 0; i < kAreaRows; ++i) area += rs[i] * std::sin(dth_a * i);
  area = area * dth_a * dph;
  const double dth = kPi / kRows;
  std::vector<double> pm(kRows);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < (int)kRows; ++i)
    pm[i] = (E.row(kRows, (uint32_t)i, nullptr) * std::sin(dth * i) * dph) / area;
  inverted_cdf(pm.data(), kRows, kPi / kRows, inv_m, kBins);
#pragma omp parallel
  {
    std::vector<double> g(kCols);
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < (int)kRows; ++i) {
      E.row(kRows, (uint32_t)i, g.data());
      const double st = std::sin(dth * i);
      for (uint32_t j = 0; j < kCols; ++j) g[j] = g[j] * st / (area * pm[i]);
      inverted_cdf(g.data(), kCols, 2 * kPi / kCols, inv_c + (size_t)i * kBins, kBins);
    }
  }
  return area;
}

}  // namespace chk
