// The environment light's device data: the image (FP64 RGB) and the inverse
// CDFs of EnvSource::initSamples (lightSources.cpp:662-782), which
// chk_env_host.cpp builds with the host's libm (see there why).  The
// estimators read them through chk_env_dev.h.
#include "chk_env_dev.h"

namespace chk {

double env_tables(const chk_env_source &src, uint32_t *inv_m, uint32_t *inv_c);

namespace {
inline size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
}  // namespace

bool build_env(chk_scene_s *sc, const chk_env_source &src, std::string &err) {
  const size_t n_px = (size_t)src.width * src.height;
  std::vector<uint32_t> inv_m(kEnvBins), inv_c((size_t)kEnvRows * kEnvBins);
  const double area = env_tables(src, inv_m.data(), inv_c.data());
  const size_t b_img = al256(n_px * 3 * sizeof(double));
  const size_t b_im = al256(inv_m.size() * 4);
  const size_t b_ic = al256(inv_c.size() * 4);
  void *mem = nullptr;
  if (hipMalloc(&mem, b_img + b_im + b_ic) != hipSuccess) {
    err = "cannot allocate the environment light's image and tables";
    return false;
  }
  char *p = (char *)mem;
  EnvDev E;
  E.img = (const double *)p;
  hipError_t e = hipMemcpy(p, src.image, n_px * 3 * sizeof(double), hipMemcpyHostToDevice);
  p += b_img;
  E.inv_m = (const uint32_t *)p;
  if (e == hipSuccess) e = hipMemcpy(p, inv_m.data(), inv_m.size() * 4, hipMemcpyHostToDevice);
  p += b_im;
  E.inv_c = (const uint32_t *)p;
  if (e == hipSuccess) e = hipMemcpy(p, inv_c.data(), inv_c.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    err = std::string("environment light upload: ") + hipGetErrorString(e);
    (void)hipFree(mem);
    return false;
  }
  for (int c = 0; c < 3; ++c) E.rad[c] = src.radiance[c];
  E.area = area;
  E.w = src.width;
  E.h = src.height;
  E.axis = src.axis;
  E.on = 1;
  sc->d.env = E;
  sc->env_mem = mem;
  return true;
}

}  // namespace chk
