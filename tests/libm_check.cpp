// Compares the restated single-precision exp/pow/log of cm1_amd/csrc/cm1g_libm.h
// (the functions the HIP kernels call) with the host C library's expf/powf.
// usage: libm_check STRIDE  -> prints "exp N BAD", "pow Y N BAD" and "log N BAD" lines
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "cm1g_libm.h"

static uint32_t bits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
static float fl(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

int main(int argc, char **argv) {
  const uint32_t stride = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 1;
  long n = 0, bad = 0;
  for (uint64_t u = 0; u < 0xffffffffull; u += stride) {
    const float x = fl((uint32_t)u);
    if (!(x > -100.0f && x < 88.0f)) continue;
    n++;
    if (bits(cm1libm::expf_host(x)) != bits(expf(x))) bad++;
  }
  // special and edge inputs
  const float ex[] = {0.0f, -0.0f, 88.0f, 88.7f, 89.0f, -103.0f, -104.0f, -150.0f, INFINITY, -INFINITY};
  for (float x : ex) { n++; if (bits(cm1libm::expf_host(x)) != bits(expf(x))) bad++; }
  std::printf("exp %ld %ld\n", n, bad);
  volatile float rd = 287.04f, cp = 1005.7f;
  const float cv = cp - rd;
  const float ys[] = {0.1346f, 0.2046f, 0.525f, 0.875f, rd / cv, cp / rd, rd / cp, 0.33333333f};
  for (float y : ys) {
    n = 0;
    bad = 0;
    for (uint64_t u = 1; u < 0x7f800000ull; u += stride) {
      const float x = fl((uint32_t)u);
      n++;
      if (bits(cm1libm::powf_host(x, y)) != bits(powf(x, y))) bad++;
    }
    const float px[] = {0.0f, -0.0f, 1.0f, INFINITY, -2.0f, 1e-45f, 3.0e38f};
    for (float x : px) {
      n++;
      const float a = cm1libm::powf_host(x, y), b = powf(x, y);
      if (bits(a) != bits(b) && !(std::isnan(a) && std::isnan(b))) bad++;
    }
    std::printf("pow %a %ld %ld\n", y, n, bad);
  }
  // log: every STRIDE-th positive float plus edge inputs
  n = 0;
  bad = 0;
  for (uint64_t u = 1; u < 0x7f800000ull; u += stride) {
    const float x = fl((uint32_t)u);
    n++;
    if (bits(cm1libm::logf_host(x)) != bits(logf(x))) bad++;
  }
  const float lx[] = {1.0f, 0.0f, -0.0f, INFINITY, 1e-45f, 0.99999994f, 1.0000001f, 3.4e38f};
  for (float x : lx) { n++; if (bits(cm1libm::logf_host(x)) != bits(logf(x))) bad++; }
  std::printf("log %ld %ld\n", n, bad);
  return 0;
}
