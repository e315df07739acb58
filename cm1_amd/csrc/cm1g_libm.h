// Single-precision exp, pow and log with the host C library's rounding.
//
// The reference's REAL exp() and ** lower to the host C library's expf and
// powf, which (glibc >= 2.28, x86-64 with FMA) implement the table-driven
// method of Szabolcs Nagy published in ARM's optimized-routines: exp via a
// 32-entry table of 2^(i/32) and a cubic in double precision; pow via a
// 16-interval log2 (table of 1/c and log2(c), degree-5 polynomial) followed
// by the same exp2 step, everything in double precision and rounded to
// float once.  These device functions restate that method, with the fused
// multiply-adds the library's FMA build forms, so the GPU returns the same
// float as the host for every input: checked on the CPU against the host
// library for every float x in (-100, 88) for exp and every positive float
// x for each exponent the model uses (0.1346, 0.2046, 0.525, 0.875, Rd/cv,
// cp/Rd, Rd/cp), with no difference.
#pragma once
#include <stdint.h>
#include <cmath>

// host and device (the CPU test compiles this header with g++)
#if defined(__HIP__) || defined(__HIPCC__)
#define CM1_HD __host__ __device__
#else
#define CM1_HD
#endif

namespace cm1libm {

// bits(2^(i/32)) - (i << 47): adding k << 47 to entry k % 32 gives 2^(k/32)
static constexpr uint64_t EXP2_TAB[32] = {
    0x3ff0000000000000ULL, 0x3fefd9b0d3158574ULL, 0x3fefb5586cf9890fULL, 0x3fef9301d0125b51ULL,
    0x3fef72b83c7d517bULL, 0x3fef54873168b9aaULL, 0x3fef387a6e756238ULL, 0x3fef1e9df51fdee1ULL,
    0x3fef06fe0a31b715ULL, 0x3feef1a7373aa9cbULL, 0x3feedea64c123422ULL, 0x3feece086061892dULL,
    0x3feebfdad5362a27ULL, 0x3feeb42b569d4f82ULL, 0x3feeab07dd485429ULL, 0x3feea47eb03a5585ULL,
    0x3feea09e667f3bcdULL, 0x3fee9f75e8ec5f74ULL, 0x3feea11473eb0187ULL, 0x3feea589994cce13ULL,
    0x3feeace5422aa0dbULL, 0x3feeb737b0cdc5e5ULL, 0x3feec49182a3f090ULL, 0x3feed503b23e255dULL,
    0x3feee89f995ad3adULL, 0x3feeff76f2fb5e47ULL, 0x3fef199bdd85529cULL, 0x3fef3720dcef9069ULL,
    0x3fef5818dcfba487ULL, 0x3fef7c97337b9b5fULL, 0x3fefa4afa2a490daULL, 0x3fefd0765b6e4540ULL};
// 2^r ~ 1 + C2 r + C1 r^2 + C0 r^3 on |r| <= 1/64
constexpr double EXP2_C0 = 0x1.c6af84b912394p-5, EXP2_C1 = 0x1.ebfce50fac4f3p-3,
                 EXP2_C2 = 0x1.62e42ff0c52d6p-1;
// log2 intervals: 1/c and log2(c) of each of the 16 subintervals of [0x3f330000, 2x)
static constexpr double LOG2_INVC[16] = {
    0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0,  0x1.3c995b0b80385p+0,
    0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0,  0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
    0x1.0953f419900a7p+0, 0x1p+0,               0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
    0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1};
static constexpr double LOG2_LOGC[16] = {
    -0x1.efec65b963019p-2, -0x1.b0b6832d4fca4p-2, -0x1.7418b0a1fb77bp-2, -0x1.39de91a6dcf7bp-2,
    -0x1.01d9bf3f2b631p-2, -0x1.97c1d1b3b7afp-3,  -0x1.2f9e393af3c9fp-3, -0x1.960cbbf788d5cp-4,
    -0x1.a6f9db6475fcep-5, 0x0p+0,                0x1.338ca9f24f53dp-4,  0x1.476a9543891bap-3,
    0x1.e840b4ac4e4d2p-3,  0x1.40645f0c6651cp-2,  0x1.88e9c2c1b9ff8p-2,  0x1.ce0a44eb17bccp-2};
// log2(1+r) ~ A4 r + A3 r^2 + A2 r^3 + A1 r^4 + A0 r^5
constexpr double LOG2_A0 = 0x1.27616c9496e0bp-2, LOG2_A1 = -0x1.71969a075c67ap-2,
                 LOG2_A2 = 0x1.ec70a6ca7baddp-2, LOG2_A3 = -0x1.7154748bef6c8p-1,
                 LOG2_A4 = 0x1.71547652ab82bp0;

CM1_HD inline uint32_t f2u(float f) { return __builtin_bit_cast(uint32_t, f); }
CM1_HD inline float u2f(uint32_t u) { return __builtin_bit_cast(float, u); }
CM1_HD inline uint64_t d2u(double d) { return __builtin_bit_cast(uint64_t, d); }
CM1_HD inline double u2d(uint64_t u) { return __builtin_bit_cast(double, u); }
CM1_HD inline double fmad(double a, double b, double c) { return std::fma(a, b, c); }

// s * (1 + C2 r + C1 r^2 + C0 r^3) with s = 2^(k/32) from the table, k the
// rounded 32*x held in the low bits of kd + shift
CM1_HD inline float exp2_finish(double r, uint64_t ki, uint64_t sign_bias) {
  uint64_t t = EXP2_TAB[ki % 32];
  t += (ki + sign_bias) << 47;
  const double s = u2d(t);
  const double z = fmad(EXP2_C0, r, EXP2_C1);
  const double r2 = r * r;
  double y = fmad(EXP2_C2, r, 1.0);
  y = fmad(z, r2, y);
  y = y * s;
  return (float)y;
}

CM1_HD inline float expf_host(float x) {
  const uint32_t abstop = (f2u(x) >> 20) & 0x7ff;
  if (abstop >= (f2u(88.0f) >> 20)) {
    if (f2u(x) == f2u(-INFINITY)) return 0.0f;
    if (abstop >= (f2u(INFINITY) >> 20)) return x + x;
    if (x > 0x1.62e42ep6f) return INFINITY;
    if (x < -0x1.9fe368p6f) return 0.0f;
  }
  const double xd = (double)x;
  const double invln2n = 0x1.71547652b82fep+0 * 32, shift = 0x1.8p+52;
  double kd = fmad(invln2n, xd, shift);
  const uint64_t ki = d2u(kd);
  kd -= shift;
  const double r = fmad(invln2n, xd, -kd);
  // r is in units of 1/32: rescale the cubic accordingly
  const uint64_t t0 = EXP2_TAB[ki % 32] + (ki << 47);
  const double s = u2d(t0);
  const double c0 = EXP2_C0 / 32 / 32 / 32, c1 = EXP2_C1 / 32 / 32, c2 = EXP2_C2 / 32;
  const double z = fmad(c0, r, c1);
  const double r2 = r * r;
  double y = fmad(c2, r, 1.0);
  y = fmad(z, r2, y);
  y = y * s;
  return (float)y;
}

// log2 of the positive normal float with bits ix (subnormals come in
// normalised, with a negative biased exponent)
CM1_HD inline double log2_inline(uint32_t ix) {
  const uint32_t tmp = ix - 0x3f330000u;
  const int i = (int)((tmp >> 19) % 16);
  const uint32_t top = tmp & 0xff800000u;
  const uint32_t iz = ix - top;
  const int k = (int32_t)top >> 23;
  const double invc = LOG2_INVC[i], logc = LOG2_LOGC[i];
  const double z = (double)u2f(iz);
  const double r = fmad(z, invc, -1.0);
  const double y0 = logc + (double)k;
  const double r2 = r * r;
  double y = fmad(LOG2_A0, r, LOG2_A1);
  const double p = fmad(LOG2_A2, r, LOG2_A3);
  const double r4 = r2 * r2;
  double q = fmad(LOG2_A4, r, y0);
  q = fmad(p, r2, q);
  y = fmad(y, r4, q);
  return y;
}

// 0 if iy is not an integer, 1 if odd, 2 if even
CM1_HD inline int checkint(uint32_t iy) {
  const int e = (int)(iy >> 23 & 0xff);
  if (e < 0x7f) return 0;
  if (e > 0x7f + 23) return 2;
  if (iy & ((1u << (0x7f + 23 - e)) - 1)) return 0;
  if (iy & (1u << (0x7f + 23 - e))) return 1;
  return 2;
}
CM1_HD inline bool zeroinfnan(uint32_t ix) { return 2 * ix - 1 >= 2u * 0x7f800000u - 1; }

CM1_HD inline float powf_host(float x, float y) {
  uint64_t sign_bias = 0;
  uint32_t ix = f2u(x);
  const uint32_t iy = f2u(y);
  if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u || zeroinfnan(iy)) {
    if (zeroinfnan(iy)) {
      if (2 * iy == 0) return 1.0f;
      if (ix == 0x3f800000u) return 1.0f;
      if (2 * ix > 2u * 0x7f800000u || 2 * iy > 2u * 0x7f800000u) return x + y;
      if (2 * ix == 2 * 0x3f800000u) return 1.0f;
      if ((2 * ix < 2 * 0x3f800000u) == !(iy & 0x80000000u)) return 0.0f;
      return y * y;
    }
    if (zeroinfnan(ix)) {
      float x2 = x * x;
      if ((ix & 0x80000000u) && checkint(iy) == 1) x2 = -x2;
      return (iy & 0x80000000u) ? 1.0f / x2 : x2;
    }
    if (ix & 0x80000000u) {
      const int yint = checkint(iy);
      if (yint == 0) return NAN;
      if (yint == 1) sign_bias = 1ull << 16;
      ix &= 0x7fffffffu;
    }
    if (ix < 0x00800000u) {
      ix = f2u(u2f(ix) * 0x1p23f);
      ix &= 0x7fffffffu;
      ix -= 23u << 23;
    }
  }
  const double logx = log2_inline(ix);
  const double ylogx = (double)y * logx;
  if (((d2u(ylogx) >> 47) & 0xffff) >= (d2u(126.0) >> 47)) {
    if (ylogx > 0x1.fffffffd1d571p+6) return sign_bias ? -INFINITY : INFINITY;
    if (ylogx <= -150.0) return sign_bias ? -0.0f : 0.0f;
  }
  // exp2(ylogx): ylogx = k/32 + r
  const double shift = 0x1.8p+52 / 32;
  double kd = ylogx + shift;
  const uint64_t ki = d2u(kd);
  kd -= shift;
  const double r = ylogx - kd;
  return exp2_finish(r, ki, sign_bias);
}

// logf: the same 16 subintervals and 1/c as pow's log2 step, with ln(c)
// and ln(1+r) ~ r + C2 r^2 + C1 r^3 + C0 r^4 in double precision, rounded
// to float once (the method of the library's logf, FMA build)
static constexpr double LOGF_LOGC[16] = {
    -0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
    -0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3,   -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4,
    -0x1.252f438e10c1ep-5, 0x0p+0,                0x1.aa5aa5df25984p-5,  0x1.c5e53aa362eb4p-4,
    0x1.526e57720db08p-3,  0x1.bc2860d22477p-3,   0x1.1058bc8a07ee1p-2,  0x1.4043057b6ee09p-2};
constexpr double LOGF_C0 = -0x1.00ea348b88334p-2, LOGF_C1 = 0x1.5575b0be00b6ap-2,
                 LOGF_C2 = -0x1.ffffef20a4123p-2, LOGF_LN2 = 0x1.62e42fefa39efp-1;

CM1_HD inline float logf_host(float x) {
  uint32_t ix = f2u(x);
  if (ix == 0x3f800000u) return 0.0f;
  if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
    if (ix * 2 == 0) return -INFINITY;
    if (ix == 0x7f800000u) return x;
    if ((ix & 0x80000000u) || ix * 2 >= 0xff000000u) return NAN;
    // subnormal: normalise
    ix = f2u(x * 0x1p23f);
    ix -= 23u << 23;
  }
  const uint32_t tmp = ix - 0x3f330000u;
  const int i = (int)((tmp >> 19) % 16);
  const int k = (int32_t)tmp >> 23;
  const uint32_t iz = ix - (tmp & 0xff800000u);
  const double z = (double)u2f(iz);
  const double r = fmad(z, LOG2_INVC[i], -1.0);
  const double y0 = fmad((double)k, LOGF_LN2, LOGF_LOGC[i]);
  const double r2 = r * r;
  double y = fmad(LOGF_C1, r, LOGF_C2);
  y = fmad(LOGF_C0, r2, y);
  y = fmad(y, r2, y0 + r);
  return (float)y;
}

}  // namespace cm1libm
