// Device helpers shared by the advection kernels: flux formulas of
// src/adv_routines.F (flx3 :5971-5984, flx5 :6006-6019, weno5 :6166-6199,
// upstrpd :6331-6348) and the vertical face fluxes of vadv_flx5 /
// vadv_weno5 with their lower-order boundary fluxes.
#pragma once
#include "cm1g_internal.h"

__device__ __forceinline__ float flx3(float s1, float s2, float s3) {
  return (-1.0f / 6.0f) * s1 + (5.0f / 6.0f) * s2 + (2.0f / 6.0f) * s3;
}
__device__ __forceinline__ float flx5(float s1, float s2, float s3, float s4, float s5) {
  return (2.0f / 60.0f) * s1 + (-13.0f / 60.0f) * s2 + (47.0f / 60.0f) * s3 +
         (27.0f / 60.0f) * s4 + (-3.0f / 60.0f) * s5;
}

// weno5 with the Borges et al. indicators (siform=2),
// src/adv_routines.F:6166-6199: the indicators are formed in single
// precision and widened, the weights and the combination are double
// precision, the result is rounded to single once.
__device__ __forceinline__ float weno5(float s1, float s2, float s3, float s4, float s5, double weps) {
  const float a1 = s1 - 2.0f * s2 + s3, e1 = s1 - 4.0f * s2 + 3.0f * s3;
  const float a2 = s2 - 2.0f * s3 + s4, e2 = s2 - s4;
  const float a3 = s3 - 2.0f * s4 + s5, e3 = 3.0f * s3 - 4.0f * s4 + s5;
  const double b1 = (double)((13.0f / 12.0f) * (a1 * a1) + 0.25f * (e1 * e1));
  const double b2 = (double)((13.0f / 12.0f) * (a2 * a2) + 0.25f * (e2 * e2));
  const double b3 = (double)((13.0f / 12.0f) * (a3 * a3) + 0.25f * (e3 * e3));
  const double d13 = fabs(b1 - b3);
  const double q1 = fmin(1.0e30, d13 / (b1 + weps));
  const double q2 = fmin(1.0e30, d13 / (b2 + weps));
  const double q3 = fmin(1.0e30, d13 / (b3 + weps));
  const double w1 = (double)0.1f * ((double)1.0f + q1 * q1);
  const double w2 = (double)0.6f * ((double)1.0f + q2 * q2);
  const double w3 = (double)0.3f * ((double)1.0f + q3 * q3);
  const float p1 = (2.0f / 6.0f) * s1 + (-7.0f / 6.0f) * s2 + (11.0f / 6.0f) * s3;
  const float p2 = (-1.0f / 6.0f) * s2 + (5.0f / 6.0f) * s3 + (2.0f / 6.0f) * s4;
  const float p3 = (2.0f / 6.0f) * s3 + (5.0f / 6.0f) * s4 + (-1.0f / 6.0f) * s5;
  return (float)((w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3));
}

// positive-definite upstream value (Beets & Koren), src/adv_routines.F:6331-6348
__device__ __forceinline__ float upstrpd(float s1, float s2, float s3, double weps) {
  const float dd = s2 - s1;
  const float rr = (s3 - s2) / (copysignf((float)weps, dd) + dd);
  const float phi = fmaxf(0.0f, fminf(rr, fminf((1.0f / 6.0f) + (2.0f / 6.0f) * rr, 1.0f)));
  return s2 + phi * (s2 - s1);
}

// the pdef magnitude test of hadv_weno5/vadv_weno5: false when all five
// stencil values are within weps of zero (the flux is then zero)
__device__ __forceinline__ bool weno_doit(float s1, float s2, float s3, float s4, float s5, double weps) {
  const float bmax = fmaxf(fmaxf(fmaxf(fmaxf(fabsf(s1), fabsf(s2)), fabsf(s3)), fabsf(s4)), fabsf(s5));
  return bmax > (float)weps;
}

// upwind face value with 5th-order fluxes or WENO5
template <bool WENO>
__device__ __forceinline__ float face5(float s1, float s2, float s3, float s4, float s5, double weps) {
  if (WENO) return weno5(s1, s2, s3, s4, s5, weps);
  return flx5(s1, s2, s3, s4, s5);
}

// upwind 5th-order face value between a(m) (left) and a(m+1) given the
// five-point stencils: positive flow uses a(m-2..m+2), negative a(m+3..m-1)
#define FLXP(A, o1, o2, o3, o4, o5) flx5(A(o1), A(o2), A(o3), A(o4), A(o5))

// Upwind 5th-order face value for the face between cells m-1 and m along one
// horizontal line, from the seven values S(m-3..m+3) (only six are used).
// (the upwind stencil is selected first, then one flux is evaluated)
template <bool WENO>
__device__ __forceinline__ float hface5(float r, float sm3, float sm2, float sm1, float s0, float sp1,
                                        float sp2, int pdef, double weps) {
  const bool up = (r >= 0.0f);
  const float s1 = up ? sm3 : sp2, s2 = up ? sm2 : sp1, s3 = up ? sm1 : s0;
  const float s4 = up ? s0 : sm1, s5 = up ? sp1 : sm2;
  if (!WENO) return r * flx5(s1, s2, s3, s4, s5);
  if (pdef == 1 && !weno_doit(s1, s2, s3, s4, s5, weps)) return 0.0f;
  return r * weno5(s1, s2, s3, s4, s5, weps);
}

__device__ __forceinline__ float wface(const KP &P, int stag, const float *rrw, int i, int j, int k) {
  if (stag == 1) return rrw[IXD(P, i, j, k)];
  if (stag == 2) return 0.5f * (rrw[IXD(P, i, j, k)] + rrw[IXD(P, i - 1, j, k)]);
  if (stag == 3) return 0.5f * (rrw[IXD(P, i, j, k)] + rrw[IXD(P, i, j - 1, k)]);
  return 0.5f * (rrw[IXD(P, i, j, k)] + rrw[IXD(P, i, j, k + 1)]);
}

// vertical face flux at level k for staggering stag (interior 5th order and
// the lower-order boundary fluxes of vadv_lwr_flx_23/35); S(kk) gives the
// advected value at level kk of the column, wb the face mass flux.
template <class SF>
__device__ __forceinline__ float vflux5_of(const KP &P, int stag, float wb, SF S, int k) {
  const int nk = P.nk;
  if (stag != 4) {
    if (k >= 4 && k <= nk - 2)
      return (wb >= 0.0f) ? wb * flx5(S(k - 3), S(k - 2), S(k - 1), S(k), S(k + 1))
                          : wb * flx5(S(k + 2), S(k + 1), S(k), S(k - 1), S(k - 2));
    if (k == 3)
      return (wb >= 0.0f) ? wb * flx3(S(k - 2), S(k - 1), S(k))
                          : wb * flx5(S(k + 2), S(k + 1), S(k), S(k - 1), S(k - 2));
    if (k == nk - 1)
      return (wb >= 0.0f) ? wb * flx5(S(k - 3), S(k - 2), S(k - 1), S(k), S(k + 1))
                          : wb * flx3(S(k + 1), S(k), S(k - 1));
    // k == 2 or k == nk
    float lin;
    if (stag == 1) lin = P.c1[k] * S(k - 1) + P.c2[k] * S(k);
    else {
      float cc2 = 0.5f * (P.c2[k] + P.c2[k]);
      float cc1 = 1.0f - cc2;
      lin = cc1 * S(k - 1) + cc2 * S(k);
    }
    if (k == 2) return (wb >= 0.0f) ? wb * lin : wb * flx3(S(k + 1), S(k), S(k - 1));
    return (wb >= 0.0f) ? wb * flx3(S(k - 2), S(k - 1), S(k)) : wb * lin;
  }
  if (k >= 3 && k <= nk - 2)
    return (wb >= 0.0f) ? wb * flx5(S(k - 2), S(k - 1), S(k), S(k + 1), S(k + 2))
                        : wb * flx5(S(k + 3), S(k + 2), S(k + 1), S(k), S(k - 1));
  if (k == 1)
    return (wb >= 0.0f) ? wb * 0.5f * (S(k) + S(k + 1)) : wb * flx3(S(k + 2), S(k + 1), S(k));
  if (k == nk)
    return (wb >= 0.0f) ? wb * flx3(S(k - 1), S(k), S(k + 1)) : wb * 0.5f * (S(k) + S(k + 1));
  if (k == 2)
    return (wb >= 0.0f) ? wb * flx3(S(k - 1), S(k), S(k + 1))
                        : wb * flx5(S(k + 3), S(k + 2), S(k + 1), S(k), S(k - 1));
  // k == nk-1
  return (wb >= 0.0f) ? wb * flx5(S(k - 2), S(k - 1), S(k), S(k + 1), S(k + 2))
                      : wb * flx3(S(k + 2), S(k + 1), S(k));
}

// vertical face flux with WENO5 (vadv_weno5, src/adv_routines.F:2359-2557)
// and its lower-order boundary fluxes (vadv_lwr_weno_23 :4038-4208,
// vadv_lwr_weno_35 :4214-4406, vadv_lwr_weno_35_b :4412-4447).  pdefweno
// (scalars only) adds the magnitude test and the one-sided upstream values
// at k=3 and k=nk-1.
template <class SF>
__device__ __forceinline__ float vfluxw_of(const KP &P, int stag, float wb, SF S, int k, int pdefweno,
                                           double weps) {
  const int nk = P.nk;
  if (stag == 4) {
    if (k == 1)
      return (wb >= 0.0f) ? wb * 0.5f * (S(k) + S(k + 1)) : wb * upstrpd(S(k + 2), S(k + 1), S(k), weps);
    if (k == nk)
      return (wb > 0.0f) ? wb * upstrpd(S(k - 1), S(k), S(k + 1), weps) : wb * 0.5f * (S(k) + S(k + 1));
    if (k == 2)
      return (wb >= 0.0f) ? wb * upstrpd(S(k - 1), S(k), S(k + 1), weps)
                          : wb * weno5(S(k + 3), S(k + 2), S(k + 1), S(k), S(k - 1), weps);
    if (k == nk - 1)
      return (wb > 0.0f) ? wb * weno5(S(k - 2), S(k - 1), S(k), S(k + 1), S(k + 2), weps)
                         : wb * upstrpd(S(k + 2), S(k + 1), S(k), weps);
    return (wb >= 0.0f) ? wb * weno5(S(k - 2), S(k - 1), S(k), S(k + 1), S(k + 2), weps)
                        : wb * weno5(S(k + 3), S(k + 2), S(k + 1), S(k), S(k - 1), weps);
  }
  if (k == 2 || k == nk) {
    float lin;
    if (stag == 1) lin = P.c1[k] * S(k - 1) + P.c2[k] * S(k);
    else {
      float cc2 = 0.5f * (P.c2[k] + P.c2[k]);
      float cc1 = 1.0f - cc2;
      lin = cc1 * S(k - 1) + cc2 * S(k);
    }
    if (k == 2) return (wb >= 0.0f) ? wb * lin : wb * upstrpd(S(k + 1), S(k), S(k - 1), weps);
    return (wb > 0.0f) ? wb * upstrpd(S(k - 2), S(k - 1), S(k), weps) : wb * lin;
  }
  if (stag == 1 && pdefweno == 1) {
    if (k == 3 && wb > 0.0f) return wb * upstrpd(S(k - 2), S(k - 1), S(k), weps);
    if (k == nk - 1 && wb < 0.0f) return wb * upstrpd(S(k + 1), S(k), S(k - 1), weps);
    float s1, s2, s3, s4, s5;
    if (wb >= 0.0f) { s1 = S(k - 3); s2 = S(k - 2); s3 = S(k - 1); s4 = S(k); s5 = S(k + 1); }
    else { s1 = S(k + 2); s2 = S(k + 1); s3 = S(k); s4 = S(k - 1); s5 = S(k - 2); }
    return weno_doit(s1, s2, s3, s4, s5, weps) ? wb * weno5(s1, s2, s3, s4, s5, weps) : 0.0f;
  }
  if (k == 3)
    return (wb >= 0.0f) ? wb * upstrpd(S(k - 2), S(k - 1), S(k), weps)
                        : wb * weno5(S(k + 2), S(k + 1), S(k), S(k - 1), S(k - 2), weps);
  if (k == nk - 1) {
    const bool up = (stag == 1) ? (wb >= 0.0f) : (wb > 0.0f);
    return up ? wb * weno5(S(k - 3), S(k - 2), S(k - 1), S(k), S(k + 1), weps)
              : wb * upstrpd(S(k + 1), S(k), S(k - 1), weps);
  }
  return (wb >= 0.0f) ? wb * weno5(S(k - 3), S(k - 2), S(k - 1), S(k), S(k + 1), weps)
                      : wb * weno5(S(k + 2), S(k + 1), S(k), S(k - 1), S(k - 2), weps);
}

__device__ __forceinline__ float vflux5(const KP &P, int stag, const float *rrw, const float *a,
                                       int i, int j, int k) {
  auto S = [&](int kk) { return a[IXD(P, i, j, kk)]; };
  return vflux5_of(P, stag, wface(P, stag, rrw, i, j, k), S, k);
}

// vadv_flx5 (src/adv_routines.F:3269-3388) or vadv_weno5 (:2359-2557) for
// the momentum equations (no pdef variant)
