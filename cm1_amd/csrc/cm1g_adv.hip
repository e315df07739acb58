// Flux-form advection on the device.
// flx3/flx5: src/adv_routines.F:5971-5984, :6006-6019.  hadv_flx5
// :1335-1476, vadv_flx5 :3269-3388 (+ vadv_lwr_flx_23 :4817-4995,
// vadv_lwr_flx_35 :5001-5169).  advs: src/adv.F:14-903, advu :911-1839,
// advv :1847-2770, advw :2778-3610; pdef: src/pdef.F:15-581.
//
// Every flux is evaluated with the reference's operation order in FP32
// without contraction, so results are bitwise those of the reference.
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

// ------------------------------------------------------------ horizontal
// stag 1: scalar; 2: u; 3: v; 4: w.  Writes dumx/dumy over the reference
// loop ranges.
template <bool WENO>
static void hadv5t(cm1g_model *m, int stag, const float *rru, const float *rrv, float *dumx,
                   float *dumy, const float *a, double weps) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (stag == 1) {
    for3(m->stream, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      auto S = [&](int ii, int jj) { return a[IXD(P, ii, jj, k)]; };
      float ru = rru[IXD(P, i, j, k)], rv = rrv[IXD(P, i, j, k)];
      dumx[IXD(P, i, j, k)] = (ru >= 0.0f)
          ? ru * face5<WENO>(S(i - 3, j), S(i - 2, j), S(i - 1, j), S(i, j), S(i + 1, j), weps)
          : ru * face5<WENO>(S(i + 2, j), S(i + 1, j), S(i, j), S(i - 1, j), S(i - 2, j), weps);
      dumy[IXD(P, i, j, k)] = (rv >= 0.0f)
          ? rv * face5<WENO>(S(i, j - 3), S(i, j - 2), S(i, j - 1), S(i, j), S(i, j + 1), weps)
          : rv * face5<WENO>(S(i, j + 2), S(i, j + 1), S(i, j), S(i, j - 1), S(i, j - 2), weps);
    });
  } else if (stag == 2) {
    int i1 = P.ibw == 1 ? 2 : 1, i2 = P.ibe == 1 ? ni : ni + 1;
    for3(m->stream, i1 - 1, i2, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      auto S = [&](int ii, int jj) { return a[IXD(P, ii, jj, k)]; };
      float ubar = 0.5f * (rru[IXD(P, i, j, k)] + rru[IXD(P, i + 1, j, k)]);
      dumx[IXD(P, i, j, k)] = (ubar >= 0.0f)
          ? ubar * face5<WENO>(S(i - 2, j), S(i - 1, j), S(i, j), S(i + 1, j), S(i + 2, j), weps)
          : ubar * face5<WENO>(S(i + 3, j), S(i + 2, j), S(i + 1, j), S(i, j), S(i - 1, j), weps);
      float vbar = 0.5f * (rrv[IXD(P, i, j, k)] + rrv[IXD(P, i - 1, j, k)]);
      dumy[IXD(P, i, j, k)] = (vbar >= 0.0f)
          ? vbar * face5<WENO>(S(i, j - 3), S(i, j - 2), S(i, j - 1), S(i, j), S(i, j + 1), weps)
          : vbar * face5<WENO>(S(i, j + 2), S(i, j + 1), S(i, j), S(i, j - 1), S(i, j - 2), weps);
    });
  } else if (stag == 3) {
    int j1 = P.ibs == 1 ? 2 : 1, j2 = P.ibn == 1 ? nj : nj + 1;
    for3(m->stream, 1, ni + 1, j1 - 1, j2, 1, nk, [=] __device__(int i, int j, int k) {
      auto S = [&](int ii, int jj) { return a[IXD(P, ii, jj, k)]; };
      float ubar = 0.5f * (rru[IXD(P, i, j, k)] + rru[IXD(P, i, j - 1, k)]);
      dumx[IXD(P, i, j, k)] = (ubar >= 0.0f)
          ? ubar * face5<WENO>(S(i - 3, j), S(i - 2, j), S(i - 1, j), S(i, j), S(i + 1, j), weps)
          : ubar * face5<WENO>(S(i + 2, j), S(i + 1, j), S(i, j), S(i - 1, j), S(i - 2, j), weps);
      float vbar = 0.5f * (rrv[IXD(P, i, j, k)] + rrv[IXD(P, i, j + 1, k)]);
      dumy[IXD(P, i, j, k)] = (vbar >= 0.0f)
          ? vbar * face5<WENO>(S(i, j - 2), S(i, j - 1), S(i, j), S(i, j + 1), S(i, j + 2), weps)
          : vbar * face5<WENO>(S(i, j + 3), S(i, j + 2), S(i, j + 1), S(i, j), S(i, j - 1), weps);
    });
  } else {
    for3(m->stream, 1, ni + 1, 1, nj + 1, 2, nk, [=] __device__(int i, int j, int k) {
      auto S = [&](int ii, int jj) { return a[IXD(P, ii, jj, k)]; };
      float cc2 = 0.5f * (P.c2[k] + P.c2[k]);
      float cc1 = 1.0f - cc2;
      float ubar = cc2 * rru[IXD(P, i, j, k)] + cc1 * rru[IXD(P, i, j, k - 1)];
      dumx[IXD(P, i, j, k)] = (ubar >= 0.0f)
          ? ubar * face5<WENO>(S(i - 3, j), S(i - 2, j), S(i - 1, j), S(i, j), S(i + 1, j), weps)
          : ubar * face5<WENO>(S(i + 2, j), S(i + 1, j), S(i, j), S(i - 1, j), S(i - 2, j), weps);
      float vbar = cc2 * rrv[IXD(P, i, j, k)] + cc1 * rrv[IXD(P, i, j, k - 1)];
      dumy[IXD(P, i, j, k)] = (vbar >= 0.0f)
          ? vbar * face5<WENO>(S(i, j - 3), S(i, j - 2), S(i, j - 1), S(i, j), S(i, j + 1), weps)
          : vbar * face5<WENO>(S(i, j + 2), S(i, j + 1), S(i, j), S(i, j - 1), S(i, j - 2), weps);
    });
  }
}

// hadv_flx5 (src/adv_routines.F:1335-1476) or hadv_weno5 (:204-430, no
// pdef test: the momentum equations)
static void hadv5(cm1g_model *m, int stag, const float *rru, const float *rrv, float *dumx,
                  float *dumy, const float *a, bool weno = false, double weps = 0.0) {
  if (weno) hadv5t<true>(m, stag, rru, rrv, dumx, dumy, a, weps);
  else hadv5t<false>(m, stag, rru, rrv, dumx, dumy, a, weps);
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
static void vadv5(cm1g_model *m, int stag, const float *rrw, float *dumz, const float *a,
                  bool weno = false, double weps = 0.0) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  int ilo = 1, ihi = ni, jlo = 1, jhi = nj, klo = 2, khi = nk;
  if (stag == 2) { ilo = P.ibw == 1 ? 2 : 1; ihi = P.ibe == 1 ? ni : ni + 1; }
  if (stag == 3) { jlo = P.ibs == 1 ? 2 : 1; jhi = P.ibn == 1 ? nj : nj + 1; }
  if (stag == 4) { klo = 1; khi = nk; }
  if (weno)
    for3(m->stream, ilo, ihi, jlo, jhi, klo, khi, [=] __device__(int i, int j, int k) {
      auto S = [&](int kk) { return a[IXD(P, i, j, kk)]; };
      dumz[IXD(P, i, j, k)] = vfluxw_of(P, stag, wface(P, stag, rrw, i, j, k), S, k, 0, weps);
    });
  else
    for3(m->stream, ilo, ihi, jlo, jhi, klo, khi, [=] __device__(int i, int j, int k) {
      dumz[IXD(P, i, j, k)] = vflux5(P, stag, rrw, a, i, j, k);
    });
}

// ------------------------------------------------------------------ pdef
// One direction of the positive-definite correction (pdefx1+pdefx2,
// pdefy1+pdefy2 or pdefz).  The reference scatters mass to i-1,i,i+1
// inside a sequential loop; here each cell gathers the same terms in the
// same order (left neighbour, itself, right neighbour), which reproduces
// the sequential sums bit for bit.
static void pdef_dir(cm1g_model *m, int dir, const float *vel, const float *adv, float *advout,
                     float *dum, const float *s0, const float *s, float dt) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  // pdefx1 / pdefy1 / first part of pdefz: dum
  for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    int c = IXD(P, i, j, k);
    float div;
    if (dir == 0) div = (vel[IXD(P, i + 1, j, k)] - vel[c]) * P.rdx * P.uh[i + 2];
    else if (dir == 1) div = (vel[IXD(P, i, j + 1, k)] - vel[c]) * P.rdy * P.vh[j + 2];
    else div = (vel[IXD(P, i, j, k + 1)] - vel[c]) * P.rdz * P.mh[k];
    float d = P.rho0[k] * s0[c] + dt * (adv[c] + s[c] * div);
    if (dir == 0) {
      if (P.wbc == 2 && P.ibw == 1 && i == 1 && vel[IXD(P, 1, j, k)] >= 0.0f) d = P.rho0[k] * s0[c];
      if (P.ebc == 2 && P.ibe == 1 && i == ni && vel[IXD(P, ni + 1, j, k)] <= 0.0f) d = P.rho0[k] * s0[c];
    } else if (dir == 1) {
      if (P.sbc == 2 && P.ibs == 1 && j == 1 && vel[IXD(P, i, 1, k)] >= 0.0f) d = P.rho0[k] * s0[c];
      if (P.nbc == 2 && P.ibn == 1 && j == nj && vel[IXD(P, i, nj + 1, k)] <= 0.0f) d = P.rho0[k] * s0[c];
    }
    dum[c] = d;
  });
  if (dir != 2) {
    bc_fill(m, BC_S, dum);
  } else {
    for2(m->stream, 1, ni, 1, nj, [=] __device__(int i, int j) {
      dum[IXD(P, i, j, 0)] = 0.0f;
      dum[IXD(P, i, j, nk + 1)] = 0.0f;
    });
  }
  // second part: gather the mass exchanged with the two neighbours
  const float rdt = 1.0f / dt;
  for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    const float smeps = P.K.smeps;
    auto D = [&](int o) -> float {
      if (dir == 0) return dum[IXD(P, i + o, j, k)];
      if (dir == 1) return dum[IXD(P, i, j + o, k)];
      return dum[IXD(P, i, j, k + o)];
    };
    // the reference visits cells 0..ni+1 (x, y) or 1..nk (z)
    auto visited = [&](int o) -> bool {
      if (dir == 0) { int ii = i + o; return ii >= 0 && ii <= ni + 1; }
      if (dir == 1) { int jj = j + o; return jj >= 0 && jj <= nj + 1; }
      int kk = k + o; return kk >= 1 && kk <= nk;
    };
    float mass = 0.0f;
    bool flag = false;
    float d0 = D(0);
    // contribution from the left neighbour (cell -1 scatters foo2*foo3 to us)
    if (visited(-1)) {
      float dm = D(-1);
      if (dm < 0.0f) {
        float foo1 = fmaxf(0.0f, D(-2)), foo2 = fmaxf(0.0f, d0);
        if (foo1 + foo2 > smeps) {
          float foo3 = fmaxf(dm, -(foo1 + foo2)) / (foo1 + foo2);
          mass = mass + foo2 * foo3;
          if (d0 > smeps) flag = true;
        }
      }
    }
    if (d0 < 0.0f) {
      float foo1 = fmaxf(0.0f, D(-1)), foo2 = fmaxf(0.0f, D(1));
      if (foo1 + foo2 > smeps) {
        float foo3 = fmaxf(d0, -(foo1 + foo2)) / (foo1 + foo2);
        mass = mass - (foo1 + foo2) * foo3;
        flag = true;
      }
    }
    if (visited(1)) {
      float dp = D(1);
      if (dp < 0.0f) {
        float foo1 = fmaxf(0.0f, d0), foo2 = fmaxf(0.0f, D(2));
        if (foo1 + foo2 > smeps) {
          float foo3 = fmaxf(dp, -(foo1 + foo2)) / (foo1 + foo2);
          mass = mass + foo1 * foo3;
          if (d0 > smeps) flag = true;
        }
      }
    }
    int c = IXD(P, i, j, k);
    float out = adv[c];
    if (flag) {
      float div;
      if (dir == 0) div = (vel[IXD(P, i + 1, j, k)] - vel[c]) * P.rdx * P.uh[i + 2];
      else if (dir == 1) div = (vel[IXD(P, i, j + 1, k)] - vel[c]) * P.rdy * P.vh[j + 2];
      else div = (vel[IXD(P, i, j, k + 1)] - vel[c]) * P.rdz * P.mh[k];
      out = (d0 + mass - P.rho0[k] * s0[c]) * rdt - s[c] * div;
    }
    advout[c] = out;
  });
}

// ------------------------------------------------------ fused advs column
// Upwind 5th-order face value for the face between cells m-1 and m along one
// horizontal line, from the seven values S(m-3..m+3) (only six are used).
template <bool WENO>
__device__ __forceinline__ float hface5(float r, float sm3, float sm2, float sm1, float s0, float sp1,
                                        float sp2, int pdef, double weps) {
  if (!WENO) return (r >= 0.0f) ? r * flx5(sm3, sm2, sm1, s0, sp1) : r * flx5(sp2, sp1, s0, sm1, sm2);
  float s1, s2, s3, s4, s5;
  if (r >= 0.0f) { s1 = sm3; s2 = sm2; s3 = sm1; s4 = s0; s5 = sp1; }
  else { s1 = sp2; s2 = sp1; s3 = s0; s4 = sm1; s5 = sm2; }
  if (pdef == 1 && !weno_doit(s1, s2, s3, s4, s5, weps)) return 0.0f;
  return r * weno5(s1, s2, s3, s4, s5, weps);
}

struct AdvsArgs {
  const float *s;                 // advected scalar
  float sadd;                     // nonzero: advect (th0(k)-th0r) + s
  const float *s0, *sten0, *t11, *thterm, *rru, *rrv, *rrw, *divx;
  float *out_sten, *out_new;      // tendency and/or s0 + dt*tendency
  float *advx, *advy, *advz;      // mode 1: the three advective tendencies
  float dt;
  int smeps_flush, mode;          // mode 0: final tendency; 1: advx/advy/advz only
  int pdef, pdefweno;             // WENO5 magnitude tests (horizontal / vertical)
  double weps;
};

// One column (i,j) of advs (src/adv.F:14-658 with hadv_flx5/vadv_flx5, or
// hadv_weno5/vadv_weno5 when WENO):
// face fluxes at i, i+1, j, j+1 with the open-boundary inflow fixes
// (src/adv.F:211-340), the vertical fluxes marched up the column with a
// six-level register window, and the tendency assembly.  Each flux is the
// same FP32 expression as the reference's, so results are bitwise equal
// to separate flux / divergence passes.
template <bool WENO>
__global__ void __launch_bounds__(256) k_advs_col(KP P, AdvsArgs a) {
  const int i = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = 1 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (i > ni || j > nj) return;
  const bool doitw = (P.wbc == 2 && P.ibw == 1), doite = (P.ebc == 2 && P.ibe == 1);
  const bool doits = (P.sbc == 2 && P.ibs == 1), doitn = (P.nbc == 2 && P.ibn == 1);
  const float *s = a.s;
  const bool sadd = (a.sadd != 0.0f);
  const int px = P.px, pxy = P.pxy;
  auto val = [&](int kk) -> float {  // column value at level kk (1..nk)
    if (kk < 1 || kk > nk) return 0.0f;
    const float v = s[IXD(P, i, j, kk)];
    return sadd ? (P.th0[kk] - P.K.th0r) + v : v;
  };
  // window: wv[0..5] = S(k-2 .. k+3)
  float wv0 = val(-1), wv1 = val(0), wv2 = val(1), wv3 = val(2), wv4 = val(3), wv5 = val(4);
  float fz_lo = 0.0f;
  for (int k = 1; k <= nk; k++) {
    const int c = IXD(P, i, j, k);
    const float tem = sadd ? (P.th0[k] - P.K.th0r) : 0.0f;
    auto H = [&](int off) -> float {
      const float v = s[c + off];
      return sadd ? tem + v : v;
    };
    const float sc = wv2;
    // x faces i and i+1
    const float xm3 = H(-3), xm2 = H(-2), xm1 = H(-1), xp1 = H(1), xp2 = H(2), xp3 = H(3);
    const float ru0 = a.rru[c], ru1 = a.rru[c + 1];
    float fx0 = hface5<WENO>(ru0, xm3, xm2, xm1, sc, xp1, xp2, a.pdef, a.weps);
    float fx1 = hface5<WENO>(ru1, xm2, xm1, sc, xp1, xp2, xp3, a.pdef, a.weps);
    const float fx0_raw = fx0;
    if (doitw && i == 1 && ru0 >= 0.0f) fx0 = fx1;
    if (doite && i == ni && ru1 <= 0.0f) fx1 = fx0_raw;
    // y faces j and j+1
    const float ym3 = H(-3 * px), ym2 = H(-2 * px), ym1 = H(-px), yp1 = H(px), yp2 = H(2 * px),
                yp3 = H(3 * px);
    const float rv0 = a.rrv[c], rv1 = a.rrv[c + px];
    float fy0 = hface5<WENO>(rv0, ym3, ym2, ym1, sc, yp1, yp2, a.pdef, a.weps);
    float fy1 = hface5<WENO>(rv1, ym2, ym1, sc, yp1, yp2, yp3, a.pdef, a.weps);
    const float fy0_raw = fy0;
    if (doits && j == 1 && rv0 >= 0.0f) fy0 = fy1;
    if (doitn && j == nj && rv1 <= 0.0f) fy1 = fy0_raw;
    float ax = -(fx1 - fx0) * P.rdx * P.uh[i + 2];
    float ay = -(fy1 - fy0) * P.rdy * P.vh[j + 2];
    if (doitw && i == 1 && ru0 >= 0.0f) ax = ax - sc * (ru1 - ru0) * P.rdx * P.uh[1 + 2];
    if (doite && i == ni && ru1 <= 0.0f) ax = ax - sc * (ru1 - ru0) * P.rdx * P.uh[ni + 2];
    if (doits && j == 1 && rv0 >= 0.0f) ay = ay - sc * (rv1 - rv0) * P.rdy * P.vh[1 + 2];
    if (doitn && j == nj && rv1 <= 0.0f) ay = ay - sc * (rv1 - rv0) * P.rdy * P.vh[nj + 2];
    // vertical face k+1 from S(k-2..k+3)
    float fz_hi = 0.0f;
    if (k + 1 <= nk) {
      const int kf = k + 1;
      auto S = [=](int kk) -> float {
        const int o = kk - kf;  // -3..2
        return o == -3 ? wv0 : o == -2 ? wv1 : o == -1 ? wv2 : o == 0 ? wv3 : o == 1 ? wv4 : wv5;
      };
      fz_hi = WENO ? vfluxw_of(P, 1, a.rrw[c + pxy], S, kf, a.pdefweno, a.weps)
                   : vflux5_of(P, 1, a.rrw[c + pxy], S, kf);
    }
    float az;
    if (k == 1) az = -fz_hi * P.rdz * P.mh[1];
    else if (k == nk) az = +fz_lo * P.rdz * P.mh[nk];
    else az = -(fz_hi - fz_lo) * P.rdz * P.mh[k];
    if (a.mode == 1) {
      a.advx[c] = ax;
      a.advy[c] = ay;
      a.advz[c] = az;
    } else {
      float st = a.sten0[c];
      if (a.t11) st = st - a.t11[c] * a.thterm[c];
      st = st + ((ax + ay) + az + sc * a.divx[c]) * P.rr0[k];
      if (a.out_sten) a.out_sten[c] = st;
      if (a.out_new) {
        float v = a.s0[c] + a.dt * st;
        if (a.smeps_flush && fabsf(v) < P.K.smeps) v = 0.0f;
        a.out_new[c] = v;
      }
    }
    fz_lo = fz_hi;
    wv0 = wv1;
    wv1 = wv2;
    wv2 = wv3;
    wv3 = wv4;
    wv4 = wv5;
    wv5 = val(k + 4);
  }
}

// ------------------------------------------------------------------ advs
// sadd: when nonzero, s(i,j,k) = (th0(k) - th0r) + s_in(i,j,k) is formed on
// the fly (the theta equation's sadv, src/solve2.F:1354-1374).
// sten0/thterm_div: initial tendency; with thterm_div != nullptr the
// initial tendency is thten1 - t11*thterm (src/solve2.F:1320-1330), where
// sten0 = thten1, thterm_div[0] = t11 and thterm_div[1] = thterm.
// out_new, if given, receives s0 + dt*tendency with the smeps flush
// (src/solve2.F:1416-1426, :1518-1528).
void advs_run(cm1g_model *m, int nrk, const float *s_in, float sadd, const float *s0,
              const float *sten0, const float *thterm_div, float *out_sten, float *out_new,
              int pdef, int pdefweno, float dt, double weps, int smeps_flush) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  // WENO5 on every stage (advwenos=1) or the last one (2), src/adv.F:75-78
  const bool weno = (m->c.advwenos == 1) || (m->c.advwenos == 2 && nrk == 3);
  AdvsArgs g;
  g.s = s_in;
  g.sadd = sadd;
  g.s0 = s0;
  g.sten0 = sten0;
  g.t11 = thterm_div;
  g.thterm = thterm_div ? m->f[CM1G_THTERM] : nullptr;
  g.rru = m->rrus;
  g.rrv = m->rrvs;
  g.rrw = m->rrws;
  g.divx = m->f[CM1G_DIVX];
  g.out_sten = out_sten;
  g.out_new = out_new;
  g.advx = m->tx;
  g.advy = m->ty;
  g.advz = m->tz;
  g.dt = dt;
  g.smeps_flush = smeps_flush;
  g.pdef = pdef;
  g.pdefweno = pdefweno;
  g.weps = weps;
  auto *kern = weno ? k_advs_col<true> : k_advs_col<false>;
  const dim3 blk(64, 4, 1), grd((ni + 63) / 64, (nj + 3) / 4, 1);
  // an in-place update (out_new aliasing the advected field) goes through
  // a partner array, since neighbouring columns still read the old values
  const bool inplace = (out_new != nullptr && out_new == s_in);
  if (inplace) g.out_new = m->thb;
  if (pdef != 1) {
    g.mode = 0;
    hipLaunchKernelGGL(kern, grd, blk, 0, m->stream, P, g);
  } else {
    g.mode = 1;
    hipLaunchKernelGGL(kern, grd, blk, 0, m->stream, P, g);
    float *advx = m->tx, *advy = m->ty, *advz = m->tz;
    float *dumx = m->sx, *dumy = m->sy, *dumz = m->sz;
    const float *rru = m->rrus, *rrv = m->rrvs, *rrw = m->rrws;
    const float *a = s_in;
    if (sadd != 0.0f) {
      // the positive-definite pass needs the advected values themselves
      float *sv = m->f[CM1G_SADV];
      for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
        float tem = P.th0[k] - P.K.th0r;
        sv[IXD(P, i, j, k)] = tem + s_in[IXD(P, i, j, k)];
      });
      a = sv;
    }
    // each direction writes its corrected tendency into dumx/dumy/dumz
    pdef_dir(m, 0, rru, advx, dumx, m->scr1, s0, a, dt);
    pdef_dir(m, 2, rrw, advz, dumz, m->scr1, s0, a, dt);
    pdef_dir(m, 1, rrv, advy, dumy, m->scr1, s0, a, dt);
    const float *ax = dumx, *ay = dumy, *az = dumz;
    const float *t11 = thterm_div, *thterm = g.thterm, *divx = g.divx;
    float *onew = g.out_new;
    for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      float st = sten0[c];
      if (t11) st = st - t11[c] * thterm[c];
      st = st + ((ax[c] + ay[c]) + az[c] + a[c] * divx[c]) * P.rr0[k];
      if (out_sten) out_sten[c] = st;
      if (onew) {
        float v = s0[c] + dt * st;
        if (smeps_flush && fabsf(v) < P.K.smeps) v = 0.0f;
        onew[c] = v;
      }
    });
  }
  if (inplace) {
    const float *src = m->thb;
    for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      out_new[c] = src[c];
    });
  }
}

// ----------------------------------------------------------- advu/v/w
// The momentum tendencies use the pre-acoustic mass fluxes rru/rrv/rrw
// (model fields) and divx; uten/vten/wten receive uten1 + advection.
void adv_uvw(cm1g_model *m, int nrk, int mask, int buoy) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float *rru = m->f[CM1G_RRU], *rrv = m->f[CM1G_RRV], *rrw = m->f[CM1G_RRW];
  const float *divx = m->f[CM1G_DIVX];
  float *dumx = m->sx, *dumy = m->sy, *dumz = m->sz;
  // WENO5 on every stage (advwenov=1) or the last one (2), src/adv.F:958-960
  const bool weno = (m->c.advwenov == 1) || (m->c.advwenov == 2 && nrk == 3);
  const double weps = (double)(100.0f * P.K.epsilon);
  const bool doitw = (P.wbc == 2 && P.ibw == 1), doite = (P.ebc == 2 && P.ibe == 1);
  const bool doits = (P.sbc == 2 && P.ibs == 1), doitn = (P.nbc == 2 && P.ibn == 1);
  // ---- u
  if (mask & 1) {
    const float *u3d = m->f[CM1G_U3D];
    float *uten = m->f[CM1G_UTEN];
    int i1 = P.ibw == 1 ? 2 : 1, i2 = P.ibe == 1 ? ni : ni + 1;
    hadv5(m, 2, rru, rrv, dumx, dumy, u3d, weno, weps);
    if (doits || doitn)
      for2(m->stream, i1, i2, 1, nk, [=] __device__(int i, int k) {
        if (doits && (rrv[IXD(P, i, 1, k)] + rrv[IXD(P, i - 1, 1, k)]) >= 0.0f)
          dumy[IXD(P, i, 1, k)] = dumy[IXD(P, i, 2, k)];
        if (doitn && (rrv[IXD(P, i, nj + 1, k)] + rrv[IXD(P, i - 1, nj + 1, k)]) <= 0.0f)
          dumy[IXD(P, i, nj + 1, k)] = dumy[IXD(P, i, nj, k)];
      });
    vadv5(m, 2, rrw, dumz, u3d, weno, weps);
    for3(m->stream, i1, i2, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      float ax = -(dumx[c] - dumx[IXD(P, i - 1, j, k)]) * P.rdx * P.uf[i + 2];
      float ay = -(dumy[IXD(P, i, j + 1, k)] - dumy[c]) * P.rdy * P.vh[j + 2];
      if (doits && j == 1 && (rrv[IXD(P, i, 1, k)] + rrv[IXD(P, i - 1, 1, k)]) >= 0.0f)
        ay = ay - u3d[c] * 0.5f * ((rrv[IXD(P, i - 1, 2, k)] - rrv[IXD(P, i - 1, 1, k)]) +
                                  (rrv[IXD(P, i, 2, k)] - rrv[IXD(P, i, 1, k)])) * P.rdy * P.vh[1 + 2];
      if (doitn && j == nj && (rrv[IXD(P, i, nj + 1, k)] + rrv[IXD(P, i - 1, nj + 1, k)]) <= 0.0f)
        ay = ay - u3d[c] * 0.5f * ((rrv[IXD(P, i - 1, nj + 1, k)] - rrv[IXD(P, i - 1, nj, k)]) +
                                  (rrv[IXD(P, i, nj + 1, k)] - rrv[IXD(P, i, nj, k)])) * P.rdy * P.vh[nj + 2];
      float az;
      if (k == 1) az = -dumz[IXD(P, i, j, 2)] * P.rdz * P.mh[1];
      else if (k == nk) az = +dumz[c] * P.rdz * P.mh[nk];
      else az = -(dumz[IXD(P, i, j, k + 1)] - dumz[c]) * P.rdz * P.mh[k];
      uten[c] = uten[c] + ((ax + ay) + az + u3d[c] * 0.5f * (divx[c] + divx[IXD(P, i - 1, j, k)])) * P.rr0[k];
    });
  }
  // ---- v
  if (mask & 2) {
    const float *v3d = m->f[CM1G_V3D];
    float *vten = m->f[CM1G_VTEN];
    int j1 = P.ibs == 1 ? 2 : 1, j2 = P.ibn == 1 ? nj : nj + 1;
    hadv5(m, 3, rru, rrv, dumx, dumy, v3d, weno, weps);
    if (doitw || doite)
      for2(m->stream, j1, j2, 1, nk, [=] __device__(int j, int k) {
        if (doitw && (rru[IXD(P, 1, j, k)] + rru[IXD(P, 1, j - 1, k)]) >= 0.0f)
          dumx[IXD(P, 1, j, k)] = dumx[IXD(P, 2, j, k)];
        if (doite && (rru[IXD(P, ni + 1, j, k)] + rru[IXD(P, ni + 1, j - 1, k)]) <= 0.0f)
          dumx[IXD(P, ni + 1, j, k)] = dumx[IXD(P, ni, j, k)];
      });
    vadv5(m, 3, rrw, dumz, v3d, weno, weps);
    for3(m->stream, 1, ni, j1, j2, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      float ax = -(dumx[IXD(P, i + 1, j, k)] - dumx[c]) * P.rdx * P.uh[i + 2];
      float ay = -(dumy[c] - dumy[IXD(P, i, j - 1, k)]) * P.rdy * P.vf[j + 2];
      if (doitw && i == 1 && (rru[IXD(P, 1, j, k)] + rru[IXD(P, 1, j - 1, k)]) >= 0.0f)
        ax = ax - v3d[c] * 0.5f * ((rru[IXD(P, 2, j - 1, k)] - rru[IXD(P, 1, j - 1, k)]) +
                                  (rru[IXD(P, 2, j, k)] - rru[IXD(P, 1, j, k)])) * P.rdx * P.uh[1 + 2];
      if (doite && i == ni && (rru[IXD(P, ni + 1, j, k)] + rru[IXD(P, ni + 1, j - 1, k)]) <= 0.0f)
        ax = ax - v3d[c] * 0.5f * ((rru[IXD(P, ni + 1, j - 1, k)] - rru[IXD(P, ni, j - 1, k)]) +
                                  (rru[IXD(P, ni + 1, j, k)] - rru[IXD(P, ni, j, k)])) * P.rdx * P.uh[ni + 2];
      float az;
      if (k == 1) az = -dumz[IXD(P, i, j, 2)] * P.rdz * P.mh[1];
      else if (k == nk) az = +dumz[c] * P.rdz * P.mh[nk];
      else az = -(dumz[IXD(P, i, j, k + 1)] - dumz[c]) * P.rdz * P.mh[k];
      vten[c] = vten[c] + ((ax + ay) + az + v3d[c] * 0.5f * (divx[c] + divx[IXD(P, i, j - 1, k)])) * P.rr0[k];
    });
  }
  // ---- w (+ buoyancy, src/solve2.F:899-907: dum8 holds buoyancy)
  if (mask & 4) {
    const float *w3d = m->f[CM1G_W3D];
    float *wten = m->f[CM1G_WTEN];
    const float *dum8 = m->f[CM1G_DUM8];
    hadv5(m, 4, rru, rrv, dumx, dumy, w3d, weno, weps);
    if (doitw || doite)
      for2(m->stream, 1, nj, 2, nk, [=] __device__(int j, int k) {
        float cc2 = 0.5f * (P.c2[k] + P.c2[k]), cc1 = 1.0f - cc2;
        if (doitw && (cc2 * rru[IXD(P, 1, j, k)] + cc1 * rru[IXD(P, 1, j, k - 1)]) >= 0.0f)
          dumx[IXD(P, 1, j, k)] = dumx[IXD(P, 2, j, k)];
        if (doite && (cc2 * rru[IXD(P, ni + 1, j, k)] + cc1 * rru[IXD(P, ni + 1, j, k - 1)]) <= 0.0f)
          dumx[IXD(P, ni + 1, j, k)] = dumx[IXD(P, ni, j, k)];
      });
    if (doits || doitn)
      for2(m->stream, 1, ni, 2, nk, [=] __device__(int i, int k) {
        float cc2 = 0.5f * (P.c2[k] + P.c2[k]), cc1 = 1.0f - cc2;
        if (doits && (cc2 * rrv[IXD(P, i, 1, k)] + cc1 * rrv[IXD(P, i, 1, k - 1)]) >= 0.0f)
          dumy[IXD(P, i, 1, k)] = dumy[IXD(P, i, 2, k)];
        if (doitn && (cc2 * rrv[IXD(P, i, nj + 1, k)] + cc1 * rrv[IXD(P, i, nj + 1, k - 1)]) <= 0.0f)
          dumy[IXD(P, i, nj + 1, k)] = dumy[IXD(P, i, nj, k)];
      });
    vadv5(m, 4, rrw, dumz, w3d, weno, weps);
    for3(m->stream, 1, ni, 1, nj, 2, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      float cc2 = 0.5f * (P.c2[k] + P.c2[k]), cc1 = 1.0f - cc2;
      float ax = -(dumx[IXD(P, i + 1, j, k)] - dumx[c]) * P.rdx * P.uh[i + 2];
      float ay = -(dumy[IXD(P, i, j + 1, k)] - dumy[c]) * P.rdy * P.vh[j + 2];
      if (doitw && i == 1 && (cc2 * rru[IXD(P, 1, j, k)] + cc1 * rru[IXD(P, 1, j, k - 1)]) >= 0.0f)
        ax = ax - w3d[c] * (P.c1[k] * (rru[IXD(P, 2, j, k - 1)] - rru[IXD(P, 1, j, k - 1)]) +
                            P.c2[k] * (rru[IXD(P, 2, j, k)] - rru[IXD(P, 1, j, k)])) * P.rdx * P.uh[1 + 2];
      if (doite && i == ni && (cc2 * rru[IXD(P, ni + 1, j, k)] + cc1 * rru[IXD(P, ni + 1, j, k - 1)]) <= 0.0f)
        ax = ax - w3d[c] * (P.c1[k] * (rru[IXD(P, ni + 1, j, k - 1)] - rru[IXD(P, ni, j, k - 1)]) +
                            P.c2[k] * (rru[IXD(P, ni + 1, j, k)] - rru[IXD(P, ni, j, k)])) * P.rdx * P.uh[ni + 2];
      if (doits && j == 1 && (cc2 * rrv[IXD(P, i, 1, k)] + cc1 * rrv[IXD(P, i, 1, k - 1)]) >= 0.0f)
        ay = ay - w3d[c] * (P.c1[k] * (rrv[IXD(P, i, 2, k - 1)] - rrv[IXD(P, i, 1, k - 1)]) +
                            P.c2[k] * (rrv[IXD(P, i, 2, k)] - rrv[IXD(P, i, 1, k)])) * P.rdy * P.vh[1 + 2];
      if (doitn && j == nj && (cc2 * rrv[IXD(P, i, nj + 1, k)] + cc1 * rrv[IXD(P, i, nj + 1, k - 1)]) <= 0.0f)
        ay = ay - w3d[c] * (P.c1[k] * (rrv[IXD(P, i, nj + 1, k - 1)] - rrv[IXD(P, i, nj, k - 1)]) +
                            P.c2[k] * (rrv[IXD(P, i, nj + 1, k)] - rrv[IXD(P, i, nj, k)])) * P.rdy * P.vh[nj + 2];
      float az = -(dumz[c] - dumz[IXD(P, i, j, k - 1)]) * P.rdz * P.mf[k];
      float t = wten[c] + ((ax + ay) + az + w3d[c] * (P.c2[k] * divx[c] + P.c1[k] * divx[IXD(P, i, j, k - 1)])) * P.rrf0[k];
      wten[c] = buoy ? t + (P.c1[k] * dum8[IXD(P, i, j, k - 1)] + P.c2[k] * dum8[c]) : t;
    });
  }
}
