/*
 * CPU restatement (oracle) of CM1's LES subgrid turbulence for the branch
 * set the hot path supports: cm1setup = 1, sgsmodel = 1 (Deardorff TKE),
 * tconfig = 1, bcturbs = 1, horizturb = 0, doimpl = 0 or 1, bbc = 3 with
 * sfcmodel = 1 (set_znt = 1: neutral log-law friction velocity) and
 * set_flx = 1 (specified surface heat flux), tbc = 1, dry, Cartesian, no
 * terrain, doubly periodic.
 *
 * TEST INFRASTRUCTURE ONLY: the checker the HIP path is compared against.
 *
 * Routines, each following the reference statement by statement:
 *   o_sfc_and_turb  src/turb.F:12-1700 (the parts this branch set runs):
 *                   u1/v1/s1 and getcecd's ust (src/sfcphys.F:18-110),
 *                   sfcflux's specified flux (:313-330), bc2d, calcnm
 *                   (src/turb.F:5432-5470), calcdef (:4711-5430), tkekm
 *                   (:2977-3120), tkebc (:3219-3300), gettau and sfcstress
 *                   (:4059-4700)
 *   o_turb_tend     src/solve1.F:315-380 (tketen, turbt) and :1118-1240
 *                   (turbu, turbv, turbw, turbs for theta and tracers),
 *                   src/turbtend.F
 *   o_tke_step      src/solve2.F:1709-1800 (TKE advection with advw, step)
 *   o_tke_copy      src/solve3.F:563-575
 */
#include "cm1o_internal.h"
#include <stdlib.h>
#include <string.h>

#define KARMAN 0.40f
#define C_M 0.10f
#define C_L 0.8165f
#define RI_C 0.25f
#define TKE_MIN 1.0e-10f
#define NM_MIN 1.0e-6f
#define SMALL_LEN_FRAC 0.001f
#define VIALPHA 1.0f
#define VIBETA (1.0f - VIALPHA)

/* single-precision natural log of the host C library (the reference's REAL
 * alog), or its double-precision value rounded once in libm mode 1 */
static inline float o_logf(float x) { return cm1o_libm_mode ? (float)log((double)x) : logf(x); }

/* periodic fill of a 2-D (i,j) array over the ngxy = 3 halo: src/bc.F:12-172
 * (bc2d, wbc = ebc = sbc = nbc = 1) */
void o_bc2d(cm1o_model *M, float *s) {
  const int ni = M->ni, nj = M->nj;
  int i, j;
  for (j = -2; j <= nj + 3; j++)
    for (i = 1; i <= 3; i++) {
      s[IX2(M, 1 - i, j)] = s[IX2(M, ni + 1 - i, j)];
      s[IX2(M, ni + i, j)] = s[IX2(M, i, j)];
    }
  for (j = 1; j <= 3; j++)
    for (i = -2; i <= ni + 3; i++) {
      s[IX2(M, i, 1 - j)] = s[IX2(M, i, nj + 1 - j)];
      s[IX2(M, i, nj + j)] = s[IX2(M, i, j)];
    }
}

/* the turbulence constants of one LES column (src/init3d.F:2672-2680):
 * cme = cm0 = c_m, ce1, ce2 */
static void tke_consts(float *cme, float *ce1, float *ce2) {
  const float pi = 3.1415926535897932384626433f;
  *cme = C_M;
  *ce1 = *cme * C_L * C_L * (1.0f / RI_C - 1.0f);
  *ce2 = fmaxf(0.0f, *cme * pi * pi - *ce1);
}

void o_sfc_and_turb(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz;
  const float *mh = M->kf[K_MH], *mf = M->kf[K_MF], *rmf = M->kf[K_RMF], *c1 = M->kf[K_C1],
              *c2 = M->kf[K_C2], *th0 = M->kf[K_TH0], *zh = M->kf[K_ZH];
  float *ua = M->f[F_UA], *va = M->f[F_VA], *wa = M->f[F_WA], *tha = M->f[F_THA];
  float *rho = M->f[F_RHO], *rf = M->f[F_RF];
  float *t11 = M->f[F_T11], *t12 = M->f[F_T12], *t13 = M->f[F_T13], *t22 = M->f[F_T22],
        *t23 = M->f[F_T23], *t33 = M->f[F_T33];
  float *nm = M->f[F_NM], *defv = M->f[F_DEFV], *defh = M->f[F_DEFH], *diss = M->f[F_DISSTEN];
  float *kmh = M->f[F_KMH], *kmv = M->f[F_KMV], *khh = M->f[F_KHH], *khv = M->f[F_KHV];
  float *tkea = M->f[F_TKEA], *tke3d = M->f[F_TKE3D], *lenscl = M->f[F_LENSCL];
  float *ust = M->d2[D_UST], *u1 = M->d2[D_U1], *v1 = M->d2[D_V1], *s1 = M->d2[D_S1],
        *thflux = M->d2[D_THFLUX], *znt = M->d2[D_ZNT];
  int i, j, k;
  float cme, ce1, ce2;
  tke_consts(&cme, &ce1, &ce2);

  for (k = 1; k <= nk + 1; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) A3(diss, i, j, k) = 0.0f;

  /* bbc = 3: lowest-level wind at scalar points, getcecd (set_znt = 1),
   * sfcflux (set_flx = 1) */
  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) {
      const size_t c = IX2(M, i, j);
      u1[c] = 0.5f * (A3(ua, i, j, 1) + A3(ua, i + 1, j, 1));
      v1[c] = 0.5f * (A3(va, i, j, 1) + A3(va, i, j + 1, 1));
      s1[c] = sqrtf(u1[c] * u1[c] + v1[c] * v1[c]);
      const float za = zh[1];
      const float rznt = 1.0f / znt[c];
      ust[c] = fmaxf(s1[c] * KARMAN / o_logf(za * rznt), 1.0e-6f);
      thflux[c] = M->c.cnst_shflx;
    }
  o_bc2d(M, ust);
  o_bc2d(M, u1);
  o_bc2d(M, v1);
  o_bc2d(M, s1);
  o_bc2d(M, znt);

  /* calcnm, dry */
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(nm, i, j, k) = o_logf((th0[k] + A3(tha, i, j, k)) / (th0[k - 1] + A3(tha, i, j, k - 1))) * K->g * rdz * mf[k];
  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) { A3(nm, i, j, 1) = 0.0f; A3(nm, i, j, nk + 1) = 0.0f; }

  /* calcdef: rho * S_ij into t11..t33, then defv, defh */
  for (k = 1; k <= nk; k++) {
    const float temz = rdz * mh[k];
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) {
        A3(t11, i, j, k) = A3(rho, i, j, k) * (A3(ua, i + 1, j, k) - A3(ua, i, j, k)) * rdx * UH(i);
        A3(t22, i, j, k) = A3(rho, i, j, k) * (A3(va, i, j + 1, k) - A3(va, i, j, k)) * rdy * VH(j);
        A3(t33, i, j, k) = A3(rho, i, j, k) * (A3(wa, i, j, k + 1) - A3(wa, i, j, k)) * temz;
        A3(t12, i, j, k) = 0.5f * ((A3(ua, i, j, k) - A3(ua, i, j - 1, k)) * rdy * VF(j) +
                                   (A3(va, i, j, k) - A3(va, i - 1, j, k)) * rdx * UF(i)) *
                           0.25f * ((A3(rho, i - 1, j - 1, k) + A3(rho, i, j, k)) + (A3(rho, i - 1, j, k) + A3(rho, i, j - 1, k)));
      }
  }
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni + 1; i++) {
        A3(t13, i, j, k) = 0.5f * ((A3(wa, i, j, k) - A3(wa, i - 1, j, k)) * rdx * UF(i) +
                                   (A3(ua, i, j, k) - A3(ua, i, j, k - 1)) * rdz * 0.5f * (mf[k] + mf[k])) *
                           0.5f * (A3(rf, i - 1, j, k) + A3(rf, i, j, k));
        A3(t23, i, j, k) = 0.5f * ((A3(wa, i, j, k) - A3(wa, i, j - 1, k)) * rdy * VF(j) +
                                   (A3(va, i, j, k) - A3(va, i, j, k - 1)) * rdz * 0.5f * (mf[k] + mf[k])) *
                           0.5f * (A3(rf, i, j - 1, k) + A3(rf, i, j, k));
      }
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
#define SQ(x) ((x) * (x))
        const float tmp11 = (c1[k] * SQ(A3(t11, i, j, k - 1)) + c2[k] * SQ(A3(t11, i, j, k)));
        const float tmp22 = (c1[k] * SQ(A3(t22, i, j, k - 1)) + c2[k] * SQ(A3(t22, i, j, k)));
        const float tmp33 = (c1[k] * SQ(A3(t33, i, j, k - 1)) + c2[k] * SQ(A3(t33, i, j, k)));
        const float tmp12 =
            0.25f * (c1[k] * ((SQ(A3(t12, i, j, k - 1)) + SQ(A3(t12, i + 1, j + 1, k - 1))) +
                              (SQ(A3(t12, i, j + 1, k - 1)) + SQ(A3(t12, i + 1, j, k - 1)))) +
                     c2[k] * ((SQ(A3(t12, i, j, k)) + SQ(A3(t12, i + 1, j + 1, k))) +
                              (SQ(A3(t12, i, j + 1, k)) + SQ(A3(t12, i + 1, j, k)))));
        const float tmp13 = 0.5f * (SQ(A3(t13, i, j, k)) + SQ(A3(t13, i + 1, j, k)));
        const float tmp23 = 0.5f * (SQ(A3(t23, i, j, k)) + SQ(A3(t23, i, j + 1, k)));
        const float rrf = 1.0f / SQ(A3(rf, i, j, k));
#undef SQ
        A3(defv, i, j, k) = 4.0f * (tmp13 + tmp23) * rrf;
        A3(defh, i, j, k) = (2.0f * ((tmp11 + tmp22) + tmp33) + 4.0f * tmp12) * rrf;
      }

  /* tkekm, tconfig = 1 */
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float grdscl = o_powf(((M->c.dx * RUH(i)) * (M->c.dy * RVH(j))) * (M->c.dz * rmf[k]), 0.33333333f);
        const float rgrdscl = 1.0f / grdscl;
        float tk, len;
        if (A3(tkea, i, j, k) <= TKE_MIN) {
          tk = TKE_MIN;
          len = SMALL_LEN_FRAC * grdscl;
        } else {
          tk = A3(tkea, i, j, k);
          len = grdscl;
          if (A3(nm, i, j, k) > NM_MIN) {
            len = C_L * sqrtf(tk / A3(nm, i, j, k));
            len = fminf(len, grdscl);
            len = fmaxf(len, SMALL_LEN_FRAC * grdscl);
          }
        }
        A3(lenscl, i, j, k) = len;
        A3(diss, i, j, k) = A3(diss, i, j, k) + (ce1 + ce2 * len * rgrdscl) * tk * sqrtf(tk) / len;
        const float km = cme * sqrtf(tk) * len;
        float prinv = 3.00f;
        if (A3(nm, i, j, k) > NM_MIN) prinv = fminf(1.0f + 2.00f * len * rgrdscl, 3.00f);
        A3(kmh, i, j, k) = km;
        A3(kmv, i, j, k) = km;
        A3(khh, i, j, k) = km * prinv;
        A3(khv, i, j, k) = km * prinv;
      }
  o_bcw(M, kmh, 1);
  o_bcw(M, kmv, 1);
  o_bcw(M, khh, 1);
  o_bcw(M, khv, 1);

  /* tkebc, bbc = 3 (zntmp = znt) */
  {
    float *tk1 = M->s1, *km1 = M->s2;  /* 2-D slices for bc2d */
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(tkea, i, j, 1) = 0.0f;
        A3(kmh, i, j, 1) = KARMAN * znt[IX2(M, i, j)] * ust[IX2(M, i, j)];
      }
    for (j = -2; j <= nj + 4; j++)
      for (i = -2; i <= ni + 4; i++) {
        tk1[IX2(M, i, j)] = A3(tkea, i, j, 1);
        km1[IX2(M, i, j)] = A3(kmh, i, j, 1);
      }
    o_bc2d(M, tk1);
    o_bc2d(M, km1);
    for (j = -2; j <= nj + 3; j++)
      for (i = -2; i <= ni + 3; i++) {
        A3(tkea, i, j, 1) = tk1[IX2(M, i, j)];
        A3(kmh, i, j, 1) = km1[IX2(M, i, j)];
      }
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) {
        A3(kmv, i, j, 1) = A3(kmh, i, j, 1);
        A3(khv, i, j, 1) = A3(kmv, i, j, 1) * (A3(khv, i, j, 2) / (1.0e-10f + A3(kmv, i, j, 2)));
        A3(khh, i, j, 1) = A3(kmh, i, j, 1) * (A3(khh, i, j, 2) / (1.0e-10f + A3(kmh, i, j, 2)));
      }
  }
  for (j = 0; j <= nj + 1; j++)
    for (i = 0; i <= ni + 1; i++) {
      A3(tke3d, i, j, 1) = A3(tkea, i, j, 1);
      A3(tke3d, i, j, nk + 1) = A3(tkea, i, j, nk + 1);
    }

  /* gettau: t_ij = 2 rho K S_ij; sfcstress (bbc = 3); tbc = 1 */
  for (k = 1; k <= nk; k++)
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) {
        const float tem = (A3(kmh, i, j, k) + A3(kmh, i, j, k + 1));
        A3(t11, i, j, k) = A3(t11, i, j, k) * tem;
        A3(t22, i, j, k) = A3(t22, i, j, k) * tem;
        A3(t33, i, j, k) = A3(t33, i, j, k) * tem;
        A3(t12, i, j, k) = A3(t12, i, j, k) * 0.25f *
                           (((A3(kmh, i - 1, j - 1, k) + A3(kmh, i, j, k)) + (A3(kmh, i - 1, j, k) + A3(kmh, i, j - 1, k))) +
                            ((A3(kmh, i - 1, j - 1, k + 1) + A3(kmh, i, j, k + 1)) +
                             (A3(kmh, i - 1, j, k + 1) + A3(kmh, i, j - 1, k + 1))));
      }
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni + 1; i++) {
        A3(t13, i, j, k) = A3(t13, i, j, k) * (A3(kmv, i - 1, j, k) + A3(kmv, i, j, k));
        A3(t23, i, j, k) = A3(t23, i, j, k) * (A3(kmv, i, j - 1, k) + A3(kmv, i, j, k));
      }
  for (j = 1; j <= nj + 1; j++)
    for (i = 1; i <= ni + 1; i++) {
      const size_t cw = IX2(M, i - 1, j), c = IX2(M, i, j), cs = IX2(M, i, j - 1);
      A3(t13, i, j, 1) = 0.25f * ((ust[cw] * ust[cw]) * (u1[cw] / fmaxf(s1[cw], 0.01f)) +
                                  (ust[c] * ust[c]) * (u1[c] / fmaxf(s1[c], 0.01f))) *
                         (A3(rf, i - 1, j, 1) + A3(rf, i, j, 1));
      A3(t23, i, j, 1) = 0.25f * ((ust[cs] * ust[cs]) * (v1[cs] / fmaxf(s1[cs], 0.01f)) +
                                  (ust[c] * ust[c]) * (v1[c] / fmaxf(s1[c], 0.01f))) *
                         (A3(rf, i, j - 1, 1) + A3(rf, i, j, 1));
    }
  for (j = 1; j <= nj + 1; j++)
    for (i = 1; i <= ni + 1; i++) {
      A3(t13, i, j, nk + 1) = A3(t13, i, j, nk);
      A3(t23, i, j, nk + 1) = A3(t23, i, j, nk);
    }
}

/* turbs (src/turbtend.F:18-440) with turbsz (:442-698): sten += the
 * subgrid flux divergence of s; iflux = 1 puts the surface flux sflux at
 * the lowest face */
static void o_turbs(cm1o_model *M, int iflux, const float *sflux, const float *s, float *sten) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz, dt = M->c.dt;
  const float *mh = M->kf[K_MH], *mf = M->kf[K_MF];
  const float *rho = M->f[F_RHO], *rr = M->f[F_RR], *rf = M->f[F_RF];
  const float *khh = M->f[F_KHH], *khv = M->f[F_KHV];
  float *dumx = M->s3, *dumy = M->s4, *dumz = M->s5, *turbx = M->s6, *turby = M->s7, *turbz = M->s8;
  float *d1 = M->f[F_DUM1], *d2 = M->f[F_DUM2];
  const int dosfcflx = (M->c.isfcflx == 1);
  int i, j, k;
  for (k = 1; k <= nk; k++) {
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni + 1; i++) {
        A3(dumx, i, j, k) = -0.125f * (A3(rho, i, j, k) + A3(rho, i - 1, j, k)) *
                            ((A3(khh, i, j, k) + A3(khh, i - 1, j, k)) + (A3(khh, i, j, k + 1) + A3(khh, i - 1, j, k + 1))) *
                            (A3(s, i, j, k) - A3(s, i - 1, j, k)) * rdx * UF(i);
        A3(dumy, i, j, k) = -0.125f * (A3(rho, i, j, k) + A3(rho, i, j - 1, k)) *
                            ((A3(khh, i, j, k) + A3(khh, i, j - 1, k)) + (A3(khh, i, j, k + 1) + A3(khh, i, j - 1, k + 1))) *
                            (A3(s, i, j, k) - A3(s, i, j - 1, k)) * rdy * VF(j);
      }
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(turbx, i, j, k) = -(A3(dumx, i + 1, j, k) - A3(dumx, i, j, k)) * rdx * UH(i);
        A3(turby, i, j, k) = -(A3(dumy, i, j + 1, k) - A3(dumy, i, j, k)) * rdy * VH(j);
      }
  }
  if (M->c.doimpl == 0) {
    for (j = 1; j <= nj; j++) {
      for (k = 2; k <= nk; k++)
        for (i = 1; i <= ni; i++)
          A3(dumz, i, j, k) = -A3(khv, i, j, k) * (A3(s, i, j, k) - A3(s, i, j, k - 1)) * rdz * mf[k] * A3(rf, i, j, k);
      for (i = 1; i <= ni; i++) { A3(dumz, i, j, 1) = 0.0f; A3(dumz, i, j, nk + 1) = 0.0f; }
      if (iflux == 1 && dosfcflx)
        for (i = 1; i <= ni; i++) A3(dumz, i, j, 1) = sflux[IX2(M, i, j)] * A3(rf, i, j, 1);
      for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++) A3(turbz, i, j, k) = -(A3(dumz, i, j, k + 1) - A3(dumz, i, j, k)) * rdz * mh[k];
    }
  } else {
    const float rdt = 1.0f / dt;
    const float tema = -1.0f * dt * VIALPHA * rdz * rdz;
    const float temb = dt * VIBETA * rdz * rdz;
    const float temc = dt * rdz;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) { A3(dumz, i, j, 1) = 0.0f; A3(dumz, i, j, nk + 1) = 0.0f; }
    if (iflux == 1 && dosfcflx)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(dumz, i, j, 1) = sflux[IX2(M, i, j)] * A3(rf, i, j, 1);
    /* dum7(k) = khv(k) mf(k) rf(k) mh(k) rr(k), dum8(k) = khv(k+1) mf(k+1) rf(k+1) mh(k) rr(k)
     * (src/solve1.F:1169-1182) */
#define D7(i, j, k) (A3(khv, i, j, k) * mf[k] * A3(rf, i, j, k) * mh[k] * A3(rr, i, j, k))
#define D8(i, j, k) (A3(khv, i, j, (k) + 1) * mf[(k) + 1] * A3(rf, i, j, (k) + 1) * mh[k] * A3(rr, i, j, k))
    k = 1;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float r2 = D8(i, j, k);
        const float cfc = tema * r2;
        const float cfb = 1.0f - cfc;
        const float cfd = A3(s, i, j, k) + temb * (-r2 * A3(s, i, j, k) + r2 * A3(s, i, j, k + 1)) +
                          temc * A3(dumz, i, j, 1) * mh[1] * A3(rr, i, j, 1);
        const float tem = 1.0f / cfb;
        A3(d1, i, j, 1) = -cfc * tem;
        A3(d2, i, j, 1) = cfd * tem;
      }
    for (k = 2; k <= nk - 1; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          const float r1 = D7(i, j, k), r2 = D8(i, j, k);
          const float cfa = tema * r1, cfc = tema * r2;
          const float cfb = 1.0f - cfa - cfc;
          const float cfd = A3(s, i, j, k) + temb * (r1 * A3(s, i, j, k - 1) - (r1 + r2) * A3(s, i, j, k) + r2 * A3(s, i, j, k + 1));
          const float tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
          A3(d1, i, j, k) = -cfc * tem;
          A3(d2, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
        }
    k = nk;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float r1 = D7(i, j, k);
        const float cfa = tema * r1;
        const float cfb = 1.0f - cfa;
        const float cfd = A3(s, i, j, k) + temb * (r1 * A3(s, i, j, k - 1) - r1 * A3(s, i, j, k)) -
                          temc * A3(dumz, i, j, nk + 1) * mh[nk] * A3(rr, i, j, nk);
        const float tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
        A3(dumz, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
        A3(turbz, i, j, k) = A3(rho, i, j, k) * (A3(dumz, i, j, k) - A3(s, i, j, k)) * rdt;
      }
    for (k = nk - 1; k >= 1; k--)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(dumz, i, j, k) = A3(d1, i, j, k) * A3(dumz, i, j, k + 1) + A3(d2, i, j, k);
          A3(turbz, i, j, k) = A3(rho, i, j, k) * (A3(dumz, i, j, k) - A3(s, i, j, k)) * rdt;
        }
#undef D7
#undef D8
  }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(sten, i, j, k) = A3(sten, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) * A3(rr, i, j, k);
}

/* turbt (src/turbtend.F:700-1050): the TKE's own subgrid transport, added
 * to tten (K multiplied by 2) */
static void o_turbt(cm1o_model *M, const float *t, float *tten) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz, dt = M->c.dt;
  const float *mh = M->kf[K_MH], *mf = M->kf[K_MF];
  const float *rho = M->f[F_RHO], *rf = M->f[F_RF];
  const float *kmh = M->f[F_KMH], *kmv = M->f[F_KMV];
  float *dumx = M->s3, *dumy = M->s4, *dumz = M->s5, *turbx = M->s6, *turby = M->s7, *turbz = M->s8;
  int i, j, k;
  for (k = 2; k <= nk; k++) {
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni + 1; i++) {
        A3(dumx, i, j, k) = -0.25f * (A3(rf, i, j, k) + A3(rf, i - 1, j, k)) * 2.0f *
                            (A3(kmh, i, j, k) + A3(kmh, i - 1, j, k)) * (A3(t, i, j, k) - A3(t, i - 1, j, k)) * rdx * UF(i);
        A3(dumy, i, j, k) = -0.25f * (A3(rf, i, j, k) + A3(rf, i, j - 1, k)) * 2.0f *
                            (A3(kmh, i, j, k) + A3(kmh, i, j - 1, k)) * (A3(t, i, j, k) - A3(t, i, j - 1, k)) * rdy * VF(j);
      }
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(turbx, i, j, k) = -(A3(dumx, i + 1, j, k) - A3(dumx, i, j, k)) * rdx * UH(i);
        A3(turby, i, j, k) = -(A3(dumy, i, j + 1, k) - A3(dumy, i, j, k)) * rdy * VH(j);
      }
  }
  if (M->c.doimpl == 0) {
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++)
          A3(dumz, i, j, k) = -(A3(kmv, i, j, k) + A3(kmv, i, j, k + 1)) * (A3(t, i, j, k + 1) - A3(t, i, j, k)) * rdz *
                              mh[k] * A3(rho, i, j, k);
    /* bbc = 3: tke is about constant in the surface layer */
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) A3(dumz, i, j, 1) = 0.0f;
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(turbz, i, j, k) = -(A3(dumz, i, j, k) - A3(dumz, i, j, k - 1)) * rdz * mf[k];
          A3(tten, i, j, k) = A3(tten, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) / A3(rf, i, j, k);
        }
  } else {
    const float rdt = 1.0f / dt;
    const float tema = -1.0f * dt * VIALPHA * rdz * rdz;
    const float temb = dt * VIBETA * rdz * rdz;
    float *cx = M->f[F_DUM1], *cy = M->f[F_DUM2];
    k = 2;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float rrf = mf[k] / A3(rf, i, j, k);
        const float r2 = (A3(kmv, i, j, k) + A3(kmv, i, j, k + 1)) * mh[k] * A3(rho, i, j, k) * rrf;
        const float cfc = tema * r2;
        const float cfb = 1.0f - cfc;
        const float cfd = A3(t, i, j, k) + temb * (r2 * A3(t, i, j, k + 1) - r2 * A3(t, i, j, k));
        const float tem = 1.0f / cfb;
        A3(cx, i, j, k) = -cfc * tem;
        A3(cy, i, j, k) = cfd * tem;
      }
    for (k = 3; k <= nk - 1; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          const float rrf = mf[k] / A3(rf, i, j, k);
          const float r1 = (A3(kmv, i, j, k - 1) + A3(kmv, i, j, k)) * mh[k - 1] * A3(rho, i, j, k - 1) * rrf;
          const float r2 = (A3(kmv, i, j, k) + A3(kmv, i, j, k + 1)) * mh[k] * A3(rho, i, j, k) * rrf;
          const float cfa = tema * r1, cfc = tema * r2;
          const float cfb = 1.0f - cfa - cfc;
          const float cfd = A3(t, i, j, k) + temb * (r2 * A3(t, i, j, k + 1) - (r1 + r2) * A3(t, i, j, k) + r1 * A3(t, i, j, k - 1));
          const float tem = 1.0f / (cfa * A3(cx, i, j, k - 1) + cfb);
          A3(cx, i, j, k) = -cfc * tem;
          A3(cy, i, j, k) = (cfd - cfa * A3(cy, i, j, k - 1)) * tem;
        }
    k = nk;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float rrf = mf[k] / A3(rf, i, j, k);
        const float r1 = (A3(kmv, i, j, k - 1) + A3(kmv, i, j, k)) * mh[k - 1] * A3(rho, i, j, k - 1) * rrf;
        const float cfa = tema * r1;
        const float cfb = 1.0f - cfa;
        const float cfd = A3(t, i, j, k) + temb * (-r1 * A3(t, i, j, k) + r1 * A3(t, i, j, k - 1));
        const float tem = 1.0f / (cfa * A3(cx, i, j, k - 1) + cfb);
        A3(dumz, i, j, k) = (cfd - cfa * A3(cy, i, j, k - 1)) * tem;
        A3(turbz, i, j, k) = A3(rf, i, j, k) * (A3(dumz, i, j, k) - A3(t, i, j, k)) * rdt;
        A3(tten, i, j, k) = A3(tten, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) / A3(rf, i, j, k);
      }
    for (k = nk - 1; k >= 2; k--)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(dumz, i, j, k) = A3(cx, i, j, k) * A3(dumz, i, j, k + 1) + A3(cy, i, j, k);
          A3(turbz, i, j, k) = A3(rf, i, j, k) * (A3(dumz, i, j, k) - A3(t, i, j, k)) * rdt;
          A3(tten, i, j, k) = A3(tten, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) / A3(rf, i, j, k);
        }
  }
}

/* turbu, turbv, turbw with turbuz/turbvz (src/turbtend.F:1052-2892) */
static void o_turbuvw(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz, dt = M->c.dt;
  const float *mh = M->kf[K_MH], *mf = M->kf[K_MF];
  const float *rho = M->f[F_RHO], *rf = M->f[F_RF];
  const float *kmv = M->f[F_KMV], *kmh = M->f[F_KMH];
  const float *t11 = M->f[F_T11], *t12 = M->f[F_T12], *t13 = M->f[F_T13], *t22 = M->f[F_T22],
              *t23 = M->f[F_T23], *t33 = M->f[F_T33];
  const float *ua = M->f[F_UA], *va = M->f[F_VA], *wa = M->f[F_WA];
  float *uten1 = M->f[F_UTEN1], *vten1 = M->f[F_VTEN1], *wten1 = M->f[F_WTEN1];
  float *turbx = M->s6, *turby = M->s7, *turbz = M->s8;
  float *d1 = M->f[F_DUM1], *d2 = M->f[F_DUM2], *d3 = M->f[F_DUM3], *d7 = M->f[F_DUM7];
  int i, j, k;
  const float rdt = 0.5f / dt;
  const float tema = -0.0625f * dt * VIALPHA * rdz * rdz;
  const float temb = 0.0625f * dt * VIBETA * rdz * rdz;
  const float temc = 0.5f * dt * rdz;

  /* ---- u ---- */
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni + 1; i++) {
        A3(turbx, i, j, k) = (A3(t11, i, j, k) - A3(t11, i - 1, j, k)) * rdx * UF(i);
        A3(turby, i, j, k) = (A3(t12, i, j + 1, k) - A3(t12, i, j, k)) * rdy * VH(j);
      }
  if (M->c.doimpl == 0) {
    const float tem = rdz * 0.5f;
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni + 1; i++)
          A3(turbz, i, j, k) = (A3(t13, i, j, k + 1) - A3(t13, i, j, k)) * tem * (mh[k] + mh[k]);
  } else {
    const float *u = ua, *w = wa;
    k = 1;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni + 1; i++) {
        const float rru0 = 1.0f / (0.5f * (A3(rho, i - 1, j, k) + A3(rho, i, j, k)));
        float tem = (mh[k] + mh[k]) * rru0;
        const float r2 = (A3(kmv, i - 1, j, k + 1) + A3(kmv, i, j, k + 1)) * (mf[k + 1] + mf[k + 1]) *
                         (A3(rf, i - 1, j, k + 1) + A3(rf, i, j, k + 1)) * tem;
        const float cfc = tema * r2;
        const float cfb = 1.0f - cfc;
        const float cfd = A3(u, i, j, k) + temb * (r2 * A3(u, i, j, k + 1) - r2 * A3(u, i, j, k)) -
                          temc * A3(t13, i, j, 1) * (mh[1] + mh[1]) * rru0;
        tem = 1.0f / cfb;
        A3(d1, i, j, 1) = -cfc * tem;
        A3(d2, i, j, 1) = cfd * tem;
        A3(d7, i, j, 1) = 0.0f;
      }
    for (k = 2; k <= nk - 1; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni + 1; i++) {
          const float rru0 = 1.0f / (0.5f * (A3(rho, i - 1, j, k) + A3(rho, i, j, k)));
          float tem = (mh[k] + mh[k]) * rru0;
          const float r1 = (A3(kmv, i - 1, j, k) + A3(kmv, i, j, k)) * (mf[k] + mf[k]) *
                           (A3(rf, i - 1, j, k) + A3(rf, i, j, k)) * tem;
          const float r2 = (A3(kmv, i - 1, j, k + 1) + A3(kmv, i, j, k + 1)) * (mf[k + 1] + mf[k + 1]) *
                           (A3(rf, i - 1, j, k + 1) + A3(rf, i, j, k + 1)) * tem;
          const float cfa = tema * r1, cfc = tema * r2;
          const float cfb = 1.0f - cfa - cfc;
          const float cfd = A3(u, i, j, k) + temb * (r2 * A3(u, i, j, k + 1) - (r1 + r2) * A3(u, i, j, k) + r1 * A3(u, i, j, k - 1));
          tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
          A3(d1, i, j, k) = -cfc * tem;
          A3(d2, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
          A3(d7, i, j, k) = (A3(w, i, j, k) - A3(w, i - 1, j, k)) * rdx * UF(i) * 0.25f *
                            (A3(kmv, i - 1, j, k) + A3(kmv, i, j, k)) * (A3(rf, i - 1, j, k) + A3(rf, i, j, k));
        }
    k = nk;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni + 1; i++) {
        const float rru0 = 1.0f / (0.5f * (A3(rho, i - 1, j, k) + A3(rho, i, j, k)));
        float tem = (mh[k] + mh[k]) * rru0;
        const float r1 = (A3(kmv, i - 1, j, k) + A3(kmv, i, j, k)) * (mf[k] + mf[k]) *
                         (A3(rf, i - 1, j, k) + A3(rf, i, j, k)) * tem;
        const float cfa = tema * r1;
        const float cfb = 1.0f - cfa;
        const float cfd = A3(u, i, j, k) + temb * (-r1 * A3(u, i, j, k) + r1 * A3(u, i, j, k - 1)) +
                          temc * A3(t13, i, j, nk + 1) * (mh[nk] + mh[nk]) * rru0;
        tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
        A3(d3, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
        A3(turbz, i, j, k) = (A3(rho, i - 1, j, k) + A3(rho, i, j, k)) * (A3(d3, i, j, k) - A3(u, i, j, k)) * rdt;
        A3(d7, i, j, k) = (A3(w, i, j, k) - A3(w, i - 1, j, k)) * rdx * UF(i) * 0.25f *
                          (A3(kmv, i - 1, j, k) + A3(kmv, i, j, k)) * (A3(rf, i - 1, j, k) + A3(rf, i, j, k));
        A3(turbz, i, j, k) = A3(turbz, i, j, k) + (0.0f - A3(d7, i, j, k)) * rdz * 0.5f * (mh[k] + mh[k]);
      }
    for (k = nk - 1; k >= 1; k--)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni + 1; i++) {
          A3(d3, i, j, k) = A3(d1, i, j, k) * A3(d3, i, j, k + 1) + A3(d2, i, j, k);
          A3(turbz, i, j, k) = (A3(rho, i - 1, j, k) + A3(rho, i, j, k)) * (A3(d3, i, j, k) - A3(u, i, j, k)) * rdt;
          A3(turbz, i, j, k) = A3(turbz, i, j, k) + (A3(d7, i, j, k + 1) - A3(d7, i, j, k)) * rdz * 0.5f * (mh[k] + mh[k]);
        }
  }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni + 1; i++) {
        const float rru0 = 1.0f / (0.5f * (A3(rho, i - 1, j, k) + A3(rho, i, j, k)));
        A3(uten1, i, j, k) = A3(uten1, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) * rru0;
      }

  /* ---- v ---- */
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni; i++) {
        A3(turbx, i, j, k) = (A3(t12, i + 1, j, k) - A3(t12, i, j, k)) * rdx * UH(i);
        A3(turby, i, j, k) = (A3(t22, i, j, k) - A3(t22, i, j - 1, k)) * rdy * VF(j);
      }
  if (M->c.doimpl == 0) {
    const float tem = rdz * 0.5f;
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni; i++)
          A3(turbz, i, j, k) = (A3(t23, i, j, k + 1) - A3(t23, i, j, k)) * tem * (mh[k] + mh[k]);
  } else {
    const float *v = va, *w = wa;
    k = 1;
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni; i++) {
        const float rrv0 = 1.0f / (0.5f * (A3(rho, i, j - 1, k) + A3(rho, i, j, k)));
        float tem = (mh[k] + mh[k]) * rrv0;
        const float r2 = (A3(kmv, i, j - 1, k + 1) + A3(kmv, i, j, k + 1)) * (mf[k + 1] + mf[k + 1]) *
                         (A3(rf, i, j - 1, k + 1) + A3(rf, i, j, k + 1)) * tem;
        const float cfc = tema * r2;
        const float cfb = 1.0f - cfc;
        const float cfd = A3(v, i, j, k) + temb * (r2 * A3(v, i, j, k + 1) - r2 * A3(v, i, j, k)) -
                          temc * A3(t23, i, j, 1) * (mh[1] + mh[1]) * rrv0;
        tem = 1.0f / cfb;
        A3(d1, i, j, 1) = -cfc * tem;
        A3(d2, i, j, 1) = cfd * tem;
        A3(d7, i, j, 1) = 0.0f;
      }
    for (k = 2; k <= nk - 1; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni; i++) {
          const float rrv0 = 1.0f / (0.5f * (A3(rho, i, j - 1, k) + A3(rho, i, j, k)));
          float tem = (mh[k] + mh[k]) * rrv0;
          const float r1 = (A3(kmv, i, j - 1, k) + A3(kmv, i, j, k)) * (mf[k] + mf[k]) *
                           (A3(rf, i, j - 1, k) + A3(rf, i, j, k)) * tem;
          const float r2 = (A3(kmv, i, j - 1, k + 1) + A3(kmv, i, j, k + 1)) * (mf[k + 1] + mf[k + 1]) *
                           (A3(rf, i, j - 1, k + 1) + A3(rf, i, j, k + 1)) * tem;
          const float cfa = tema * r1, cfc = tema * r2;
          const float cfb = 1.0f - cfa - cfc;
          const float cfd = A3(v, i, j, k) + temb * (r2 * A3(v, i, j, k + 1) - (r1 + r2) * A3(v, i, j, k) + r1 * A3(v, i, j, k - 1));
          tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
          A3(d1, i, j, k) = -cfc * tem;
          A3(d2, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
          A3(d7, i, j, k) = (A3(w, i, j, k) - A3(w, i, j - 1, k)) * rdy * VF(j) * 0.25f *
                            (A3(kmv, i, j - 1, k) + A3(kmv, i, j, k)) * (A3(rf, i, j - 1, k) + A3(rf, i, j, k));
        }
    k = nk;
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni; i++) {
        const float rrv0 = 1.0f / (0.5f * (A3(rho, i, j - 1, k) + A3(rho, i, j, k)));
        float tem = (mh[k] + mh[k]) * rrv0;
        const float r1 = (A3(kmv, i, j - 1, k) + A3(kmv, i, j, k)) * (mf[k] + mf[k]) *
                         (A3(rf, i, j - 1, k) + A3(rf, i, j, k)) * tem;
        const float cfa = tema * r1;
        const float cfb = 1.0f - cfa;
        const float cfd = A3(v, i, j, k) + temb * (-r1 * A3(v, i, j, k) + r1 * A3(v, i, j, k - 1)) +
                          temc * A3(t23, i, j, nk + 1) * (mh[nk] + mh[nk]) * rrv0;
        tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
        A3(d3, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
        A3(turbz, i, j, k) = (A3(rho, i, j - 1, k) + A3(rho, i, j, k)) * (A3(d3, i, j, k) - A3(v, i, j, k)) * rdt;
        A3(d7, i, j, k) = (A3(w, i, j, k) - A3(w, i, j - 1, k)) * rdy * VF(j) * 0.25f *
                          (A3(kmv, i, j - 1, k) + A3(kmv, i, j, k)) * (A3(rf, i, j - 1, k) + A3(rf, i, j, k));
        A3(turbz, i, j, k) = A3(turbz, i, j, k) + (0.0f - A3(d7, i, j, k)) * rdz * 0.5f * (mh[k] + mh[k]);
      }
    for (k = nk - 1; k >= 1; k--)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni; i++) {
          A3(d3, i, j, k) = A3(d1, i, j, k) * A3(d3, i, j, k + 1) + A3(d2, i, j, k);
          A3(turbz, i, j, k) = (A3(rho, i, j - 1, k) + A3(rho, i, j, k)) * (A3(d3, i, j, k) - A3(v, i, j, k)) * rdt;
          A3(turbz, i, j, k) = A3(turbz, i, j, k) + (A3(d7, i, j, k + 1) - A3(d7, i, j, k)) * rdz * 0.5f * (mh[k] + mh[k]);
        }
  }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni; i++) {
        const float rrv0 = 1.0f / (0.5f * (A3(rho, i, j - 1, k) + A3(rho, i, j, k)));
        A3(vten1, i, j, k) = A3(vten1, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) * rrv0;
      }

  /* ---- w ---- */
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(turbx, i, j, k) = (A3(t13, i + 1, j, k) - A3(t13, i, j, k)) * rdx * UH(i);
        A3(turby, i, j, k) = (A3(t23, i, j + 1, k) - A3(t23, i, j, k)) * rdy * VH(j);
      }
  if (M->c.doimpl == 0) {
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(turbz, i, j, k) = (A3(t33, i, j, k) - A3(t33, i, j, k - 1)) * rdz * mf[k];
  } else {
    /* src/turbtend.F:2658-2800: implicit in w with kmh, explicit t33
     * pieces at the two ends */
    const float rdtw = 1.0f / dt;
    const float temaw = -1.0f * dt * VIALPHA * rdz * rdz;
    const float tembw = dt * VIBETA * rdz * rdz;
    const float temcw = dt * rdz;
    const float *w = wa;
    k = 2;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float rrf = mf[k] / A3(rf, i, j, k);
        const float r2 = (A3(kmh, i, j, k) + A3(kmh, i, j, k + 1)) * mh[k] * A3(rho, i, j, k) * rrf;
        const float cfc = temaw * r2;
        const float cfb = 1.0f - cfc;
        const float cfd = A3(w, i, j, k) + tembw * (r2 * A3(w, i, j, k + 1) - r2 * A3(w, i, j, k)) -
                          temcw * A3(t33, i, j, k - 1) * rrf;
        const float tem = 1.0f / cfb;
        A3(d1, i, j, k) = -cfc * tem;
        A3(d2, i, j, k) = cfd * tem;
      }
    for (k = 3; k <= nk - 1; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          const float rrf = mf[k] / A3(rf, i, j, k);
          const float r1 = (A3(kmh, i, j, k - 1) + A3(kmh, i, j, k)) * mh[k - 1] * A3(rho, i, j, k - 1) * rrf;
          const float r2 = (A3(kmh, i, j, k) + A3(kmh, i, j, k + 1)) * mh[k] * A3(rho, i, j, k) * rrf;
          const float cfa = temaw * r1, cfc = temaw * r2;
          const float cfb = 1.0f - cfa - cfc;
          const float cfd = A3(w, i, j, k) + tembw * (r2 * A3(w, i, j, k + 1) - (r1 + r2) * A3(w, i, j, k) + r1 * A3(w, i, j, k - 1));
          const float tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
          A3(d1, i, j, k) = -cfc * tem;
          A3(d2, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
        }
    k = nk;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        const float rrf = mf[k] / A3(rf, i, j, k);
        const float r1 = (A3(kmh, i, j, k - 1) + A3(kmh, i, j, k)) * mh[k - 1] * A3(rho, i, j, k - 1) * rrf;
        const float cfa = temaw * r1;
        const float cfb = 1.0f - cfa;
        const float cfd = A3(w, i, j, k) + tembw * (-r1 * A3(w, i, j, k) + r1 * A3(w, i, j, k - 1)) +
                          temcw * A3(t33, i, j, k) * rrf;
        const float tem = 1.0f / (cfa * A3(d1, i, j, k - 1) + cfb);
        A3(d3, i, j, k) = (cfd - cfa * A3(d2, i, j, k - 1)) * tem;
        A3(turbz, i, j, k) = A3(rf, i, j, k) * (A3(d3, i, j, k) - A3(w, i, j, k)) * rdtw;
      }
    for (k = nk - 1; k >= 2; k--)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(d3, i, j, k) = A3(d1, i, j, k) * A3(d3, i, j, k + 1) + A3(d2, i, j, k);
          A3(turbz, i, j, k) = A3(rf, i, j, k) * (A3(d3, i, j, k) - A3(w, i, j, k)) * rdtw;
        }
  }
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(wten1, i, j, k) = A3(wten1, i, j, k) + ((A3(turbx, i, j, k) + A3(turby, i, j, k)) + A3(turbz, i, j, k)) / A3(rf, i, j, k);
}

void o_turb_tend(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float *th0 = M->kf[K_TH0];
  float *tketen = M->f[F_TKETEN], *tkea = M->f[F_TKEA];
  const float *khv = M->f[F_KHV], *kmv = M->f[F_KMV], *nm = M->f[F_NM], *diss = M->f[F_DISSTEN],
              *defv = M->f[F_DEFV], *defh = M->f[F_DEFH];
  float *sadv = M->f[F_SADV], *tha = M->f[F_THA];
  int i, j, k, n;
  /* RHS of the TKE equation, tconfig = 1 (src/solve1.F:320-332) */
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(tketen, i, j, k) = -A3(khv, i, j, k) * A3(nm, i, j, k) - A3(diss, i, j, k) +
                              A3(kmv, i, j, k) * fmaxf(0.0f, (A3(defv, i, j, k) + A3(defh, i, j, k)));
  o_turbt(M, tkea, tketen);
  o_turbuvw(M);
  /* theta relative to th0r (src/solve1.F:1184-1206), then turbs */
  for (k = 1; k <= nk; k++) {
    const float tem = th0[k] - K->th0r;
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) A3(sadv, i, j, k) = tem + A3(tha, i, j, k);
  }
  o_turbs(M, 1, M->d2[D_THFLUX], sadv, M->f[F_THTEN1]);
  for (n = 0; n < M->c.npt; n++) o_turbs(M, 0, M->d2[D_THFLUX], M->f[F_PTA0 + n], M->f[F_PTTEN0 + n]);
}

void o_tke_step(cm1o_model *M, int nrk, float dttmp) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  float *wten = M->f[F_WTEN], *tketen = M->f[F_TKETEN], *tkea = M->f[F_TKEA], *tke3d = M->f[F_TKE3D];
  int i, j, k;
  const int weno = (M->c.advwenos == 1) || (M->c.advwenos == 2 && nrk == 3);
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) A3(wten, i, j, k) = A3(tketen, i, j, k);
  o_advw_w(M, nrk, weno, M->f[F_RRU], M->f[F_RRV], M->f[F_RRW], tke3d, wten, M->f[F_DIVX]);
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(tke3d, i, j, k) = A3(tkea, i, j, k) + dttmp * A3(wten, i, j, k);
        if (A3(tke3d, i, j, k) < 1.0e-10f) A3(tke3d, i, j, k) = 0.0f;
      }
  o_bcw(M, tke3d, 1);
}

void o_tke_copy(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) A3(M->f[F_TKEA], i, j, k) = A3(M->f[F_TKE3D], i, j, k);
}
