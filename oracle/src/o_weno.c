/* Oracle: WENO5 horizontal and vertical fluxes (test infrastructure only).
 * hadv_weno5: src/adv_routines.F:204-430 (stag 1 with the pdef magnitude
 * check :230-295).  vadv_weno5: :2359-2557, with vadv_lwr_weno_23
 * :4038-4208, vadv_lwr_weno_35 :4214-4406 and vadv_lwr_weno_35_b
 * :4412-4447.  upstrpd: :6331-6348.  weno5 itself is o_weno5 (o_adv.c). */
#include "cm1o_internal.h"

/* positive-definite upstream value (Beets & Koren), single precision:
 * src/adv_routines.F:6331-6348 */
float o_upstrpd(float s1, float s2, float s3, double weps) {
  float dd = s2 - s1;
  float rr = (s3 - s2) / (copysignf((float)weps, dd) + dd);
  float phi = fmaxf(0.0f, fminf(rr, fminf((1.0f / 6.0f) + (2.0f / 6.0f) * rr, 1.0f)));
  return s2 + phi * (s2 - s1);
}

#define S(i, j, k) A3(a, i, j, k)

/* |s| of the five stencil points all below weps -> zero flux (pdef=1) */
static int weno_doit(float s1, float s2, float s3, float s4, float s5, double weps) {
  float bmin = (float)weps;
  float bmax = fmaxf(fmaxf(fmaxf(fmaxf(fabsf(s1), fabsf(s2)), fabsf(s3)), fabsf(s4)), fabsf(s5));
  return bmax > bmin;
}

void o_hadv_weno5(cm1o_model *M, int stag, const float *rru, const float *rrv,
                  float *dumx, float *dumy, const float *a, int pdef, double weps) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *c2 = M->kf[K_C2];
  int i, j, k;
  if (stag == 1) {
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          float s1, s2, s3, s4, s5;
          float ru = A3(rru, i, j, k), rv = A3(rrv, i, j, k);
          if (ru >= 0.0f) { s1 = S(i-3,j,k); s2 = S(i-2,j,k); s3 = S(i-1,j,k); s4 = S(i,j,k); s5 = S(i+1,j,k); }
          else            { s1 = S(i+2,j,k); s2 = S(i+1,j,k); s3 = S(i,j,k); s4 = S(i-1,j,k); s5 = S(i-2,j,k); }
          if (pdef == 1 && !weno_doit(s1, s2, s3, s4, s5, weps)) A3(dumx, i, j, k) = 0.0f;
          else A3(dumx, i, j, k) = ru * o_weno5(s1, s2, s3, s4, s5, weps);
          if (rv >= 0.0f) { s1 = S(i,j-3,k); s2 = S(i,j-2,k); s3 = S(i,j-1,k); s4 = S(i,j,k); s5 = S(i,j+1,k); }
          else            { s1 = S(i,j+2,k); s2 = S(i,j+1,k); s3 = S(i,j,k); s4 = S(i,j-1,k); s5 = S(i,j-2,k); }
          if (pdef == 1 && !weno_doit(s1, s2, s3, s4, s5, weps)) A3(dumy, i, j, k) = 0.0f;
          else A3(dumy, i, j, k) = rv * o_weno5(s1, s2, s3, s4, s5, weps);
        }
  } else if (stag == 2) {
    int i1 = M->ibw == 1 ? 2 : 1, i2 = M->ibe == 1 ? ni : ni + 1;
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = i1 - 1; i <= i2; i++) {
          float ubar = 0.5f * (A3(rru, i, j, k) + A3(rru, i + 1, j, k));
          if (ubar >= 0.0f) A3(dumx, i, j, k) = ubar * o_weno5(S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k), S(i+2,j,k), weps);
          else              A3(dumx, i, j, k) = ubar * o_weno5(S(i+3,j,k), S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k), weps);
          float vbar = 0.5f * (A3(rrv, i, j, k) + A3(rrv, i - 1, j, k));
          if (vbar >= 0.0f) A3(dumy, i, j, k) = vbar * o_weno5(S(i,j-3,k), S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k), weps);
          else              A3(dumy, i, j, k) = vbar * o_weno5(S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k), S(i,j-2,k), weps);
        }
  } else if (stag == 3) {
    int j1 = M->ibs == 1 ? 2 : 1, j2 = M->ibn == 1 ? nj : nj + 1;
    for (k = 1; k <= nk; k++)
      for (j = j1 - 1; j <= j2; j++)
        for (i = 1; i <= ni + 1; i++) {
          float ubar = 0.5f * (A3(rru, i, j, k) + A3(rru, i, j - 1, k));
          if (ubar >= 0.0f) A3(dumx, i, j, k) = ubar * o_weno5(S(i-3,j,k), S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k), weps);
          else              A3(dumx, i, j, k) = ubar * o_weno5(S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k), S(i-2,j,k), weps);
          float vbar = 0.5f * (A3(rrv, i, j, k) + A3(rrv, i, j + 1, k));
          if (vbar >= 0.0f) A3(dumy, i, j, k) = vbar * o_weno5(S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k), S(i,j+2,k), weps);
          else              A3(dumy, i, j, k) = vbar * o_weno5(S(i,j+3,k), S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k), weps);
        }
  } else {
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          float cc2 = 0.5f * (c2[k] + c2[k]);
          float cc1 = 1.0f - cc2;
          float ubar = cc2 * A3(rru, i, j, k) + cc1 * A3(rru, i, j, k - 1);
          if (ubar >= 0.0f) A3(dumx, i, j, k) = ubar * o_weno5(S(i-3,j,k), S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k), weps);
          else              A3(dumx, i, j, k) = ubar * o_weno5(S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k), S(i-2,j,k), weps);
          float vbar = cc2 * A3(rrv, i, j, k) + cc1 * A3(rrv, i, j, k - 1);
          if (vbar >= 0.0f) A3(dumy, i, j, k) = vbar * o_weno5(S(i,j-3,k), S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k), weps);
          else              A3(dumy, i, j, k) = vbar * o_weno5(S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k), S(i,j-2,k), weps);
        }
  }
}

static inline float wface_w(cm1o_model *M, int stag, const float *rrw, int i, int j, int k) {
  if (stag == 1) return A3(rrw, i, j, k);
  if (stag == 2) return 0.5f * (A3(rrw, i, j, k) + A3(rrw, i - 1, j, k));
  if (stag == 3) return 0.5f * (A3(rrw, i, j, k) + A3(rrw, i, j - 1, k));
  return 0.5f * (A3(rrw, i, j, k) + A3(rrw, i, j, k + 1));
}

void o_vadv_weno5(cm1o_model *M, int stag, const float *rrw, float *dumz, const float *a,
                  int pdef, int pdefweno, double weps) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  int i, j, k;
  int ilo = 1, ihi = ni, jlo = 1, jhi = nj;
  (void)pdef;
  if (stag == 2) { ilo = M->ibw == 1 ? 2 : 1; ihi = M->ibe == 1 ? ni : ni + 1; }
  if (stag == 3) { jlo = M->ibs == 1 ? 2 : 1; jhi = M->ibn == 1 ? nj : nj + 1; }
  if (stag == 1 && pdefweno == 1) {
    /* interior with the magnitude check, k = 3..nk-1 (:2384-2412) */
    for (k = 3; k <= nk - 1; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          float s1, s2, s3, s4, s5;
          float w = A3(rrw, i, j, k);
          if (w >= 0.0f) { s1 = S(i,j,k-3); s2 = S(i,j,k-2); s3 = S(i,j,k-1); s4 = S(i,j,k); s5 = S(i,j,k+1); }
          else           { s1 = S(i,j,k+2); s2 = S(i,j,k+1); s3 = S(i,j,k); s4 = S(i,j,k-1); s5 = S(i,j,k-2); }
          int doit = weno_doit(s1, s2, s3, s4, s5, weps);
          if (k == 3 && w > 0.0f) doit = 0;
          if (k == nk - 1 && w < 0.0f) doit = 0;
          A3(dumz, i, j, k) = doit ? w * o_weno5(s1, s2, s3, s4, s5, weps) : 0.0f;
        }
  } else if (stag != 4) {
    for (k = 4; k <= nk - 2; k++)
      for (j = jlo; j <= jhi; j++)
        for (i = ilo; i <= ihi; i++) {
          float wb = wface_w(M, stag, rrw, i, j, k);
          if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k-3), S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), weps);
          else            A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1), S(i,j,k-2), weps);
        }
  } else {
    for (k = 3; k <= nk - 2; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          float wb = wface_w(M, 4, rrw, i, j, k);
          if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), S(i,j,k+2), weps);
          else            A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k+3), S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1), weps);
        }
  }
  /* vadv_lwr_weno_23 */
  for (j = jlo; j <= jhi; j++)
    for (i = ilo; i <= ihi; i++) {
      if (stag != 4) {
        k = 2;
        float wb = wface_w(M, stag, rrw, i, j, k);
        if (wb >= 0.0f) {
          if (stag == 1) A3(dumz, i, j, k) = wb * (c1[k] * S(i,j,k-1) + c2[k] * S(i,j,k));
          else {
            float cc2 = 0.5f * (c2[k] + c2[k]);
            float cc1 = 1.0f - cc2;
            A3(dumz, i, j, k) = wb * (cc1 * S(i,j,k-1) + cc2 * S(i,j,k));
          }
        } else A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k+1), S(i,j,k), S(i,j,k-1), weps);
        k = nk;
        wb = wface_w(M, stag, rrw, i, j, k);
        if (wb > 0.0f) A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k-2), S(i,j,k-1), S(i,j,k), weps);
        else {
          if (stag == 1) A3(dumz, i, j, k) = wb * (c1[k] * S(i,j,k-1) + c2[k] * S(i,j,k));
          else {
            float cc2 = 0.5f * (c2[k] + c2[k]);
            float cc1 = 1.0f - cc2;
            A3(dumz, i, j, k) = wb * (cc1 * S(i,j,k-1) + cc2 * S(i,j,k));
          }
        }
      } else {
        k = 1;
        float wb = wface_w(M, 4, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * 0.5f * (S(i,j,k) + S(i,j,k+1));
        else A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k+2), S(i,j,k+1), S(i,j,k), weps);
        k = nk;
        wb = wface_w(M, 4, rrw, i, j, k);
        if (wb > 0.0f) A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k-1), S(i,j,k), S(i,j,k+1), weps);
        else A3(dumz, i, j, k) = wb * 0.5f * (S(i,j,k) + S(i,j,k+1));
      }
    }
  if (stag == 1 && pdefweno == 1) {
    /* vadv_lwr_weno_35_b: only the upstream-side faces are replaced */
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        k = 3;
        if (A3(rrw, i, j, k) > 0.0f)
          A3(dumz, i, j, k) = A3(rrw, i, j, k) * o_upstrpd(S(i,j,k-2), S(i,j,k-1), S(i,j,k), weps);
        k = nk - 1;
        if (A3(rrw, i, j, k) < 0.0f)
          A3(dumz, i, j, k) = A3(rrw, i, j, k) * o_upstrpd(S(i,j,k+1), S(i,j,k), S(i,j,k-1), weps);
      }
    return;
  }
  /* vadv_lwr_weno_35 */
  for (j = jlo; j <= jhi; j++)
    for (i = ilo; i <= ihi; i++) {
      if (stag != 4) {
        k = 3;
        float wb = wface_w(M, stag, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k-2), S(i,j,k-1), S(i,j,k), weps);
        else A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1), S(i,j,k-2), weps);
        k = nk - 1;
        wb = wface_w(M, stag, rrw, i, j, k);
        if (stag == 1) {
          if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k-3), S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), weps);
          else A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k+1), S(i,j,k), S(i,j,k-1), weps);
        } else {
          if (wb > 0.0f) A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k-3), S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), weps);
          else A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k+1), S(i,j,k), S(i,j,k-1), weps);
        }
      } else {
        k = 2;
        float wb = wface_w(M, 4, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k-1), S(i,j,k), S(i,j,k+1), weps);
        else A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k+3), S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1), weps);
        k = nk - 1;
        wb = wface_w(M, 4, rrw, i, j, k);
        if (wb > 0.0f) A3(dumz, i, j, k) = wb * o_weno5(S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), S(i,j,k+2), weps);
        else A3(dumz, i, j, k) = wb * o_upstrpd(S(i,j,k+2), S(i,j,k+1), S(i,j,k), weps);
      }
    }
}
#undef S
