/* Oracle: flux-form advection (test infrastructure only).
 * flx3/flx5: src/adv_routines.F:5971-5984, :6006-6019.
 * hadv_flx5: src/adv_routines.F:1335-1476.  vadv_flx5: :3269-3388 with
 * vadv_lwr_flx_23 :4817-4995 and vadv_lwr_flx_35 :5001-5169.
 * advs: src/adv.F:14-903; advu :911-1839; advv :1847-2770; advw :2778-3610.
 * pdefx1/x2/y1/y2/z: src/pdef.F:15-581 (pdscheme=1). */
#include "cm1o_internal.h"
#include <stdlib.h>
#include <string.h>

float o_flx3(float s1, float s2, float s3) {
  return (-1.0f / 6.0f) * s1 + (5.0f / 6.0f) * s2 + (2.0f / 6.0f) * s3;
}

float o_flx5(float s1, float s2, float s3, float s4, float s5) {
  return (2.0f / 60.0f) * s1 + (-13.0f / 60.0f) * s2 + (47.0f / 60.0f) * s3 +
         (27.0f / 60.0f) * s4 + (-3.0f / 60.0f) * s5;
}

/* weno5 (siform=2, Borges et al. indicators): src/adv_routines.F:6166-6199.
 * The indicators are formed in single precision and widened; weights and
 * the final combination are double precision. */
float o_weno5(float s1, float s2, float s3, float s4, float s5, double weps) {
  float t;
  float a1 = s1 - 2.0f * s2 + s3, c1 = s1 - 4.0f * s2 + 3.0f * s3;
  float a2 = s2 - 2.0f * s3 + s4, c2 = s2 - s4;
  float a3 = s3 - 2.0f * s4 + s5, c3 = 3.0f * s3 - 4.0f * s4 + s5;
  double b1 = (double)((13.0f / 12.0f) * (a1 * a1) + 0.25f * (c1 * c1));
  double b2 = (double)((13.0f / 12.0f) * (a2 * a2) + 0.25f * (c2 * c2));
  double b3 = (double)((13.0f / 12.0f) * (a3 * a3) + 0.25f * (c3 * c3));
  double d13 = fabs(b1 - b3);
  double q1 = fmin(1.0e30, d13 / (b1 + weps));
  double q2 = fmin(1.0e30, d13 / (b2 + weps));
  double q3 = fmin(1.0e30, d13 / (b3 + weps));
  double w1 = (double)0.1f * ((double)1.0f + q1 * q1);
  double w2 = (double)0.6f * ((double)1.0f + q2 * q2);
  double w3 = (double)0.3f * ((double)1.0f + q3 * q3);
  float p1 = (2.0f / 6.0f) * s1 + (-7.0f / 6.0f) * s2 + (11.0f / 6.0f) * s3;
  float p2 = (-1.0f / 6.0f) * s2 + (5.0f / 6.0f) * s3 + (2.0f / 6.0f) * s4;
  float p3 = (2.0f / 6.0f) * s3 + (5.0f / 6.0f) * s4 + (-1.0f / 6.0f) * s5;
  t = (float)((w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3));
  return t;
}

#define S(i, j, k) A3(a, i, j, k)

/* hadv_flx5 */
static void hadv_flx5(cm1o_model *M, int stag, const float *rru, const float *rrv,
                      float *dumx, float *dumy, const float *a) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  const float *c2 = M->kf[K_C2];
  if (stag == 1) {
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          float ru = A3(rru, i, j, k), rv = A3(rrv, i, j, k);
          if (ru >= 0.0f) A3(dumx, i, j, k) = ru * o_flx5(S(i-3,j,k), S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k));
          else            A3(dumx, i, j, k) = ru * o_flx5(S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k), S(i-2,j,k));
          if (rv >= 0.0f) A3(dumy, i, j, k) = rv * o_flx5(S(i,j-3,k), S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k));
          else            A3(dumy, i, j, k) = rv * o_flx5(S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k), S(i,j-2,k));
        }
  } else if (stag == 2) {
    int i1 = M->ibw == 1 ? 2 : 1, i2 = M->ibe == 1 ? ni : ni + 1;
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = i1 - 1; i <= i2; i++) {
          float ubar = 0.5f * (A3(rru, i, j, k) + A3(rru, i + 1, j, k));
          if (ubar >= 0.0f) A3(dumx, i, j, k) = ubar * o_flx5(S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k), S(i+2,j,k));
          else              A3(dumx, i, j, k) = ubar * o_flx5(S(i+3,j,k), S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k));
          float vbar = 0.5f * (A3(rrv, i, j, k) + A3(rrv, i - 1, j, k));
          if (vbar >= 0.0f) A3(dumy, i, j, k) = vbar * o_flx5(S(i,j-3,k), S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k));
          else              A3(dumy, i, j, k) = vbar * o_flx5(S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k), S(i,j-2,k));
        }
  } else if (stag == 3) {
    int j1 = M->ibs == 1 ? 2 : 1, j2 = M->ibn == 1 ? nj : nj + 1;
    for (k = 1; k <= nk; k++)
      for (j = j1 - 1; j <= j2; j++)
        for (i = 1; i <= ni + 1; i++) {
          float ubar = 0.5f * (A3(rru, i, j, k) + A3(rru, i, j - 1, k));
          if (ubar >= 0.0f) A3(dumx, i, j, k) = ubar * o_flx5(S(i-3,j,k), S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k));
          else              A3(dumx, i, j, k) = ubar * o_flx5(S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k), S(i-2,j,k));
          float vbar = 0.5f * (A3(rrv, i, j, k) + A3(rrv, i, j + 1, k));
          if (vbar >= 0.0f) A3(dumy, i, j, k) = vbar * o_flx5(S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k), S(i,j+2,k));
          else              A3(dumy, i, j, k) = vbar * o_flx5(S(i,j+3,k), S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k));
        }
  } else {
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          /* c2 is horizontally uniform without terrain: 0.5*(c2+c2) */
          float cc2 = 0.5f * (c2[k] + c2[k]);
          float cc1 = 1.0f - cc2;
          float ubar = cc2 * A3(rru, i, j, k) + cc1 * A3(rru, i, j, k - 1);
          if (ubar >= 0.0f) A3(dumx, i, j, k) = ubar * o_flx5(S(i-3,j,k), S(i-2,j,k), S(i-1,j,k), S(i,j,k), S(i+1,j,k));
          else              A3(dumx, i, j, k) = ubar * o_flx5(S(i+2,j,k), S(i+1,j,k), S(i,j,k), S(i-1,j,k), S(i-2,j,k));
          float vbar = cc2 * A3(rrv, i, j, k) + cc1 * A3(rrv, i, j, k - 1);
          if (vbar >= 0.0f) A3(dumy, i, j, k) = vbar * o_flx5(S(i,j-3,k), S(i,j-2,k), S(i,j-1,k), S(i,j,k), S(i,j+1,k));
          else              A3(dumy, i, j, k) = vbar * o_flx5(S(i,j+2,k), S(i,j+1,k), S(i,j,k), S(i,j-1,k), S(i,j-2,k));
        }
  }
}

/* vertical face velocity for staggering stag at (i,j,k) */
static inline float wface(cm1o_model *M, int stag, const float *rrw, int i, int j, int k) {
  if (stag == 1) return A3(rrw, i, j, k);
  if (stag == 2) return 0.5f * (A3(rrw, i, j, k) + A3(rrw, i - 1, j, k));
  if (stag == 3) return 0.5f * (A3(rrw, i, j, k) + A3(rrw, i, j - 1, k));
  return 0.5f * (A3(rrw, i, j, k) + A3(rrw, i, j, k + 1));
}

static void vadv_flx5(cm1o_model *M, int stag, const float *rrw, float *dumz, const float *a) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  int i, j, k;
  int ilo = 1, ihi = ni, jlo = 1, jhi = nj;
  if (stag == 2) { ilo = M->ibw == 1 ? 2 : 1; ihi = M->ibe == 1 ? ni : ni + 1; }
  if (stag == 3) { jlo = M->ibs == 1 ? 2 : 1; jhi = M->ibn == 1 ? nj : nj + 1; }
  if (stag != 4) {
    for (k = 4; k <= nk - 2; k++)
      for (j = jlo; j <= jhi; j++)
        for (i = ilo; i <= ihi; i++) {
          float wb = wface(M, stag, rrw, i, j, k);
          if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k-3), S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1));
          else            A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1), S(i,j,k-2));
        }
    /* vadv_lwr_flx_23 */
    for (j = jlo; j <= jhi; j++)
      for (i = ilo; i <= ihi; i++) {
        k = 2;
        float wb = wface(M, stag, rrw, i, j, k);
        if (wb >= 0.0f) {
          if (stag == 1) A3(dumz, i, j, k) = wb * (c1[k] * S(i,j,k-1) + c2[k] * S(i,j,k));
          else {
            float cc2 = 0.5f * (c2[k] + c2[k]);
            float cc1 = 1.0f - cc2;
            A3(dumz, i, j, k) = wb * (cc1 * S(i,j,k-1) + cc2 * S(i,j,k));
          }
        } else A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k+1), S(i,j,k), S(i,j,k-1));
        k = nk;
        wb = wface(M, stag, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k-2), S(i,j,k-1), S(i,j,k));
        else {
          if (stag == 1) A3(dumz, i, j, k) = wb * (c1[k] * S(i,j,k-1) + c2[k] * S(i,j,k));
          else {
            float cc2 = 0.5f * (c2[k] + c2[k]);
            float cc1 = 1.0f - cc2;
            A3(dumz, i, j, k) = wb * (cc1 * S(i,j,k-1) + cc2 * S(i,j,k));
          }
        }
      }
    /* vadv_lwr_flx_35 */
    for (j = jlo; j <= jhi; j++)
      for (i = ilo; i <= ihi; i++) {
        k = 3;
        float wb = wface(M, stag, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k-2), S(i,j,k-1), S(i,j,k));
        else A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1), S(i,j,k-2));
        k = nk - 1;
        wb = wface(M, stag, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k-3), S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1));
        else A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k+1), S(i,j,k), S(i,j,k-1));
      }
  } else {
    for (k = 3; k <= nk - 2; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          float wb = wface(M, 4, rrw, i, j, k);
          if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), S(i,j,k+2));
          else            A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k+3), S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1));
        }
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        k = 1;
        float wb = wface(M, 4, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * 0.5f * (S(i,j,k) + S(i,j,k+1));
        else A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k+2), S(i,j,k+1), S(i,j,k));
        k = nk;
        wb = wface(M, 4, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k-1), S(i,j,k), S(i,j,k+1));
        else A3(dumz, i, j, k) = wb * 0.5f * (S(i,j,k) + S(i,j,k+1));
      }
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        k = 2;
        float wb = wface(M, 4, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k-1), S(i,j,k), S(i,j,k+1));
        else A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k+3), S(i,j,k+2), S(i,j,k+1), S(i,j,k), S(i,j,k-1));
        k = nk - 1;
        wb = wface(M, 4, rrw, i, j, k);
        if (wb >= 0.0f) A3(dumz, i, j, k) = wb * o_flx5(S(i,j,k-2), S(i,j,k-1), S(i,j,k), S(i,j,k+1), S(i,j,k+2));
        else A3(dumz, i, j, k) = wb * o_flx3(S(i,j,k+2), S(i,j,k+1), S(i,j,k));
      }
  }
}
#undef S

/* weno5 flux routines are provided by o_weno.c */
void o_hadv_weno5(cm1o_model *M, int stag, const float *rru, const float *rrv,
                  float *dumx, float *dumy, const float *a, int pdef, double weps);
void o_vadv_weno5(cm1o_model *M, int stag, const float *rrw, float *dumz, const float *a,
                  int pdef, int pdefweno, double weps);

/* ---------------------------------------------------------------- pdef */
static void pdefx(cm1o_model *M, const float *rru, float *advx, float *dum, float *mass,
                  const float *s0, const float *s, float dt) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *rho0 = M->kf[K_RHO0];
  const float smeps = M->k.smeps, rdx = M->rdx;
  unsigned char *flag = M->flag;
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(dum, i, j, k) = rho0[k] * A3(s0, i, j, k) +
            dt * (A3(advx, i, j, k) + A3(s, i, j, k) * (A3(rru, i + 1, j, k) - A3(rru, i, j, k)) * rdx * UH(i));
  if (M->c.wbc == 2 && M->ibw == 1)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        if (A3(rru, 1, j, k) >= 0.0f) A3(dum, 1, j, k) = rho0[k] * A3(s0, 1, j, k);
  if (M->c.ebc == 2 && M->ibe == 1)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        if (A3(rru, ni + 1, j, k) <= 0.0f) A3(dum, ni, j, k) = rho0[k] * A3(s0, ni, j, k);
  o_bcs(M, dum);
  /* pdefx2 */
  const float rdt = 1.0f / dt;
  for (k = 1; k <= nk; k++) {
    for (j = 1; j <= nj; j++)
      for (i = -1; i <= ni + 2; i++) { A3(mass, i, j, k) = 0.0f; flag[IX(M, i, j, k)] = 0; }
    for (j = 1; j <= nj; j++)
      for (i = 0; i <= ni + 1; i++) {
        if (A3(dum, i, j, k) < 0.0f) {
          float foo1 = fmaxf(0.0f, A3(dum, i - 1, j, k));
          float foo2 = fmaxf(0.0f, A3(dum, i + 1, j, k));
          if (foo1 + foo2 > smeps) {
            float foo3 = fmaxf(A3(dum, i, j, k), -(foo1 + foo2)) / (foo1 + foo2);
            A3(mass, i - 1, j, k) = A3(mass, i - 1, j, k) + foo1 * foo3;
            A3(mass, i, j, k) = A3(mass, i, j, k) - (foo1 + foo2) * foo3;
            A3(mass, i + 1, j, k) = A3(mass, i + 1, j, k) + foo2 * foo3;
            if (A3(dum, i - 1, j, k) > smeps) flag[IX(M, i - 1, j, k)] = 1;
            flag[IX(M, i, j, k)] = 1;
            if (A3(dum, i + 1, j, k) > smeps) flag[IX(M, i + 1, j, k)] = 1;
          }
        }
      }
  }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        if (flag[IX(M, i, j, k)])
          A3(advx, i, j, k) = (A3(dum, i, j, k) + A3(mass, i, j, k) - rho0[k] * A3(s0, i, j, k)) * rdt
              - A3(s, i, j, k) * (A3(rru, i + 1, j, k) - A3(rru, i, j, k)) * rdx * UH(i);
}

static void pdefy(cm1o_model *M, const float *rrv, float *advy, float *dum, float *mass,
                  const float *s0, const float *s, float dt) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *rho0 = M->kf[K_RHO0];
  const float smeps = M->k.smeps, rdy = M->rdy;
  unsigned char *flag = M->flag;
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(dum, i, j, k) = rho0[k] * A3(s0, i, j, k) +
            dt * (A3(advy, i, j, k) + A3(s, i, j, k) * (A3(rrv, i, j + 1, k) - A3(rrv, i, j, k)) * rdy * VH(j));
  if (M->c.sbc == 2 && M->ibs == 1)
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        if (A3(rrv, i, 1, k) >= 0.0f) A3(dum, i, 1, k) = rho0[k] * A3(s0, i, 1, k);
  if (M->c.nbc == 2 && M->ibn == 1)
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        if (A3(rrv, i, nj + 1, k) <= 0.0f) A3(dum, i, nj, k) = rho0[k] * A3(s0, i, nj, k);
  o_bcs(M, dum);
  const float rdt = 1.0f / dt;
  for (k = 1; k <= nk; k++) {
    for (j = -1; j <= nj + 2; j++)
      for (i = 1; i <= ni; i++) { A3(mass, i, j, k) = 0.0f; flag[IX(M, i, j, k)] = 0; }
    for (j = 0; j <= nj + 1; j++)
      for (i = 1; i <= ni; i++) {
        if (A3(dum, i, j, k) < 0.0f) {
          float foo1 = fmaxf(0.0f, A3(dum, i, j - 1, k));
          float foo2 = fmaxf(0.0f, A3(dum, i, j + 1, k));
          if (foo1 + foo2 > smeps) {
            float foo3 = fmaxf(A3(dum, i, j, k), -(foo1 + foo2)) / (foo1 + foo2);
            A3(mass, i, j - 1, k) = A3(mass, i, j - 1, k) + foo1 * foo3;
            A3(mass, i, j, k) = A3(mass, i, j, k) - (foo1 + foo2) * foo3;
            A3(mass, i, j + 1, k) = A3(mass, i, j + 1, k) + foo2 * foo3;
            if (A3(dum, i, j - 1, k) > smeps) flag[IX(M, i, j - 1, k)] = 1;
            flag[IX(M, i, j, k)] = 1;
            if (A3(dum, i, j + 1, k) > smeps) flag[IX(M, i, j + 1, k)] = 1;
          }
        }
      }
  }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        if (flag[IX(M, i, j, k)])
          A3(advy, i, j, k) = (A3(dum, i, j, k) + A3(mass, i, j, k) - rho0[k] * A3(s0, i, j, k)) * rdt
              - A3(s, i, j, k) * (A3(rrv, i, j + 1, k) - A3(rrv, i, j, k)) * rdy * VH(j);
}

static void pdefz(cm1o_model *M, const float *rrw, float *advz, float *dum, float *mass,
                  const float *s0, const float *s, float dt) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *rho0 = M->kf[K_RHO0], *mh = M->kf[K_MH];
  const float smeps = M->k.smeps, rdz = M->rdz;
  unsigned char *flag = M->flag;
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(dum, i, j, k) = rho0[k] * A3(s0, i, j, k) +
            dt * (A3(advz, i, j, k) + A3(s, i, j, k) * (A3(rrw, i, j, k + 1) - A3(rrw, i, j, k)) * rdz * mh[k]);
  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) { A3(dum, i, j, 0) = 0.0f; A3(dum, i, j, nk + 1) = 0.0f; }
  for (k = 0; k <= nk + 1; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) { A3(mass, i, j, k) = 0.0f; flag[IX(M, i, j, k)] = 0; }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        if (A3(dum, i, j, k) < 0.0f) {
          float foo1 = fmaxf(0.0f, A3(dum, i, j, k - 1));
          float foo2 = fmaxf(0.0f, A3(dum, i, j, k + 1));
          if (foo1 + foo2 > smeps) {
            float foo3 = fmaxf(A3(dum, i, j, k), -(foo1 + foo2)) / (foo1 + foo2);
            A3(mass, i, j, k - 1) = A3(mass, i, j, k - 1) + foo1 * foo3;
            A3(mass, i, j, k) = A3(mass, i, j, k) - (foo1 + foo2) * foo3;
            A3(mass, i, j, k + 1) = A3(mass, i, j, k + 1) + foo2 * foo3;
            if (A3(dum, i, j, k - 1) > smeps) flag[IX(M, i, j, k - 1)] = 1;
            flag[IX(M, i, j, k)] = 1;
            if (A3(dum, i, j, k + 1) > smeps) flag[IX(M, i, j, k + 1)] = 1;
          }
        }
      }
  const float rdt = 1.0f / dt;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        if (flag[IX(M, i, j, k)])
          A3(advz, i, j, k) = (A3(dum, i, j, k) + A3(mass, i, j, k) - rho0[k] * A3(s0, i, j, k)) * rdt
              - A3(s, i, j, k) * (A3(rrw, i, j, k + 1) - A3(rrw, i, j, k)) * rdz * mh[k];
}

/* ---------------------------------------------------------------- advs */
void o_advs(cm1o_model *M, int nrk, const float *rru, const float *rrv, const float *rrw,
            const float *divx, const float *s0, const float *s, float *sten,
            int pdef, int pdefweno, float dt, double weps) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz;
  const float *mh = M->kf[K_MH], *rr0 = M->kf[K_RR0];
  float *advx = M->s1, *advy = M->s2, *advz = M->s3, *dumx = M->s4, *dumy = M->s5,
        *dumz = M->s6, *mass = M->s7;
  int i, j, k;
  int doweno = (M->c.advwenos == 1) || (M->c.advwenos == 2 && nrk == 3);
  int doitw = (M->c.wbc == 2 && M->ibw == 1), doite = (M->c.ebc == 2 && M->ibe == 1);
  int doits = (M->c.sbc == 2 && M->ibs == 1), doitn = (M->c.nbc == 2 && M->ibn == 1);

  if (doweno) o_hadv_weno5(M, 1, rru, rrv, dumx, dumy, s, pdef, weps);
  else hadv_flx5(M, 1, rru, rrv, dumx, dumy, s);

  if (doitw)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        if (A3(rru, 1, j, k) >= 0.0f) A3(dumx, 1, j, k) = A3(dumx, 2, j, k);
  if (doite)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        if (A3(rru, ni + 1, j, k) <= 0.0f) A3(dumx, ni + 1, j, k) = A3(dumx, ni, j, k);
  if (doits)
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        if (A3(rrv, i, 1, k) >= 0.0f) A3(dumy, i, 1, k) = A3(dumy, i, 2, k);
  if (doitn)
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        if (A3(rrv, i, nj + 1, k) <= 0.0f) A3(dumy, i, nj + 1, k) = A3(dumy, i, nj, k);

  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(advx, i, j, k) = -(A3(dumx, i + 1, j, k) - A3(dumx, i, j, k)) * rdx * UH(i);
        A3(advy, i, j, k) = -(A3(dumy, i, j + 1, k) - A3(dumy, i, j, k)) * rdy * VH(j);
      }
  if (doitw)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        if (A3(rru, 1, j, k) >= 0.0f)
          A3(advx, 1, j, k) = A3(advx, 1, j, k) - A3(s, 1, j, k) * (A3(rru, 2, j, k) - A3(rru, 1, j, k)) * rdx * UH(1);
  if (doite)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        if (A3(rru, ni + 1, j, k) <= 0.0f)
          A3(advx, ni, j, k) = A3(advx, ni, j, k) - A3(s, ni, j, k) * (A3(rru, ni + 1, j, k) - A3(rru, ni, j, k)) * rdx * UH(ni);
  if (doits)
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        if (A3(rrv, i, 1, k) >= 0.0f)
          A3(advy, i, 1, k) = A3(advy, i, 1, k) - A3(s, i, 1, k) * (A3(rrv, i, 2, k) - A3(rrv, i, 1, k)) * rdy * VH(1);
  if (doitn)
    for (k = 1; k <= nk; k++)
      for (i = 1; i <= ni; i++)
        if (A3(rrv, i, nj + 1, k) <= 0.0f)
          A3(advy, i, nj, k) = A3(advy, i, nj, k) - A3(s, i, nj, k) * (A3(rrv, i, nj + 1, k) - A3(rrv, i, nj, k)) * rdy * VH(nj);

  if (pdef == 1) pdefx(M, rru, advx, dumx, mass, s0, s, dt);
  /* pdefy1 builds its own dum (dumy) before pdefx2 runs in the reference;
   * both only read advx/advy and s0/s, so running x then y is equivalent. */

  if (doweno) o_vadv_weno5(M, 1, rrw, dumz, s, pdef, pdefweno, weps);
  else vadv_flx5(M, 1, rrw, dumz, s);

  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) {
      A3(advz, i, j, 1) = -A3(dumz, i, j, 2) * rdz * mh[1];
      A3(advz, i, j, nk) = +A3(dumz, i, j, nk) * rdz * mh[nk];
    }
  for (k = 2; k <= nk - 1; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(advz, i, j, k) = -(A3(dumz, i, j, k + 1) - A3(dumz, i, j, k)) * rdz * mh[k];

  if (pdef == 1) {
    pdefz(M, rrw, advz, dumz, mass, s0, s, dt);
    pdefy(M, rrv, advy, dumy, mass, s0, s, dt);
  }

  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(sten, i, j, k) = A3(sten, i, j, k) +
            ((A3(advx, i, j, k) + A3(advy, i, j, k)) + A3(advz, i, j, k) + A3(s, i, j, k) * A3(divx, i, j, k)) * rr0[k];
}

/* ---------------------------------------------------------------- advu */
void o_advu(cm1o_model *M, int nrk, const float *rru, const float *u3d, float *uten,
            const float *rrv, const float *rrw, const float *divx) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz;
  const float *mh = M->kf[K_MH], *rr0 = M->kf[K_RR0];
  float *dumx = M->s1, *dumy = M->s2, *dumz = M->s3, *advx = M->s4, *advy = M->s5, *advz = M->s6;
  int i, j, k;
  int i1 = M->ibw == 1 ? 2 : 1, i2 = M->ibe == 1 ? ni : ni + 1;
  int doits = (M->c.sbc == 2 && M->ibs == 1), doitn = (M->c.nbc == 2 && M->ibn == 1);
  int doweno = (M->c.advwenov == 1) || (M->c.advwenov == 2 && nrk == 3);
  double weps = (double)(100.0f * M->k.epsilon);
  (void)nrk;
  if (doweno) o_hadv_weno5(M, 2, rru, rrv, dumx, dumy, u3d, 0, weps);
  else hadv_flx5(M, 2, rru, rrv, dumx, dumy, u3d);
  if (doits)
    for (k = 1; k <= nk; k++)
      for (i = i1; i <= i2; i++)
        if ((A3(rrv, i, 1, k) + A3(rrv, i - 1, 1, k)) >= 0.0f) A3(dumy, i, 1, k) = A3(dumy, i, 2, k);
  if (doitn)
    for (k = 1; k <= nk; k++)
      for (i = i1; i <= i2; i++)
        if ((A3(rrv, i, nj + 1, k) + A3(rrv, i - 1, nj + 1, k)) <= 0.0f) A3(dumy, i, nj + 1, k) = A3(dumy, i, nj, k);
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = i1; i <= i2; i++) {
        A3(advx, i, j, k) = -(A3(dumx, i, j, k) - A3(dumx, i - 1, j, k)) * rdx * UF(i);
        A3(advy, i, j, k) = -(A3(dumy, i, j + 1, k) - A3(dumy, i, j, k)) * rdy * VH(j);
      }
  if (doits)
    for (k = 1; k <= nk; k++)
      for (i = i1; i <= i2; i++)
        if ((A3(rrv, i, 1, k) + A3(rrv, i - 1, 1, k)) >= 0.0f)
          A3(advy, i, 1, k) = A3(advy, i, 1, k) - A3(u3d, i, 1, k) * 0.5f *
              ((A3(rrv, i - 1, 2, k) - A3(rrv, i - 1, 1, k)) + (A3(rrv, i, 2, k) - A3(rrv, i, 1, k))) * rdy * VH(1);
  if (doitn)
    for (k = 1; k <= nk; k++)
      for (i = i1; i <= i2; i++)
        if ((A3(rrv, i, nj + 1, k) + A3(rrv, i - 1, nj + 1, k)) <= 0.0f)
          A3(advy, i, nj, k) = A3(advy, i, nj, k) - A3(u3d, i, nj, k) * 0.5f *
              ((A3(rrv, i - 1, nj + 1, k) - A3(rrv, i - 1, nj, k)) + (A3(rrv, i, nj + 1, k) - A3(rrv, i, nj, k))) * rdy * VH(nj);
  if (doweno) o_vadv_weno5(M, 2, rrw, dumz, u3d, 0, 0, weps);
  else vadv_flx5(M, 2, rrw, dumz, u3d);
  for (j = 1; j <= nj; j++)
    for (i = i1; i <= i2; i++) {
      A3(advz, i, j, 1) = -A3(dumz, i, j, 2) * rdz * mh[1];
      A3(advz, i, j, nk) = +A3(dumz, i, j, nk) * rdz * mh[nk];
    }
  for (k = 2; k <= nk - 1; k++)
    for (j = 1; j <= nj; j++)
      for (i = i1; i <= i2; i++)
        A3(advz, i, j, k) = -(A3(dumz, i, j, k + 1) - A3(dumz, i, j, k)) * rdz * mh[k];
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = i1; i <= i2; i++)
        A3(uten, i, j, k) = A3(uten, i, j, k) +
            ((A3(advx, i, j, k) + A3(advy, i, j, k)) + A3(advz, i, j, k) +
             A3(u3d, i, j, k) * 0.5f * (A3(divx, i, j, k) + A3(divx, i - 1, j, k))) * rr0[k];
}

/* ---------------------------------------------------------------- advv */
void o_advv(cm1o_model *M, int nrk, const float *rru, const float *rrv, const float *v3d,
            float *vten, const float *rrw, const float *divx) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz;
  const float *mh = M->kf[K_MH], *rr0 = M->kf[K_RR0];
  float *dumx = M->s1, *dumy = M->s2, *dumz = M->s3, *advx = M->s4, *advy = M->s5, *advz = M->s6;
  int i, j, k;
  int j1 = M->ibs == 1 ? 2 : 1, j2 = M->ibn == 1 ? nj : nj + 1;
  int doitw = (M->c.wbc == 2 && M->ibw == 1), doite = (M->c.ebc == 2 && M->ibe == 1);
  int doweno = (M->c.advwenov == 1) || (M->c.advwenov == 2 && nrk == 3);
  double weps = (double)(100.0f * M->k.epsilon);
  if (doweno) o_hadv_weno5(M, 3, rru, rrv, dumx, dumy, v3d, 0, weps);
  else hadv_flx5(M, 3, rru, rrv, dumx, dumy, v3d);
  if (doitw)
    for (k = 1; k <= nk; k++)
      for (j = j1; j <= j2; j++)
        if ((A3(rru, 1, j, k) + A3(rru, 1, j - 1, k)) >= 0.0f) A3(dumx, 1, j, k) = A3(dumx, 2, j, k);
  if (doite)
    for (k = 1; k <= nk; k++)
      for (j = j1; j <= j2; j++)
        if ((A3(rru, ni + 1, j, k) + A3(rru, ni + 1, j - 1, k)) <= 0.0f) A3(dumx, ni + 1, j, k) = A3(dumx, ni, j, k);
  for (k = 1; k <= nk; k++)
    for (j = j1; j <= j2; j++)
      for (i = 1; i <= ni; i++) {
        A3(advx, i, j, k) = -(A3(dumx, i + 1, j, k) - A3(dumx, i, j, k)) * rdx * UH(i);
        A3(advy, i, j, k) = -(A3(dumy, i, j, k) - A3(dumy, i, j - 1, k)) * rdy * VF(j);
      }
  if (doitw)
    for (k = 1; k <= nk; k++)
      for (j = j1; j <= j2; j++)
        if ((A3(rru, 1, j, k) + A3(rru, 1, j - 1, k)) >= 0.0f)
          A3(advx, 1, j, k) = A3(advx, 1, j, k) - A3(v3d, 1, j, k) * 0.5f *
              ((A3(rru, 2, j - 1, k) - A3(rru, 1, j - 1, k)) + (A3(rru, 2, j, k) - A3(rru, 1, j, k))) * rdx * UH(1);
  if (doite)
    for (k = 1; k <= nk; k++)
      for (j = j1; j <= j2; j++)
        if ((A3(rru, ni + 1, j, k) + A3(rru, ni + 1, j - 1, k)) <= 0.0f)
          A3(advx, ni, j, k) = A3(advx, ni, j, k) - A3(v3d, ni, j, k) * 0.5f *
              ((A3(rru, ni + 1, j - 1, k) - A3(rru, ni, j - 1, k)) + (A3(rru, ni + 1, j, k) - A3(rru, ni, j, k))) * rdx * UH(ni);
  if (doweno) o_vadv_weno5(M, 3, rrw, dumz, v3d, 0, 0, weps);
  else vadv_flx5(M, 3, rrw, dumz, v3d);
  for (j = j1; j <= j2; j++)
    for (i = 1; i <= ni; i++) {
      A3(advz, i, j, 1) = -A3(dumz, i, j, 2) * rdz * mh[1];
      A3(advz, i, j, nk) = +A3(dumz, i, j, nk) * rdz * mh[nk];
    }
  for (k = 2; k <= nk - 1; k++)
    for (j = j1; j <= j2; j++)
      for (i = 1; i <= ni; i++)
        A3(advz, i, j, k) = -(A3(dumz, i, j, k + 1) - A3(dumz, i, j, k)) * rdz * mh[k];
  for (k = 1; k <= nk; k++)
    for (j = j1; j <= j2; j++)
      for (i = 1; i <= ni; i++)
        A3(vten, i, j, k) = A3(vten, i, j, k) +
            ((A3(advx, i, j, k) + A3(advy, i, j, k)) + A3(advz, i, j, k) +
             A3(v3d, i, j, k) * 0.5f * (A3(divx, i, j, k) + A3(divx, i, j - 1, k))) * rr0[k];
}

/* ---------------------------------------------------------------- advw */
void o_advw(cm1o_model *M, int nrk, const float *rru, const float *rrv, const float *rrw,
            const float *w3d, float *wten, const float *divx) {
  const int doweno = (M->c.advwenov == 1) || (M->c.advwenov == 2 && nrk == 3);
  o_advw_w(M, nrk, doweno, rru, rrv, rrw, w3d, wten, divx);
}

/* advw with the WENO switch given: the velocity's (advwenov) for w, the
 * scalars' (advwenos) for the TKE (src/solve2.F:1746-1749) */
void o_advw_w(cm1o_model *M, int nrk, int doweno, const float *rru, const float *rrv, const float *rrw,
              const float *w3d, float *wten, const float *divx) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz;
  const float *mf = M->kf[K_MF], *rrf0 = M->kf[K_RRF0], *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  float *dumx = M->s1, *dumy = M->s2, *dumz = M->s3, *advx = M->s4, *advy = M->s5, *advz = M->s6;
  int i, j, k;
  int doitw = (M->c.wbc == 2 && M->ibw == 1), doite = (M->c.ebc == 2 && M->ibe == 1);
  int doits = (M->c.sbc == 2 && M->ibs == 1), doitn = (M->c.nbc == 2 && M->ibn == 1);
  double weps = (double)(100.0f * M->k.epsilon);
  if (doweno) o_hadv_weno5(M, 4, rru, rrv, dumx, dumy, w3d, 0, weps);
  else hadv_flx5(M, 4, rru, rrv, dumx, dumy, w3d);
  if (doitw)
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float ubar = cc2 * A3(rru, 1, j, k) + cc1 * A3(rru, 1, j, k - 1);
        if (ubar >= 0.0f) A3(dumx, 1, j, k) = A3(dumx, 2, j, k);
      }
  if (doite)
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float ubar = cc2 * A3(rru, ni + 1, j, k) + cc1 * A3(rru, ni + 1, j, k - 1);
        if (ubar <= 0.0f) A3(dumx, ni + 1, j, k) = A3(dumx, ni, j, k);
      }
  if (doits)
    for (k = 2; k <= nk; k++)
      for (i = 1; i <= ni; i++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float vbar = cc2 * A3(rrv, i, 1, k) + cc1 * A3(rrv, i, 1, k - 1);
        if (vbar >= 0.0f) A3(dumy, i, 1, k) = A3(dumy, i, 2, k);
      }
  if (doitn)
    for (k = 2; k <= nk; k++)
      for (i = 1; i <= ni; i++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float vbar = cc2 * A3(rrv, i, nj + 1, k) + cc1 * A3(rrv, i, nj + 1, k - 1);
        if (vbar <= 0.0f) A3(dumy, i, nj + 1, k) = A3(dumy, i, nj, k);
      }
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(advx, i, j, k) = -(A3(dumx, i + 1, j, k) - A3(dumx, i, j, k)) * rdx * UH(i);
        A3(advy, i, j, k) = -(A3(dumy, i, j + 1, k) - A3(dumy, i, j, k)) * rdy * VH(j);
      }
  if (doitw)
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float ubar = cc2 * A3(rru, 1, j, k) + cc1 * A3(rru, 1, j, k - 1);
        if (ubar >= 0.0f)
          A3(advx, 1, j, k) = A3(advx, 1, j, k) - A3(w3d, 1, j, k) *
              (c1[k] * (A3(rru, 2, j, k - 1) - A3(rru, 1, j, k - 1)) + c2[k] * (A3(rru, 2, j, k) - A3(rru, 1, j, k))) * rdx * UH(1);
      }
  if (doite)
    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float ubar = cc2 * A3(rru, ni + 1, j, k) + cc1 * A3(rru, ni + 1, j, k - 1);
        if (ubar <= 0.0f)
          A3(advx, ni, j, k) = A3(advx, ni, j, k) - A3(w3d, ni, j, k) *
              (c1[k] * (A3(rru, ni + 1, j, k - 1) - A3(rru, ni, j, k - 1)) + c2[k] * (A3(rru, ni + 1, j, k) - A3(rru, ni, j, k))) * rdx * UH(ni);
      }
  if (doits)
    for (k = 2; k <= nk; k++)
      for (i = 1; i <= ni; i++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float vbar = cc2 * A3(rrv, i, 1, k) + cc1 * A3(rrv, i, 1, k - 1);
        if (vbar >= 0.0f)
          A3(advy, i, 1, k) = A3(advy, i, 1, k) - A3(w3d, i, 1, k) *
              (c1[k] * (A3(rrv, i, 2, k - 1) - A3(rrv, i, 1, k - 1)) + c2[k] * (A3(rrv, i, 2, k) - A3(rrv, i, 1, k))) * rdy * VH(1);
      }
  if (doitn)
    for (k = 2; k <= nk; k++)
      for (i = 1; i <= ni; i++) {
        float cc2 = 0.5f * (c2[k] + c2[k]), cc1 = 1.0f - cc2;
        float vbar = cc2 * A3(rrv, i, nj + 1, k) + cc1 * A3(rrv, i, nj + 1, k - 1);
        if (vbar <= 0.0f)
          A3(advy, i, nj, k) = A3(advy, i, nj, k) - A3(w3d, i, nj, k) *
              (c1[k] * (A3(rrv, i, nj + 1, k - 1) - A3(rrv, i, nj, k - 1)) + c2[k] * (A3(rrv, i, nj + 1, k) - A3(rrv, i, nj, k))) * rdy * VH(nj);
      }
  if (doweno) o_vadv_weno5(M, 4, rrw, dumz, w3d, 0, 0, weps);
  else vadv_flx5(M, 4, rrw, dumz, w3d);
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(advz, i, j, k) = -(A3(dumz, i, j, k) - A3(dumz, i, j, k - 1)) * rdz * mf[k];
  for (k = 2; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++)
        A3(wten, i, j, k) = A3(wten, i, j, k) +
            ((A3(advx, i, j, k) + A3(advy, i, j, k)) + A3(advz, i, j, k) +
             A3(w3d, i, j, k) * (c2[k] * A3(divx, i, j, k) + c1[k] * A3(divx, i, j, k - 1))) * rrf0[k];
}
