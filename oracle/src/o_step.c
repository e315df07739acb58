/* Oracle: time-step orchestration and C API (test infrastructure only).
 * solve1: src/solve1.F:414-454,1700-1780 (irdamp=0, no turbulence).
 * solve2: src/solve2.F:296-2052 (psolver=2 or 3, no terrain, Cartesian).
 * solve3: src/solve3.F:246-705.  getset: src/init3d.F:2032-2872.
 * Constants: src/constants.F:69-170; metrics: src/param.F:7790-8215,
 * :8590-8622, :8980-9060, :9182-9218; base state: src/base.F:2085-2198. */
#include "cm1o_internal.h"

int cm1o_libm_mode = 0;
void cm1o_set_libm(int mode) { cm1o_libm_mode = mode; }
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void set_constants(cm1o_const *K) {
  K->g = 9.81f; K->rd = 287.04f; K->cp = 1005.7f; K->cpv = 1870.0f; K->xlv = 2501000.0f;
  K->rv = 461.5f;
  K->cv = K->cp - K->rd;
  K->cvv = K->cpv - K->rv;
  K->cpdcv = K->cp / K->cv;
  K->rovcp = K->rd / K->cp;
  K->rddcv = K->rd / K->cv;
  K->cpdrd = K->cp / K->rd;
  K->eps = K->rd / K->rv;
  K->reps = K->rv / K->rd;
  K->repsm1 = K->rv / K->rd - 1.0f;
  K->cpl = 4190.0f;
  K->cpi = 2106.0f;
  K->lv1 = K->xlv + (K->cpl - K->cpv) * 273.15f;
  K->lv2 = K->cpl - K->cpv;
  K->p00 = 1.0e5f;
  K->rp00 = 1.0f / K->p00;
  K->th0r = 300.0f;
  K->smeps = 1.0e-30f;
  K->tsmall = 0.0001f;
  K->qsmall = 1.0e-14f;
  K->epsilon = 1.0e-18f;
}

cm1o_model *cm1o_create(const cm1o_config *cfg) {
  cm1o_model *M = (cm1o_model *)calloc(1, sizeof(cm1o_model));
  int i, j, k;
  M->c = *cfg;
  if (M->c.csmax == 0.0f) M->c.csmax = 350.0f;
  set_constants(&M->k);
  M->ni = cfg->ni; M->nj = cfg->nj; M->nk = cfg->nk;
  M->px = cfg->ni + 7; M->py = cfg->nj + 7; M->pz = cfg->nk + 3;
  M->pxy = M->px * M->py; M->ntot = M->pxy * M->pz;
  M->ibw = cfg->wbc != 1; M->ibe = cfg->ebc != 1; M->ibs = cfg->sbc != 1; M->ibn = cfg->nbc != 1;
  M->rdx = 1.0f / cfg->dx; M->rdy = 1.0f / cfg->dy; M->rdz = 1.0f / cfg->dz;
  for (i = 0; i < F_NFIELDS; i++) M->f[i] = (float *)calloc(M->ntot, sizeof(float));
  for (i = 0; i < K_NKFIELDS; i++) M->kf[i] = (float *)calloc(M->pz + 1, sizeof(float));
  for (i = 0; i < 3; i++) M->h[i] = (float *)calloc(M->px, sizeof(float));
  for (i = 3; i < 6; i++) M->h[i] = (float *)calloc(M->py, sizeof(float));
  for (i = 0; i < D_N2D; i++) M->d2[i] = (float *)calloc(M->pxy + M->px * M->pz + M->py * M->pz, sizeof(float));
  M->dtu = (float *)calloc(M->pxy, sizeof(float)); M->dtv = (float *)calloc(M->pxy, sizeof(float));
  M->dtu0 = (float *)calloc(M->pxy, sizeof(float)); M->dtv0 = (float *)calloc(M->pxy, sizeof(float));
  M->s1 = (float *)calloc(M->ntot, sizeof(float)); M->s2 = (float *)calloc(M->ntot, sizeof(float));
  M->s3 = (float *)calloc(M->ntot, sizeof(float)); M->s4 = (float *)calloc(M->ntot, sizeof(float));
  M->s5 = (float *)calloc(M->ntot, sizeof(float)); M->s6 = (float *)calloc(M->ntot, sizeof(float));
  M->s7 = (float *)calloc(M->ntot, sizeof(float)); M->s8 = (float *)calloc(M->ntot, sizeof(float));
  M->flag = (unsigned char *)calloc(M->ntot, 1);
  M->qmag[0] = M->qmag[1] = M->qmag[2] = 0.01f;
  /* roughness length, specified (src/init_surface.F:251) */
  if (M->c.cm1setup == 1)
    for (i = 0; i < M->pxy; i++) M->d2[D_ZNT][i] = fmaxf(1.0e-12f, cfg->cnst_znt);
  /* horizontal metrics of an unstretched grid: uh=dx/(xf(i+1)-xf(i)) = 1 */
  for (i = 0; i < M->px; i++) { M->h[0][i] = 1.0f; M->h[1][i] = 1.0f; M->h[2][i] = 1.0f; }
  for (j = 0; j < M->py; j++) { M->h[3][j] = 1.0f; M->h[4][j] = 1.0f; M->h[5][j] = 1.0f; }
  /* vertical metrics of an unstretched grid (src/param.F:7790-7810, 8140-8165, 8596-8615) */
  {
    const int nk = M->nk;
    float *zf = M->kf[K_ZF], *zh = M->kf[K_ZH], *mh = M->kf[K_MH], *mf = M->kf[K_MF];
    for (k = 0; k <= nk + 2; k++) { zf[k] = cfg->dz * (float)(k - 1); mf[k] = 1.0f; }
    for (k = 0; k <= nk + 1; k++) { zh[k] = 0.5f * (zf[k] + zf[k + 1]); mh[k] = 1.0f; }
    cm1o_set_extrap(M, NULL);
  }
  /* dtu0/dtv0: src/param.F:9182-9218 */
  for (j = 1; j <= M->nj; j++)
    for (i = 1; i <= M->ni + 1; i++) M->dtu0[IX2(M, i, j)] = 1.0f;
  if (M->ibw) for (j = 1; j <= M->nj; j++) M->dtu0[IX2(M, 1, j)] = 0.0f;
  if (M->ibe) for (j = 1; j <= M->nj; j++) M->dtu0[IX2(M, M->ni + 1, j)] = 0.0f;
  for (j = 1; j <= M->nj + 1; j++)
    for (i = 1; i <= M->ni; i++) M->dtv0[IX2(M, i, j)] = 1.0f;
  if (M->ibs) for (i = 1; i <= M->ni; i++) M->dtv0[IX2(M, i, 1)] = 0.0f;
  if (M->ibn) for (i = 1; i <= M->ni; i++) M->dtv0[IX2(M, i, M->nj + 1)] = 0.0f;
  return M;
}

/* Recompute rmh/rmf/c1/c2 and the 2nd-order extrapolation coefficients
 * from zh/zf/mh/mf (src/param.F:8590-8622, :8980-9060).  cgs6, if given,
 * overrides (cgs1..3, cgt1..3). */
void cm1o_set_extrap(cm1o_model *M, const float *cgs6) {
  const int nk = M->nk;
  float *zf = M->kf[K_ZF], *zh = M->kf[K_ZH], *mh = M->kf[K_MH], *mf = M->kf[K_MF];
  float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  int k;
  for (k = 0; k <= nk + 1; k++) M->kf[K_RMH][k] = 1.0f / mh[k];
  for (k = 0; k <= nk + 2; k++) M->kf[K_RMF][k] = 1.0f / mf[k];
  for (k = 1; k <= nk + 1; k++) {
    c2[k] = (zf[k] - zh[k - 1]) / (zh[k] - zh[k - 1]);
    c1[k] = 1.0f - c2[k];
  }
  /* Rayleigh-damping coefficients (src/param.F:8633-8668): zero below
   * zd, 0.5*(1-cos(pi*(z-zd)/(zf(nk+1)-zd))) above, with the host's
   * single-precision cos as the reference evaluates it. */
  {
    float *tauh = M->kf[K_TAUH], *tauf = M->kf[K_TAUF];
    const float pi = 3.1415926535897932384626433f, zd = M->c.zd, maxz = zf[nk + 1];
    for (k = 0; k <= nk + 2; k++) { tauh[k] = 0.0f; tauf[k] = 0.0f; }
    if (M->c.irdamp >= 1 && zd < maxz) {
      for (k = 1; k <= nk; k++)
        if (zh[k] > zd) tauh[k] = 0.5f * (1.0f - cosf(pi * (zh[k] - zd) / (zf[nk + 1] - zd)));
      for (k = 1; k <= nk + 1; k++)
        if (zf[k] > zd) tauf[k] = 0.5f * (1.0f - cosf(pi * (zf[k] - zd) / (zf[nk + 1] - zd)));
    }
  }
  if (cgs6) {
    M->cgs1 = cgs6[0]; M->cgs2 = cgs6[1]; M->cgs3 = cgs6[2];
    M->cgt1 = cgs6[3]; M->cgt2 = cgs6[4]; M->cgt3 = cgs6[5];
  } else {
    /* 2nd-order extrapolation weights to the surface zf(1) from zh(1..3)
     * and to the top zf(nk+1) from zh(nk..nk-2): Fornberg's recurrence as
     * src/param.F:8988-9050 runs it (nn = 1, 2), with x0 and alpha the
     * REAL quotients of the heights by the end-level thickness, the
     * recurrence in double precision, the m = 0 weights stored as REAL.
     * sigmaf(k) = zf(k) and sigma(k) = zh(k) for k = 1..nk+1 (no terrain). */
    for (int nn = 1; nn <= 2; nn++) {
      float dsf;
      double x0, al[3], w[3];
      if (nn == 1) {
        dsf = zf[2] - zf[1];
        x0 = (double)(zf[1] / dsf);
        al[0] = (double)(zh[1] / dsf); al[1] = (double)(zh[2] / dsf); al[2] = (double)(zh[3] / dsf);
      } else {
        dsf = zf[nk + 1] - zf[nk];
        x0 = (double)(zf[nk + 1] / dsf);
        al[0] = (double)(zh[nk] / dsf); al[1] = (double)(zh[nk - 1] / dsf); al[2] = (double)(zh[nk - 2] / dsf);
      }
      {
        /* delta(n,0,nu) for n = 1, 2 (src/param.F:9013-9034) */
        double b1 = 1.0, b2, b3, d10, d11;
        b2 = 1.0;
        b3 = al[1] - al[0]; b2 = b2 * b3;
        d10 = ((al[1] - x0) * 1.0 - 0.0) / b3;
        d11 = (b1 / b2) * (0.0 - (al[0] - x0) * 1.0);
        b1 = b2;
        b2 = 1.0;
        b3 = al[2] - al[0]; b2 = b2 * b3;
        w[0] = ((al[2] - x0) * d10 - 0.0) / b3;
        b3 = al[2] - al[1]; b2 = b2 * b3;
        w[1] = ((al[2] - x0) * d11 - 0.0) / b3;
        w[2] = (b1 / b2) * (0.0 - (al[1] - x0) * d11);
      }
      if (nn == 1) { M->cgs1 = (float)w[0]; M->cgs2 = (float)w[1]; M->cgs3 = (float)w[2]; }
      else { M->cgt1 = (float)w[0]; M->cgt2 = (float)w[1]; M->cgt3 = (float)w[2]; }
    }
  }
}

void cm1o_destroy(cm1o_model *M) {
  int i;
  if (!M) return;
  for (i = 0; i < F_NFIELDS; i++) free(M->f[i]);
  for (i = 0; i < K_NKFIELDS; i++) free(M->kf[i]);
  for (i = 0; i < 6; i++) free(M->h[i]);
  for (i = 0; i < D_N2D; i++) free(M->d2[i]);
  free(M->dtu); free(M->dtv); free(M->dtu0); free(M->dtv0);
  free(M->s1); free(M->s2); free(M->s3); free(M->s4); free(M->s5); free(M->s6);
  free(M->s7); free(M->s8); free(M->flag);
  free(M);
}

float *cm1o_field(cm1o_model *M, int id) { return (id >= 0 && id < F_NFIELDS) ? M->f[id] : NULL; }
float *cm1o_kfield(cm1o_model *M, int id) { return (id >= 0 && id < K_NKFIELDS) ? M->kf[id] : NULL; }
float *cm1o_2dfield(cm1o_model *M, int id) { return (id >= 0 && id < D_N2D) ? M->d2[id] : NULL; }
float *cm1o_hfield(cm1o_model *M, int which) { return (which >= 0 && which < 6) ? M->h[which] : NULL; }
void cm1o_dims(const cm1o_model *M, int64_t *px, int64_t *py, int64_t *pz) { *px = M->px; *py = M->py; *pz = M->pz; }
static double g_scalars[8];
double *cm1o_scalars(cm1o_model *M) {
  g_scalars[0] = M->mass1; g_scalars[1] = M->train; g_scalars[2] = M->cflmax; g_scalars[3] = M->mass2;
  return g_scalars;
}

/* derived base state from pi0, th0, qv0 (src/base.F:2085-2198) */
void cm1o_base_derive(cm1o_model *M) {
  const int nk = M->nk;
  const cm1o_const *K = &M->k;
  float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0], *qv0 = M->kf[K_QV0], *prs0 = M->kf[K_PRS0];
  float *rho0 = M->kf[K_RHO0], *rr0 = M->kf[K_RR0], *rf0 = M->kf[K_RF0], *rrf0 = M->kf[K_RRF0];
  float *thv0 = M->kf[K_THV0], *rth0 = M->kf[K_RTH0], *qc0 = M->kf[K_QC0], *qi0 = M->kf[K_QI0];
  float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  int k;
  for (k = 1; k <= nk; k++) {
    prs0[k] = K->p00 * powf(pi0[k], K->cpdrd);
    rho0[k] = prs0[k] / (K->rd * th0[k] * pi0[k] * (1.0f + qv0[k] * K->reps));
    if (M->c.imoist == 1)
      thv0[k] = (th0[k] + 0.0f) * (1.0f + K->reps * fmaxf(0.0f, qv0[k])) /
                (1.0f + fmaxf(0.0f, qv0[k]) + fmaxf(0.0f, qc0[k] + qi0[k]));
    else
      thv0[k] = th0[k];
  }
  /* extrapbcs(rho0) */
  rho0[0] = M->cgs1 * rho0[1] + M->cgs2 * rho0[2] + M->cgs3 * rho0[3];
  rho0[nk + 1] = M->cgt1 * rho0[nk] + M->cgt2 * rho0[nk - 1] + M->cgt3 * rho0[nk - 2];
  for (k = 0; k <= nk + 1; k++) rr0[k] = 1.0f / rho0[k];
  for (k = 2; k <= nk; k++) rf0[k] = c1[k] * rho0[k - 1] + c2[k] * rho0[k];
  rf0[1] = M->cgs1 * rho0[1] + M->cgs2 * rho0[2] + M->cgs3 * rho0[3];
  rf0[0] = rf0[1];
  rf0[nk + 1] = M->cgt1 * rho0[nk] + M->cgt2 * rho0[nk - 1] + M->cgt3 * rho0[nk - 2];
  for (k = 0; k <= nk + 1; k++) rrf0[k] = 1.0f / rf0[k];
  th0[0] = th0[1];
  th0[nk + 1] = th0[nk];
  for (k = 0; k <= nk + 1; k++) rth0[k] = 1.0f / th0[k];
}

/* ------------------------------------------------------------- getset */
void cm1o_getset(cm1o_model *M, int fresh_start) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0], *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  const float *rmh = M->kf[K_RMH];
  float *rho = M->f[F_RHO], *prs = M->f[F_PRS], *rr = M->f[F_RR], *rf = M->f[F_RF];
  int i, j, k, n;
  size_t t;
  o_bcu(M, M->f[F_UA]);
  o_bcv(M, M->f[F_VA]);
  o_bcw(M, M->f[F_WA], 1);
  o_bcs(M, M->f[F_PPI]);
  o_bcs(M, M->f[F_THA]);
  if (M->c.imoist == 1)
    for (n = 0; n < 3; n++) o_bcs(M, M->f[F_QA0 + n]);
  for (n = 0; n < M->c.npt; n++) o_bcs(M, M->f[F_PTA0 + n]);
  for (t = 0; t < (size_t)M->ntot; t++) {
    M->f[F_U3D][t] = M->f[F_UA][t];
    M->f[F_V3D][t] = M->f[F_VA][t];
    M->f[F_W3D][t] = M->f[F_WA][t];
    M->f[F_PP3D][t] = M->f[F_PPI][t];
    M->f[F_TH3D][t] = M->f[F_THA][t];
    for (n = 0; n < 3; n++) M->f[F_Q3D0 + n][t] = M->f[F_QA0 + n][t];
    for (n = 0; n < M->c.npt; n++) M->f[F_PT3D0 + n][t] = M->f[F_PTA0 + n][t];
  }
  if (M->c.cm1setup == 1) {
    /* subgrid TKE (src/init3d.F:2160, 2250-2266, 2322-2331): w-level
     * boundaries, tkea(1) = tkea(2), tke3d = tkea */
    float *tkea = M->f[F_TKEA];
    o_bcw(M, tkea, 1);
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) A3(tkea, i, j, 1) = A3(tkea, i, j, 2);
    o_bc2d(M, tkea + M->pxy);
    for (t = 0; t < (size_t)M->ntot; t++) M->f[F_TKE3D][t] = tkea[t];
  }
  if (fresh_start) {
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          float pp = A3(M->f[F_PP3D], i, j, k), th = A3(M->f[F_TH3D], i, j, k);
          A3(prs, i, j, k) = K->p00 * o_powf(pi0[k] + pp, K->cpdrd);
          if (M->c.imoist == 1)
            A3(rho, i, j, k) = A3(prs, i, j, k) /
                ((th0[k] + th) * (pi0[k] + pp) * (K->rd + fmaxf(0.0f, A3(M->f[F_Q3D0], i, j, k)) * K->rv));
          else
            A3(rho, i, j, k) = A3(prs, i, j, k) / (K->rd * (th0[k] + th) * (pi0[k] + pp));
        }
  }
  o_bcs(M, rho);
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(rr, i, j, k) = 1.0f / A3(rho, i, j, k);
        A3(rf, i, j, k) = c1[k] * A3(rho, i, j, k - 1) + c2[k] * A3(rho, i, j, k);
      }
  o_bcs(M, rr);
  o_bcs(M, rf);
  for (j = 0; j <= nj + 1; j++)
    for (i = 0; i <= ni + 1; i++) {
      A3(rf, i, j, 1) = M->cgs1 * A3(rho, i, j, 1) + M->cgs2 * A3(rho, i, j, 2) + M->cgs3 * A3(rho, i, j, 3);
      A3(rf, i, j, nk + 1) = M->cgt1 * A3(rho, i, j, nk) + M->cgt2 * A3(rho, i, j, nk - 1) + M->cgt3 * A3(rho, i, j, nk - 2);
    }
  if (fresh_start) {
    quad temq = 0;
    for (k = 1; k <= nk; k++) {
      double dk = 0.0;
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) dk = dk + (double)(A3(rho, i, j, k) * RUH(i) * RVH(j) * rmh[k]);
      temq = temq + (quad)dk;
    }
    M->mass1 = (double)temq;
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(M->f[F_PPX], i, j, k) = A3(M->f[F_PPI], i, j, k);
  }
  o_bcs(M, M->f[F_PPX]);
}

/* ------------------------------------------------------------- solve1 */
void o_solve1(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float dt = M->c.dt;
  const float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0];
  float *ppten = M->f[F_PPTEN], *thten1 = M->f[F_THTEN1];
  int i, j, k, n;
  if (M->c.irdamp == 1) {
    /* Rayleigh damping toward the base state, src/solve1.F:414-426
     * (no terrain: tauh, tauf, u0, v0 are columns; taus = tauh) */
    const float ra = M->c.rdalpha;
    const float *tauh = M->kf[K_TAUH], *tauf = M->kf[K_TAUF], *u0 = M->kf[K_U0], *v0 = M->kf[K_V0];
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          A3(M->f[F_UTEN1], i, j, k) = -ra * 0.5f * (tauh[k] + tauh[k]) * (A3(M->f[F_UA], i, j, k) - u0[k]);
          A3(M->f[F_VTEN1], i, j, k) = -ra * 0.5f * (tauh[k] + tauh[k]) * (A3(M->f[F_VA], i, j, k) - v0[k]);
          A3(M->f[F_WTEN1], i, j, k) = -ra * tauf[k] * A3(M->f[F_WA], i, j, k);
          A3(thten1, i, j, k) = -ra * tauh[k] * A3(M->f[F_THA], i, j, k);
        }
  } else {
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          A3(M->f[F_UTEN1], i, j, k) = 0.0f;
          A3(M->f[F_VTEN1], i, j, k) = 0.0f;
          A3(M->f[F_WTEN1], i, j, k) = 0.0f;
          A3(thten1, i, j, k) = 0.0f;
        }
  }
  if (M->c.imoist == 1)
    for (n = 0; n < 3; n++)
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) A3(M->f[F_QTEN0 + n], i, j, k) = 0.0f;
  for (n = 0; n < M->c.npt; n++)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(M->f[F_PTTEN0 + n], i, j, k) = 0.0f;
  /* subgrid turbulence tendencies (src/solve1.F:315-380, 1118-1240) */
  if (M->c.cm1setup == 1) o_turb_tend(M);
  {
    const float rdt = 1.0f / dt;
    const float tem = 0.0001f * K->tsmall;
    float *rho = M->f[F_RHO], *tha = M->f[F_THA], *ppi = M->f[F_PPI];
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          if (M->c.imoist == 1) {
            float qvt = A3(M->f[F_QTEN0], i, j, k);
            if (fabsf(dt * A3(thten1, i, j, k)) > tem || fabsf(dt * qvt) > K->qsmall) {
              float thnew = A3(tha, i, j, k) + dt * A3(thten1, i, j, k);
              float qvnew = A3(M->f[F_QA0], i, j, k) + dt * qvt;
              float pinew = o_powf(A3(rho, i, j, k) * (th0[k] + thnew) * (K->rd + K->rv * qvnew) * K->rp00, K->rddcv) - pi0[k];
              A3(ppten, i, j, k) = (pinew - A3(ppi, i, j, k)) * rdt;
            } else A3(ppten, i, j, k) = 0.0f;
          } else {
            if (fabsf(dt * A3(thten1, i, j, k)) > tem) {
              float thnew = A3(tha, i, j, k) + dt * A3(thten1, i, j, k);
              float pinew = o_powf(A3(rho, i, j, k) * (th0[k] + thnew) * K->rd * K->rp00, K->rddcv) - pi0[k];
              A3(ppten, i, j, k) = (pinew - A3(ppi, i, j, k)) * rdt;
            } else A3(ppten, i, j, k) = 0.0f;
          }
        }
    if (M->c.imoist == 1 && M->c.ptype >= 1)
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            A3(ppten, i, j, k) = A3(ppten, i, j, k) + A3(M->f[F_QPTEN], i, j, k);
            A3(thten1, i, j, k) = A3(thten1, i, j, k) + A3(M->f[F_QTTEN], i, j, k);
            A3(M->f[F_QTEN0], i, j, k) = A3(M->f[F_QTEN0], i, j, k) + A3(M->f[F_QVTEN], i, j, k);
            A3(M->f[F_QTEN1], i, j, k) = A3(M->f[F_QTEN1], i, j, k) + A3(M->f[F_QCTEN], i, j, k);
          }
  }
}

/* ------------------------------------------------------------- solve2 */
void o_solve2(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float dt = M->c.dt;
  const float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0], *rth0 = M->kf[K_RTH0], *qv0 = M->kf[K_QV0];
  const float *qc0 = M->kf[K_QC0], *qi0 = M->kf[K_QI0], *rho0 = M->kf[K_RHO0], *rf0 = M->kf[K_RF0];
  const float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2], *mh = M->kf[K_MH], *rmh = M->kf[K_RMH];
  float *u3d = M->f[F_U3D], *v3d = M->f[F_V3D], *w3d = M->f[F_W3D];
  float *uten = M->f[F_UTEN], *vten = M->f[F_VTEN], *wten = M->f[F_WTEN];
  float *rru = M->f[F_RRU], *rrv = M->f[F_RRV], *rrw = M->f[F_RRW], *divx = M->f[F_DIVX];
  float *th3d = M->f[F_TH3D], *tha = M->f[F_THA], *pp3d = M->f[F_PP3D], *sadv = M->f[F_SADV];
  float *thten = M->f[F_THTEN], *thterm = M->f[F_THTERM], *sten = M->f[F_STEN];
  float *dum2 = M->f[F_DUM2], *dum3 = M->f[F_DUM3], *dum8 = M->f[F_DUM8];
  float *t11 = M->f[F_T11], *t22 = M->f[F_T22];
  float *rho = M->f[F_RHO], *prs = M->f[F_PRS];
  int i, j, k, n, nrk;
  double dumk3[1024], dumk4[1024];
  int dopf = 0;

  for (nrk = 1; nrk <= 3; nrk++) {
    const float dttmp = dt / (float)(3 + 1 - nrk);

    if (M->c.imoist == 1) {
      float *qc = M->f[F_Q3D1], *qr = M->f[F_Q3D2];
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) A3(dum2, i, j, k) = A3(qc, i, j, k);
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) A3(dum2, i, j, k) = A3(dum2, i, j, k) + A3(qr, i, j, k);
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) A3(dum3, i, j, k) = 0.0f;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            float qv = fmaxf(A3(M->f[F_Q3D0], i, j, k), 0.0f);
            float qli = fmaxf(0.0f, A3(dum2, i, j, k) + A3(dum3, i, j, k));
            float cpli = K->cpl * fmaxf(0.0f, A3(dum2, i, j, k)) + K->cpi * fmaxf(0.0f, A3(dum3, i, j, k));
            A3(dum8, i, j, k) = K->g * (A3(th3d, i, j, k) * rth0[k] + K->repsm1 * (qv - qv0[k]) - (qli - qc0[k] - qi0[k]));
            A3(t11, i, j, k) = (th0[k] + A3(th3d, i, j, k)) * (1.0f + K->reps * qv) / (1.0f + qv + qli);
            float cpm = K->cp + K->cpv * qv + cpli;
            float cvm = 1.0f / (K->cv + K->cvv * qv + cpli);
            A3(thterm, i, j, k) = (th0[k] + A3(th3d, i, j, k)) * (K->rd + K->rv * qv - K->rovcp * cpm) * cvm;
            A3(t22, i, j, k) = (pi0[k] + A3(pp3d, i, j, k)) * K->rovcp * cpm * cvm;
          }
    } else {
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            A3(dum8, i, j, k) = K->g * A3(th3d, i, j, k) * rth0[k];
            A3(t11, i, j, k) = th0[k] + A3(th3d, i, j, k);
            A3(t22, i, j, k) = (pi0[k] + A3(pp3d, i, j, k)) * K->rddcv;
          }
    }
    o_bcp(M, t11);

    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj + 1; j++)
        for (i = 1; i <= ni + 1; i++) {
          A3(uten, i, j, k) = A3(M->f[F_UTEN1], i, j, k);
          A3(vten, i, j, k) = A3(M->f[F_VTEN1], i, j, k);
          A3(wten, i, j, k) = A3(M->f[F_WTEN1], i, j, k);
        }

    for (k = 1; k <= nk; k++) {
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 2; i++) A3(rru, i, j, k) = rho0[k] * A3(u3d, i, j, k);
      for (j = 0; j <= nj + 2; j++)
        for (i = 0; i <= ni + 1; i++) A3(rrv, i, j, k) = rho0[k] * A3(v3d, i, j, k);
      if (k == 1) {
        for (j = 0; j <= nj + 1; j++)
          for (i = 0; i <= ni + 1; i++) { A3(rrw, i, j, 1) = 0.0f; A3(rrw, i, j, nk + 1) = 0.0f; }
      } else {
        for (j = 0; j <= nj + 1; j++)
          for (i = 0; i <= ni + 1; i++) A3(rrw, i, j, k) = rf0[k] * A3(w3d, i, j, k);
      }
    }
    for (k = 1; k <= nk; k++)
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) {
          float d = ((A3(rru, i + 1, j, k) - A3(rru, i, j, k)) * M->rdx * UH(i)
                   + (A3(rrv, i, j + 1, k) - A3(rrv, i, j, k)) * M->rdy * VH(j))
                   + (A3(rrw, i, j, k + 1) - A3(rrw, i, j, k)) * M->rdz * mh[k];
          if (fabsf(d) < K->smeps) d = 0.0f;
          A3(divx, i, j, k) = d;
        }

    o_advu(M, nrk, rru, u3d, uten, rrv, rrw, divx);
    o_advv(M, nrk, rru, rrv, v3d, vten, rrw, divx);
    o_advw(M, nrk, rru, rrv, rrw, w3d, wten, divx);

    for (k = 2; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++)
          A3(wten, i, j, k) = A3(wten, i, j, k) + (c1[k] * A3(dum8, i, j, k - 1) + c2[k] * A3(dum8, i, j, k));

    for (k = 1; k <= nk; k++) {
      float tem = pi0[k];
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) A3(sadv, i, j, k) = tem + A3(pp3d, i, j, k);
    }

    if (M->c.psolver == 2) o_sounde(M, nrk, dttmp);
    else o_sound(M, nrk, dttmp);

    if (nrk == 3) o_calccflquick(M, dt, u3d, v3d, w3d);

    if (M->c.irbc == 4) {
      if (M->ibw == 1 || M->ibe == 1)
        o_radbcew4(M, M->d2[D_RADBCW], M->d2[D_RADBCE], M->f[F_UA], u3d, dttmp);
      if (M->ibs == 1 || M->ibn == 1)
        o_radbcns4(M, M->d2[D_RADBCS], M->d2[D_RADBCN], M->f[F_VA], v3d, dttmp);
    }

    if (M->c.imoist == 1) o_getdiv(M, rru, rrv, rrw, t11);
    o_getdivx(M, rru, rrv, rrw, divx);

    /* theta equation */
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          if (M->c.imoist == 1)
            A3(thten, i, j, k) = A3(M->f[F_THTEN1], i, j, k) - A3(t11, i, j, k) * A3(thterm, i, j, k);
          else
            A3(thten, i, j, k) = A3(M->f[F_THTEN1], i, j, k);
        }
    for (k = 1; k <= nk; k++) {
      float tem = th0[k] - K->th0r;
      for (j = -2; j <= nj + 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(sadv, i, j, k) = tem + A3(th3d, i, j, k);
    }
    o_advs(M, nrk, rru, rrv, rrw, divx, tha, sadv, thten, 0, 0, dttmp, (double)(10.0f * K->epsilon));
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(th3d, i, j, k) = A3(tha, i, j, k) + dttmp * A3(thten, i, j, k);
          if (fabsf(A3(th3d, i, j, k)) < K->smeps) A3(th3d, i, j, k) = 0.0f;
        }

    if (M->c.imoist == 1) {
      for (n = 0; n < 3; n++) {
        float *qa = M->f[F_QA0 + n], *q3d = M->f[F_Q3D0 + n], *qten = M->f[F_QTEN0 + n];
        for (k = 1; k <= nk; k++)
          for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) A3(sten, i, j, k) = A3(qten, i, j, k);
        int pdef = (nrk == 3) ? 1 : 0;
        double weps = (double)(M->qmag[n] * K->epsilon);
        o_advs(M, nrk, rru, rrv, rrw, divx, qa, q3d, sten, pdef, (n == 0) ? 0 : 1, dttmp, weps);
        for (k = 1; k <= nk; k++)
          for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
              A3(q3d, i, j, k) = A3(qa, i, j, k) + dttmp * A3(sten, i, j, k);
              if (fabsf(A3(q3d, i, j, k)) < K->smeps) A3(q3d, i, j, k) = 0.0f;
            }
      }
    }

    /* TKE advection and step (src/solve2.F:1709-1800) */
    if (M->c.cm1setup == 1) o_tke_step(M, nrk, dttmp);

    /* pressure and density */
    if (M->c.imoist == 1) {
      if (nrk == 3 && M->c.ptype >= 1) {
        float *qv = M->f[F_Q3D0], *qc = M->f[F_Q3D1];
        float *qpten = M->f[F_QPTEN], *qtten = M->f[F_QTTEN], *qvten = M->f[F_QVTEN], *qcten = M->f[F_QCTEN];
        for (k = 1; k <= nk; k++)
          for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
              A3(pp3d, i, j, k) = A3(pp3d, i, j, k) - dt * A3(qpten, i, j, k);
              A3(qpten, i, j, k) = A3(pp3d, i, j, k);
              A3(th3d, i, j, k) = A3(th3d, i, j, k) - dt * A3(qtten, i, j, k);
              A3(qtten, i, j, k) = A3(th3d, i, j, k);
              A3(qv, i, j, k) = A3(qv, i, j, k) - dt * A3(qvten, i, j, k);
              A3(qvten, i, j, k) = A3(qv, i, j, k);
              A3(qc, i, j, k) = A3(qc, i, j, k) - dt * A3(qcten, i, j, k);
              A3(qcten, i, j, k) = A3(qc, i, j, k);
            }
      }
    }
    if (nrk == 3) {
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            A3(prs, i, j, k) = K->p00 * o_powf(pi0[k] + A3(pp3d, i, j, k), K->cpdrd);
            if (M->c.imoist == 1)
              A3(rho, i, j, k) = A3(prs, i, j, k) / ((th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k)) *
                                                      (K->rd + fmaxf(0.0f, A3(M->f[F_Q3D0], i, j, k)) * K->rv));
            else
              A3(rho, i, j, k) = A3(prs, i, j, k) / (K->rd * (th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k)));
          }
      if (M->c.apmasscon == 1) {
        for (k = 1; k <= nk; k++) {
          dumk3[k] = 0.0; dumk4[k] = 0.0;
          for (j = 1; j <= nj; j++)
            for (i = 1; i <= ni; i++) {
              dumk3[k] = dumk3[k] + (double)(A3(rho, i, j, k) * RUH(i) * RVH(j) * rmh[k]);
              dumk4[k] = dumk4[k] + (double)(pi0[k] + A3(pp3d, i, j, k));
            }
        }
        dopf = 1;
      }
    }

    o_bcu(M, u3d);
    o_bcv(M, v3d);
    o_bcw(M, w3d, 1);
    if (nrk < 3) {
      o_bcp(M, rho);
      o_bcp(M, pp3d);
      o_bcs(M, th3d);
      if (M->c.imoist == 1)
        for (n = 0; n < 3; n++) o_bcs(M, M->f[F_Q3D0 + n]);
    }

    /* passive tracers: src/solve2.F:1902-1970 */
    for (n = 0; n < M->c.npt; n++) {
      float *pta = M->f[F_PTA0 + n], *pt3d = M->f[F_PT3D0 + n], *ptten = M->f[F_PTTEN0 + n];
      int pdef = (nrk == 3 && M->c.pdtra == 1) ? 1 : 0;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) A3(sten, i, j, k) = A3(ptten, i, j, k);
      o_advs(M, nrk, rru, rrv, rrw, divx, pta, pt3d, sten, pdef, 1, dttmp, (double)(1.0f * K->epsilon));
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) A3(pt3d, i, j, k) = A3(pta, i, j, k) + dttmp * A3(sten, i, j, k);
      if (nrk <= 2) o_bcs(M, pt3d);
    }

    if (dopf && nrk == 3) {
      quad temq1 = 0, temq2 = 0;
      for (k = 1; k <= nk; k++) { temq1 = temq1 + (quad)dumk3[k]; temq2 = temq2 + (quad)dumk4[k]; }
      double mass2 = (double)temq1, p2 = (double)temq2;
      M->mass2 = mass2;
      double ntot = (double)((long long)M->ni * (long long)M->nj * (long long)M->nk);
      float tem = (float)((pow(M->mass1 / mass2, (double)K->rddcv) - 1.0) * p2 / ntot);
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            A3(pp3d, i, j, k) = A3(pp3d, i, j, k) + tem;
            A3(prs, i, j, k) = K->p00 * o_powf(pi0[k] + A3(pp3d, i, j, k), K->cpdrd);
            if (M->c.imoist == 1)
              A3(rho, i, j, k) = A3(prs, i, j, k) / ((th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k)) *
                                                      (K->rd + fmaxf(0.0f, A3(M->f[F_Q3D0], i, j, k)) * K->rv));
            else
              A3(rho, i, j, k) = A3(prs, i, j, k) / (K->rd * (th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k)));
          }
    }
  }

  for (n = 0; n < M->c.npt; n++) {
    if (M->c.pdtra == 1) o_pdefq(M, 0.0f, rho, M->f[F_PT3D0 + n]);
    o_bcs(M, M->f[F_PT3D0 + n]);
  }
}

/* ------------------------------------------------------------- solve3 */
void o_solve3(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float dt = M->c.dt;
  const float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  float *pp3d = M->f[F_PP3D], *th3d = M->f[F_TH3D], *ppx = M->f[F_PPX], *rho = M->f[F_RHO];
  float *rr = M->f[F_RR], *rf = M->f[F_RF];
  int i, j, k, n;
  o_bcp(M, pp3d);
  o_bcs(M, th3d);
  if (M->c.imoist == 1)
    for (n = 0; n < 3; n++) o_bcs(M, M->f[F_Q3D0 + n]);
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) A3(ppx, i, j, k) = A3(pp3d, i, j, k) + A3(ppx, i, j, k);
  o_bcp(M, ppx);
  o_bcp(M, rho);
  if (M->c.imoist == 1 && M->c.ptype >= 1) {
    const float rdt = 1.0f / dt;
    float *qpten = M->f[F_QPTEN], *qtten = M->f[F_QTTEN], *qvten = M->f[F_QVTEN], *qcten = M->f[F_QCTEN];
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(qpten, i, j, k) = (A3(pp3d, i, j, k) - A3(qpten, i, j, k)) * rdt;
          A3(qtten, i, j, k) = (A3(th3d, i, j, k) - A3(qtten, i, j, k)) * rdt;
          A3(qvten, i, j, k) = (A3(M->f[F_Q3D0], i, j, k) - A3(qvten, i, j, k)) * rdt;
          A3(qcten, i, j, k) = (A3(M->f[F_Q3D1], i, j, k) - A3(qcten, i, j, k)) * rdt;
        }
  }
  for (k = 1; k <= nk; k++)
    for (j = 0; j <= nj + 2; j++)
      for (i = 0; i <= ni + 2; i++) {
        A3(M->f[F_UA], i, j, k) = A3(M->f[F_U3D], i, j, k);
        A3(M->f[F_VA], i, j, k) = A3(M->f[F_V3D], i, j, k);
        A3(M->f[F_WA], i, j, k) = A3(M->f[F_W3D], i, j, k);
      }
  for (n = 0; n < M->c.npt; n++)
    for (k = 1; k <= nk; k++)
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) A3(M->f[F_PTA0 + n], i, j, k) = A3(M->f[F_PT3D0 + n], i, j, k);
  if (M->c.cm1setup == 1) o_tke_copy(M);
  for (k = 1; k <= nk; k++)
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) {
        A3(M->f[F_PPI], i, j, k) = A3(pp3d, i, j, k);
        A3(M->f[F_THA], i, j, k) = A3(th3d, i, j, k);
      }
  if (M->c.imoist == 1)
    for (n = 0; n < 3; n++)
      for (k = 1; k <= nk; k++)
        for (j = 0; j <= nj + 1; j++)
          for (i = 0; i <= ni + 1; i++) A3(M->f[F_QA0 + n], i, j, k) = A3(M->f[F_Q3D0 + n], i, j, k);
  for (k = 2; k <= nk; k++)
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) {
        A3(rr, i, j, k) = 1.0f / A3(rho, i, j, k);
        A3(rf, i, j, k) = c1[k] * A3(rho, i, j, k - 1) + c2[k] * A3(rho, i, j, k);
      }
  for (j = 0; j <= nj + 1; j++)
    for (i = 0; i <= ni + 1; i++) {
      A3(rr, i, j, 1) = 1.0f / A3(rho, i, j, 1);
      A3(rf, i, j, 1) = M->cgs1 * A3(rho, i, j, 1) + M->cgs2 * A3(rho, i, j, 2) + M->cgs3 * A3(rho, i, j, 3);
      A3(rf, i, j, nk + 1) = M->cgt1 * A3(rho, i, j, nk) + M->cgt2 * A3(rho, i, j, nk - 1) + M->cgt3 * A3(rho, i, j, nk - 2);
    }
}

void cm1o_step(cm1o_model *M) {
  if (M->c.cm1setup == 1) o_sfc_and_turb(M);
  o_solve1(M);
  o_solve2(M);
  if (M->c.imoist == 1) o_mp_driver(M);
  o_solve3(M);
}

/* ----------------------------------------------------- unit entry points */
void cm1o_solve1(cm1o_model *M) { o_solve1(M); }
void cm1o_sfc_and_turb(cm1o_model *M) { o_sfc_and_turb(M); }
void cm1o_solve2(cm1o_model *M) { o_solve2(M); }
void cm1o_mp_driver(cm1o_model *M) { if (M->c.imoist == 1) o_mp_driver(M); }
void cm1o_solve3(cm1o_model *M) { o_solve3(M); }
void cm1o_bc(cm1o_model *M, int id, int kind) {
  float *a = M->f[id];
  switch (kind) {
    case 0: o_bcs(M, a); break;
    case 1: o_bcp(M, a); break;
    case 2: o_bcu(M, a); break;
    case 3: o_bcv(M, a); break;
    case 4: o_bcw(M, a, 1); break;
    default: o_bcw(M, a, 0); break;
  }
}
void cm1o_advs(cm1o_model *M, int nrk, int s_id, int s0_id, int sten_id, int pdef, int pdefweno, double weps) {
  const float dttmp = M->c.dt / (float)(3 + 1 - nrk);
  o_advs(M, nrk, M->f[F_RRU], M->f[F_RRV], M->f[F_RRW], M->f[F_DIVX], M->f[s0_id], M->f[s_id],
         M->f[sten_id], pdef, pdefweno, dttmp, weps);
}
void cm1o_advu(cm1o_model *M, int nrk) { o_advu(M, nrk, M->f[F_RRU], M->f[F_U3D], M->f[F_UTEN], M->f[F_RRV], M->f[F_RRW], M->f[F_DIVX]); }
void cm1o_advv(cm1o_model *M, int nrk) { o_advv(M, nrk, M->f[F_RRU], M->f[F_RRV], M->f[F_V3D], M->f[F_VTEN], M->f[F_RRW], M->f[F_DIVX]); }
void cm1o_advw(cm1o_model *M, int nrk) { o_advw(M, nrk, M->f[F_RRU], M->f[F_RRV], M->f[F_RRW], M->f[F_W3D], M->f[F_WTEN], M->f[F_DIVX]); }
void cm1o_sound(cm1o_model *M, int nrk, float dttmp) {
  if (M->c.psolver == 2) o_sounde(M, nrk, dttmp);
  else o_sound(M, nrk, dttmp);
}
void cm1o_kessler_branch(cm1o_model *M) { o_mp_driver(M); }
void cm1o_getdiv(cm1o_model *M, int div_id) {
  o_getdiv(M, M->f[F_RRU], M->f[F_RRV], M->f[F_RRW], M->f[div_id]);
}
void cm1o_getdivx(cm1o_model *M) { o_getdivx(M, M->f[F_RRU], M->f[F_RRV], M->f[F_RRW], M->f[F_DIVX]); }
void cm1o_calccflquick(cm1o_model *M, float dt, int u_id, int v_id, int w_id) {
  o_calccflquick(M, dt, M->f[u_id], M->f[v_id], M->f[w_id]);
}
void cm1o_pdefq(cm1o_model *M, float rmax, int q_id) { o_pdefq(M, rmax, M->f[F_RHO], M->f[q_id]); }
