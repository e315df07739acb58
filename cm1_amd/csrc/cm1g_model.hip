// cm1gpu model: device residency, C ABI and the large-step orchestration.
// solve1: src/solve1.F:414-454,1700-1780; solve2: src/solve2.F:296-2052;
// solve3: src/solve3.F:246-705; getset: src/init3d.F:2032-2872;
// constants: src/constants.F:69-170; metrics: src/param.F:7790-8215,
// :8590-8622, :8980-9060, :9182-9218; base state: src/base.F:2085-2198.
#include "cm1g_internal.h"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

static thread_local std::string g_err;
void cm1g_set_error(const std::string &s) { g_err = s; }
extern "C" const char *cm1g_last_error(void) { return g_err.c_str(); }

#define ARGCHK(cond, msg)              \
  do {                                 \
    if (!(cond)) {                     \
      cm1g_set_error(msg);             \
      return CM1G_E_ARG;               \
    }                                  \
  } while (0)
// fields read or written by id see the 3-D arrays equal the a-arrays
#define UNSWAP(m)                      \
  do {                                 \
    const int rc_ = unswap(m);         \
    if (rc_) return rc_;               \
  } while (0)

static void set_constants(Consts &K) {
  K.g = 9.81f; K.rd = 287.04f; K.cp = 1005.7f; K.cpv = 1870.0f; K.xlv = 2501000.0f;
  K.rv = 461.5f;
  K.cv = K.cp - K.rd;
  K.cvv = K.cpv - K.rv;
  K.cpdcv = K.cp / K.cv;
  K.rovcp = K.rd / K.cp;
  K.rddcv = K.rd / K.cv;
  K.cpdrd = K.cp / K.rd;
  K.eps = K.rd / K.rv;
  K.reps = K.rv / K.rd;
  K.repsm1 = K.rv / K.rd - 1.0f;
  K.cpl = 4190.0f;
  K.cpi = 2106.0f;
  K.lv1 = K.xlv + (K.cpl - K.cpv) * 273.15f;
  K.lv2 = K.cpl - K.cpv;
  K.p00 = 1.0e5f;
  K.rp00 = 1.0f / K.p00;
  K.th0r = 300.0f;
  K.smeps = 1.0e-30f;
  K.tsmall = 0.0001f;
  K.qsmall = 1.0e-14f;
  K.epsilon = 1.0e-18f;
}

int dalloc(float **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(float));
  if (e != hipSuccess) {
    cm1g_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return CM1G_E_NOMEM;
  }
  e = hipMemset(*p, 0, n * sizeof(float));
  if (e != hipSuccess) {
    cm1g_set_error(std::string("hipMemset: ") + hipGetErrorString(e));
    return CM1G_E_HIP;
  }
  return 0;
}

static int upload_column(cm1g_model *m, int id) {
  HIPCHK(hipMemcpy(m->kf[id], m->hk[id], (size_t)(m->c.nk + 3) * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_dtsmall(const cm1g_config *c, float min_dz, int32_t *nsound) {
  ARGCHK(c && nsound, "cm1g_dtsmall: null argument");
  float dtsm;
  if (c->psolver == 2) {
    // min( min_dx , min_dy , min_dz ): the smallest level thickness on a
    // stretched grid (src/param.F:8522-8541); the namelist dz on a uniform one
    const float mdz = (min_dz > 0.0f) ? min_dz : c->dz;
    float mn = fminf(fminf(c->dx, c->dy), mdz);
    dtsm = 0.50f * mn / 360.0f;
  } else {
    dtsm = 0.60f * fminf(c->dx, c->dy) / 360.0f;
  }
  double dbldt = (double)c->dt;
  int ns = (int)lround(dbldt / (double)dtsm);
  if (ns < 4) ns = 4;
  if (ns % 2 != 0) ns = ns + 1;
  if (dbldt / (double)(float)ns > (double)dtsm) ns = ns + 2;
  if (ns > 192) {
    cm1g_set_error("cm1g_dtsmall: nsound > 192 (src/misclibs.F:887-909)");
    return CM1G_E_ARG;
  }
  *nsound = ns;
  return 0;
}

extern "C" int cm1g_create(const cm1g_config *cfg, cm1g_model **out) {
  ARGCHK(cfg && out, "cm1g_create: null argument");
  ARGCHK(cfg->ni >= 8 && cfg->nj >= 8 && cfg->nk >= 6, "cm1g_create: grid too small (ni,nj >= 8, nk >= 6)");
  ARGCHK(cfg->ni <= 8192 && cfg->nj <= 8192 && cfg->nk <= 510, "cm1g_create: grid too large");
  ARGCHK((cfg->wbc == 1 || cfg->wbc == 2) && (cfg->ebc == 1 || cfg->ebc == 2) &&
             (cfg->sbc == 1 || cfg->sbc == 2) && (cfg->nbc == 1 || cfg->nbc == 2),
         "cm1g_create: lateral boundaries must be 1 (periodic) or 2 (open)");
  ARGCHK((cfg->wbc == 1) == (cfg->ebc == 1) && (cfg->sbc == 1) == (cfg->nbc == 1),
         "cm1g_create: periodic boundaries must be set on both sides");
  ARGCHK(cfg->psolver == 2 || cfg->psolver == 3,
         "cm1g_create: psolver must be 2 (sounde.F) or 3 (sound.F)");
  ARGCHK(cfg->advwenos >= 0 && cfg->advwenos <= 2 && cfg->advwenov >= 0 && cfg->advwenov <= 2,
         "cm1g_create: advwenos/advwenov must be 0, 1 or 2 (weno_order=5)");
  ARGCHK(cfg->imoist == 0 || cfg->imoist == 1, "cm1g_create: imoist must be 0 or 1");
  ARGCHK(cfg->ptype == 0 || cfg->ptype == 1, "cm1g_create: ptype must be 0 or 1 (Kessler)");
  ARGCHK(cfg->nsound >= 1 && cfg->nsound <= 192, "cm1g_create: nsound out of range");
  ARGCHK(cfg->npt >= 0 && cfg->npt <= 8, "cm1g_create: npt must be 0..8");
  ARGCHK(cfg->dx > 0 && cfg->dy > 0 && cfg->dz > 0 && cfg->dt > 0, "cm1g_create: dx,dy,dz,dt must be > 0");
  ARGCHK(cfg->irdamp == 0 || cfg->irdamp == 1,
         "cm1g_create: irdamp must be 0 or 1 (damping toward the base state)");
  cm1g_model *m = new cm1g_model();  // value-initialized: pointers null, counters zero
  m->c = *cfg;
  if (m->c.csmax == 0.0f) m->c.csmax = 350.0f;
  m->tile_rows = 4;
  KP &P = m->P;
  P.ni = cfg->ni; P.nj = cfg->nj; P.nk = cfg->nk;
  P.px = ((cfg->ni + CM1G_IOFF + 5) + 63) / 64 * 64;
  P.py = cfg->nj + 7;
  P.pxy = P.px * P.py;
  P.wbc = cfg->wbc; P.ebc = cfg->ebc; P.sbc = cfg->sbc; P.nbc = cfg->nbc;
  P.ibw = cfg->wbc != 1; P.ibe = cfg->ebc != 1; P.ibs = cfg->sbc != 1; P.ibn = cfg->nbc != 1;
  P.dx = cfg->dx; P.dy = cfg->dy; P.dz = cfg->dz;
  P.rdx = 1.0f / cfg->dx; P.rdy = 1.0f / cfg->dy; P.rdz = 1.0f / cfg->dz;
  // rdalpha = 0 without damping (src/param.F:724)
  P.irdamp = cfg->irdamp;
  P.rdalpha = cfg->irdamp ? cfg->rdalpha : 0.0f;
  set_constants(P.K);
  const int pz = cfg->nk + 3;
  m->ntot = (size_t)P.pxy * pz;
  int rc = 0;
  if (hipStreamCreate(&m->stream) != hipSuccess) {
    cm1g_set_error("hipStreamCreate failed");
    delete m;
    return CM1G_E_HIP;
  }
  for (int i = 0; i < CM1G_NFIELDS && !rc; i++) {
    bool need = true;
    if (i >= CM1G_PTA0 && i < CM1G_PTTEN0 + 8) {
      int t = (i - CM1G_PTA0) % 8;
      need = t < cfg->npt;
    }
    if (i >= CM1G_TKEA) need = false;  // cm1g_set_turbulence allocates these
    if (need) rc = dalloc(&m->f[i], m->ntot);
  }
  for (int i = 0; i < CM1G_K_NFIELDS && !rc; i++) {
    rc = dalloc(&m->kf[i], pz + 1);
    m->hk[i] = (float *)calloc(pz + 1, sizeof(float));
  }
  for (int i = 0; i < 6 && !rc; i++) rc = dalloc(&m->h[i], (i < 3 ? P.px : P.py));
  for (int i = 0; i < CM1G_D_NFIELDS && !rc; i++) rc = dalloc(&m->d2[i], (size_t)P.pxy + (size_t)(P.px + P.py) * pz);
  if (!rc) rc = dalloc(&m->scoef, 8 * (size_t)pz);
  if (!rc) rc = dalloc(&m->dtu0, P.pxy);
  if (!rc) rc = dalloc(&m->dtv0, P.pxy);
  float **scr[] = {&m->sx, &m->sy, &m->sz, &m->tx, &m->ty, &m->tz,
                   &m->scr1, &m->scr2, &m->thb};
  for (auto p : scr)
    if (!rc) rc = dalloc(p, m->ntot);
  if (!rc) {
    size_t nd = (size_t)cfg->ni * cfg->nj + (size_t)cfg->nj + 2 * 32 * (size_t)pz + 4 * (size_t)pz + 64;
    if (hipMalloc((void **)&m->dred, nd * sizeof(double)) != hipSuccess) rc = CM1G_E_NOMEM;
    if (!rc && hipMalloc((void **)&m->dscal, 16 * sizeof(double)) != hipSuccess) rc = CM1G_E_NOMEM;
    if (!rc && hipMalloc((void **)&m->fred, 32 * (pz + 8) * sizeof(float)) != hipSuccess) rc = CM1G_E_NOMEM;
    if (!rc && hipMalloc((void **)&m->iflag, 16 * sizeof(int)) != hipSuccess) rc = CM1G_E_NOMEM;
    if (!rc && hipMalloc((void **)&m->dcount, 8 * sizeof(unsigned long long)) != hipSuccess) rc = CM1G_E_NOMEM;
    if (!rc && hipMalloc((void **)&m->colflag, (size_t)cfg->ni * cfg->nj * sizeof(int)) != hipSuccess)
      rc = CM1G_E_NOMEM;
    if (!rc) {
      hipMemset(m->dscal, 0, 16 * sizeof(double));
      hipMemset(m->iflag, 0, 16 * sizeof(int));
      hipMemset(m->dcount, 0, 8 * sizeof(unsigned long long));
      hipMemset(m->colflag, 0, (size_t)cfg->ni * cfg->nj * sizeof(int));
    }
    if (rc) cm1g_set_error("device allocation failed");
  }
  if (rc) {
    cm1g_destroy(m);
    return rc;
  }
  // horizontal metrics of an unstretched grid (uh = dx/(xf(i+1)-xf(i)) = 1)
  {
    std::vector<float> one((size_t)(P.px > P.py ? P.px : P.py), 1.0f);
    for (int i = 0; i < 6; i++)
      hipMemcpy(m->h[i], one.data(), (i < 3 ? P.px : P.py) * sizeof(float), hipMemcpyHostToDevice);
  }
  // vertical metrics of an unstretched grid (src/param.F:7790-7810)
  {
    const int nk = cfg->nk;
    float *zf = m->hk[CM1G_K_ZF], *zh = m->hk[CM1G_K_ZH], *mh = m->hk[CM1G_K_MH], *mf = m->hk[CM1G_K_MF];
    for (int k = 0; k <= nk + 2; k++) { zf[k] = cfg->dz * (float)(k - 1); mf[k] = 1.0f; }
    for (int k = 0; k <= nk + 1; k++) { zh[k] = 0.5f * (zf[k] + zf[k + 1]); mh[k] = 1.0f; }
  }
  // open-boundary masks (src/param.F:9182-9218)
  {
    std::vector<float> du((size_t)P.pxy, 0.0f), dv((size_t)P.pxy, 0.0f);
    for (int j = 1; j <= P.nj; j++)
      for (int i = 1; i <= P.ni + 1; i++) du[IX2D(P, i, j)] = 1.0f;
    if (P.ibw) for (int j = 1; j <= P.nj; j++) du[IX2D(P, 1, j)] = 0.0f;
    if (P.ibe) for (int j = 1; j <= P.nj; j++) du[IX2D(P, P.ni + 1, j)] = 0.0f;
    for (int j = 1; j <= P.nj + 1; j++)
      for (int i = 1; i <= P.ni; i++) dv[IX2D(P, i, j)] = 1.0f;
    if (P.ibs) for (int i = 1; i <= P.ni; i++) dv[IX2D(P, i, 1)] = 0.0f;
    if (P.ibn) for (int i = 1; i <= P.ni; i++) dv[IX2D(P, i, P.nj + 1)] = 0.0f;
    hipMemcpy(m->dtu0, du.data(), du.size() * sizeof(float), hipMemcpyHostToDevice);
    hipMemcpy(m->dtv0, dv.data(), dv.size() * sizeof(float), hipMemcpyHostToDevice);
  }
  P.mh = m->kf[CM1G_K_MH]; P.rmh = m->kf[CM1G_K_RMH]; P.mf = m->kf[CM1G_K_MF];
  P.c1 = m->kf[CM1G_K_C1]; P.c2 = m->kf[CM1G_K_C2]; P.pi0 = m->kf[CM1G_K_PI0];
  P.th0 = m->kf[CM1G_K_TH0]; P.qv0 = m->kf[CM1G_K_QV0]; P.rho0 = m->kf[CM1G_K_RHO0];
  P.rr0 = m->kf[CM1G_K_RR0]; P.rf0 = m->kf[CM1G_K_RF0]; P.rrf0 = m->kf[CM1G_K_RRF0];
  P.thv0 = m->kf[CM1G_K_THV0]; P.rth0 = m->kf[CM1G_K_RTH0]; P.qc0 = m->kf[CM1G_K_QC0];
  P.qi0 = m->kf[CM1G_K_QI0]; P.u0 = m->kf[CM1G_K_U0]; P.v0 = m->kf[CM1G_K_V0];
  P.tauh = m->kf[CM1G_K_TAUH]; P.tauf = m->kf[CM1G_K_TAUF];
  P.uh = m->h[0]; P.ruh = m->h[1]; P.uf = m->h[2]; P.vh = m->h[3]; P.rvh = m->h[4]; P.vf = m->h[5];
  rc = cm1g_metrics(m);
  if (rc) {
    cm1g_destroy(m);
    return rc;
  }
  *out = m;
  return 0;
}

extern "C" int cm1g_destroy(cm1g_model *m) {
  if (!m) return 0;
  if (m->stream) hipStreamSynchronize(m->stream);
  graphs_clear(m);
  for (int i = 0; i < CM1G_NFIELDS; i++) if (m->f[i]) hipFree(m->f[i]);
  for (int i = 0; i < CM1G_K_NFIELDS; i++) { if (m->kf[i]) hipFree(m->kf[i]); free(m->hk[i]); }
  for (int i = 0; i < 6; i++) if (m->h[i]) hipFree(m->h[i]);
  for (int i = 0; i < CM1G_D_NFIELDS; i++) if (m->d2[i]) hipFree(m->d2[i]);
  if (m->tcol) hipFree(m->tcol);
  if (m->pin) hipHostFree(m->pin);
  float *scr[] = {m->scoef, m->dtu0, m->dtv0, m->sx, m->sy, m->sz, m->tx, m->ty, m->tz,
                  m->scr1, m->scr2, m->thb, m->fred};
  for (auto p : scr) if (p) hipFree(p);
  if (m->dred) hipFree(m->dred);
  if (m->dscal) hipFree(m->dscal);
  if (m->iflag) hipFree(m->iflag);
  if (m->dcount) hipFree(m->dcount);
  if (m->colflag) hipFree(m->colflag);
  if (m->actlist) hipFree(m->actlist);
  io_free(m);
  for (hipEvent_t e : m->kt_ev) hipEventDestroy(e);
  if (m->stream) hipStreamDestroy(m->stream);
  delete m;
  return 0;
}

extern "C" int cm1g_dims(const cm1g_model *m, int64_t *px, int64_t *py, int64_t *pz) {
  ARGCHK(m && px && py && pz, "cm1g_dims: null argument");
  *px = m->P.px; *py = m->P.py; *pz = m->P.nk + 3;
  return 0;
}

extern "C" int cm1g_put_column(cm1g_model *m, int id, const float *host) {
  ARGCHK(m && host && id >= 0 && id < CM1G_K_NFIELDS, "cm1g_put_column: bad argument");
  memcpy(m->hk[id], host, (size_t)(m->c.nk + 3) * sizeof(float));
  return upload_column(m, id);
}

extern "C" int cm1g_get_column(cm1g_model *m, int id, float *host) {
  ARGCHK(m && host && id >= 0 && id < CM1G_K_NFIELDS, "cm1g_get_column: bad argument");
  memcpy(host, m->hk[id], (size_t)(m->c.nk + 3) * sizeof(float));
  return 0;
}

extern "C" int cm1g_metrics(cm1g_model *m) {
  ARGCHK(m, "cm1g_metrics: null model");
  // the metric scalars are launch arguments of the captured steps
  graphs_clear(m);
  const int nk = m->c.nk;
  float *zf = m->hk[CM1G_K_ZF], *zh = m->hk[CM1G_K_ZH], *mh = m->hk[CM1G_K_MH], *mf = m->hk[CM1G_K_MF];
  float *c1 = m->hk[CM1G_K_C1], *c2 = m->hk[CM1G_K_C2];
  for (int k = 0; k <= nk + 1; k++) m->hk[CM1G_K_RMH][k] = 1.0f / mh[k];
  for (int k = 0; k <= nk + 2; k++) m->hk[CM1G_K_RMF][k] = 1.0f / mf[k];
  for (int k = 1; k <= nk + 1; k++) {
    c2[k] = (zf[k] - zh[k - 1]) / (zh[k] - zh[k - 1]);
    c1[k] = 1.0f - c2[k];
  }
  // three-point extrapolation weights to zf(1) from zh(1..3) and to
  // zf(nk+1) from zh(nk..nk-2): Fornberg's recurrence as
  // src/param.F:8988-9050 runs it (nn = 1, 2): x0 and alpha are REAL
  // quotients of the heights by the end-level thickness, the recurrence
  // is double precision, the m = 0 weights are stored as REAL.
  float w6[6];
  for (int nn = 0; nn < 2; nn++) {
    float dsf;
    double x0, al[3];
    if (nn == 0) {
      dsf = zf[2] - zf[1];
      x0 = (double)(zf[1] / dsf);
      al[0] = (double)(zh[1] / dsf); al[1] = (double)(zh[2] / dsf); al[2] = (double)(zh[3] / dsf);
    } else {
      dsf = zf[nk + 1] - zf[nk];
      x0 = (double)(zf[nk + 1] / dsf);
      al[0] = (double)(zh[nk] / dsf); al[1] = (double)(zh[nk - 1] / dsf); al[2] = (double)(zh[nk - 2] / dsf);
    }
    // delta(n,0,nu) for n = 1, 2 (src/param.F:9013-9034)
    double b1 = 1.0, b2 = 1.0, b3;
    b3 = al[1] - al[0]; b2 = b2 * b3;
    const double d10 = ((al[1] - x0) * 1.0 - 0.0) / b3;
    const double d11 = (b1 / b2) * (0.0 - (al[0] - x0) * 1.0);
    b1 = b2;
    b2 = 1.0;
    b3 = al[2] - al[0]; b2 = b2 * b3;
    w6[nn * 3 + 0] = (float)(((al[2] - x0) * d10 - 0.0) / b3);
    b3 = al[2] - al[1]; b2 = b2 * b3;
    w6[nn * 3 + 1] = (float)(((al[2] - x0) * d11 - 0.0) / b3);
    w6[nn * 3 + 2] = (float)((b1 / b2) * (0.0 - (al[1] - x0) * d11));
  }
  m->P.cgs1 = w6[0]; m->P.cgs2 = w6[1]; m->P.cgs3 = w6[2];
  m->P.cgt1 = w6[3]; m->P.cgt2 = w6[4]; m->P.cgt3 = w6[5];
  // Rayleigh-damping coefficients (src/param.F:8633-8668): zero below zd,
  // 0.5*(1-cos(pi*(z-zd)/(zf(nk+1)-zd))) above, with the host's
  // single-precision cos as the reference evaluates it; taus = tauh
  {
    float *tauh = m->hk[CM1G_K_TAUH], *tauf = m->hk[CM1G_K_TAUF];
    const float pi = 3.1415926535897932384626433f, zd = m->c.zd, maxz = zf[nk + 1];
    for (int k = 0; k <= nk + 2; k++) tauh[k] = tauf[k] = 0.0f;
    if (m->c.irdamp >= 1 && zd < maxz) {
      ARGCHK(zd >= 0.5f * maxz,
             "cm1g_metrics: Rayleigh damping would cover more than half the domain (zd < maxz/2)");
      for (int k = 1; k <= nk; k++)
        if (zh[k] > zd) tauh[k] = 0.5f * (1.0f - cosf(pi * (zh[k] - zd) / (zf[nk + 1] - zd)));
      for (int k = 1; k <= nk + 1; k++)
        if (zf[k] > zd) tauf[k] = 0.5f * (1.0f - cosf(pi * (zf[k] - zd) / (zf[nk + 1] - zd)));
    }
  }
  int ids[] = {CM1G_K_ZF, CM1G_K_ZH, CM1G_K_MH, CM1G_K_MF, CM1G_K_RMH, CM1G_K_RMF, CM1G_K_C1, CM1G_K_C2,
               CM1G_K_TAUH, CM1G_K_TAUF};
  for (int id : ids) {
    int rc = upload_column(m, id);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int cm1g_base_derive(cm1g_model *m) {
  ARGCHK(m, "cm1g_base_derive: null model");
  graphs_clear(m);
  const int nk = m->c.nk;
  const Consts &K = m->P.K;
  float *pi0 = m->hk[CM1G_K_PI0], *th0 = m->hk[CM1G_K_TH0], *qv0 = m->hk[CM1G_K_QV0];
  float *prs0 = m->hk[CM1G_K_PRS0], *rho0 = m->hk[CM1G_K_RHO0], *rr0 = m->hk[CM1G_K_RR0];
  float *rf0 = m->hk[CM1G_K_RF0], *rrf0 = m->hk[CM1G_K_RRF0], *thv0 = m->hk[CM1G_K_THV0];
  float *rth0 = m->hk[CM1G_K_RTH0], *qc0 = m->hk[CM1G_K_QC0], *qi0 = m->hk[CM1G_K_QI0];
  float *c1 = m->hk[CM1G_K_C1], *c2 = m->hk[CM1G_K_C2];
  const KP &P = m->P;
  for (int k = 1; k <= nk; k++) {
    prs0[k] = K.p00 * powf(pi0[k], K.cpdrd);  // host libm, as the reference
    rho0[k] = prs0[k] / (K.rd * th0[k] * pi0[k] * (1.0f + qv0[k] * K.reps));
    if (m->c.imoist == 1)
      thv0[k] = (th0[k] + 0.0f) * (1.0f + K.reps * fmaxf(0.0f, qv0[k])) /
                (1.0f + fmaxf(0.0f, qv0[k]) + fmaxf(0.0f, qc0[k] + qi0[k]));
    else
      thv0[k] = th0[k];
  }
  rho0[0] = P.cgs1 * rho0[1] + P.cgs2 * rho0[2] + P.cgs3 * rho0[3];
  rho0[nk + 1] = P.cgt1 * rho0[nk] + P.cgt2 * rho0[nk - 1] + P.cgt3 * rho0[nk - 2];
  for (int k = 0; k <= nk + 1; k++) rr0[k] = 1.0f / rho0[k];
  for (int k = 2; k <= nk; k++) rf0[k] = c1[k] * rho0[k - 1] + c2[k] * rho0[k];
  rf0[1] = P.cgs1 * rho0[1] + P.cgs2 * rho0[2] + P.cgs3 * rho0[3];
  rf0[0] = rf0[1];
  rf0[nk + 1] = P.cgt1 * rho0[nk] + P.cgt2 * rho0[nk - 1] + P.cgt3 * rho0[nk - 2];
  for (int k = 0; k <= nk + 1; k++) rrf0[k] = 1.0f / rf0[k];
  th0[0] = th0[1];
  th0[nk + 1] = th0[nk];
  for (int k = 0; k <= nk + 1; k++) rth0[k] = 1.0f / th0[k];
  int ids[] = {CM1G_K_PI0, CM1G_K_TH0, CM1G_K_QV0, CM1G_K_PRS0, CM1G_K_RHO0, CM1G_K_RR0, CM1G_K_RF0,
               CM1G_K_RRF0, CM1G_K_THV0, CM1G_K_RTH0, CM1G_K_QC0, CM1G_K_QI0};
  for (int id : ids) {
    int rc = upload_column(m, id);
    if (rc) return rc;
  }
  return 0;
}

bool field_ok(const cm1g_model *m, int id) {
  return m && id >= 0 && id < CM1G_NFIELDS && m->f[id] != nullptr;
}

// the logical box (i = -2..ni+4) as rows of the padded device box
static float *box_origin(const cm1g_model *m, int id) { return m->f[id] + (CM1G_IOFF - 2); }

extern "C" int cm1g_put_field(cm1g_model *m, int id, const float *host) {
  ARGCHK(field_ok(m, id) && host, "cm1g_put_field: bad field id or null pointer");
  UNSWAP(m);
  m->synced = 0;
  const size_t w = (size_t)m->P.ni + 7, rows = (size_t)m->P.py * (size_t)(m->P.nk + 3);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy2D(box_origin(m, id), (size_t)m->P.px * sizeof(float), host, w * sizeof(float),
                     w * sizeof(float), rows, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_get_field(cm1g_model *m, int id, float *host) {
  ARGCHK(field_ok(m, id) && host, "cm1g_get_field: bad field id or null pointer");
  UNSWAP(m);
  const size_t w = (size_t)m->P.ni + 7, rows = (size_t)m->P.py * (size_t)(m->P.nk + 3);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy2D(host, w * sizeof(float), box_origin(m, id), (size_t)m->P.px * sizeof(float),
                     w * sizeof(float), rows, hipMemcpyDeviceToHost));
  return 0;
}

// Host arrays in the reference's own shapes (src/cm1.F:747-777): scalars
// (ib:ie,jb:je,kb:ke) with ib=-2, jb=-2, kb=0; u adds one point in i, v
// one in j, w one in k.  One strided copy per k plane into the padded
// device box.
static int array_shape(const cm1g_model *m, int stag, size_t *nx, size_t *ny, size_t *nz) {
  if (stag < CM1G_STAG_S || stag > CM1G_STAG_W) return 1;
  *nx = (size_t)m->P.ni + 6 + (stag == CM1G_STAG_U ? 1 : 0);
  *ny = (size_t)m->P.nj + 6 + (stag == CM1G_STAG_V ? 1 : 0);
  *nz = (size_t)m->P.nk + 2 + (stag == CM1G_STAG_W ? 1 : 0);
  return 0;
}

// pinned staging of the array moves (allocated on first use, nx*ny*nz
// floats of the largest staggering)
static int pinned_stage(cm1g_model *m, size_t n, float **out) {
  if (m->pin_cap < n) {
    if (m->pin) HIPCHK(hipHostFree(m->pin));
    m->pin = nullptr;
    m->pin_cap = 0;
    HIPCHK(hipHostMalloc((void **)&m->pin, n * sizeof(float), hipHostMallocDefault));
    m->pin_cap = n;
  }
  *out = m->pin;
  return 0;
}

extern "C" int cm1g_put_array(cm1g_model *m, int id, int stag, const float *host) {
  ARGCHK(field_ok(m, id) && host, "cm1g_put_array: bad field id or null pointer");
  UNSWAP(m);
  m->synced = 0;
  size_t nx, ny, nz;
  ARGCHK(array_shape(m, stag, &nx, &ny, &nz) == 0, "cm1g_put_array: bad staggering");
  HIPCHK(hipStreamSynchronize(m->stream));
  // staged through pinned memory, then one pitched 3-D copy: host rows of
  // nx, ny rows a level; device rows of px floats, py rows a level
  float *stage = nullptr;
  {
    const int rc = pinned_stage(m, (size_t)(m->P.ni + 8) * (m->P.nj + 8) * (m->P.nk + 3), &stage);
    if (rc) return rc;
  }
  std::memcpy(stage, host, nx * ny * nz * sizeof(float));
  hipMemcpy3DParms p = {};
  p.srcPtr = make_hipPitchedPtr((void *)stage, nx * sizeof(float), nx, ny);
  p.dstPtr = make_hipPitchedPtr((void *)box_origin(m, id), (size_t)m->P.px * sizeof(float), (size_t)m->P.px,
                                (size_t)m->P.py);
  p.extent = make_hipExtent(nx * sizeof(float), ny, nz);
  p.kind = hipMemcpyHostToDevice;
  HIPCHK(hipMemcpy3D(&p));
  return 0;
}

extern "C" int cm1g_get_array(cm1g_model *m, int id, int stag, float *host) {
  ARGCHK(field_ok(m, id) && host, "cm1g_get_array: bad field id or null pointer");
  UNSWAP(m);
  size_t nx, ny, nz;
  ARGCHK(array_shape(m, stag, &nx, &ny, &nz) == 0, "cm1g_get_array: bad staggering");
  HIPCHK(hipStreamSynchronize(m->stream));
  float *stage = nullptr;
  {
    const int rc = pinned_stage(m, (size_t)(m->P.ni + 8) * (m->P.nj + 8) * (m->P.nk + 3), &stage);
    if (rc) return rc;
  }
  hipMemcpy3DParms p = {};
  p.srcPtr = make_hipPitchedPtr((void *)box_origin(m, id), (size_t)m->P.px * sizeof(float), (size_t)m->P.px,
                                (size_t)m->P.py);
  p.dstPtr = make_hipPitchedPtr((void *)stage, nx * sizeof(float), nx, ny);
  p.extent = make_hipExtent(nx * sizeof(float), ny, nz);
  p.kind = hipMemcpyDeviceToHost;
  HIPCHK(hipMemcpy3D(&p));
  std::memcpy(host, stage, nx * ny * nz * sizeof(float));
  return 0;
}

extern "C" int cm1g_put_2d(cm1g_model *m, int id, const float *host) {
  ARGCHK(m && host && id >= 0 && id < CM1G_D_NFIELDS, "cm1g_put_2d: bad argument");
  size_t n = (size_t)m->P.pxy + (size_t)(m->P.px + m->P.py) * (m->c.nk + 3);
  HIPCHK(hipMemcpy(m->d2[id], host, n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_get_2d(cm1g_model *m, int id, float *host) {
  ARGCHK(m && host && id >= 0 && id < CM1G_D_NFIELDS, "cm1g_get_2d: bad argument");
  size_t n = (size_t)m->P.pxy + (size_t)(m->P.px + m->P.py) * (m->c.nk + 3);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(host, m->d2[id], n * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

// 2-D fields in the reference's own shapes: rain, prate and the surface
// layer's ust..znt (ib:ie,jb:je), radbcw/radbce (jb:je,kb:ke),
// radbcs/radbcn (ib:ie,kb:ke), ngxy = 3, ngz = 1, first index fastest.
// Device positions: (i,j) fields at IX2D(i,j); radbcw/e at (j+2) + k*py;
// radbcs/n at (i+2) + k*px.
static void ref2d_map(const cm1g_model *m, int id, std::vector<size_t> &dev) {
  const KP &P = m->P;
  dev.clear();
  if (id == CM1G_D_RAIN || id == CM1G_D_PRATE || id >= CM1G_D_UST) {
    for (int j = -2; j <= P.nj + 3; j++)
      for (int i = -2; i <= P.ni + 3; i++) dev.push_back((size_t)IX2D(P, i, j));
  } else if (id == CM1G_D_RADBCW || id == CM1G_D_RADBCE) {
    for (int k = 0; k <= P.nk + 1; k++)
      for (int j = -2; j <= P.nj + 3; j++) dev.push_back((size_t)(j + 2) + (size_t)k * P.py);
  } else {
    for (int k = 0; k <= P.nk + 1; k++)
      for (int i = -2; i <= P.ni + 3; i++) dev.push_back((size_t)(i + 2) + (size_t)k * P.px);
  }
}

extern "C" int cm1g_put_2d_array(cm1g_model *m, int id, const float *host) {
  ARGCHK(m && host && id >= 0 && id < CM1G_D_NFIELDS, "cm1g_put_2d_array: bad argument");
  const size_t n = (size_t)m->P.pxy + (size_t)(m->P.px + m->P.py) * (m->c.nk + 3);
  std::vector<float> buf(n);
  std::vector<size_t> dev;
  ref2d_map(m, id, dev);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(buf.data(), m->d2[id], n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < dev.size(); q++) buf[dev[q]] = host[q];
  HIPCHK(hipMemcpy(m->d2[id], buf.data(), n * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_get_2d_array(cm1g_model *m, int id, float *host) {
  ARGCHK(m && host && id >= 0 && id < CM1G_D_NFIELDS, "cm1g_get_2d_array: bad argument");
  const size_t n = (size_t)m->P.pxy + (size_t)(m->P.px + m->P.py) * (m->c.nk + 3);
  std::vector<float> buf(n);
  std::vector<size_t> dev;
  ref2d_map(m, id, dev);
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(buf.data(), m->d2[id], n * sizeof(float), hipMemcpyDeviceToHost));
  for (size_t q = 0; q < dev.size(); q++) host[q] = buf[dev[q]];
  return 0;
}

// ------------------------------------------------------------ checkpoint
// Arrays a step carries into the next: the time-level "a" arrays and their
// "3d" copies (halos included), ppx, the previous step's diabatic
// tendencies, rho/prs/rr/rf, the tracers, the subgrid TKE, the scalar updates' partner
// buffer (it rotates into the scalar fields, carrying the ghost cells no
// kernel writes), and the 2-D fields (rain, prate, and the
// radiative-boundary speeds of the last stage, which the next step's first
// acoustic substeps use).
struct StateHeader {
  char magic[8];
  int32_t ni, nj, nk, px, py, npt, imoist, n3d, n2d, reserved;
  double dscal[4];
  float cflmax, pad;
};
static_assert(sizeof(StateHeader) == 88, "checkpoint header layout");

static int state_ids(const cm1g_model *m, int *ids) {
  static const int base[] = {CM1G_UA, CM1G_VA, CM1G_WA, CM1G_THA, CM1G_PPI, CM1G_U3D, CM1G_V3D,
                             CM1G_W3D, CM1G_TH3D, CM1G_PP3D, CM1G_PPX, CM1G_RHO, CM1G_PRS,
                             CM1G_RR, CM1G_RF};
  int n = 0;
  for (int id : base) ids[n++] = id;
  if (m->c.imoist == 1) {
    const int moist[] = {CM1G_QA0, CM1G_QA1, CM1G_QA2, CM1G_Q3D0, CM1G_Q3D1, CM1G_Q3D2,
                         CM1G_QPTEN, CM1G_QTTEN, CM1G_QVTEN, CM1G_QCTEN};
    for (int id : moist) ids[n++] = id;
  }
  for (int t = 0; t < m->c.npt; t++) {
    ids[n++] = CM1G_PTA0 + t;
    ids[n++] = CM1G_PT3D0 + t;
  }
  // the subgrid TKE (the viscosities and the surface set-up are formed again
  // from the state at the next step's start; the surface fields are among
  // the 2-D fields)
  if (m->turb.cm1setup == 1) {
    ids[n++] = CM1G_TKEA;
    ids[n++] = CM1G_TKE3D;
  }
  return n;
}

static size_t state_box(const cm1g_model *m) {
  return ((size_t)m->P.ni + 7) * (size_t)m->P.py * (size_t)(m->P.nk + 3) * sizeof(float);
}
static size_t state_2d(const cm1g_model *m) {
  return ((size_t)m->P.pxy + (size_t)(m->P.px + m->P.py) * (m->c.nk + 3)) * sizeof(float);
}

extern "C" int cm1g_state_bytes(cm1g_model *m, int64_t *nbytes) {
  ARGCHK(m && nbytes, "cm1g_state_bytes: null argument");
  int ids[64];
  const int n3 = state_ids(m, ids);
  *nbytes = (int64_t)(sizeof(StateHeader) + (n3 + 1) * state_box(m) + CM1G_D_NFIELDS * state_2d(m));
  return 0;
}

// the logical box of the partner buffer, to or from the host
static int partner_box(cm1g_model *m, void *host, bool to_host) {
  const size_t w = (size_t)m->P.ni + 7, rows = (size_t)m->P.py * (size_t)(m->P.nk + 3);
  float *d = m->thb + (CM1G_IOFF - 2);
  const size_t dp = (size_t)m->P.px * sizeof(float), hp = w * sizeof(float);
  HIPCHK(hipStreamSynchronize(m->stream));
  if (to_host) HIPCHK(hipMemcpy2D(host, hp, d, dp, hp, rows, hipMemcpyDeviceToHost));
  else HIPCHK(hipMemcpy2D(d, dp, host, hp, hp, rows, hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_state_save(cm1g_model *m, void *host, int64_t nbytes) {
  ARGCHK(m && host, "cm1g_state_save: null argument");
  UNSWAP(m);
  int64_t need = 0;
  cm1g_state_bytes(m, &need);
  ARGCHK(nbytes >= need, "cm1g_state_save: buffer too small");
  int ids[64];
  const int n3 = state_ids(m, ids);
  StateHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, "CM1GST02", 8);
  h.ni = m->P.ni; h.nj = m->P.nj; h.nk = m->P.nk; h.px = m->P.px; h.py = m->P.py;
  h.npt = m->c.npt; h.imoist = m->c.imoist; h.n3d = n3; h.n2d = CM1G_D_NFIELDS;
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(h.dscal, m->dscal, 4 * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&h.cflmax, m->fred, sizeof(float), hipMemcpyDeviceToHost));
  char *p = (char *)host;
  memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  for (int n = 0; n < n3; n++, p += state_box(m)) {
    int rc = cm1g_get_field(m, ids[n], (float *)p);
    if (rc) return rc;
  }
  {
    int rc = partner_box(m, p, true);
    if (rc) return rc;
    p += state_box(m);
  }
  for (int n = 0; n < CM1G_D_NFIELDS; n++, p += state_2d(m))
    HIPCHK(hipMemcpy(p, m->d2[n], state_2d(m), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int cm1g_state_load(cm1g_model *m, const void *host, int64_t nbytes) {
  ARGCHK(m && host, "cm1g_state_load: null argument");
  UNSWAP(m);  // the load then writes every field
  m->synced = 0;
  int64_t need = 0;
  cm1g_state_bytes(m, &need);
  ARGCHK(nbytes >= need, "cm1g_state_load: buffer too small");
  StateHeader h;
  memcpy(&h, host, sizeof(h));
  int ids[64];
  const int n3 = state_ids(m, ids);
  ARGCHK(memcmp(h.magic, "CM1GST02", 8) == 0, "cm1g_state_load: not a cm1gpu state");
  ARGCHK(h.ni == m->P.ni && h.nj == m->P.nj && h.nk == m->P.nk && h.px == m->P.px &&
             h.py == m->P.py && h.npt == m->c.npt && h.imoist == m->c.imoist && h.n3d == n3 &&
             h.n2d == CM1G_D_NFIELDS,
         "cm1g_state_load: the state was saved from a model of another configuration");
  const char *p = (const char *)host + sizeof(h);
  for (int n = 0; n < n3; n++, p += state_box(m)) {
    int rc = cm1g_put_field(m, ids[n], (const float *)p);
    if (rc) return rc;
  }
  {
    int rc = partner_box(m, (void *)p, false);
    if (rc) return rc;
    p += state_box(m);
  }
  for (int n = 0; n < CM1G_D_NFIELDS; n++, p += state_2d(m))
    HIPCHK(hipMemcpy(m->d2[n], p, state_2d(m), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->dscal, h.dscal, 4 * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(m->fred, &h.cflmax, sizeof(float), hipMemcpyHostToDevice));
  m->mass1 = h.dscal[0];
  return 0;
}

extern "C" int cm1g_field_ptr(cm1g_model *m, int id, float **dptr) {
  ARGCHK(field_ok(m, id) && dptr, "cm1g_field_ptr: bad field id");
  UNSWAP(m);
  m->synced = 0;
  // writes through the address are invisible to the library: the first
  // acoustic stage stops relying on the a-arrays equalling the 3d-arrays
  if (id == CM1G_UA || id == CM1G_VA || id == CM1G_WA || id == CM1G_PPI || id == CM1G_U3D ||
      id == CM1G_V3D || id == CM1G_W3D || id == CM1G_PP3D) {
    if (!m->ext_writers) graphs_clear(m);
    m->ext_writers = 1;
  }
  *dptr = m->f[id];
  return 0;
}

// satadj's non-convergence flag (set by the kernels, never cleared: the
// reference stops the run there, src/kessler.F:343-356, so every later
// check of a model that hit it reports it again)
int satadj_check(cm1g_model *m) {
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, m->iflag, sizeof(int), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  if (flag) {
    cm1g_set_error("satadj: convergence cannot be reached (src/kessler.F:343-356)");
    return CM1G_E_SATADJ;
  }
  return 0;
}

extern "C" int cm1g_sync(cm1g_model *m) {
  ARGCHK(m, "cm1g_sync: null model");
  HIPCHK(hipStreamSynchronize(m->stream));
  return satadj_check(m);
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    cm1g_set_error(std::string("kernel launch: ") + hipGetErrorString(e));
    return CM1G_E_HIP;
  }
  return 0;
}

// ------------------------------------------------------------ reductions
// Per-level partial sums of the dry-air mass and of pi for the apmasscon
// adjustment (src/solve2.F:1617-1656), in double precision.
#define MASS_NB 32
__device__ __forceinline__ void mass_level_sums(const KP &P, const float *rho, const float *pp3d, double *out,
                                                int bx, int by) {
  __shared__ double r3[256], r4[256];
  const int k = 1 + bx, b = by;
  double a3 = 0.0, a4 = 0.0;
  for (int j = 1 + b; j <= P.nj; j += MASS_NB)
    for (int i = 1 + (int)threadIdx.x; i <= P.ni; i += blockDim.x) {
      int c = IXD(P, i, j, k);
      a3 += (double)(rho[c] * P.ruh[i + 2] * P.rvh[j + 2] * P.rmh[k]);
      a4 += (double)(P.pi0[k] + pp3d[c]);
    }
  a3 = block_reduce256(a3, r3, OpAdd());
  a4 = block_reduce256(a4, r4, OpAdd());
  if (threadIdx.x == 0) {
    out[2 * (k * MASS_NB + b)] = a3;
    out[2 * (k * MASS_NB + b) + 1] = a4;
  }
}

// two-sum accumulation of the nk level sums (the reference accumulates them
// in real(kind=qp)); then the pi' shift tem (src/solve2.F:1976-1995).
__device__ inline void dd_add(double &hi, double &lo, double x) {
  double s = hi + x;
  double bb = s - hi;
  double err = (hi - (s - bb)) + (x - bb);
  hi = s;
  lo += err;
}

__device__ __forceinline__ void mass_final(int nk, int mode, float rddcv, double ntot, const double *lev,
                                           double *dscal) {
  double h3 = 0.0, l3 = 0.0, h4 = 0.0, l4 = 0.0;
  for (int k = 1; k <= nk; k++) {
    double s3 = 0.0, s4 = 0.0;
    for (int b = 0; b < MASS_NB; b++) {
      s3 += lev[2 * (k * MASS_NB + b)];
      s4 += lev[2 * (k * MASS_NB + b) + 1];
    }
    dd_add(h3, l3, s3);
    dd_add(h4, l4, s4);
  }
  double mass = h3 + l3, p2 = h4 + l4;
  if (mode == 0) {
    dscal[0] = mass;  // mass1 (getset)
  } else {
    dscal[1] = mass;  // mass2
    double tem = (pow(dscal[0] / mass, (double)rddcv) - 1.0) * p2 / ntot;
    dscal[2] = (double)(float)tem;
  }
}

static void mass_levels(cm1g_model *m, const float *rho, int mode) {
  const KP P = m->P;
  const float *pp3d = m->f[CM1G_PP3D];
  double *lev = m->dred, *dscal = m->dscal;
  launch<256>(m, dim3(P.nk, MASS_NB), dim3(256),
              [=] __device__(int bx, int by, int) { mass_level_sums(P, rho, pp3d, lev, bx, by); });
  const double ntot = (double)((long long)P.ni * P.nj * P.nk);
  const int nk = P.nk;
  const float rddcv = P.K.rddcv;
  launch<64>(m, dim3(1), dim3(1),
             [=] __device__(int, int, int) { mass_final(nk, mode, rddcv, ntot, lev, dscal); });
}

// domain rain-accumulation budget (src/kessler.F:470-476): per-row sums in
// i order, then rows in j order, as the reference does.
__device__ __forceinline__ void train_rows(int ni, int nj, const double *colbud, double *rows, int bx) {
  const int j = bx * (int)blockDim.x + (int)threadIdx.x;
  if (j >= nj) return;
  double b = 0.0;
  for (int i = 0; i < ni; i++) b = b + colbud[i + j * ni];
  rows[j] = b;
}
__device__ __forceinline__ void train_sum(int nj, double tem, const double *rows, double *dscal) {
  double tr = dscal[3];
  for (int j = 0; j < nj; j++) tr = tr + rows[j] * tem;
  dscal[3] = tr;
}
// train += the fallout's column budgets of this call (the end of
// src/kessler.F:fallout)
void train_accumulate(cm1g_model *m) {
  const int ni = m->P.ni, nj = m->P.nj;
  const double *colbud = m->dred;
  double *rows = m->dred + (size_t)ni * nj, *dscal = m->dscal;
  const double tem = (double)(m->c.dx * m->c.dy);
  launch<64>(m, dim3((nj + 63) / 64), dim3(64),
             [=] __device__(int bx, int, int) { train_rows(ni, nj, colbud, rows, bx); });
  launch<64>(m, dim3(1), dim3(1), [=] __device__(int, int, int) { train_sum(nj, tem, rows, dscal); });
}

// ------------------------------------------------------------------ getset
extern "C" int cm1g_getset(cm1g_model *m, int fresh_start) {
  ARGCHK(m, "cm1g_getset: null model");
  UNSWAP(m);
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  float **f = m->f;
  bc_fill(m, BC_U, f[CM1G_UA]);
  bc_fill(m, BC_V, f[CM1G_VA]);
  bc_fill(m, BC_W1, f[CM1G_WA]);
  bc_fill(m, BC_S, f[CM1G_PPI]);
  bc_fill(m, BC_S, f[CM1G_THA]);
  if (m->c.imoist == 1)
    for (int n = 0; n < 3; n++) bc_fill(m, BC_S, f[CM1G_QA0 + n]);
  for (int n = 0; n < m->c.npt; n++) bc_fill(m, BC_S, f[CM1G_PTA0 + n]);
  const size_t bytes = m->ntot * sizeof(float);
  HIPCHK(hipMemcpyAsync(f[CM1G_U3D], f[CM1G_UA], bytes, hipMemcpyDeviceToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(f[CM1G_V3D], f[CM1G_VA], bytes, hipMemcpyDeviceToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(f[CM1G_W3D], f[CM1G_WA], bytes, hipMemcpyDeviceToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(f[CM1G_PP3D], f[CM1G_PPI], bytes, hipMemcpyDeviceToDevice, m->stream));
  HIPCHK(hipMemcpyAsync(f[CM1G_TH3D], f[CM1G_THA], bytes, hipMemcpyDeviceToDevice, m->stream));
  if (m->c.imoist == 1)
    for (int n = 0; n < 3; n++)
      HIPCHK(hipMemcpyAsync(f[CM1G_Q3D0 + n], f[CM1G_QA0 + n], bytes, hipMemcpyDeviceToDevice, m->stream));
  for (int n = 0; n < m->c.npt; n++)
    HIPCHK(hipMemcpyAsync(f[CM1G_PT3D0 + n], f[CM1G_PTA0 + n], bytes, hipMemcpyDeviceToDevice, m->stream));
  if (m->turb.cm1setup == 1) {
    const int rc = tke_init(m);
    if (rc) return rc;
  }
  float *rho = f[CM1G_RHO], *prs = f[CM1G_PRS], *rr = f[CM1G_RR], *rf = f[CM1G_RF];
  const float *pp3d = f[CM1G_PP3D], *th3d = f[CM1G_TH3D], *qv = f[CM1G_Q3D0];
  const int moist = m->c.imoist;
  if (fresh_start) {
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      const Consts &K = P.K;
      int c = IXD(P, i, j, k);
      float p = K.p00 * pow_f(P.pi0[k] + pp3d[c], K.cpdrd);
      prs[c] = p;
      if (moist)
        rho[c] = p / ((P.th0[k] + th3d[c]) * (P.pi0[k] + pp3d[c]) * (K.rd + fmaxf(0.0f, qv[c]) * K.rv));
      else
        rho[c] = p / (K.rd * (P.th0[k] + th3d[c]) * (P.pi0[k] + pp3d[c]));
    });
  }
  bc_fill(m, BC_S, rho);
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    int c = IXD(P, i, j, k);
    rr[c] = 1.0f / rho[c];
    rf[c] = P.c1[k] * rho[IXD(P, i, j, k - 1)] + P.c2[k] * rho[c];
  });
  bc_fill(m, BC_S, rr);
  bc_fill(m, BC_S, rf);
  for2(m, 0, ni + 1, 0, nj + 1, [=] __device__(int i, int j) {
    rf[IXD(P, i, j, 1)] = P.cgs1 * rho[IXD(P, i, j, 1)] + P.cgs2 * rho[IXD(P, i, j, 2)] + P.cgs3 * rho[IXD(P, i, j, 3)];
    rf[IXD(P, i, j, nk + 1)] = P.cgt1 * rho[IXD(P, i, j, nk)] + P.cgt2 * rho[IXD(P, i, j, nk - 1)] +
                               P.cgt3 * rho[IXD(P, i, j, nk - 2)];
  });
  if (fresh_start) {
    mass_levels(m, rho, 0);
    float *ppx = f[CM1G_PPX];
    const float *ppi = f[CM1G_PPI];
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      ppx[IXD(P, i, j, k)] = ppi[IXD(P, i, j, k)];
    });
  }
  bc_fill(m, BC_S, f[CM1G_PPX]);
  int rc = check_launch();
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(&m->mass1, m->dscal, sizeof(double), hipMemcpyDeviceToHost));
  m->synced = !m->ext_writers;
  return 0;
}

// ------------------------------------------------------------------ solve1
// rru = rho0*u, rrv = rho0*v, rrw = rf0*w (0 at k = 1 and nk+1) over
// i,j = 0..n+2 (src/solve2.F:490-545) and divx over 0..n+1 (:587-625),
// marching up each column: the fluxes divx needs at i+1, j+1 are formed
// from the velocities with the same products, w(k+1) is carried to the next
// level, and each level's loads are issued before the previous level's
// stores.
__device__ __forceinline__ void massflux_div(const KP &P, const float *u3d, const float *v3d, const float *w3d,
                                             float *rru, float *rrv, float *rrw, float *divx,
                                             const ColGrid &cg, int bx) {
  int i, j;
  if (!col_tile_b(cg, bx, i, j)) return;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (i < 0 || i > ni + 2 || j > nj + 2) return;
  const bool du = (j <= nj + 1), dv = (i <= ni + 1), dd = du && dv;
  const int px = P.px, pxy = P.pxy;
  int c = IXD(P, i, j, 1);
  float u0 = u3d[c], u1 = dd ? u3d[c + 1] : 0.0f, v0 = v3d[c], v1 = dd ? v3d[c + px] : 0.0f;
  float wk = 0.0f, wk1 = (dd && nk >= 2) ? w3d[c + pxy] : 0.0f;
  for (int k = 1; k <= nk; k++, c += pxy) {
    const float cu0 = u0, cu1 = u1, cv0 = v0, cv1 = v1, cw0 = wk, cw1 = wk1;
    if (k < nk) {
      const int cn = c + pxy;
      u0 = u3d[cn];
      v0 = v3d[cn];
      if (dd) {
        u1 = u3d[cn + 1];
        v1 = v3d[cn + px];
        wk = cw1;
        wk1 = (k + 1 < nk) ? w3d[cn + pxy] : 0.0f;
      }
    }
    const float r0 = P.rho0[k];
    const float ru0 = r0 * cu0, rv0 = r0 * cv0;
    if (du) rru[c] = ru0;
    if (dv) rrv[c] = rv0;
    if (dd) {
      const float rw0 = (k == 1) ? 0.0f : P.rf0[k] * cw0;
      const float rw1 = (k == nk) ? 0.0f : P.rf0[k + 1] * cw1;
      rrw[c] = rw0;
      if (k == nk) rrw[c + pxy] = 0.0f;
      if (i <= ni + 1 && j <= nj + 1) {
        const float ru1 = r0 * cu1, rv1 = r0 * cv1;
        float d = ((ru1 - ru0) * P.rdx * P.uh[i + 2] + (rv1 - rv0) * P.rdy * P.vh[j + 2]) +
                  (rw1 - rw0) * P.rdz * P.mh[k];
        if (fabsf(d) < P.K.smeps) d = 0.0f;
        divx[c] = d;
      }
    }
  }
}

// the (a-array, 3-D array) pairs solve3 hands over
static std::vector<const int *> swap_pairs(const cm1g_model *m) {
  static const int base[5][2] = {{CM1G_UA, CM1G_U3D}, {CM1G_VA, CM1G_V3D}, {CM1G_WA, CM1G_W3D},
                                 {CM1G_PPI, CM1G_PP3D}, {CM1G_THA, CM1G_TH3D}};
  static const int q[3][2] = {{CM1G_QA0, CM1G_Q3D0}, {CM1G_QA1, CM1G_Q3D1}, {CM1G_QA2, CM1G_Q3D2}};
  static const int pt[8][2] = {{CM1G_PTA0 + 0, CM1G_PT3D0 + 0}, {CM1G_PTA0 + 1, CM1G_PT3D0 + 1},
                               {CM1G_PTA0 + 2, CM1G_PT3D0 + 2}, {CM1G_PTA0 + 3, CM1G_PT3D0 + 3},
                               {CM1G_PTA0 + 4, CM1G_PT3D0 + 4}, {CM1G_PTA0 + 5, CM1G_PT3D0 + 5},
                               {CM1G_PTA0 + 6, CM1G_PT3D0 + 6}, {CM1G_PTA0 + 7, CM1G_PT3D0 + 7}};
  std::vector<const int *> v;
  for (auto &p : base) v.push_back(p);
  if (m->c.imoist == 1)
    for (auto &p : q) v.push_back(p);
  for (int n = 0; n < m->c.npt; n++) v.push_back(pt[n]);
  return v;
}

int unswap(cm1g_model *m) {
  if (!m->swapped) return 0;
  const size_t bytes = m->ntot * sizeof(float);
  for (const int *pr : swap_pairs(m))
    HIPCHK(hipMemcpyAsync(m->f[pr[1]], m->f[pr[0]], bytes, hipMemcpyDeviceToDevice, m->stream));
  m->swapped = 0;
  return 0;
}

extern "C" int cm1g_solve1(cm1g_model *m) {
  ARGCHK(m, "cm1g_solve1: null model");
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  float **f = m->f;
  float *uten1 = f[CM1G_UTEN1], *vten1 = f[CM1G_VTEN1], *wten1 = f[CM1G_WTEN1], *thten1 = f[CM1G_THTEN1];
  for (int n = 0; n < m->c.npt; n++) {
    float *pt = f[CM1G_PTTEN0 + n];
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) { pt[IXD(P, i, j, k)] = 0.0f; });
  }
  float *ppten = f[CM1G_PPTEN];
  const float *rho = f[CM1G_RHO], *tha = f[CM1G_THA], *ppi = f[CM1G_PPI];
  const int moist = m->c.imoist, diab = (m->c.imoist == 1 && m->c.ptype >= 1);
  float *qten0 = f[CM1G_QTEN0], *qten1 = f[CM1G_QTEN1], *qten2 = f[CM1G_QTEN2];
  const float *qa0 = f[CM1G_QA0];
  const float *qpten = f[CM1G_QPTEN], *qtten = f[CM1G_QTTEN], *qvten = f[CM1G_QVTEN], *qcten = f[CM1G_QCTEN];
  const float rdt = 1.0f / dt;
  // ten1 over the staggered ranges: Rayleigh damping toward the base state
  // (irdamp = 1, src/solve1.F:414-426; no terrain, so tauh(i-1) = tauh(i)
  // and taus = tauh) or zero (:442-453); then ppten and the diabatic ten1
  // on the interior (:1716-1756)
  const float *ua = f[CM1G_UA], *va = f[CM1G_VA], *wa = f[CM1G_WA];
  const int rdamp = P.irdamp;
  // with the subgrid turbulence its tendencies come between the ten1 set-up
  // and ppten (src/solve1.F:1118-1240): this pass then stops after ten1
  const int turb = (m->turb.cm1setup == 1);
  for3(m, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
    const Consts &K = P.K;
    const float tem = 0.0001f * K.tsmall;
    int c = IXD(P, i, j, k);
    float u1 = 0.0f, v1 = 0.0f, w1 = 0.0f, th1 = 0.0f;
    if (rdamp) {
      const float ra = P.rdalpha, th = P.tauh[k];
      u1 = -ra * 0.5f * (th + th) * (ua[c] - P.u0[k]);
      v1 = -ra * 0.5f * (th + th) * (va[c] - P.v0[k]);
      w1 = -ra * P.tauf[k] * wa[c];
      th1 = -ra * th * tha[c];
    }
    uten1[c] = u1;
    vten1[c] = v1;
    wten1[c] = w1;
    if (i > ni || j > nj || turb) {
      thten1[c] = th1;
      return;
    }
    float pt;
    if (moist) {
      float qvt = 0.0f;     // qten(nqv) was just zeroed
      if (fabsf(dt * th1) > tem || fabsf(dt * qvt) > K.qsmall) {
        float thnew = tha[c] + dt * th1;
        float qvnew = qa0[c] + dt * qvt;
        float pinew = pow_f(rho[c] * (P.th0[k] + thnew) * (K.rd + K.rv * qvnew) * K.rp00, K.rddcv) - P.pi0[k];
        pt = (pinew - ppi[c]) * rdt;
      } else pt = 0.0f;
      float q0 = 0.0f, q1 = 0.0f;
      if (diab) {
        pt = pt + qpten[c];
        th1 = th1 + qtten[c];
        q0 = q0 + qvten[c];
        q1 = q1 + qcten[c];
      }
      qten0[c] = q0;
      qten1[c] = q1;
      qten2[c] = 0.0f;
    } else {
      if (fabsf(dt * th1) > tem) {
        float thnew = tha[c] + dt * th1;
        float pinew = pow_f(rho[c] * (P.th0[k] + thnew) * K.rd * K.rp00, K.rddcv) - P.pi0[k];
        pt = (pinew - ppi[c]) * rdt;
      } else pt = 0.0f;
    }
    ppten[c] = pt;
    thten1[c] = th1;
  });
  if (turb) {
    // the turbulence tendencies, then ppten from the final thten1 (dry:
    // the turbulence is only set on dry models)
    turb_tend(m);
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      const Consts &K = P.K;
      const float tem = 0.0001f * K.tsmall;
      const int c = IXD(P, i, j, k);
      const float th1 = thten1[c];
      float pt;
      if (fabsf(dt * th1) > tem) {
        float thnew = tha[c] + dt * th1;
        float pinew = pow_f(rho[c] * (P.th0[k] + thnew) * K.rd * K.rp00, K.rddcv) - P.pi0[k];
        pt = (pinew - ppi[c]) * rdt;
      } else pt = 0.0f;
      ppten[c] = pt;
    });
  }
  return check_launch();
}

// ------------------------------------------------------------------ solve2
extern "C" int cm1g_solve2(cm1g_model *m) {
  ARGCHK(m, "cm1g_solve2: null model");
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  float **f = m->f;
  const int moist = m->c.imoist;
  float *u3d = f[CM1G_U3D], *v3d = f[CM1G_V3D], *w3d = f[CM1G_W3D];
  float *uten = f[CM1G_UTEN], *vten = f[CM1G_VTEN], *wten = f[CM1G_WTEN];
  const float *uten1 = f[CM1G_UTEN1], *vten1 = f[CM1G_VTEN1], *wten1 = f[CM1G_WTEN1];
  float *rru = f[CM1G_RRU], *rrv = f[CM1G_RRV], *rrw = f[CM1G_RRW], *divx = f[CM1G_DIVX];
  float *pp3d = f[CM1G_PP3D], *sadv = f[CM1G_SADV];
  float *thterm = f[CM1G_THTERM], *dum8 = f[CM1G_DUM8], *t11 = f[CM1G_T11], *t22 = f[CM1G_T22];
  float *rho = f[CM1G_RHO], *prs = f[CM1G_PRS];
  for (int nrk = 1; nrk <= 3; nrk++) {
    // the scalar updates swap th3d / q3d with a partner array: fetch the
    // current buffers at the top of the stage and again after the updates
    float *th3d = f[CM1G_TH3D];
    const float *qv = f[CM1G_Q3D0], *qc = f[CM1G_Q3D1], *qr = f[CM1G_Q3D2];
    const float dttmp = dt / (float)(3 + 1 - nrk);
    // after an exchanging solve3 the first stage reads the state from the
    // a-arrays wherever it would read the (equal) 3-D arrays before writing
    // them; its writes go to the 3-D arrays, which it rewrites completely
    const bool from_a = (nrk == 1 && m->swapped);
    const float *th_r = from_a ? f[CM1G_THA] : th3d, *pp_r = from_a ? f[CM1G_PPI] : pp3d;
    const float *qv_r = from_a ? f[CM1G_QA0] : qv, *qc_r = from_a ? f[CM1G_QA1] : qc,
                *qr_r = from_a ? f[CM1G_QA2] : qr;
    const float *u_r = from_a ? f[CM1G_UA] : u3d, *v_r = from_a ? f[CM1G_VA] : v3d,
                *w_r = from_a ? f[CM1G_WA] : w3d;
    // thermodynamic prep (src/solve2.F:302-450), with sadv = pi0 + pi'
    // (src/solve2.F:1027-1050) over the interior and its 1-wide ring
    for3(m, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      const Consts &K = P.K;
      int c = IXD(P, i, j, k);
      sadv[c] = P.pi0[k] + pp_r[c];
      if (i < 1 || i > ni || j < 1 || j > nj) return;
      float th = th_r[c];
      if (moist) {
        float ql = qc_r[c];
        ql = ql + qr_r[c];
        float qvv = fmaxf(qv_r[c], 0.0f);
        float qli = fmaxf(0.0f, ql + 0.0f);
        float cpli = K.cpl * fmaxf(0.0f, ql) + K.cpi * fmaxf(0.0f, 0.0f);
        dum8[c] = K.g * (th * P.rth0[k] + K.repsm1 * (qvv - P.qv0[k]) - (qli - P.qc0[k] - P.qi0[k]));
        t11[c] = (P.th0[k] + th) * (1.0f + K.reps * qvv) / (1.0f + qvv + qli);
        float cpm = K.cp + K.cpv * qvv + cpli;
        float cvm = 1.0f / (K.cv + K.cvv * qvv + cpli);
        thterm[c] = (P.th0[k] + th) * (K.rd + K.rv * qvv - K.rovcp * cpm) * cvm;
        t22[c] = (P.pi0[k] + pp_r[c]) * K.rovcp * cpm * cvm;
      } else {
        dum8[c] = K.g * th * P.rth0[k];
        t11[c] = P.th0[k] + th;
        t22[c] = (P.pi0[k] + pp_r[c]) * K.rddcv;
      }
    });
    bc_fill(m, BC_P, t11);
    // rru, rrv, rrw (src/solve2.F:490-545) and divx (:587-625) in one pass
    {
      const ColGrid cg = col_grid_box(0, ni + 2, 0, nj + 2, m->tile_rows);
      launch<256>(m, col_blocks(cg), col_threads(cg), [=] __device__(int bx, int, int) {
        massflux_div(P, u_r, v_r, w_r, rru, rrv, rrw, divx, cg, bx);
      });
    }
    // tendencies: the pre-RK values (src/solve2.F:459-471) plus advection,
    // each component's own mass flux formed from its velocity
    {
      const float *src3[3] = {u_r, v_r, w_r};
      adv_uvw(m, nrk, 7, 1, 1, 1, from_a ? src3 : nullptr);
    }
    if (m->c.psolver == 2) sounde_run(m, nrk);
    else sound_run(m, nrk, nrk == 1 && m->synced);
    if (nrk == 3) cfl_run(m, dt, u3d, v3d, w3d);
    if (m->c.irbc == 4 && (P.ibw || P.ibe || P.ibs || P.ibn))
      radbc4(m, f[CM1G_UA], u3d, f[CM1G_VA], v3d, dttmp);
    getdiv_both(m, moist);
    // theta (src/solve2.F:1310-1428); from the a-arrays, the update writes
    // the 3-D array directly instead of a partner
    advs_run(m, nrk, from_a ? f[CM1G_THA] : th3d, 1.0f, f[CM1G_THA], f[CM1G_THTEN1], moist ? t11 : nullptr,
             nullptr, th3d, 0, 0, dttmp, (double)(10.0f * P.K.epsilon), 1);
    if (moist) {
      for (int n = 0; n < 3; n++) {
        advs_run(m, nrk, from_a ? f[CM1G_QA0 + n] : f[CM1G_Q3D0 + n], 0.0f, f[CM1G_QA0 + n], f[CM1G_QTEN0 + n],
                 nullptr, nullptr, f[CM1G_Q3D0 + n], nrk == 3 ? 1 : 0, n == 0 ? 0 : 1, dttmp,
                 (double)(0.01f * P.K.epsilon), 1);
      }
    }
    th3d = f[CM1G_TH3D];
    // subgrid TKE: advection and update (src/solve2.F:1709-1800)
    if (m->turb.cm1setup == 1) tke_stage(m, nrk, dttmp);
    // pressure and density (src/solve2.F:1536-1656)
    if (nrk == 3) {
      float *qvw = f[CM1G_Q3D0], *qcw = f[CM1G_Q3D1];
      float *qpten = f[CM1G_QPTEN], *qtten = f[CM1G_QTTEN], *qvten = f[CM1G_QVTEN], *qcten = f[CM1G_QCTEN];
      const int diab = (moist && m->c.ptype >= 1);
      // with the dry-mass adjustment the pass below recomputes prs from the
      // shifted pi', so only rho (for the mass sums) is stored here
      const int keep_prs = (m->c.apmasscon != 1);
      for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
        const Consts &K = P.K;
        int c = IXD(P, i, j, k);
        float pp = pp3d[c], th = th3d[c];
        if (diab) {
          pp = pp - dt * qpten[c];
          qpten[c] = pp;
          th = th - dt * qtten[c];
          qtten[c] = th;
          float q0 = qvw[c] - dt * qvten[c];
          qvten[c] = q0;
          qvw[c] = q0;
          float q1 = qcw[c] - dt * qcten[c];
          qcten[c] = q1;
          qcw[c] = q1;
          pp3d[c] = pp;
          th3d[c] = th;
        }
        float p = K.p00 * pow_f(P.pi0[k] + pp, K.cpdrd);
        if (keep_prs) prs[c] = p;
        if (moist)
          rho[c] = p / ((P.th0[k] + th) * (P.pi0[k] + pp) * (K.rd + fmaxf(0.0f, qvw[c]) * K.rv));
        else
          rho[c] = p / (K.rd * (P.th0[k] + th) * (P.pi0[k] + pp));
      });
      if (m->c.apmasscon == 1) mass_levels(m, rho, 1);
    }
    bc_fill(m, BC_U, u3d);
    bc_fill(m, BC_V, v3d);
    bc_fill(m, BC_W1, w3d);
    if (nrk < 3) {
      float *arr[5] = {rho, pp3d, nullptr, nullptr, nullptr};
      bc_fill_n(m, BC_P, arr, 2);
      float *sarr[4] = {th3d, f[CM1G_Q3D0], f[CM1G_Q3D1], f[CM1G_Q3D2]};
      bc_fill_n(m, BC_S, sarr, moist ? 4 : 1);
    }
    // passive tracers (src/solve2.F:1902-1970)
    for (int n = 0; n < m->c.npt; n++) {
      advs_run(m, nrk, from_a ? f[CM1G_PTA0 + n] : f[CM1G_PT3D0 + n], 0.0f, f[CM1G_PTA0 + n],
               f[CM1G_PTTEN0 + n], nullptr, nullptr, f[CM1G_PT3D0 + n],
               (nrk == 3 && m->c.pdtra == 1) ? 1 : 0, 1, dttmp, (double)(1.0f * P.K.epsilon), 0);
      if (nrk <= 2) bc_fill(m, BC_S, f[CM1G_PT3D0 + n]);
    }
    if (from_a) m->swapped = 0;
    // finish the dry-mass adjustment (src/solve2.F:1973-2030)
    if (nrk == 3 && m->c.apmasscon == 1) {
      const double *dscal = m->dscal;
      float *qvw = f[CM1G_Q3D0];
      for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
        const Consts &K = P.K;
        int c = IXD(P, i, j, k);
        const float tem = (float)dscal[2];
        float pp = pp3d[c] + tem;
        pp3d[c] = pp;
        float p = K.p00 * pow_f(P.pi0[k] + pp, K.cpdrd);
        prs[c] = p;
        if (moist)
          rho[c] = p / ((P.th0[k] + th3d[c]) * (P.pi0[k] + pp) * (K.rd + fmaxf(0.0f, qvw[c]) * K.rv));
        else
          rho[c] = p / (K.rd * (P.th0[k] + th3d[c]) * (P.pi0[k] + pp));
      });
    }
  }
  for (int n = 0; n < m->c.npt; n++) {
    if (m->c.pdtra == 1) pdefq_run(m, 0.0f, rho, f[CM1G_PT3D0 + n]);
    bc_fill(m, BC_S, f[CM1G_PT3D0 + n]);
  }
  m->synced = 0;
  return check_launch();
}

// --------------------------------------------------------------- mp_driver
extern "C" int cm1g_mp_driver(cm1g_model *m) {
  ARGCHK(m, "cm1g_mp_driver: null model");
  UNSWAP(m);
  m->synced = 0;
  if (m->c.imoist != 1) return 0;
  if (m->c.ptype == 0) {
    pdefq_run(m, 0.0f, m->f[CM1G_RHO], m->f[CM1G_Q3D0]);
    return check_launch();
  }
  kessler_branch_run(m);
  train_accumulate(m);
  return check_launch();
}

// ------------------------------------------------------------------ solve3
extern "C" int cm1g_solve3(cm1g_model *m) {
  ARGCHK(m, "cm1g_solve3: null model");
  UNSWAP(m);
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  float **f = m->f;
  const int moist = m->c.imoist;
  float *pp3d = f[CM1G_PP3D], *th3d = f[CM1G_TH3D], *ppx = f[CM1G_PPX], *rho = f[CM1G_RHO];
  float *rr = f[CM1G_RR], *rf = f[CM1G_RF];
  bc_fill(m, BC_P, pp3d);
  {
    float *sarr[4] = {th3d, f[CM1G_Q3D0], f[CM1G_Q3D1], f[CM1G_Q3D2]};
    bc_fill_n(m, BC_S, sarr, moist ? 4 : 1);
  }
  // rho's halo first (rf below reads it); ppx's halo after the update
  bc_fill(m, BC_P, rho);
  {
    // one pass over the step's final copies (src/solve3.F:246-705):
    // ppx = pp3d + ppx and the diabatic tendencies on the interior,
    // ua/va/wa = u3d/v3d/w3d over i,j = 0..n+2, ppi/tha/qa = pp3d/th3d/q3d,
    // rr = 1/rho and rf (extrapolated at k = 1 and nk+1) over 0..n+1
    const int diab = (moist && m->c.ptype >= 1);
    const float rdt = 1.0f / dt;
    float *qpten = f[CM1G_QPTEN], *qtten = f[CM1G_QTTEN], *qvten = f[CM1G_QVTEN], *qcten = f[CM1G_QCTEN];
    float *ua = f[CM1G_UA], *va = f[CM1G_VA], *wa = f[CM1G_WA];
    const float *u3d = f[CM1G_U3D], *v3d = f[CM1G_V3D], *w3d = f[CM1G_W3D];
    float *ppi = f[CM1G_PPI], *tha = f[CM1G_THA];
    float *qa0 = f[CM1G_QA0], *qa1 = f[CM1G_QA1], *qa2 = f[CM1G_QA2];
    const float *q0 = f[CM1G_Q3D0], *q1 = f[CM1G_Q3D1], *q2 = f[CM1G_Q3D2];
    // exchange instead of copying unless a caller holds field addresses
    const bool copy = m->ext_writers != 0;
    for3(m, 0, ni + 2, 0, nj + 2, 1, nk, [=] __device__(int i, int j, int k) {
      const int c = IXD(P, i, j, k);
      if (copy) {
        ua[c] = u3d[c];
        va[c] = v3d[c];
        wa[c] = w3d[c];
      }
      if (i > ni + 1 || j > nj + 1) return;
      const float pp = pp3d[c], th = th3d[c];
      float qv = 0.0f, qc = 0.0f;
      if (moist) {
        qv = q0[c];
        qc = q1[c];
      }
      if (copy) {
        ppi[c] = pp;
        tha[c] = th;
        if (moist) {
          qa0[c] = qv;
          qa1[c] = qc;
          qa2[c] = q2[c];
        }
      }
      const float rh = rho[c];
      rr[c] = 1.0f / rh;
      if (k >= 2) rf[c] = P.c1[k] * rho[IXD(P, i, j, k - 1)] + P.c2[k] * rh;
      else {
        rf[c] = P.cgs1 * rh + P.cgs2 * rho[IXD(P, i, j, 2)] + P.cgs3 * rho[IXD(P, i, j, 3)];
        rf[IXD(P, i, j, nk + 1)] = P.cgt1 * rho[IXD(P, i, j, nk)] + P.cgt2 * rho[IXD(P, i, j, nk - 1)] +
                                   P.cgt3 * rho[IXD(P, i, j, nk - 2)];
      }
      if (i < 1 || i > ni || j < 1 || j > nj) return;
      ppx[c] = pp + ppx[c];
      if (diab) {
        qpten[c] = (pp - qpten[c]) * rdt;
        qtten[c] = (th - qtten[c]) * rdt;
        qvten[c] = (qv - qvten[c]) * rdt;
        qcten[c] = (qc - qcten[c]) * rdt;
      }
    });
  }
  bc_fill(m, BC_P, ppx);
  if (m->ext_writers) {
    for (int n = 0; n < m->c.npt; n++) {
      float *pta = f[CM1G_PTA0 + n];
      const float *pt3d = f[CM1G_PT3D0 + n];
      for3(m, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
        pta[IXD(P, i, j, k)] = pt3d[IXD(P, i, j, k)];
      });
    }
  } else {
    // the step's end state becomes the a-arrays by exchanging buffers
    // (src/solve3.F's copies ua = u3d ... pta = pt3d); the next first
    // stage reads them in place of the 3-D arrays
    for (const int *pr : swap_pairs(m)) std::swap(f[pr[0]], f[pr[1]]);
    m->swapped = 1;
  }
  // ua..qa, pta now equal u3d..q3d, pt3d where the next stage reads them
  m->synced = !m->ext_writers;
  if (m->turb.cm1setup == 1) tke_copy(m);
  return check_launch();
}

int step_body(cm1g_model *m) {
  // the subgrid turbulence set-up of the step (src/cm1.F:3737)
  if (m->turb.cm1setup == 1) turb_sfc(m, 1);
  int rc = cm1g_solve1(m);
  if (!rc) rc = cm1g_solve2(m);
  if (!rc) rc = cm1g_mp_driver(m);
  if (!rc) rc = cm1g_solve3(m);
  return rc;
}

// the buffer arrangement a step starts from: every field pointer, the
// scalar updates' partner, and whether the a-arrays equal the 3d-arrays
std::vector<float *> field_map(const cm1g_model *m) {
  std::vector<float *> v(m->f, m->f + CM1G_NFIELDS);
  v.push_back(m->thb);
  v.push_back((float *)(uintptr_t)m->synced);  // decides the first acoustic stage's launches
  v.push_back((float *)(uintptr_t)m->swapped);  // decides what the first stage reads
  return v;
}
void set_field_map(cm1g_model *m, const std::vector<float *> &v) {
  for (int i = 0; i < CM1G_NFIELDS; i++) m->f[i] = v[i];
  m->thb = v[CM1G_NFIELDS];
  m->synced = (int)(uintptr_t)v[CM1G_NFIELDS + 1];
  m->swapped = (int)(uintptr_t)v[CM1G_NFIELDS + 2];
}

void graphs_clear(cm1g_model *m) {
  for (auto &g : m->graphs) hipGraphExecDestroy(g.exec);
  m->graphs.clear();
}

// One step as a graph: replay the one captured from this arrangement, or
// capture it now (capturing runs the host side of the step, which leaves
// the arrangement after the step, but launches nothing) and replay it.
static int step_graph(cm1g_model *m) {
  const std::vector<float *> key = field_map(m);
  for (auto &g : m->graphs)
    if (g.before == key) {
      HIPCHK(hipGraphLaunch(g.exec, m->stream));
      set_field_map(m, g.after);
      return 0;
    }
  if (m->graphs.size() >= 16) graphs_clear(m);
  HIPCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
  const int rc = step_body(m);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(m->stream, &graph);
  if (rc || e != hipSuccess) {
    // nothing ran: the device still holds the arrangement before the step
    if (graph) hipGraphDestroy(graph);
    set_field_map(m, key);
    if (rc) return rc;
    cm1g_set_error(std::string("cm1g_step: graph capture: ") + hipGetErrorString(e));
    return CM1G_E_HIP;
  }
  cm1g_model::StepGraph g;
  g.before = key;
  g.after = field_map(m);
  const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) {
    set_field_map(m, key);
    cm1g_set_error(std::string("cm1g_step: graph instantiate: ") + hipGetErrorString(ei));
    return CM1G_E_HIP;
  }
  m->graphs.push_back(g);
  HIPCHK(hipGraphLaunch(g.exec, m->stream));
  return 0;
}

extern "C" int cm1g_set_graphs(cm1g_model *m, int on) {
  ARGCHK(m && (on == 0 || on == 1), "cm1g_set_graphs: bad argument");
  m->no_graphs = !on;
  if (!on) graphs_clear(m);
  return 0;
}

extern "C" int cm1g_step(cm1g_model *m, float *cflmax) {
  ARGCHK(m, "cm1g_step: null model");
  // a kernel timer records events between launches: time kernel by kernel
  const bool graph = !m->no_graphs && m->kt_which == CM1G_KT_NONE;
  int rc = graph ? step_graph(m) : step_body(m);
  if (rc) return rc;
  if (cflmax) {
    HIPCHK(hipMemcpyAsync(&m->cflmax, m->fred, sizeof(float), hipMemcpyDeviceToHost, m->stream));
    rc = satadj_check(m);
    *cflmax = m->cflmax;
    if (rc) return rc;
    // the reference's MPI build stops once the Courant number reaches 1.5
    // (src/solve2.F:2082, src/cm1.F:4385-4397); the step's results stay in
    // place, as the reference writes its output before stopping
    if (m->cflmax >= 1.50f) {
      cm1g_set_error("cm1g_step: Courant number has exceeded 1.5 (src/cm1.F:4385-4397)");
      return CM1G_E_CFL;
    }
  }
  return 0;
}

// --------------------------------------------------- individual routines
extern "C" int cm1g_advs(cm1g_model *m, int nrk, int s_id, int s0_id, int sten_id, int pdef,
                         int pdefweno, float dt, double weps) {
  ARGCHK(field_ok(m, s_id) && field_ok(m, s0_id) && field_ok(m, sten_id), "cm1g_advs: bad field id");
  UNSWAP(m);
  m->synced = 0;
  ARGCHK(nrk >= 1 && nrk <= 3, "cm1g_advs: nrk must be 1..3");
  // the model's rru/rrv/rrw are used as given (no rho0 scaling here)
  float *sten = m->f[sten_id];
  m->advs_unscaled = 1;
  advs_run(m, nrk, m->f[s_id], 0.0f, m->f[s0_id], sten, nullptr, sten, nullptr, pdef, pdefweno, dt,
           weps, 0);
  m->advs_unscaled = 0;
  return check_launch();
}
extern "C" int cm1g_advu(cm1g_model *m, int nrk) {
  ARGCHK(m && nrk >= 1 && nrk <= 3, "cm1g_advu: bad argument");
  UNSWAP(m);
  adv_uvw(m, nrk, 1, 0);
  return check_launch();
}
extern "C" int cm1g_advv(cm1g_model *m, int nrk) {
  ARGCHK(m && nrk >= 1 && nrk <= 3, "cm1g_advv: bad argument");
  UNSWAP(m);
  adv_uvw(m, nrk, 2, 0);
  return check_launch();
}
extern "C" int cm1g_advw(cm1g_model *m, int nrk) {
  ARGCHK(m && nrk >= 1 && nrk <= 3, "cm1g_advw: bad argument");
  UNSWAP(m);
  adv_uvw(m, nrk, 4, 0);
  return check_launch();
}
extern "C" int cm1g_sound(cm1g_model *m, int nrk, float dttmp) {
  ARGCHK(m && nrk >= 1 && nrk <= 3, "cm1g_sound: bad argument");
  UNSWAP(m);
  m->synced = 0;
  (void)dttmp;
  if (m->c.psolver == 2) sounde_run(m, nrk);
  else sound_run(m, nrk);
  return check_launch();
}
extern "C" int cm1g_kessler_branch(cm1g_model *m) {
  ARGCHK(m, "null model");
  UNSWAP(m);
  m->synced = 0;
  kessler_branch_run(m);
  train_accumulate(m);
  return check_launch();
}
extern "C" int cm1g_pdefq(cm1g_model *m, float rmax, int q_id) {
  ARGCHK(field_ok(m, q_id), "cm1g_pdefq: bad field id");
  UNSWAP(m);
  pdefq_run(m, rmax, m->f[CM1G_RHO], m->f[q_id]);
  return check_launch();
}
extern "C" int cm1g_bc(cm1g_model *m, int id, int kind) {
  ARGCHK(field_ok(m, id) && kind >= 0 && kind <= 5, "cm1g_bc: bad argument");
  UNSWAP(m);
  m->synced = 0;
  bc_fill(m, kind, m->f[id]);
  return check_launch();
}
extern "C" int cm1g_calccflquick(cm1g_model *m, float dt, int u_id, int v_id, int w_id, float *cflmax) {
  ARGCHK(field_ok(m, u_id) && field_ok(m, v_id) && field_ok(m, w_id) && cflmax, "cm1g_calccflquick: bad argument");
  UNSWAP(m);
  cfl_run(m, dt, m->f[u_id], m->f[v_id], m->f[w_id]);
  HIPCHK(hipMemcpyAsync(cflmax, m->fred, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return check_launch();
}
extern "C" int cm1g_getdiv(cm1g_model *m, int div_id) {
  ARGCHK(field_ok(m, div_id) && div_id != CM1G_RRU && div_id != CM1G_RRV && div_id != CM1G_RRW,
         "cm1g_getdiv: bad field id");
  UNSWAP(m);
  getdiv_both(m, 1, 0, m->f[div_id]);
  return check_launch();
}
extern "C" int cm1g_getdivx(cm1g_model *m) {
  ARGCHK(m, "null model");
  UNSWAP(m);
  getdiv_both(m, 0, 1);
  scale_mass_fluxes(m);
  return check_launch();
}
extern "C" int cm1g_scalars(cm1g_model *m, double *out4) {
  ARGCHK(m && out4, "cm1g_scalars: bad argument");
  double d[4];
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(d, m->dscal, 4 * sizeof(double), hipMemcpyDeviceToHost));
  out4[0] = d[0];
  out4[1] = d[3];
  float c;
  HIPCHK(hipMemcpy(&c, m->fred, sizeof(float), hipMemcpyDeviceToHost));
  out4[2] = c;
  out4[3] = d[1];
  return 0;
}

// totals of dry-air mass and total water (calcmass/totmois analogues)
__global__ void __launch_bounds__(256) k_budget(KP P, const float *rho, const float *qv, const float *qc,
                                                const float *qr, int moist, double *out) {
  __shared__ double r1[256], r2[256];
  const int k = 1 + (int)blockIdx.x;
  double a1 = 0.0, a2 = 0.0;
  for (int t = threadIdx.x; t < P.ni * P.nj; t += blockDim.x) {
    int i = 1 + t % P.ni, j = 1 + t / P.ni;
    int c = IXD(P, i, j, k);
    double vol = (double)P.ruh[i + 2] * (double)P.rvh[j + 2] * (double)P.rmh[k];
    a1 += (double)rho[c] * vol;
    if (moist) a2 += (double)rho[c] * vol * ((double)qv[c] + (double)qc[c] + (double)qr[c]);
  }
  a1 = block_reduce256(a1, r1, OpAdd());
  a2 = block_reduce256(a2, r2, OpAdd());
  if (threadIdx.x == 0) {
    out[2 * k] = a1;
    out[2 * k + 1] = a2;
  }
}

extern "C" int cm1g_budgets(cm1g_model *m, double *tmass, double *tmois) {
  ARGCHK(m && tmass && tmois, "cm1g_budgets: bad argument");
  const int nk = m->P.nk;
  double *buf = m->dred + (size_t)m->P.ni * m->P.nj;
  hipLaunchKernelGGL(k_budget, dim3(nk), dim3(256), 0, m->stream, m->P, m->f[CM1G_RHO], m->f[CM1G_QA0],
                     m->f[CM1G_QA1], m->f[CM1G_QA2], m->c.imoist, buf);
  int rc = check_launch();
  if (rc) return rc;
  std::vector<double> h(2 * (nk + 1));
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(h.data(), buf, h.size() * sizeof(double), hipMemcpyDeviceToHost));
  double vol = (double)m->c.dx * (double)m->c.dy * (double)m->c.dz;
  double s1 = 0.0, s2 = 0.0;
  for (int k = 1; k <= nk; k++) { s1 += h[2 * k]; s2 += h[2 * k + 1]; }
  *tmass = s1 * vol;
  *tmois = s2 * vol;
  return 0;
}

// ------------------------------------------------------------ kernel timer
void ktimer_mark(cm1g_model *m, int which) {
  if (m->kt_which != which || which == CM1G_KT_NONE) return;
  if (m->kt_used == m->kt_ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return;
    m->kt_ev.push_back(e);
  }
  hipEventRecord(m->kt_ev[m->kt_used++], m->stream);
}

extern "C" int cm1g_ktimer_start(cm1g_model *m, int which) {
  ARGCHK(m, "cm1g_ktimer_start: null model");
  ARGCHK(which == CM1G_KT_NONE || which == CM1G_KT_SOUND_SUBSTEP, "cm1g_ktimer_start: unknown kernel id");
  m->kt_which = which;
  m->kt_used = 0;
  return 0;
}

extern "C" int cm1g_ktimer_stop(cm1g_model *m, double *total_ms, int64_t *launches,
                                double *bytes_per_launch, char *name, int name_len) {
  ARGCHK(m && total_ms && launches && bytes_per_launch, "cm1g_ktimer_stop: null argument");
  HIPCHK(hipStreamSynchronize(m->stream));
  double tot = 0.0;
  const size_t npair = m->kt_used / 2;
  for (size_t p = 0; p < npair; p++) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, m->kt_ev[2 * p], m->kt_ev[2 * p + 1]));
    tot += ms;
  }
  const double ncell = (double)m->P.ni * m->P.nj * m->P.nk;
  const char *nm = "";
  double fields = 0.0;
  if (m->kt_which == CM1G_KT_SOUND_SUBSTEP) {
    // one substep of sound (src/sound.F:307-835): u, v, w, ppd, pp3d, rru,
    // rrv, rrw read and written (16); uten, vten, wten, thv, ppten, ppterm,
    // piadv, pk1, pk2, qk, mm read (11).  The kernel evaluates pk1, pk2, qk
    // and mm on the fly instead of reading them, so it touches 23 of them.
    fields = 27.0;
    nm = "sound_sub";
  }
  *total_ms = tot;
  *launches = (int64_t)npair;
  *bytes_per_launch = fields * 4.0 * ncell;
  if (name && name_len > 0) {
    snprintf(name, (size_t)name_len, "%s", nm);
  }
  m->kt_which = CM1G_KT_NONE;
  m->kt_used = 0;
  return 0;
}
