// Kessler warm-rain microphysics on the device: kessler :8-138, satadj
// :146-360 (eqtset=2), k_fallout :368-396, fallout :404-477 of
// src/kessler.F; driver branch ptype=1 of src/mp_driver.F:100-232; pdefq
// of src/misclibs.F:1037-1117.  Column routines run one column per thread
// with the reference's k order (so their double-precision column sums are
// the reference's); pointwise routines run one cell per thread.
#include "cm1g_internal.h"
#include "cm1g_mlp.h"

void pdefq_run(cm1g_model *m, float rmax, const float *rho, float *q) {
  const KP P = m->P;
  const int nk = P.nk;
  for2(m, 1, P.ni, 1, P.nj, [=] __device__(int i, int j) {
    double t1 = 0.0, t2 = 0.0;
    for (int k = 1; k <= nk; k++) {
      int c = IXD(P, i, j, k);
      float qq = q[c], r = rho[c];
      t1 = t1 + (double)(r * qq);
      // a value that is +0 already is not stored again
      if (qq < rmax) {
        if (__float_as_uint(qq) != 0u) q[c] = 0.0f;
        qq = 0.0f;
      }
      t2 = t2 + (double)(r * qq);
    }
    double t3 = (t1 + 1.0e-20) / (t2 + 1.0e-20);
    if (t3 < 0.0) t3 = 1.0;
    // a factor of exactly 1 leaves every value unchanged bit for bit
    if (t3 == 1.0) return;
    for (int k = 1; k <= nk; k++) {
      int c = IXD(P, i, j, k);
      q[c] = (float)(t3 * (double)q[c]);
    }
  });
}

// pdefq of qv (rmax 0), qc and qr (rmax qsmall) in one column pass: the
// three species' double-precision sums run side by side, each in its own k
// order, so every species gets exactly its own pdefq's result while rho is
// read once
static void pdefq3_run(cm1g_model *m, const float *rho, float *qv, float *qc, float *qr) {
  const KP P = m->P;
  const int nk = P.nk;
  const float rc = P.K.qsmall;
  for2(m, 1, P.ni, 1, P.nj, [=] __device__(int i, int j) {
    double a1 = 0.0, a2 = 0.0, b1 = 0.0, b2 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int k = 1; k <= nk; k++) {
      const int c = IXD(P, i, j, k);
      const float r = rho[c];
      float v = qv[c], w = qc[c], x = qr[c];
      a1 = a1 + (double)(r * v);
      b1 = b1 + (double)(r * w);
      c1 = c1 + (double)(r * x);
      // a value that is +0 already is not stored again
      if (v < 0.0f) { v = 0.0f; qv[c] = 0.0f; }
      if (w < rc) {
        if (__float_as_uint(w) != 0u) qc[c] = 0.0f;
        w = 0.0f;
      }
      if (x < rc) {
        if (__float_as_uint(x) != 0u) qr[c] = 0.0f;
        x = 0.0f;
      }
      a2 = a2 + (double)(r * v);
      b2 = b2 + (double)(r * w);
      c2 = c2 + (double)(r * x);
    }
    double ta = (a1 + 1.0e-20) / (a2 + 1.0e-20), tb = (b1 + 1.0e-20) / (b2 + 1.0e-20),
           tc = (c1 + 1.0e-20) / (c2 + 1.0e-20);
    if (ta < 0.0) ta = 1.0;
    if (tb < 0.0) tb = 1.0;
    if (tc < 0.0) tc = 1.0;
    // a factor of exactly 1 (nothing clipped in the column, the usual case)
    // leaves every value unchanged bit for bit: that species is skipped
    const bool sa = ta != 1.0, sb = tb != 1.0, sc = tc != 1.0;
    if (!(sa || sb || sc)) return;
    for (int k = 1; k <= nk; k++) {
      const int c = IXD(P, i, j, k);
      if (sa) qv[c] = (float)(ta * (double)qv[c]);
      if (sb) qc[c] = (float)(tb * (double)qc[c]);
      if (sc) qr[c] = (float)(tc * (double)qr[c]);
    }
  });
}

// esl with the 171023 cap, and the saturation mixing ratio
__device__ __forceinline__ float esl_of(float t, float prs) {
  float esl = 611.2f * exp_f(17.67f * (t - 273.15f) / (t - 29.65f));
  return fminf(esl, prs * 0.5f);
}

// kessler (src/kessler.F:8-138) at one cell
__device__ __forceinline__ void kessler_cell(const KP &P, float dt, int c, int k, float *th3d, float *pp3d,
                                             float *prs, float *qv, float *qc, float *qr, const float *rho) {
  const Consts &K = P.K;
  float th = th3d[c], pp = pp3d[c];
  float t = (P.th0[k] + th) * (P.pi0[k] + pp);
  float qvnew = qv[c], qcnew = qc[c], qrnew = qr[c];
  float ar = 0.0f, cr = 0.0f, er = 0.0f;
  if (qcnew > 0.0f) {
    ar = fmaxf(0.001f * (qcnew - 0.001f), 0.0f);
    ar = ar * dt;
  }
  if (qcnew > 0.0f && qrnew > 0.0f) {
    cr = 2.2f * qcnew * pow_f(qrnew, 0.875f);
    cr = cr * dt;
  }
  float p = prs[c];
  if (qrnew > 0.0f) {
    float esl = esl_of(t, p);
    float qvs = K.eps * esl / (p - esl);
    if (qvnew < qvs) {
      float rq = rho[c] * qrnew;
      er = (1.6f + 30.3922f * pow_f(rq, 0.2046f)) * (1.0f - (qvnew / qvs)) * pow_f(rq, 0.525f) /
           ((2.03e4f + 9.584e6f / (qvs * p)) * rho[c]);
      er = fminf(er * dt, qrnew);
      if ((qvnew + er) > qvs) er = qvs - qvnew;
    }
  }
  if ((ar + cr) > qcnew) {
    float term1 = ar + cr;
    ar = qcnew * ar / term1;
    cr = qcnew * cr / term1;
  }
  qvnew = qvnew + er;
  qcnew = qcnew - (ar + cr);
  qrnew = qrnew + (ar + cr - er);
  if (er > 1.0e-7f) {
    float cpml = K.cp + K.cpv * qvnew + K.cpl * (qcnew + qrnew);
    float cvml = K.cv + K.cvv * qvnew + K.cpl * (qcnew + qrnew);
    float rm = K.rd + K.rv * qvnew;
    th = th - er * ((K.lv1 - K.lv2 * t) / (K.cpdcv * cvml * (P.pi0[k] + pp)) -
                    (P.th0[k] + th) * (K.rv / cvml) * (1.0f - K.rovcp * cpml / rm));
    pp = pow_f(rho[c] * rm * (P.th0[k] + th) * K.rp00, K.rddcv) - P.pi0[k];
    p = K.p00 * pow_f(P.pi0[k] + pp, K.cpdrd);
    th3d[c] = th;
    pp3d[c] = pp;
    prs[c] = p;
  }
  qv[c] = qvnew;
  qc[c] = qcnew;
  qr[c] = qrnew;
}

// satadj (src/kessler.F:146-360, eqtset=2) at one cell
__device__ __forceinline__ void satadj_cell(const KP &P, int c, int k, float *th3d, float *pp3d, float *prs,
                                            float *qv, float *qc, float *qr, const float *rho, int *iflag) {
  const Consts &K = P.K;
  const float converge = 2.0f * K.tsmall;
  float th = th3d[c], pp = pp3d[c], p = prs[c];
  float tnew = (P.th0[k] + th) * (P.pi0[k] + pp);
  float esl = esl_of(tnew, p);
  float qvs = K.eps * esl / (p - esl);
  float qv0 = qv[c], qc0 = qc[c];
  if (qc0 > 1.0e-12f || qv0 > qvs) {
    float qvnew = qv0, qcnew = qc0;
    float ql = 0.0f;
    ql = ql + qc0;
    ql = ql + qr[c];
    ql = fmaxf(0.0f, ql);
    float qi = 0.0f;
    float cvml = K.cv + K.cvv * qvnew + K.cpl * ql + K.cpi * qi;
    float lhv = K.lv1 - K.lv2 * tnew;
    float t1 = (lhv - K.rv * tnew) / cvml;
    float d1 = t1 * 17.67f * 243.5f;
    int n = 0;
    float tlast = tnew, pnew = 0.0f;
    bool doit = true;
    while (doit) {
      n = n + 1;
      float tm = tnew - 29.65f;
      float dqv = (qvs - qvnew) / (1.0f + d1 * qvs / (tm * tm));
      dqv = fminf(dqv, qcnew);
      if ((qvnew + dqv) < 1.0e-20f) dqv = 1.0e-20f - qvnew;
      qvnew = qvnew + dqv;
      qcnew = qcnew - dqv;
      tnew = tnew - dqv * t1;
      pnew = rho[c] * (K.rd + K.rv * qvnew) * tnew;
      doit = false;
      if (fabsf(tnew - tlast) > converge) {
        tlast = tnew;
        esl = esl_of(tnew, p);
        qvs = K.eps * esl / (pnew - esl);
        doit = true;
      }
      if (n == 100) {
        atomicOr(iflag, 1);
        doit = false;
      }
    }
    prs[c] = pnew;
    float ppn = pow_f(pnew * K.rp00, K.rovcp) - P.pi0[k];
    pp3d[c] = ppn;
    th3d[c] = tnew / (P.pi0[k] + ppn) - P.th0[k];
    qc[c] = qcnew;
    qv[c] = qvnew;
  }
}

// ---------------------------------------------------------------- learned surrogate
// The network replaces kessler (src/kessler.F:8-138: autoconversion,
// accretion, rain evaporation); sedimentation before it and the saturation
// adjustment after it stay exact.  Its 8 inputs at a cell, raw: T, p, rho,
// qv, qc, qr, dt, z/ztop; its 4 outputs: the changes of theta, qv, qc, qr
// over kessler.  (No counterpart in the reference.)
struct KessAct {
  bool ar, cr, er;  // autoconversion, accretion, evaporation possible
  float t;          // temperature
};
// kessler's own conditions for each process (src/kessler.F:48-75), from a
// cell's theta', pi', p, qv, qc, qr at level k
__device__ __forceinline__ KessAct kess_active(const KP &P, int k, float th, float pp, float p, float qv,
                                               float qc, float qr) {
  KessAct a;
  a.t = (P.th0[k] + th) * (P.pi0[k] + pp);
  a.ar = qc > 0.001f;
  a.cr = qc > 0.0f && qr > 0.0f;
  a.er = false;
  if (qr > 0.0f) {
    const float esl = esl_of(a.t, p);
    a.er = qv < P.K.eps * esl / (p - esl);
  }
  return a;
}

// Projection of the proposed changes onto what kessler can do at the cell:
// cloud only loses (autoconversion, accretion) and at most all of it,
// vapour only gains (evaporation) and only where rain evaporates, and rain
// takes exactly the difference, dqr = -(dqv + dqc), so qv + qc + qr is kept
// by construction; evaporation is limited so that no rain goes negative.
// The network's own dqr (and dtheta) serve as diagnostics.  Returns false
// (the column then falls back to plain kessler) for a non-finite proposal.
__device__ __forceinline__ bool sur_project(const KessAct &a, float qv, float qc, float qr, const float *d,
                                            float &nqv, float &nqc, float &nqr) {
  for (int q = 1; q < 4; q++)
    if (!isfinite(d[q])) return false;
  float dc = (a.ar || a.cr) ? fminf(d[2], 0.0f) : 0.0f;
  dc = fmaxf(dc, -qc);
  float dv = a.er ? fmaxf(d[1], 0.0f) : 0.0f;
  dv = fminf(dv, qr - dc);
  nqv = qv + dv;
  nqc = qc + dc;
  nqr = qr - (dv + dc);
  if (nqr < 0.0f) nqr = 0.0f;
  return true;
}

// The surrogate in place of kessler, followed by satadj (nrk = 4).  Three
// launches: (1) a pass that lists the cells where one of kessler's
// processes can act (wavefront ballots, one atomic add per wavefront; the
// list order does not matter, every cell's result is its own); (2) the
// fused network over the listed cells only, writing projected proposals and
// flagging columns that must fall back; (3) one pointwise pass that applies
// the proposal (or plain kessler in a flagged column) and then satadj, as
// the plain path's fused kessler + satadj pass does.
static void surrogate_kessler_satadj(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  float *th3d = m->f[CM1G_TH3D], *pp3d = m->f[CM1G_PP3D], *rho = m->f[CM1G_RHO], *prs = m->f[CM1G_PRS];
  float *qv = m->f[CM1G_Q3D0], *qc = m->f[CM1G_Q3D1], *qr = m->f[CM1G_Q3D2];
  float *pqv = m->sy, *pqc = m->sz, *pqr = m->tx;  // the projected proposals
  int *colflag = m->colflag, *list = m->actlist, *iflag = m->iflag;
  unsigned long long *cnt = m->dcount, *nact = m->dcount + 6;
  const float *zh = m->kf[CM1G_K_ZH];
  const float rztop = 1.0f / m->hk[CM1G_K_ZF][nk + 1];
  const float *prm = mlp_params(m->sur), *nm = mlp_norm(m->sur);
  for2(m, 1, ni, 1, nj, [=] __device__(int i, int j) {
    colflag[(i - 1) + (j - 1) * ni] = 0;
    if (i == 1 && j == 1) {
      cnt[1] += 1ULL;
      *nact = 0ULL;
    }
  });
  const int64_t n = (int64_t)ni * nj * nk;
  // (n < 2^31: cm1g_set_surrogate checks it); cell s: i fastest, then j, k
  auto cell = [=] __device__(int s, int &i, int &j, int &k) {
    const int r = s / ni;
    i = 1 + (s - r * ni);
    j = 1 + (r % nj);
    k = 1 + r / nj;
  };
  launch<256>(m, dim3((unsigned)((n + 255) / 256)), dim3(256), [=] __device__(int bx, int, int) {
    const int s = bx * 256 + (int)threadIdx.x;
    bool act = false;
    if (s < n) {
      int i, j, k;
      cell(s, i, j, k);
      const int c = IXD(P, i, j, k);
      const float c_ = qc[c], r_ = qr[c];
      act = c_ > 0.001f || (c_ > 0.0f && r_ > 0.0f);
      if (!act && r_ > 0.0f) {
        const KessAct a = kess_active(P, k, th3d[c], pp3d[c], prs[c], qv[c], c_, r_);
        act = a.er;
      }
    }
    const unsigned long long mask = __ballot(act);
    const int lane = (int)(threadIdx.x & 63);
    unsigned long long base = 0ULL;
    if (lane == 0 && mask) base = atomicAdd(nact, (unsigned long long)__popcll(mask));
    base = __shfl(base, 0, 64);
    if (act) list[base + (unsigned long long)__popcll(mask & ((1ULL << lane) - 1ULL))] = s;
  });
  // raw values of a listed cell: theta, pi', p, rho, qv, qc, qr
  auto load = [=] __device__(int64_t t, float *r) {
    int i, j, k;
    cell(list[t], i, j, k);
    const int c = IXD(P, i, j, k);
    r[0] = th3d[c];
    r[1] = pp3d[c];
    r[2] = prs[c];
    r[3] = rho[c];
    r[4] = qv[c];
    r[5] = qc[c];
    r[6] = qr[c];
  };
  auto prep = [=] __device__(int64_t t, const float *r, float *x) {
    int i, j, k;
    cell(list[t], i, j, k);
    const float raw[MLP_IN] = {(P.th0[k] + r[0]) * (P.pi0[k] + r[1]), r[2], r[3], r[4], r[5], r[6], dt,
                               zh[k] * rztop};
    for (int q = 0; q < MLP_IN; q++) x[q] = (raw[q] - nm[q]) / nm[MLP_IN + q];
  };
  // outside the trusted input range, or with an unusable proposal, the
  // column falls back
  auto out = [=] __device__(int64_t t, const float *y, const float *r) {
    int i, j, k;
    cell(list[t], i, j, k);
    const int c = IXD(P, i, j, k);
    const KessAct a = kess_active(P, k, r[0], r[1], r[2], r[4], r[5], r[6]);
    if (!(a.ar || a.cr || a.er)) return;
    const float raw[MLP_IN] = {a.t, r[2], r[3], r[4], r[5], r[6], dt, zh[k] * rztop};
    bool inside = true;
    for (int q = 0; q < MLP_IN; q++) {
      const float x = (raw[q] - nm[q]) / nm[MLP_IN + q];
      inside = inside && x >= nm[NM_XLO + q] && x <= nm[NM_XHI + q];
    }
    float d[MLP_OUT];
    for (int q = 0; q < MLP_OUT; q++) d[q] = y[q] * nm[2 * MLP_IN + MLP_OUT + q] + nm[2 * MLP_IN + q];
    float v, w, rr;
    const bool ok = inside && sur_project(a, r[4], r[5], r[6], d, v, w, rr);
    if (ok) {
      pqv[c] = v;
      pqc[c] = w;
      pqr[c] = rr;
    } else {
      atomicOr(&colflag[(i - 1) + (j - 1) * ni], 1);
      atomicAdd(&cnt[inside ? 5 : 4], 1ULL);
    }
  };
  const int64_t groups = (n + 16 * MLP_G * MLP_WAVES - 1) / (16 * MLP_G * MLP_WAVES);
  const int nbx = (int)(groups < 1024 ? groups : 1024);
  launch<512>(m, dim3(nbx), dim3(512),
              [=] __device__(int bx, int, int) { mlp_tiles(prm, n, nact, load, prep, out, bx, nbx); });
  // apply the projected proposal as kessler applies its own result (the
  // theta, pi', p update after evaporation above 1e-7), or plain kessler in
  // flagged columns; then satadj
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    const Consts &K = P.K;
    const int c = IXD(P, i, j, k);
    if (colflag[(i - 1) + (j - 1) * ni]) {
      kessler_cell(P, dt, c, k, th3d, pp3d, prs, qv, qc, qr, rho);
      atomicAdd(&cnt[2], 1ULL);
    } else {
      const KessAct a = kess_active(P, k, th3d[c], pp3d[c], prs[c], qv[c], qc[c], qr[c]);
      if (a.ar || a.cr || a.er) {
        const float qvnew = pqv[c], qcnew = pqc[c], qrnew = pqr[c];
        const float er = qvnew - qv[c];
        if (er > 1.0e-7f) {
          float th = th3d[c], pp = pp3d[c];
          const float cpml = K.cp + K.cpv * qvnew + K.cpl * (qcnew + qrnew);
          const float cvml = K.cv + K.cvv * qvnew + K.cpl * (qcnew + qrnew);
          const float rm = K.rd + K.rv * qvnew;
          th = th - er * ((K.lv1 - K.lv2 * a.t) / (K.cpdcv * cvml * (P.pi0[k] + pp)) -
                          (P.th0[k] + th) * (K.rv / cvml) * (1.0f - K.rovcp * cpml / rm));
          pp = pow_f(rho[c] * rm * (P.th0[k] + th) * K.rp00, K.rddcv) - P.pi0[k];
          th3d[c] = th;
          pp3d[c] = pp;
          prs[c] = K.p00 * pow_f(P.pi0[k] + pp, K.cpdrd);
        }
        qv[c] = qvnew;
        qc[c] = qcnew;
        qr[c] = qrnew;
      }
    }
    satadj_cell(P, c, k, th3d, pp3d, prs, qv, qc, qr, rho, iflag);
  });
}

// (state entering kessler, its change over kessler) pairs of every cell
// with cloud or rain or any change, and of one in `harv_keep` of the others,
// appended to the harvest buffer while it has room
static void harvest_pairs(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  const float *th3d = m->f[CM1G_TH3D], *rho = m->f[CM1G_RHO];
  const float *qv = m->f[CM1G_Q3D0], *qc = m->f[CM1G_Q3D1], *qr = m->f[CM1G_Q3D2];
  const float *b0 = m->sx, *b1 = m->sy, *b2 = m->sz, *b3 = m->tx, *b4 = m->ty, *b5 = m->tz;
  float *buf = m->harv;
  const int64_t cap = m->harv_cap;
  const unsigned keep = (unsigned)m->harv_keep;
  unsigned long long *cnt = m->dcount;
  const float *zh = m->kf[CM1G_K_ZH];
  const float rztop = 1.0f / m->hk[CM1G_K_ZF][nk + 1];
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k);
    const float d0 = th3d[c] - b0[c], d1 = qv[c] - b1[c], d2 = qc[c] - b2[c], d3 = qr[c] - b3[c];
    const bool active = b2[c] > 0.0f || b3[c] > 0.0f || d0 != 0.0f || d1 != 0.0f || d2 != 0.0f || d3 != 0.0f;
    unsigned h = (unsigned)c * 2654435761u ^ (unsigned)cnt[3] * 2246822519u;
    h ^= h >> 15;
    h *= 2654435761u;
    h ^= h >> 13;
    if (!active && (h % keep) != 0u) return;
    const unsigned long long idx = atomicAdd(&cnt[0], 1ULL);
    if ((int64_t)idx >= cap) return;
    float *o = buf + (int64_t)idx * 12;
    o[0] = (P.th0[k] + b0[c]) * (P.pi0[k] + b4[c]);
    o[1] = b5[c];
    o[2] = rho[c];
    o[3] = b1[c];
    o[4] = b2[c];
    o[5] = b3[c];
    o[6] = dt;
    o[7] = zh[k] * rztop;
    o[8] = d0;
    o[9] = d1;
    o[10] = d2;
    o[11] = d3;
  });
  launch<64>(m, dim3(1), dim3(1), [=] __device__(int, int, int) { cnt[3] += 1ULL; });
}

// fallout: per column, one downward pass for the fall speed (k_fallout,
// src/kessler.F:368-396), its maximum Courant number (taken over the speeds
// before the fill, as the reference does; the maximum does not depend on the
// order) and the fill of near-empty levels from the level above
// (src/kessler.F:446-448); then a variable number of sedimentation substeps
// (src/kessler.F:404-477).  Writes qr, rain, prate, the filled speeds into
// qten(nqr) and each column's rain budget into the reduction workspace.
static void fallout_run(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  const float *rho = m->f[CM1G_RHO];
  float *qr = m->f[CM1G_Q3D2];
  float *vr = m->f[CM1G_QTEN2];
  float *rain = m->d2[CM1G_D_RAIN], *prate = m->d2[CM1G_D_PRATE];
  double *colbud = m->dred;
  for2(m, 1, ni, 1, nj, [=] __device__(int i, int j) {
    const float rdz = P.rdz;
    float crmax = 0.0f, vabove = 0.0f;
    for (int k = nk; k >= 1; k--) {
      const int c = IXD(P, i, j, k);
      const float q = qr[c], r = rho[c];
      float v = 0.0f;
      if (q > 1.0e-12f) v = 14.34f * pow_f(r * q, 0.1346f) * sqrtf(1.15f / r);
      crmax = fmaxf(crmax, v * dt * rdz * P.mh[k]);
      if (k < nk && q < 1.0e-8f) v = vabove;
      vr[c] = v;
      vabove = v;
    }
    int nfall = (int)(crmax + 1.0f);
    if (nfall < 1) nfall = 1;
    float dtfall = dt / (float)nfall;
    double bud = 0.0;
    float rn = rain[IX2D(P, i, j)], pr = prate[IX2D(P, i, j)];
    for (int n = 1; n <= nfall; n++) {
      // rq(k) for the whole column is needed before any update; recompute
      // rq(k+1) from the not-yet-updated level above while sweeping up.
      float rq_k = rho[IXD(P, i, j, 1)] * vr[IXD(P, i, j, 1)] * fmaxf(0.0f, qr[IXD(P, i, j, 1)]);
      const float rq1 = rq_k;
      for (int k = 1; k <= nk; k++) {
        int c = IXD(P, i, j, k);
        float rq_kp1 = 0.0f;
        if (k < nk) {
          int cp = IXD(P, i, j, k + 1);
          rq_kp1 = rho[cp] * vr[cp] * fmaxf(0.0f, qr[cp]);
        }
        qr[c] = qr[c] + dtfall * (rq_kp1 - rq_k) * rdz * P.mh[k] / rho[c];
        rq_k = rq_kp1;
      }
      rn = rn + 0.1f * dtfall * rq1;
      pr = rq1;
      bud = bud + (double)(dtfall * rho[IXD(P, i, j, 1)] * vr[IXD(P, i, j, 1)] *
                           fmaxf(0.0f, qr[IXD(P, i, j, 1)]) * P.ruh[i + 2] * P.rvh[j + 2]);
    }
    rain[IX2D(P, i, j)] = rn;
    prate[IX2D(P, i, j)] = pr;
    colbud[(i - 1) + (j - 1) * ni] = bud;
  });
}

void kessler_branch_run(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  float *th3d = m->f[CM1G_TH3D], *pp3d = m->f[CM1G_PP3D], *rho = m->f[CM1G_RHO];
  float *prs = m->f[CM1G_PRS];
  float *qv = m->f[CM1G_Q3D0], *qc = m->f[CM1G_Q3D1], *qr = m->f[CM1G_Q3D2];
  int *iflag = m->iflag;

  // The reference's prelude (src/mp_driver.F:100-148) stores T, a copy of
  // T in thten, a copy of qv in qten(nqv) and a copy of rho in dum3; in the
  // Kessler branch these feed only budget diagnostics (geterain, the rr
  // argument of kessler/satadj) and the other schemes' tendency checks, and
  // kessler recomputes T itself, so nothing on this path reads them.
  pdefq3_run(m, rho, qv, qc, qr);
  fallout_run(m);
  if (m->harv) {
    // the state entering kessler (after sedimentation)
    float *b0 = m->sx, *b1 = m->sy, *b2 = m->sz, *b3 = m->tx, *b4 = m->ty, *b5 = m->tz;
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      const int c = IXD(P, i, j, k);
      b0[c] = th3d[c];
      b1[c] = qv[c];
      b2[c] = qc[c];
      b3[c] = qr[c];
      b4[c] = pp3d[c];
      b5[c] = prs[c];
    });
  }
  if (!m->sur && !m->harv) {
    // kessler then satadj (nrk=4) at each cell: both are pointwise, so one
    // pass reads the cell's state once
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      const int c = IXD(P, i, j, k);
      kessler_cell(P, dt, c, k, th3d, pp3d, prs, qv, qc, qr, rho);
      satadj_cell(P, c, k, th3d, pp3d, prs, qv, qc, qr, rho, iflag);
    });
    return;
  }
  if (m->sur) {
    surrogate_kessler_satadj(m);
    return;
  } else {
    // kessler (pointwise)
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      kessler_cell(P, dt, IXD(P, i, j, k), k, th3d, pp3d, prs, qv, qc, qr, rho);
    });
  }
  if (m->harv) harvest_pairs(m);
  // satadj (nrk=4), iterative saturation adjustment
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    satadj_cell(P, IXD(P, i, j, k), k, th3d, pp3d, prs, qv, qc, qr, rho, iflag);
  });
}

// ---------------------------------------------------------------- C ABI
extern "C" int cm1g_set_surrogate(cm1g_model *m, cm1g_mlp *net) {
  if (!m) {
    cm1g_set_error("cm1g_set_surrogate: null model");
    return CM1G_E_ARG;
  }
  if (net && (m->c.imoist != 1 || m->c.ptype != 1)) {
    cm1g_set_error("cm1g_set_surrogate: the surrogate replaces Kessler (imoist=1, ptype=1)");
    return CM1G_E_ARG;
  }
  if (net && m->harv) {
    cm1g_set_error("cm1g_set_surrogate: stop the harvest first");
    return CM1G_E_ARG;
  }
  if (net && (int64_t)m->c.ni * m->c.nj * m->c.nk >= (1LL << 31)) {
    cm1g_set_error("cm1g_set_surrogate: more than 2^31 cells");
    return CM1G_E_ARG;
  }
  if (net && !m->actlist) {
    // the list of cells the network runs on: at most every cell
    HIPCHK(hipStreamSynchronize(m->stream));
    if (hipMalloc((void **)&m->actlist, (size_t)m->c.ni * m->c.nj * m->c.nk * sizeof(int)) != hipSuccess) {
      m->actlist = nullptr;
      cm1g_set_error("cm1g_set_surrogate: cannot allocate the cell list");
      return CM1G_E_NOMEM;
    }
  }
  m->sur = net;
  graphs_clear(m);
  return 0;
}

extern "C" int cm1g_set_harvest(cm1g_model *m, float *buf, int64_t capacity, int keep_one_in) {
  if (!m || (buf && (capacity < 1 || keep_one_in < 1))) {
    cm1g_set_error("cm1g_set_harvest: bad argument");
    return CM1G_E_ARG;
  }
  if (buf && m->sur) {
    cm1g_set_error("cm1g_set_harvest: the harvest records plain Kessler; remove the surrogate first");
    return CM1G_E_ARG;
  }
  HIPCHK(hipStreamSynchronize(m->stream));
  m->harv = buf;
  m->harv_cap = buf ? capacity : 0;
  m->harv_keep = keep_one_in;
  if (buf) HIPCHK(hipMemset(m->dcount, 0, sizeof(unsigned long long)));
  graphs_clear(m);
  return 0;
}

extern "C" int cm1g_surrogate_counts(cm1g_model *m, int64_t *out6) {
  if (!m || !out6) {
    cm1g_set_error("cm1g_surrogate_counts: null argument");
    return CM1G_E_ARG;
  }
  unsigned long long h[6];
  HIPCHK(hipStreamSynchronize(m->stream));
  HIPCHK(hipMemcpy(h, m->dcount, sizeof(h), hipMemcpyDeviceToHost));
  for (int q = 0; q < 6; q++) out6[q] = (int64_t)h[q];
  return 0;
}

// ------------------------------------------------- routine-level entries
#define ARGCHK(cond, msg)  \
  do {                     \
    if (!(cond)) {         \
      cm1g_set_error(msg); \
      return CM1G_E_ARG;   \
    }                      \
  } while (0)

// The routines of kessler_module one by one, for a host that calls them
// itself (src/mp_driver.F:192-232): each on the model's fields th3d, pp3d,
// prs, rho, q3d(nqv..nqr) and qten(nqr) as the fall speed.
extern "C" int cm1g_k_fallout(cm1g_model *m) {
  ARGCHK(m && m->c.imoist == 1, "cm1g_k_fallout: needs a moist model");
  { const int rc_ = unswap(m); if (rc_) return rc_; }
  m->synced = 0;
  const KP P = m->P;
  const float *rho = m->f[CM1G_RHO], *qr = m->f[CM1G_Q3D2];
  float *vr = m->f[CM1G_QTEN2];
  // src/kessler.F:382-390
  for3(m, 1, P.ni, 1, P.nj, 1, P.nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k);
    float v = 0.0f;
    if (qr[c] > 1.0e-12f) v = 14.34f * pow_f(rho[c] * qr[c], 0.1346f) * sqrtf(1.15f / rho[c]);
    vr[c] = v;
  });
  return check_launch();
}

extern "C" int cm1g_fallout(cm1g_model *m, double *train) {
  ARGCHK(m && train && m->c.imoist == 1, "cm1g_fallout: bad argument");
  { const int rc_ = unswap(m); if (rc_) return rc_; }
  m->synced = 0;
  HIPCHK(hipMemcpyAsync(m->dscal + 3, train, sizeof(double), hipMemcpyHostToDevice, m->stream));
  fallout_run(m);
  train_accumulate(m);
  HIPCHK(hipMemcpyAsync(train, m->dscal + 3, sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  return check_launch();
}

extern "C" int cm1g_geterain(cm1g_model *m, float dt, float cpx, float lx1, int t_id, int rr_id, int q_id,
                             double *erain) {
  ARGCHK(m && erain && field_ok(m, t_id) && field_ok(m, rr_id) && field_ok(m, q_id),
         "cm1g_geterain: bad argument");
  { const int rc_ = unswap(m); if (rc_) return rc_; }
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj;
  const float *t = m->f[t_id], *rr = m->f[rr_id], *q = m->f[q_id], *vr = m->f[CM1G_QTEN2];
  double *bud = m->dred;
  // src/kessler.F:545-552: per row j, a double sum over i of REAL products
  launch<64>(m, dim3((nj + 63) / 64), dim3(64), [=] __device__(int bx, int, int) {
    const int j = 1 + bx * 64 + (int)threadIdx.x;
    if (j > nj) return;
    double b = 0.0;
    for (int i = 1; i <= ni; i++) {
      const int c = IXD(P, i, j, 1);
      b = b + (double)(dt * vr[c] * rr[c] * q[c] * (cpx * t[c] - lx1) * P.ruh[i + 2] * P.rvh[j + 2]);
    }
    bud[j - 1] = b;
  });
  std::vector<double> h(nj);
  HIPCHK(hipMemcpyAsync(h.data(), bud, nj * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  double e = *erain;
  for (int j = 0; j < nj; j++) e = e + h[j] * (double)P.dx * (double)P.dy;
  *erain = e;
  return check_launch();
}

extern "C" int cm1g_kessler(cm1g_model *m) {
  ARGCHK(m && m->c.imoist == 1, "cm1g_kessler: needs a moist model");
  { const int rc_ = unswap(m); if (rc_) return rc_; }
  m->synced = 0;
  const KP P = m->P;
  const float dt = m->c.dt;
  float *th3d = m->f[CM1G_TH3D], *pp3d = m->f[CM1G_PP3D], *prs = m->f[CM1G_PRS];
  float *qv = m->f[CM1G_Q3D0], *qc = m->f[CM1G_Q3D1], *qr = m->f[CM1G_Q3D2];
  const float *rho = m->f[CM1G_RHO];
  for3(m, 1, P.ni, 1, P.nj, 1, P.nk, [=] __device__(int i, int j, int k) {
    kessler_cell(P, dt, IXD(P, i, j, k), k, th3d, pp3d, prs, qv, qc, qr, rho);
  });
  return check_launch();
}

extern "C" int cm1g_satadj(cm1g_model *m) {
  ARGCHK(m && m->c.imoist == 1, "cm1g_satadj: needs a moist model");
  { const int rc_ = unswap(m); if (rc_) return rc_; }
  m->synced = 0;
  const KP P = m->P;
  float *th3d = m->f[CM1G_TH3D], *pp3d = m->f[CM1G_PP3D], *prs = m->f[CM1G_PRS];
  float *qv = m->f[CM1G_Q3D0], *qc = m->f[CM1G_Q3D1], *qr = m->f[CM1G_Q3D2];
  const float *rho = m->f[CM1G_RHO];
  int *iflag = m->iflag;
  for3(m, 1, P.ni, 1, P.nj, 1, P.nk, [=] __device__(int i, int j, int k) {
    satadj_cell(P, IXD(P, i, j, k), k, th3d, pp3d, prs, qv, qc, qr, rho, iflag);
  });
  int rc = check_launch();
  if (rc) return rc;
  return satadj_check(m);
}
