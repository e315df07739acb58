// Kessler warm-rain microphysics on the device: kessler :8-138, satadj
// :146-360 (eqtset=2), k_fallout :368-396, fallout :404-477 of
// src/kessler.F; driver branch ptype=1 of src/mp_driver.F:100-232; pdefq
// of src/misclibs.F:1037-1117.  Column routines run one column per thread
// with the reference's k order (so their double-precision column sums are
// the reference's); pointwise routines run one cell per thread.
#include "cm1g_internal.h"

void pdefq_run(cm1g_model *m, float rmax, const float *rho, float *q) {
  const KP P = m->P;
  const int nk = P.nk;
  for2(m, 1, P.ni, 1, P.nj, [=] __device__(int i, int j) {
    double t1 = 0.0, t2 = 0.0;
    for (int k = 1; k <= nk; k++) {
      int c = IXD(P, i, j, k);
      float qq = q[c], r = rho[c];
      t1 = t1 + (double)(r * qq);
      if (qq < rmax) { qq = 0.0f; q[c] = 0.0f; }
      t2 = t2 + (double)(r * qq);
    }
    double t3 = (t1 + 1.0e-20) / (t2 + 1.0e-20);
    if (t3 < 0.0) t3 = 1.0;
    for (int k = 1; k <= nk; k++) {
      int c = IXD(P, i, j, k);
      q[c] = (float)(t3 * (double)q[c]);
    }
  });
}

// esl with the 171023 cap, and the saturation mixing ratio
__device__ __forceinline__ float esl_of(float t, float prs) {
  float esl = 611.2f * exp_f(17.67f * (t - 273.15f) / (t - 29.65f));
  return fminf(esl, prs * 0.5f);
}

void kessler_branch_run(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt;
  float *th3d = m->f[CM1G_TH3D], *pp3d = m->f[CM1G_PP3D], *rho = m->f[CM1G_RHO];
  float *prs = m->f[CM1G_PRS];
  float *qv = m->f[CM1G_Q3D0], *qc = m->f[CM1G_Q3D1], *qr = m->f[CM1G_Q3D2];
  float *vr = m->f[CM1G_QTEN2];
  float *rain = m->d2[CM1G_D_RAIN], *prate = m->d2[CM1G_D_PRATE];
  double *colbud = m->dred;
  int *iflag = m->iflag;

  // The reference's prelude (src/mp_driver.F:100-148) stores T, a copy of
  // T in thten, a copy of qv in qten(nqv) and a copy of rho in dum3; in the
  // Kessler branch these feed only budget diagnostics (geterain, the rr
  // argument of kessler/satadj) and the other schemes' tendency checks, and
  // kessler recomputes T itself, so nothing on this path reads them.
  pdefq_run(m, 0.0f, rho, qv);
  pdefq_run(m, P.K.qsmall, rho, qc);
  pdefq_run(m, P.K.qsmall, rho, qr);
  // k_fallout
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    int c = IXD(P, i, j, k);
    float v = 0.0f;
    if (qr[c] > 1.0e-12f) v = 14.34f * pow_f(rho[c] * qr[c], 0.1346f) * sqrtf(1.15f / rho[c]);
    vr[c] = v;
  });
  // fallout: per column, variable number of sedimentation substeps
  for2(m, 1, ni, 1, nj, [=] __device__(int i, int j) {
    const float rdz = P.rdz;
    float crmax = 0.0f;
    for (int k = 1; k <= nk; k++) crmax = fmaxf(crmax, vr[IXD(P, i, j, k)] * dt * rdz * P.mh[k]);
    int nfall = (int)(crmax + 1.0f);
    if (nfall < 1) nfall = 1;
    for (int k = nk - 1; k >= 1; k--)
      if (qr[IXD(P, i, j, k)] < 1.0e-8f) vr[IXD(P, i, j, k)] = vr[IXD(P, i, j, k + 1)];
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
  // kessler (pointwise)
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    const Consts &K = P.K;
    int c = IXD(P, i, j, k);
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
  });
  // satadj (nrk=4), iterative saturation adjustment
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    const Consts &K = P.K;
    const float converge = 2.0f * K.tsmall;
    int c = IXD(P, i, j, k);
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
  });
}
