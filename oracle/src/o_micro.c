/* Oracle: Kessler warm-rain microphysics (test infrastructure only).
 * kessler :8-138, satadj :146-360 (eqtset=2), k_fallout :368-396,
 * fallout :404-477 of src/kessler.F; driver branch ptype=1 of
 * src/mp_driver.F:100-232.  exp/pow go through o_expf/o_powf (the C
 * library's expf/powf by default, what the reference's REAL exp() and **
 * lower to). */
#include "cm1o_internal.h"
#include <stdio.h>

void o_k_fallout(cm1o_model *M, const float *rho, const float *qr3d, float *vr) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(vr, i, j, k) = 0.0f;
        if (A3(qr3d, i, j, k) > 1.0e-12f)
          A3(vr, i, j, k) = 14.34f * o_powf(A3(rho, i, j, k) * A3(qr3d, i, j, k), 0.1346f) *
                            sqrtf(1.15f / A3(rho, i, j, k));
      }
}

void o_fallout(cm1o_model *M, float dt, const float *rho, float *q3d, float *vq) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *mh = M->kf[K_MH];
  const float rdz = M->rdz;
  float *rain = M->d2[D_RAIN], *prate = M->d2[D_PRATE];
  float rq[512 + 4];
  int i, j, k, n;
  const double tem = (double)(M->c.dx * M->c.dy);
  for (j = 1; j <= nj; j++) {
    double bud = 0.0;
    for (i = 1; i <= ni; i++) {
      float crmax = 0.0f;
      for (k = 1; k <= nk; k++) crmax = fmaxf(crmax, A3(vq, i, j, k) * dt * rdz * mh[k]);
      int nfall = (int)(crmax + 1.0f);
      if (nfall < 1) nfall = 1;
      for (k = nk - 1; k >= 1; k--)
        if (A3(q3d, i, j, k) < 1.0e-8f) A3(vq, i, j, k) = A3(vq, i, j, k + 1);
      float dtfall = dt / (float)nfall;
      for (n = 1; n <= nfall; n++) {
        for (k = 1; k <= nk; k++)
          rq[k] = A3(rho, i, j, k) * A3(vq, i, j, k) * fmaxf(0.0f, A3(q3d, i, j, k));
        rq[nk + 1] = 0.0f;
        for (k = 1; k <= nk; k++)
          A3(q3d, i, j, k) = A3(q3d, i, j, k) + dtfall * (rq[k + 1] - rq[k]) * rdz * mh[k] / A3(rho, i, j, k);
        rain[IX2(M, i, j)] = rain[IX2(M, i, j)] + 0.1f * dtfall * rq[1];
        prate[IX2(M, i, j)] = rq[1];
        bud = bud + (double)(dtfall * A3(rho, i, j, 1) * A3(vq, i, j, 1) *
                             fmaxf(0.0f, A3(q3d, i, j, 1)) * RUH(i) * RVH(j));
      }
    }
    /* budget: fallout's rr argument is the rho copy (src/mp_driver.F:227) */
    M->train = M->train + bud * tem;
  }
}

void o_kessler(cm1o_model *M, float dt, float *tmp, float *rho, float *pp3d, float *th3d,
               float *prs, float *qv3d, float *qc3d, float *qr3d) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0];
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(tmp, i, j, k) = (th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k));
        float qvnew = A3(qv3d, i, j, k), qcnew = A3(qc3d, i, j, k), qrnew = A3(qr3d, i, j, k);
        float ar = 0.0f, cr = 0.0f, er = 0.0f;
        if (qcnew > 0.0f) {
          ar = fmaxf(0.001f * (qcnew - 0.001f), 0.0f);
          ar = ar * dt;
        }
        if (qcnew > 0.0f && qrnew > 0.0f) {
          cr = 2.2f * qcnew * o_powf(qrnew, 0.875f);
          cr = cr * dt;
        }
        if (qrnew > 0.0f) {
          float t = A3(tmp, i, j, k);
          float esl = 611.2f * o_expf(17.67f * (t - 273.15f) / (t - 29.65f));
          esl = fminf(esl, A3(prs, i, j, k) * 0.5f);
          float qvs = K->eps * esl / (A3(prs, i, j, k) - esl);
          if (qvnew < qvs) {
            float rq = A3(rho, i, j, k) * qrnew;
            er = (1.6f + 30.3922f * o_powf(rq, 0.2046f)) * (1.0f - (qvnew / qvs)) * o_powf(rq, 0.525f) /
                 ((2.03e4f + 9.584e6f / (qvs * A3(prs, i, j, k))) * A3(rho, i, j, k));
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
          float cpml = K->cp + K->cpv * qvnew + K->cpl * (qcnew + qrnew);
          float cvml = K->cv + K->cvv * qvnew + K->cpl * (qcnew + qrnew);
          float rm = K->rd + K->rv * qvnew;
          A3(th3d, i, j, k) = A3(th3d, i, j, k) - er * (
              (K->lv1 - K->lv2 * A3(tmp, i, j, k)) / (K->cpdcv * cvml * (pi0[k] + A3(pp3d, i, j, k)))
              - (th0[k] + A3(th3d, i, j, k)) * (K->rv / cvml) * (1.0f - K->rovcp * cpml / rm));
          A3(pp3d, i, j, k) = o_powf(A3(rho, i, j, k) * rm * (th0[k] + A3(th3d, i, j, k)) * K->rp00, K->rddcv) - pi0[k];
          A3(prs, i, j, k) = K->p00 * o_powf(pi0[k] + A3(pp3d, i, j, k), K->cpdrd);
        }
        A3(qv3d, i, j, k) = qvnew;
        A3(qc3d, i, j, k) = qcnew;
        A3(qr3d, i, j, k) = qrnew;
      }
}

/* returns nonzero if the iteration failed to converge (stopcm1 path) */
int o_satadj(cm1o_model *M, int nrk, float dt, float *rho, float *pp3d, float *prs,
             float *th3d, float *qv, float *qc, float *qr) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const cm1o_const *K = &M->k;
  const float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0];
  const float converge = (nrk == 4) ? 2.0f * K->tsmall : 20.0f * K->tsmall;
  int i, j, k, iflag = 0;
  (void)dt;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        float tnew = (th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k));
        float esl = 611.2f * o_expf(17.67f * (tnew - 273.15f) / (tnew - 29.65f));
        esl = fminf(esl, A3(prs, i, j, k) * 0.5f);
        float qvs = K->eps * esl / (A3(prs, i, j, k) - esl);
        if (A3(qc, i, j, k) > 1.0e-12f || A3(qv, i, j, k) > qvs) {
          float qvnew = A3(qv, i, j, k), qcnew = A3(qc, i, j, k);
          float ql = 0.0f;
          ql = ql + A3(qc, i, j, k);
          ql = ql + A3(qr, i, j, k);
          ql = fmaxf(0.0f, ql);
          float qi = fmaxf(0.0f, 0.0f);
          float cvml = K->cv + K->cvv * qvnew + K->cpl * ql + K->cpi * qi;
          float lhv = K->lv1 - K->lv2 * tnew;
          float t1 = (lhv - K->rv * tnew) / cvml;
          float d1 = t1 * 17.67f * 243.5f;
          int n = 0;
          float tlast = tnew, pnew = 0.0f;
          int doit = 1;
          while (doit) {
            n = n + 1;
            float tm = tnew - 29.65f;
            float dqv = (qvs - qvnew) / (1.0f + d1 * qvs / (tm * tm));
            dqv = fminf(dqv, qcnew);
            if ((qvnew + dqv) < 1.0e-20f) dqv = 1.0e-20f - qvnew;
            qvnew = qvnew + dqv;
            qcnew = qcnew - dqv;
            tnew = tnew - dqv * t1;
            pnew = A3(rho, i, j, k) * (K->rd + K->rv * qvnew) * tnew;
            doit = 0;
            if (fabsf(tnew - tlast) > converge) {
              tlast = tnew;
              esl = 611.2f * o_expf(17.67f * (tnew - 273.15f) / (tnew - 29.65f));
              esl = fminf(esl, A3(prs, i, j, k) * 0.5f);
              qvs = K->eps * esl / (pnew - esl);
              doit = 1;
            }
            if (n == 100) { iflag = 1; doit = 0; }
          }
          A3(prs, i, j, k) = pnew;
          A3(pp3d, i, j, k) = o_powf(pnew * K->rp00, K->rovcp) - pi0[k];
          A3(th3d, i, j, k) = tnew / (pi0[k] + A3(pp3d, i, j, k)) - th0[k];
          A3(qc, i, j, k) = qcnew;
          A3(qv, i, j, k) = qvnew;
        }
      }
  return iflag;
}

/* mp_driver ptype=1 branch, src/mp_driver.F:100-232 */
void o_mp_driver(cm1o_model *M) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float dt = M->c.dt;
  const float *pi0 = M->kf[K_PI0], *th0 = M->kf[K_TH0];
  float *dum1 = M->f[F_DUM1], *dum3 = M->f[F_DUM3], *thten = M->f[F_THTEN];
  float *th3d = M->f[F_TH3D], *pp3d = M->f[F_PP3D], *rho = M->f[F_RHO], *prs = M->f[F_PRS];
  float *qv = M->f[F_Q3D0], *qc = M->f[F_Q3D1], *qr = M->f[F_Q3D2];
  float *qtenv = M->f[F_QTEN0], *qtenr = M->f[F_QTEN2];
  int i, j, k;
  if (M->c.imoist != 1) return;
  if (M->c.ptype == 0) {
    o_pdefq(M, 0.0f, rho, qv);
    return;
  }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(dum1, i, j, k) = (th0[k] + A3(th3d, i, j, k)) * (pi0[k] + A3(pp3d, i, j, k));
        A3(thten, i, j, k) = A3(dum1, i, j, k);
        A3(qtenv, i, j, k) = A3(qv, i, j, k);
      }
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) A3(dum3, i, j, k) = A3(rho, i, j, k);
  o_pdefq(M, 0.0f, rho, qv);
  o_pdefq(M, M->k.qsmall, rho, qc);
  o_pdefq(M, M->k.qsmall, rho, qr);
  o_k_fallout(M, rho, qr, qtenr);
  o_fallout(M, dt, rho, qr, qtenr);
  o_kessler(M, dt, dum1, rho, pp3d, th3d, prs, qv, qc, qr);
  if (o_satadj(M, 4, dt, rho, pp3d, prs, th3d, qv, qc, qr))
    fprintf(stderr, "cm1o: satadj did not converge\n");
}
