/* Oracle: split-explicit acoustic substeps with the vertically implicit
 * w-pi solve, psolver=3, and the explicit solver sounde, psolver=2; no
 * terrain, Cartesian (test infrastructure only).  Follows src/sound.F:10-859; the sound() argument mapping used by
 * src/solve2.F:1122-1136 is: mm=dum5, ppd=dum7, fpk=dum8, qk=t12,
 * pk1=t13, pk2=t23, piadv=sadv, thv=t11, ppterm=t22.
 * Also getdiv/getdivx/calccflquick/pdefq from src/misclibs.F. */
#include "cm1o_internal.h"

void o_sound(cm1o_model *M, int nrk, float dttmp) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float dt = M->c.dt;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz, cp = M->k.cp;
  const float alph = M->c.alph, kdiv = M->c.kdiv, smeps = M->k.smeps;
  const float *mh = M->kf[K_MH], *mf = M->kf[K_MF], *rr0 = M->kf[K_RR0], *rf0 = M->kf[K_RF0];
  const float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  float *dum1 = M->f[F_DUM1], *mm = M->f[F_DUM5], *ppd = M->f[F_DUM7], *fpk = M->f[F_DUM8];
  float *qk = M->f[F_T12], *pk1 = M->f[F_T13], *pk2 = M->f[F_T23];
  float *piadv = M->f[F_SADV], *thv = M->f[F_T11], *ppterm = M->f[F_T22];
  float *u3d = M->f[F_U3D], *v3d = M->f[F_V3D], *w3d = M->f[F_W3D];
  float *ua = M->f[F_UA], *va = M->f[F_VA], *wa = M->f[F_WA];
  float *uten = M->f[F_UTEN], *vten = M->f[F_VTEN], *wten = M->f[F_WTEN];
  float *rru = M->f[F_RRU], *rrv = M->f[F_RRV], *rrw = M->f[F_RRW];
  float *ppi = M->f[F_PPI], *pp3d = M->f[F_PP3D], *ppten = M->f[F_PPTEN], *ppx = M->f[F_PPX];
  float *radbcw = M->d2[D_RADBCW], *radbce = M->d2[D_RADBCE];
  float *radbcs = M->d2[D_RADBCS], *radbcn = M->d2[D_RADBCN];
  int i, j, k, n, nloop;
  float dts;
  const float beta = 1.0f - alph;
  const float asq = alph * alph;
  const int nsound = M->c.nsound;
  (void)dttmp;

  if (nrk == 1) {
    nloop = (int)lroundf((float)nsound / 3.0f);
    dts = dt / ((float)nloop * 3.0f);
    if (dts > (dt / (float)nsound)) {
      nloop = nloop + 1;
      dts = dt / ((float)nloop * 3.0f);
    }
  } else if (nrk == 2) {
    nloop = (int)(0.5f * (float)nsound);
    dts = dt / (float)nsound;
  } else {
    nloop = nsound;
    dts = dt / (float)nsound;
  }

  for (j = 1; j <= nj + 1; j++)
    for (i = 1; i <= ni + 1; i++) {
      M->dtu[IX2(M, i, j)] = dts * M->dtu0[IX2(M, i, j)];
      M->dtv[IX2(M, i, j)] = dts * M->dtv0[IX2(M, i, j)];
    }

  for (k = 1; k <= nk; k++) {
    float tem = dts * rdz * mh[k];
    float r2 = dts * rdz * mh[k] * rr0[k] * rf0[k + 1];
    float r1 = dts * rdz * mh[k] * rr0[k] * rf0[k];
    float tem1 = dts * rdz * cp * mf[k];
    float c1a = c1[k], c1b = c1[k + 1], c2a = c2[k], c2b = c2[k + 1];
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(qk, i, j, k) = tem * A3(ppterm, i, j, k);
        A3(pk2, i, j, k) = r2 * (-c2b * A3(piadv, i, j, k + 1) + (1.0f - c1b) * A3(piadv, i, j, k));
        A3(pk1, i, j, k) = r1 * (+c1a * A3(piadv, i, j, k - 1) + (c2a - 1.0f) * A3(piadv, i, j, k));
        A3(mm, i, j, k) = tem1 * (c2a * A3(thv, i, j, k) + c1a * A3(thv, i, j, k - 1));
      }
  }
  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) {
      A3(pk1, i, j, 1) = 0.0f;
      A3(pk2, i, j, nk) = 0.0f;
    }

  if (nrk == 1) {
    for (k = 1; k <= nk; k++)
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) {
          A3(ppd, i, j, k) = A3(ppx, i, j, k);
          A3(rru, i, j, k) = 0.0f;
          A3(rrv, i, j, k) = 0.0f;
          A3(rrw, i, j, k) = A3(wa, i, j, k) * beta;
        }
  } else {
    for (k = 1; k <= nk; k++)
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) {
          A3(u3d, i, j, k) = A3(ua, i, j, k);
          A3(rru, i, j, k) = 0.0f;
          A3(v3d, i, j, k) = A3(va, i, j, k);
          A3(rrv, i, j, k) = 0.0f;
          A3(w3d, i, j, k) = A3(wa, i, j, k);
          A3(rrw, i, j, k) = A3(wa, i, j, k) * beta;
          A3(pp3d, i, j, k) = A3(ppi, i, j, k);
          A3(ppd, i, j, k) = A3(ppx, i, j, k);
        }
  }

  for (n = 1; n <= nloop; n++) {
    /* open boundary tendencies: ssopenbcw/e/s/n, src/bc.F:2392-2684 */
    if (M->c.wbc == 2 && M->ibw == 1)
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          A3(u3d, 1, j, k) = A3(u3d, 1, j, k) + dts * (-radbcw[(j + 2) + (size_t)k * M->py] *
              (A3(u3d, 2, j, k) - A3(u3d, 1, j, k)) * rdx * UH(1) + A3(uten, 1, j, k));
    if (M->c.ebc == 2 && M->ibe == 1)
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          A3(u3d, ni + 1, j, k) = A3(u3d, ni + 1, j, k) + dts * (-radbce[(j + 2) + (size_t)k * M->py] *
              (A3(u3d, ni + 1, j, k) - A3(u3d, ni, j, k)) * rdx * UH(ni) + A3(uten, ni + 1, j, k));
    if (M->c.sbc == 2 && M->ibs == 1)
      for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
          A3(v3d, i, 1, k) = A3(v3d, i, 1, k) + dts * (-radbcs[(i + 2) + (size_t)k * M->px] *
              (A3(v3d, i, 2, k) - A3(v3d, i, 1, k)) * rdy * VH(1) + A3(vten, i, 1, k));
    if (M->c.nbc == 2 && M->ibn == 1)
      for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
          A3(v3d, i, nj + 1, k) = A3(v3d, i, nj + 1, k) + dts * (-radbcn[(i + 2) + (size_t)k * M->px] *
              (A3(v3d, i, nj + 1, k) - A3(v3d, i, nj, k)) * rdy * VH(nj) + A3(vten, i, nj + 1, k));

    {
      const float tem1 = rdx * cp * 0.5f, tem2 = rdy * cp * 0.5f;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj + 1; j++)
          for (i = 1; i <= ni + 1; i++) {
            A3(u3d, i, j, k) = A3(u3d, i, j, k) + M->dtu[IX2(M, i, j)] * (A3(uten, i, j, k)
                - tem1 * (A3(ppd, i, j, k) - A3(ppd, i - 1, j, k)) * UF(i) * (A3(thv, i, j, k) + A3(thv, i - 1, j, k)));
            A3(v3d, i, j, k) = A3(v3d, i, j, k) + M->dtv[IX2(M, i, j)] * (A3(vten, i, j, k)
                - tem2 * (A3(ppd, i, j, k) - A3(ppd, i, j - 1, k)) * VF(j) * (A3(thv, i, j, k) + A3(thv, i, j - 1, k)));
          }
    }

    {
      const float temx = dts * 0.5f * rdx, temy = dts * 0.5f * rdy;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            float div = (A3(u3d, i + 1, j, k) - A3(u3d, i, j, k)) * rdx * UH(i)
                      + (A3(v3d, i, j + 1, k) - A3(v3d, i, j, k)) * rdy * VH(j);
            if (fabsf(div) < smeps) div = 0.0f;
            A3(fpk, i, j, k) = dts * (A3(ppten, i, j, k) - A3(ppterm, i, j, k) * div)
                + (-(A3(u3d, i, j, k) * (A3(piadv, i, j, k) - A3(piadv, i - 1, j, k))
                     + A3(u3d, i + 1, j, k) * (A3(piadv, i + 1, j, k) - A3(piadv, i, j, k))) * temx * UH(i)
                   - (A3(v3d, i, j, k) * (A3(piadv, i, j, k) - A3(piadv, i, j - 1, k))
                     + A3(v3d, i, j + 1, k) * (A3(piadv, i, j + 1, k) - A3(piadv, i, j, k))) * temy * VH(j))
                + beta * (A3(w3d, i, j, k) * (A3(pk1, i, j, k) + A3(qk, i, j, k))
                        + A3(w3d, i, j, k + 1) * (A3(pk2, i, j, k) - A3(qk, i, j, k)));
          }
    }

    /* vertically implicit solve for w, src/sound.F:659-733 */
    {
      float *dum2 = M->f[F_DUM2];
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          k = 2;
          float cc = -asq * (-A3(mm, i, j, k) * (A3(pk2, i, j, k) - A3(qk, i, j, k)));
          float bb = 1.0f - asq * (+A3(mm, i, j, k) * (-(A3(pk1, i, j, k) + A3(qk, i, j, k))
                                                 + (A3(pk2, i, j, k - 1) - A3(qk, i, j, k - 1))));
          float fwk = dts * A3(wten, i, j, k) + beta * (-A3(mm, i, j, k) * (A3(pp3d, i, j, k) - A3(pp3d, i, j, k - 1)));
          float dd = A3(w3d, i, j, k) + fwk + alph * (-A3(mm, i, j, k) *
                     ((A3(pp3d, i, j, k) + A3(fpk, i, j, k)) - (A3(pp3d, i, j, k - 1) + A3(fpk, i, j, k - 1))));
          float r1 = 1.0f / bb;
          A3(dum1, i, j, k) = -cc * r1;
          A3(dum2, i, j, k) = dd * r1;
        }
      for (k = 3; k <= nk - 1; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            float aa = -asq * (+A3(mm, i, j, k) * (A3(pk1, i, j, k - 1) + A3(qk, i, j, k - 1)));
            float cc = -asq * (-A3(mm, i, j, k) * (A3(pk2, i, j, k) - A3(qk, i, j, k)));
            float bb = 1.0f - asq * (+A3(mm, i, j, k) * (-(A3(pk1, i, j, k) + A3(qk, i, j, k))
                                                   + (A3(pk2, i, j, k - 1) - A3(qk, i, j, k - 1))));
            float fwk = dts * A3(wten, i, j, k) + beta * (-A3(mm, i, j, k) * (A3(pp3d, i, j, k) - A3(pp3d, i, j, k - 1)));
            float dd = A3(w3d, i, j, k) + fwk + alph * (-A3(mm, i, j, k) *
                       ((A3(pp3d, i, j, k) + A3(fpk, i, j, k)) - (A3(pp3d, i, j, k - 1) + A3(fpk, i, j, k - 1))));
            float r1 = 1.0f / (aa * A3(dum1, i, j, k - 1) + bb);
            A3(dum1, i, j, k) = -cc * r1;
            A3(dum2, i, j, k) = (dd - aa * A3(dum2, i, j, k - 1)) * r1;
          }
      k = nk;
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          float aa = -asq * (+A3(mm, i, j, k) * (A3(pk1, i, j, k - 1) + A3(qk, i, j, k - 1)));
          float bb = 1.0f - asq * (+A3(mm, i, j, k) * (-(A3(pk1, i, j, k) + A3(qk, i, j, k))
                                                 + (A3(pk2, i, j, k - 1) - A3(qk, i, j, k - 1))));
          float fwk = dts * A3(wten, i, j, k) + beta * (-A3(mm, i, j, k) * (A3(pp3d, i, j, k) - A3(pp3d, i, j, k - 1)));
          float dd = A3(w3d, i, j, k) + fwk + alph * (-A3(mm, i, j, k) *
                     ((A3(pp3d, i, j, k) + A3(fpk, i, j, k)) - (A3(pp3d, i, j, k - 1) + A3(fpk, i, j, k - 1))));
          float r1 = 1.0f / (aa * A3(dum1, i, j, k - 1) + bb);
          A3(w3d, i, j, k) = (dd - aa * A3(dum2, i, j, k - 1)) * r1;
        }
      for (k = nk - 1; k >= 2; k--)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++)
            A3(w3d, i, j, k) = A3(dum1, i, j, k) * A3(w3d, i, j, k + 1) + A3(dum2, i, j, k);
    }

    /* new pi, divergence damping, src/sound.F:739-765 */
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(ppd, i, j, k) = A3(pp3d, i, j, k);
          A3(pp3d, i, j, k) = A3(pp3d, i, j, k) + A3(fpk, i, j, k)
              + alph * (A3(w3d, i, j, k) * (A3(pk1, i, j, k) + A3(qk, i, j, k))
                      + A3(w3d, i, j, k + 1) * (A3(pk2, i, j, k) - A3(qk, i, j, k)));
          if (fabsf(A3(pp3d, i, j, k)) < smeps) A3(pp3d, i, j, k) = 0.0f;
          A3(dum1, i, j, k) = kdiv * (A3(pp3d, i, j, k) - A3(ppd, i, j, k));
          A3(ppd, i, j, k) = A3(pp3d, i, j, k) + A3(dum1, i, j, k);
        }
    if (n < nloop) o_bcp(M, ppd);

    if (n < nloop) {
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj + 1; j++)
          for (i = 1; i <= ni + 1; i++) {
            A3(rru, i, j, k) = A3(rru, i, j, k) + A3(u3d, i, j, k);
            A3(rrv, i, j, k) = A3(rrv, i, j, k) + A3(v3d, i, j, k);
            A3(rrw, i, j, k) = A3(rrw, i, j, k) + A3(w3d, i, j, k);
          }
    } else {
      const float tavg = 1.0f / (float)nloop;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj + 1; j++)
          for (i = 1; i <= ni + 1; i++) {
            A3(rru, i, j, k) = (A3(rru, i, j, k) + A3(u3d, i, j, k)) * tavg;
            A3(rrv, i, j, k) = (A3(rrv, i, j, k) + A3(v3d, i, j, k)) * tavg;
            A3(rrw, i, j, k) = (A3(rrw, i, j, k) + A3(w3d, i, j, k) * alph) * tavg;
          }
    }
  }

  if (nrk == 3)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(ppx, i, j, k) = A3(dum1, i, j, k);
}

/* Explicit acoustic substeps, psolver=2, no terrain, Cartesian:
 * src/sounde.F:10-1026 (w :548-563, pi :775-790).  Same argument mapping
 * as o_sound (src/solve2.F:1100-1117); pk1=t13, pk2=t23. */
void o_sounde(cm1o_model *M, int nrk, float dttmp) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float dt = M->c.dt;
  const float rdx = M->rdx, rdy = M->rdy, rdz = M->rdz, cp = M->k.cp;
  const float kdiv = M->c.kdiv, smeps = M->k.smeps;
  const float *mh = M->kf[K_MH], *mf = M->kf[K_MF], *rr0 = M->kf[K_RR0], *rf0 = M->kf[K_RF0];
  const float *c1 = M->kf[K_C1], *c2 = M->kf[K_C2];
  float *dum1 = M->f[F_DUM1], *ppd = M->f[F_DUM7];
  float *pk1 = M->f[F_T13], *pk2 = M->f[F_T23];
  float *piadv = M->f[F_SADV], *thv = M->f[F_T11], *ppterm = M->f[F_T22];
  float *u3d = M->f[F_U3D], *v3d = M->f[F_V3D], *w3d = M->f[F_W3D];
  float *ua = M->f[F_UA], *va = M->f[F_VA], *wa = M->f[F_WA];
  float *uten = M->f[F_UTEN], *vten = M->f[F_VTEN], *wten = M->f[F_WTEN];
  float *rru = M->f[F_RRU], *rrv = M->f[F_RRV], *rrw = M->f[F_RRW];
  float *ppi = M->f[F_PPI], *pp3d = M->f[F_PP3D], *ppten = M->f[F_PPTEN], *ppx = M->f[F_PPX];
  float *radbcw = M->d2[D_RADBCW], *radbce = M->d2[D_RADBCE];
  float *radbcs = M->d2[D_RADBCS], *radbcn = M->d2[D_RADBCN];
  int i, j, k, n, nloop;
  float dts;
  const int nsound = M->c.nsound;
  (void)dttmp;

  if (nrk == 1) {
    nloop = (int)lroundf((float)nsound / 3.0f);
    dts = dt / ((float)nloop * 3.0f);
    if (dts > (dt / (float)nsound)) {
      nloop = nloop + 1;
      dts = dt / ((float)nloop * 3.0f);
    }
  } else if (nrk == 2) {
    nloop = (int)(0.5f * (float)nsound);
    dts = dt / (float)nsound;
  } else {
    nloop = nsound;
    dts = dt / (float)nsound;
  }

  for (j = 1; j <= nj + 1; j++)
    for (i = 1; i <= ni + 1; i++) {
      M->dtu[IX2(M, i, j)] = dts * M->dtu0[IX2(M, i, j)];
      M->dtv[IX2(M, i, j)] = dts * M->dtv0[IX2(M, i, j)];
    }

  /* pi0 vertical-advection coefficients (src/sounde.F:117-136, 160-170) */
  for (k = 1; k <= nk; k++) {
    float p1 = dts * rdz * mh[k] * rr0[k] * rf0[k + 1] * (1.0f - c1[k + 1]);
    float p2 = -(dts * rdz * mh[k] * rr0[k] * rf0[k + 1] * c2[k + 1]);
    float p3 = dts * rdz * mh[k] * rr0[k] * rf0[k] * c1[k];
    float p4 = dts * rdz * mh[k] * rr0[k] * rf0[k] * (c2[k] - 1.0f);
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        A3(pk2, i, j, k) = p2 * A3(piadv, i, j, k + 1) + p1 * A3(piadv, i, j, k);
        A3(pk1, i, j, k) = p3 * A3(piadv, i, j, k - 1) + p4 * A3(piadv, i, j, k);
      }
  }
  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) {
      A3(pk1, i, j, 1) = 0.0f;
      A3(pk2, i, j, nk) = 0.0f;
    }

  if (nrk == 1) {
    for (k = 1; k <= nk; k++)
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) {
          A3(ppd, i, j, k) = A3(ppx, i, j, k);
          A3(rru, i, j, k) = 0.0f;
          A3(rrv, i, j, k) = 0.0f;
          A3(rrw, i, j, k) = 0.0f;
        }
  } else {
    for (k = 1; k <= nk; k++)
      for (j = 0; j <= nj + 1; j++)
        for (i = 0; i <= ni + 1; i++) {
          A3(u3d, i, j, k) = A3(ua, i, j, k);
          A3(rru, i, j, k) = 0.0f;
          A3(v3d, i, j, k) = A3(va, i, j, k);
          A3(rrv, i, j, k) = 0.0f;
          A3(w3d, i, j, k) = A3(wa, i, j, k);
          A3(rrw, i, j, k) = 0.0f;
          A3(pp3d, i, j, k) = A3(ppi, i, j, k);
          A3(ppd, i, j, k) = A3(ppx, i, j, k);
        }
  }

  for (n = 1; n <= nloop; n++) {
    if (M->c.wbc == 2 && M->ibw == 1)
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          A3(u3d, 1, j, k) = A3(u3d, 1, j, k) + dts * (-radbcw[(j + 2) + (size_t)k * M->py] *
              (A3(u3d, 2, j, k) - A3(u3d, 1, j, k)) * rdx * UH(1) + A3(uten, 1, j, k));
    if (M->c.ebc == 2 && M->ibe == 1)
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          A3(u3d, ni + 1, j, k) = A3(u3d, ni + 1, j, k) + dts * (-radbce[(j + 2) + (size_t)k * M->py] *
              (A3(u3d, ni + 1, j, k) - A3(u3d, ni, j, k)) * rdx * UH(ni) + A3(uten, ni + 1, j, k));
    if (M->c.sbc == 2 && M->ibs == 1)
      for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
          A3(v3d, i, 1, k) = A3(v3d, i, 1, k) + dts * (-radbcs[(i + 2) + (size_t)k * M->px] *
              (A3(v3d, i, 2, k) - A3(v3d, i, 1, k)) * rdy * VH(1) + A3(vten, i, 1, k));
    if (M->c.nbc == 2 && M->ibn == 1)
      for (k = 1; k <= nk; k++)
        for (i = 1; i <= ni; i++)
          A3(v3d, i, nj + 1, k) = A3(v3d, i, nj + 1, k) + dts * (-radbcn[(i + 2) + (size_t)k * M->px] *
              (A3(v3d, i, nj + 1, k) - A3(v3d, i, nj, k)) * rdy * VH(nj) + A3(vten, i, nj + 1, k));

    /* u, v (src/sounde.F:333-352) */
    {
      const float tem1 = rdx * cp * 0.5f, tem2 = rdy * cp * 0.5f;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj + 1; j++)
          for (i = 1; i <= ni + 1; i++) {
            A3(u3d, i, j, k) = A3(u3d, i, j, k) + M->dtu[IX2(M, i, j)] * (A3(uten, i, j, k)
                - tem1 * UF(i) * (A3(ppd, i, j, k) - A3(ppd, i - 1, j, k)) * (A3(thv, i, j, k) + A3(thv, i - 1, j, k)));
            A3(v3d, i, j, k) = A3(v3d, i, j, k) + M->dtv[IX2(M, i, j)] * (A3(vten, i, j, k)
                - tem2 * VF(j) * (A3(ppd, i, j, k) - A3(ppd, i, j - 1, k)) * (A3(thv, i, j, k) + A3(thv, i, j - 1, k)));
          }
    }
    /* w (src/sounde.F:548-563) */
    for (k = 2; k <= nk; k++) {
      const float tem1 = rdz * cp * mf[k];
      const float c1a = c1[k], c2a = c2[k];
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++)
          A3(w3d, i, j, k) = A3(w3d, i, j, k) + dts * (A3(wten, i, j, k)
              - tem1 * (A3(ppd, i, j, k) - A3(ppd, i, j, k - 1))
                     * (c2a * A3(thv, i, j, k) + c1a * A3(thv, i, j, k - 1)));
    }
    /* pi (src/sounde.F:770-795) */
    {
      const float temx = dts * 0.5f * rdx, temy = dts * 0.5f * rdy;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj; j++)
          for (i = 1; i <= ni; i++) {
            float div = ((A3(u3d, i + 1, j, k) - A3(u3d, i, j, k)) * rdx * UH(i)
                       + (A3(v3d, i, j + 1, k) - A3(v3d, i, j, k)) * rdy * VH(j))
                       + (A3(w3d, i, j, k + 1) - A3(w3d, i, j, k)) * rdz * mh[k];
            if (fabsf(div) < smeps) div = 0.0f;
            float delp = dts * (A3(ppten, i, j, k) - A3(ppterm, i, j, k) * div)
                + (-(A3(u3d, i, j, k) * (A3(piadv, i, j, k) - A3(piadv, i - 1, j, k))
                     + A3(u3d, i + 1, j, k) * (A3(piadv, i + 1, j, k) - A3(piadv, i, j, k))) * temx * UH(i)
                   - (A3(v3d, i, j, k) * (A3(piadv, i, j, k) - A3(piadv, i, j - 1, k))
                     + A3(v3d, i, j + 1, k) * (A3(piadv, i, j + 1, k) - A3(piadv, i, j, k))) * temy * VH(j))
                + (A3(w3d, i, j, k) * A3(pk1, i, j, k) + A3(w3d, i, j, k + 1) * A3(pk2, i, j, k));
            A3(pp3d, i, j, k) = A3(pp3d, i, j, k) + delp;
            if (fabsf(A3(pp3d, i, j, k)) < smeps) A3(pp3d, i, j, k) = 0.0f;
            A3(dum1, i, j, k) = kdiv * delp;
            A3(ppd, i, j, k) = A3(pp3d, i, j, k) + A3(dum1, i, j, k);
          }
    }
    if (n < nloop) o_bcp(M, ppd);

    if (n < nloop) {
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj + 1; j++)
          for (i = 1; i <= ni + 1; i++) {
            A3(rru, i, j, k) = A3(rru, i, j, k) + A3(u3d, i, j, k);
            A3(rrv, i, j, k) = A3(rrv, i, j, k) + A3(v3d, i, j, k);
            A3(rrw, i, j, k) = A3(rrw, i, j, k) + A3(w3d, i, j, k);
          }
    } else {
      const float tavg = 1.0f / (float)nloop;
      for (k = 1; k <= nk; k++)
        for (j = 1; j <= nj + 1; j++)
          for (i = 1; i <= ni + 1; i++) {
            A3(rru, i, j, k) = (A3(rru, i, j, k) + A3(u3d, i, j, k)) * tavg;
            A3(rrv, i, j, k) = (A3(rrv, i, j, k) + A3(v3d, i, j, k)) * tavg;
            A3(rrw, i, j, k) = (A3(rrw, i, j, k) + A3(w3d, i, j, k)) * tavg;
          }
    }
  }

  if (nrk == 3)
    for (k = 1; k <= nk; k++)
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(ppx, i, j, k) = A3(dum1, i, j, k);
}

/* getdiv: src/misclibs.F:108-214 (Cartesian, no terrain) */
void o_getdiv(cm1o_model *M, const float *u, const float *v, const float *w, float *div) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *mh = M->kf[K_MH];
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        float d = ((A3(u, i + 1, j, k) - A3(u, i, j, k)) * M->rdx * UH(i)
                 + (A3(v, i, j + 1, k) - A3(v, i, j, k)) * M->rdy * VH(j))
                 + (A3(w, i, j, k + 1) - A3(w, i, j, k)) * M->rdz * mh[k];
        if (fabsf(d) < M->k.smeps) d = 0.0f;
        A3(div, i, j, k) = d;
      }
}

/* getdivx: src/misclibs.F:222-370 (Cartesian, no terrain) */
void o_getdivx(cm1o_model *M, float *rru, float *rrv, float *rrw, float *divx) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *rho0 = M->kf[K_RHO0], *rf0 = M->kf[K_RF0];
  int i, j, k;
  for (k = 1; k <= nk; k++) {
    float tem = rho0[k];
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni + 1; i++) A3(rru, i, j, k) = A3(rru, i, j, k) * tem;
    for (j = 1; j <= nj + 1; j++)
      for (i = 1; i <= ni; i++) A3(rrv, i, j, k) = A3(rrv, i, j, k) * tem;
    if (k == 1) {
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) {
          A3(rrw, i, j, 1) = 0.0f;
          A3(rrw, i, j, nk + 1) = 0.0f;
        }
    } else {
      tem = rf0[k];
      for (j = 1; j <= nj; j++)
        for (i = 1; i <= ni; i++) A3(rrw, i, j, k) = A3(rrw, i, j, k) * tem;
    }
  }
  o_getdiv(M, rru, rrv, rrw, divx);
}

/* calccflquick: src/misclibs.F:2726-2814 (nx>1 and ny>1 branch) */
void o_calccflquick(cm1o_model *M, float dt, const float *ua, const float *va, const float *wa) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float *mh = M->kf[K_MH];
  const float dtdx = 0.5f * dt * M->rdx, dtdy = 0.5f * dt * M->rdy, dtdz = 0.5f * dt * M->rdz;
  int i, j, k;
  float fmax = -99999999.0f;
  for (k = 1; k <= nk; k++) {
    float cfl = 0.0f;
    for (j = 1; j <= nj; j++)
      for (i = 1; i <= ni; i++) {
        float a = (A3(ua, i, j, k) + A3(ua, i + 1, j, k)) * dtdx * UH(i);
        float b = (A3(va, i, j, k) + A3(va, i, j + 1, k)) * dtdy * VH(j);
        float c = (A3(wa, i, j, k) + A3(wa, i, j, k + 1)) * dtdz * mh[k];
        float wsp = sqrtf((a * a + b * b) + c * c);
        cfl = fmaxf(cfl, wsp);
      }
    fmax = fmaxf(fmax, cfl);
  }
  M->cflmax = fmax;
}

/* pdefq (pdscheme=1): src/misclibs.F:1037-1117.  Column mass rescale in
 * double precision. */
void o_pdefq(cm1o_model *M, float rmax, const float *rho, float *q3d) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (j = 1; j <= nj; j++)
    for (i = 1; i <= ni; i++) {
      double t1 = 0.0, t2 = 0.0, t3;
      for (k = 1; k <= nk; k++) {
        t1 = t1 + (double)(A3(rho, i, j, k) * A3(q3d, i, j, k));
        if (A3(q3d, i, j, k) < rmax) A3(q3d, i, j, k) = 0.0f;
        t2 = t2 + (double)(A3(rho, i, j, k) * A3(q3d, i, j, k));
      }
      t3 = (t1 + 1.0e-20) / (t2 + 1.0e-20);
      if (t3 < 0.0) t3 = 1.0;
      for (k = 1; k <= nk; k++) A3(q3d, i, j, k) = (float)(t3 * (double)A3(q3d, i, j, k));
    }
}
