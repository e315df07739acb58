// Split-explicit acoustic substeps with the vertically implicit w-pi solve
// (psolver=3, no terrain, Cartesian): src/sound.F:10-859, called from
// src/solve2.F:1122-1136 with thv=t11, ppterm=t22, piadv=sadv, mm=dum5,
// ppd=dum7, fpk=dum8, qk=t12, pk1=t13, pk2=t23.
// Also getdiv/getdivx (src/misclibs.F:108-370) and calccflquick
// (src/misclibs.F:2726-2814).
#include "cm1g_internal.h"

static void sound_nloop(int nrk, int nsound, float dt, int *nloop_out, float *dts_out) {
  int nloop;
  float dts;
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
  *nloop_out = nloop;
  *dts_out = dts;
}

void sound_run(cm1g_model *m, int nrk) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt, alph = m->c.alph, kdiv = m->c.kdiv;
  int nloop;
  float dts;
  sound_nloop(nrk, m->c.nsound, dt, &nloop, &dts);
  const float beta = 1.0f - alph;
  const float asq = alph * alph;
  float *dum1 = m->f[CM1G_DUM1], *dum2 = m->f[CM1G_DUM2], *mm = m->f[CM1G_DUM5];
  float *ppd = m->f[CM1G_DUM7], *fpk = m->f[CM1G_DUM8];
  float *qk = m->f[CM1G_T12], *pk1 = m->f[CM1G_T13], *pk2 = m->f[CM1G_T23];
  const float *piadv = m->f[CM1G_SADV], *thv = m->f[CM1G_T11], *ppterm = m->f[CM1G_T22];
  float *u3d = m->f[CM1G_U3D], *v3d = m->f[CM1G_V3D], *w3d = m->f[CM1G_W3D];
  const float *ua = m->f[CM1G_UA], *va = m->f[CM1G_VA], *wa = m->f[CM1G_WA];
  const float *uten = m->f[CM1G_UTEN], *vten = m->f[CM1G_VTEN], *wten = m->f[CM1G_WTEN];
  float *rru = m->f[CM1G_RRU], *rrv = m->f[CM1G_RRV], *rrw = m->f[CM1G_RRW];
  const float *ppi = m->f[CM1G_PPI], *ppten = m->f[CM1G_PPTEN];
  float *pp3d = m->f[CM1G_PP3D], *ppx = m->f[CM1G_PPX];
  const float *radbcw = m->d2[CM1G_D_RADBCW], *radbce = m->d2[CM1G_D_RADBCE];
  const float *radbcs = m->d2[CM1G_D_RADBCS], *radbcn = m->d2[CM1G_D_RADBCN];
  const float *dtu0 = m->dtu0, *dtv0 = m->dtv0;
  const float rdx = P.rdx, rdy = P.rdy, rdz = P.rdz, cp = P.K.cp, smeps = P.K.smeps;

  // coefficients of the implicit solver (src/sound.F:117-180)
  for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    float tem = dts * rdz * P.mh[k];
    float r2 = dts * rdz * P.mh[k] * P.rr0[k] * P.rf0[k + 1];
    float r1 = dts * rdz * P.mh[k] * P.rr0[k] * P.rf0[k];
    float tem1 = dts * rdz * cp * P.mf[k];
    float c1a = P.c1[k], c1b = P.c1[k + 1], c2a = P.c2[k], c2b = P.c2[k + 1];
    int c = IXD(P, i, j, k);
    qk[c] = tem * ppterm[c];
    pk2[c] = (k == nk) ? 0.0f : r2 * (-c2b * piadv[IXD(P, i, j, k + 1)] + (1.0f - c1b) * piadv[c]);
    pk1[c] = (k == 1) ? 0.0f : r1 * (+c1a * piadv[IXD(P, i, j, k - 1)] + (c2a - 1.0f) * piadv[c]);
    mm[c] = tem1 * (c2a * thv[c] + c1a * thv[IXD(P, i, j, k - 1)]);
  });

  // prep copies (src/sound.F:191-250)
  if (nrk == 1) {
    for3(m->stream, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      ppd[c] = ppx[c];
      rru[c] = 0.0f;
      rrv[c] = 0.0f;
      rrw[c] = wa[c] * beta;
    });
  } else {
    for3(m->stream, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      u3d[c] = ua[c];
      rru[c] = 0.0f;
      v3d[c] = va[c];
      rrv[c] = 0.0f;
      w3d[c] = wa[c];
      rrw[c] = wa[c] * beta;
      pp3d[c] = ppi[c];
      ppd[c] = ppx[c];
    });
  }

  const bool opw = (P.wbc == 2 && P.ibw == 1), ope = (P.ebc == 2 && P.ibe == 1);
  const bool ops = (P.sbc == 2 && P.ibs == 1), opn = (P.nbc == 2 && P.ibn == 1);
  for (int n = 1; n <= nloop; n++) {
    // open-boundary normal velocities (ssopenbcw/e/s/n, src/bc.F:2392-2684)
    if (opw || ope)
      for2(m->stream, 1, nj, 1, nk, [=] __device__(int j, int k) {
        if (opw)
          u3d[IXD(P, 1, j, k)] = u3d[IXD(P, 1, j, k)] + dts * (-radbcw[(j + 2) + k * P.py] *
              (u3d[IXD(P, 2, j, k)] - u3d[IXD(P, 1, j, k)]) * rdx * P.uh[1 + 2] + uten[IXD(P, 1, j, k)]);
        if (ope)
          u3d[IXD(P, ni + 1, j, k)] = u3d[IXD(P, ni + 1, j, k)] + dts * (-radbce[(j + 2) + k * P.py] *
              (u3d[IXD(P, ni + 1, j, k)] - u3d[IXD(P, ni, j, k)]) * rdx * P.uh[ni + 2] + uten[IXD(P, ni + 1, j, k)]);
      });
    if (ops || opn)
      for2(m->stream, 1, ni, 1, nk, [=] __device__(int i, int k) {
        if (ops)
          v3d[IXD(P, i, 1, k)] = v3d[IXD(P, i, 1, k)] + dts * (-radbcs[(i + 2) + k * P.px] *
              (v3d[IXD(P, i, 2, k)] - v3d[IXD(P, i, 1, k)]) * rdy * P.vh[1 + 2] + vten[IXD(P, i, 1, k)]);
        if (opn)
          v3d[IXD(P, i, nj + 1, k)] = v3d[IXD(P, i, nj + 1, k)] + dts * (-radbcn[(i + 2) + k * P.px] *
              (v3d[IXD(P, i, nj + 1, k)] - v3d[IXD(P, i, nj, k)]) * rdy * P.vh[nj + 2] + vten[IXD(P, i, nj + 1, k)]);
      });
    // u, v forward (src/sound.F:307-345)
    {
      const float tem1 = rdx * cp * 0.5f, tem2 = rdy * cp * 0.5f;
      for3(m->stream, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
        int c = IXD(P, i, j, k);
        float dtu = dts * dtu0[IX2D(P, i, j)];
        float dtv = dts * dtv0[IX2D(P, i, j)];
        u3d[c] = u3d[c] + dtu * (uten[c] - tem1 * (ppd[c] - ppd[IXD(P, i - 1, j, k)]) * P.uf[i + 2] *
                                              (thv[c] + thv[IXD(P, i - 1, j, k)]));
        v3d[c] = v3d[c] + dtv * (vten[c] - tem2 * (ppd[c] - ppd[IXD(P, i, j - 1, k)]) * P.vf[j + 2] *
                                              (thv[c] + thv[IXD(P, i, j - 1, k)]));
      });
    }
    // explicit pi forcing (src/sound.F:484-505)
    {
      const float temx = dts * 0.5f * rdx, temy = dts * 0.5f * rdy;
      for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
        int c = IXD(P, i, j, k);
        float div = (u3d[IXD(P, i + 1, j, k)] - u3d[c]) * rdx * P.uh[i + 2] +
                    (v3d[IXD(P, i, j + 1, k)] - v3d[c]) * rdy * P.vh[j + 2];
        if (fabsf(div) < smeps) div = 0.0f;
        fpk[c] = dts * (ppten[c] - ppterm[c] * div) +
                 (-(u3d[c] * (piadv[c] - piadv[IXD(P, i - 1, j, k)]) +
                    u3d[IXD(P, i + 1, j, k)] * (piadv[IXD(P, i + 1, j, k)] - piadv[c])) * temx * P.uh[i + 2] -
                  (v3d[c] * (piadv[c] - piadv[IXD(P, i, j - 1, k)]) +
                   v3d[IXD(P, i, j + 1, k)] * (piadv[IXD(P, i, j + 1, k)] - piadv[c])) * temy * P.vh[j + 2]) +
                 beta * (w3d[c] * (pk1[c] + qk[c]) + w3d[IXD(P, i, j, k + 1)] * (pk2[c] - qk[c]));
      });
    }
    // vertically implicit solve for w (src/sound.F:659-733)
    for2(m->stream, 1, ni, 1, nj, [=] __device__(int i, int j) {
      float d1p = 0.0f, d2p = 0.0f;
      for (int k = 2; k <= nk; k++) {
        int c = IXD(P, i, j, k), cm = IXD(P, i, j, k - 1);
        float m_ = mm[c];
        float aa = -asq * (+m_ * (pk1[cm] + qk[cm]));
        float cc = -asq * (-m_ * (pk2[c] - qk[c]));
        float bb = 1.0f - asq * (+m_ * (-(pk1[c] + qk[c]) + (pk2[cm] - qk[cm])));
        float fwk = dts * wten[c] + beta * (-m_ * (pp3d[c] - pp3d[cm]));
        float dd = w3d[c] + fwk + alph * (-m_ * ((pp3d[c] + fpk[c]) - (pp3d[cm] + fpk[cm])));
        if (k == 2) {
          float r1 = 1.0f / bb;
          d1p = -cc * r1;
          d2p = dd * r1;
          dum1[c] = d1p;
          dum2[c] = d2p;
        } else if (k < nk) {
          float r1 = 1.0f / (aa * d1p + bb);
          d1p = -cc * r1;
          d2p = (dd - aa * d2p) * r1;
          dum1[c] = d1p;
          dum2[c] = d2p;
        } else {
          float r1 = 1.0f / (aa * d1p + bb);
          w3d[c] = (dd - aa * d2p) * r1;
        }
      }
      float wk1 = w3d[IXD(P, i, j, nk)];
      for (int k = nk - 1; k >= 2; k--) {
        int c = IXD(P, i, j, k);
        wk1 = dum1[c] * wk1 + dum2[c];
        w3d[c] = wk1;
      }
    });
    // new pi and divergence damping (src/sound.F:739-765)
    for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      float p_old = pp3d[c];
      float pn = p_old + fpk[c] + alph * (w3d[c] * (pk1[c] + qk[c]) + w3d[IXD(P, i, j, k + 1)] * (pk2[c] - qk[c]));
      if (fabsf(pn) < smeps) pn = 0.0f;
      pp3d[c] = pn;
      float d = kdiv * (pn - p_old);
      dum1[c] = d;
      ppd[c] = pn + d;
    });
    if (n < nloop) bc_fill(m, BC_P, ppd);
    if (n < nloop) {
      for3(m->stream, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
        int c = IXD(P, i, j, k);
        rru[c] = rru[c] + u3d[c];
        rrv[c] = rrv[c] + v3d[c];
        rrw[c] = rrw[c] + w3d[c];
      });
    } else {
      const float tavg = 1.0f / (float)nloop;
      for3(m->stream, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
        int c = IXD(P, i, j, k);
        rru[c] = (rru[c] + u3d[c]) * tavg;
        rrv[c] = (rrv[c] + v3d[c]) * tavg;
        rrw[c] = (rrw[c] + w3d[c] * alph) * tavg;
      });
    }
  }
  if (nrk == 3)
    for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      ppx[c] = dum1[c];
    });
}

// getdiv (t11 from the time-averaged velocities) and getdivx (mass fluxes
// scaled by rho0/rf0 into rrus/rrvs/rrws, and divx from them).
void getdiv_both(cm1g_model *m, int want_t11) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float *rru = m->f[CM1G_RRU], *rrv = m->f[CM1G_RRV], *rrw = m->f[CM1G_RRW];
  float *rrus = m->rrus, *rrvs = m->rrvs, *rrws = m->rrws;
  float *t11 = m->f[CM1G_T11], *divx = m->f[CM1G_DIVX];
  const float smeps = P.K.smeps;
  for3(m->stream, 0, ni + 2, 0, nj + 2, 1, nk + 1, [=] __device__(int i, int j, int k) {
    int c = IXD(P, i, j, k);
    if (k <= nk) {
      rrus[c] = rru[c] * P.rho0[k];
      rrvs[c] = rrv[c] * P.rho0[k];
    }
    rrws[c] = (k == 1 || k == nk + 1) ? 0.0f : rrw[c] * P.rf0[k];
  });
  for3(m->stream, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    int c = IXD(P, i, j, k);
    if (want_t11) {
      float d = ((rru[IXD(P, i + 1, j, k)] - rru[c]) * P.rdx * P.uh[i + 2] +
                 (rrv[IXD(P, i, j + 1, k)] - rrv[c]) * P.rdy * P.vh[j + 2]) +
                (rrw[IXD(P, i, j, k + 1)] - rrw[c]) * P.rdz * P.mh[k];
      if (fabsf(d) < smeps) d = 0.0f;
      t11[c] = d;
    }
    float ru1 = rru[IXD(P, i + 1, j, k)] * P.rho0[k], ru0 = rru[c] * P.rho0[k];
    float rv1 = rrv[IXD(P, i, j + 1, k)] * P.rho0[k], rv0 = rrv[c] * P.rho0[k];
    float rw1 = (k == nk) ? 0.0f : rrw[IXD(P, i, j, k + 1)] * P.rf0[k + 1];
    float rw0 = (k == 1) ? 0.0f : rrw[c] * P.rf0[k];
    float d = ((ru1 - ru0) * P.rdx * P.uh[i + 2] + (rv1 - rv0) * P.rdy * P.vh[j + 2]) +
              (rw1 - rw0) * P.rdz * P.mh[k];
    if (fabsf(d) < smeps) d = 0.0f;
    divx[c] = d;
  });
}

// calccflquick: per-level maximum of the Courant magnitude, then the max.
#define CFL_NB 32
__global__ void __launch_bounds__(256) k_cfl(KP P, float dt, const float *ua, const float *va,
                                             const float *wa, float *out) {
  __shared__ float red[256];
  const int k = 1 + (int)blockIdx.x, b = (int)blockIdx.y;
  const float dtdx = 0.5f * dt * P.rdx, dtdy = 0.5f * dt * P.rdy, dtdz = 0.5f * dt * P.rdz;
  float cfl = 0.0f;
  for (int j = 1 + b; j <= P.nj; j += CFL_NB)
    for (int i = 1 + (int)threadIdx.x; i <= P.ni; i += blockDim.x) {
      float a = (ua[IXD(P, i, j, k)] + ua[IXD(P, i + 1, j, k)]) * dtdx * P.uh[i + 2];
      float bb = (va[IXD(P, i, j, k)] + va[IXD(P, i, j + 1, k)]) * dtdy * P.vh[j + 2];
      float c = (wa[IXD(P, i, j, k)] + wa[IXD(P, i, j, k + 1)]) * dtdz * P.mh[k];
      float wsp = sqrtf((a * a + bb * bb) + c * c);
      cfl = fmaxf(cfl, wsp);
    }
  red[threadIdx.x] = cfl;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[1 + (k - 1) * CFL_NB + b] = red[0];
}

// maximum over the per-level partial maxima (exact in any order)
__global__ void k_cfl_final(int n, float *out) {
  float fmax = -99999999.0f;
  for (int t = 1; t <= n; t++) fmax = fmaxf(fmax, out[t]);
  out[0] = fmax;
}

void cfl_run(cm1g_model *m, float dt, const float *u, const float *v, const float *w) {
  hipLaunchKernelGGL(k_cfl, dim3(m->P.nk, CFL_NB), dim3(256), 0, m->stream, m->P, dt, u, v, w, m->fred);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(1), 0, m->stream, m->P.nk * CFL_NB, m->fred);
}
