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

// Arguments of one fused acoustic substep.
struct SubArgs {
  float dts, alph, beta, asq, kdiv, tavg;
  float tem1, tem2, temx, temy;   // cp*rdx/2, cp*rdy/2, dts*rdx/2, dts*rdy/2
  int last, opw, ope, ops, opn;
  const float *uX, *vX;           // u, v at the start of the substep
  float *uY, *vY;                 // u, v after the substep
  const float *ppdX;              // pi' + divergence damping from the previous substep
  float *ppdY;                    // ... written by this substep
  const float *uten, *vten, *wten, *thv, *piadv, *ppterm, *ppten, *dtu0, *dtv0;
  const float *radbcw, *radbce, *radbcs, *radbcn;
  float *w3d, *pp3d, *fpk, *dum1, *dum2, *rru, *rrv, *rrw;
  // first substep: w and pi' are read from wR/ppR (wa/ppi after the first
  // RK stage) and the averages start from 0 (u, v) and beta*wa (w), which
  // is what the reference's copies before the substep loop set
  // (src/sound.F:191-250)
  const float *wR, *ppR, *wa;
  // per level k: tem, r1, r2, temm, c1(k), c1(k+1), c2(k), c2(k+1) of the
  // solver coefficients (src/sound.F:117-180), 8 floats at coef + 8k
  const float *coef;
  int first;
  int nring;                      // trailing workgroups that do the averaging ring
};

// new u at (i,j,k): radiative open-boundary update (ssopenbcw/e,
// src/bc.F:2392-2530) then the pressure-gradient step (src/sound.F:307-325)
__device__ __forceinline__ float sub_u(const KP &P, const SubArgs &a, int i, int j, int k) {
  const int c = IXD(P, i, j, k);
  float u = a.uX[c];
  if (j >= 1 && j <= P.nj) {
    if (a.opw && i == 1)
      u = u + a.dts * (-a.radbcw[(j + 2) + k * P.py] * (a.uX[IXD(P, 2, j, k)] - u) * P.rdx * P.uh[1 + 2] +
                       a.uten[c]);
    if (a.ope && i == P.ni + 1)
      u = u + a.dts * (-a.radbce[(j + 2) + k * P.py] * (u - a.uX[IXD(P, P.ni, j, k)]) * P.rdx * P.uh[P.ni + 2] +
                       a.uten[c]);
  }
  const float dtu = a.dts * a.dtu0[IX2D(P, i, j)];
  return u + dtu * (a.uten[c] - a.tem1 * (a.ppdX[c] - a.ppdX[c - 1]) * P.uf[i + 2] * (a.thv[c] + a.thv[c - 1]));
}

// new v at (i,j,k) (ssopenbcs/n, src/bc.F:2538-2684; src/sound.F:327-345)
__device__ __forceinline__ float sub_v(const KP &P, const SubArgs &a, int i, int j, int k) {
  const int c = IXD(P, i, j, k);
  float v = a.vX[c];
  if (i >= 1 && i <= P.ni) {
    if (a.ops && j == 1)
      v = v + a.dts * (-a.radbcs[(i + 2) + k * P.px] * (a.vX[IXD(P, i, 2, k)] - v) * P.rdy * P.vh[1 + 2] +
                       a.vten[c]);
    if (a.opn && j == P.nj + 1)
      v = v + a.dts * (-a.radbcn[(i + 2) + k * P.px] * (v - a.vX[IXD(P, i, P.nj, k)]) * P.rdy * P.vh[P.nj + 2] +
                       a.vten[c]);
  }
  const float dtv = a.dts * a.dtv0[IX2D(P, i, j)];
  const int cs = c - P.px;
  return v + dtv * (a.vten[c] - a.tem2 * (a.ppdX[c] - a.ppdX[cs]) * P.vf[j + 2] * (a.thv[c] + a.thv[cs]));
}

// store a new u (v) and fold it into the substep average (src/sound.F:806-835)
__device__ __forceinline__ void put_u(const SubArgs &a, int c, float u) {
  a.uY[c] = u;
  const float r = a.first ? 0.0f : a.rru[c];
  a.rru[c] = a.last ? (r + u) * a.tavg : r + u;
}
__device__ __forceinline__ void put_v(const SubArgs &a, int c, float v) {
  a.vY[c] = v;
  const float r = a.first ? 0.0f : a.rrv[c];
  a.rrv[c] = a.last ? (r + v) * a.tavg : r + v;
}
__device__ __forceinline__ void put_w_avg(const SubArgs &a, int c, float w) {
  const float r = a.first ? a.wa[c] * a.beta : a.rrw[c];
  a.rrw[c] = a.last ? (r + w * a.alph) * a.tavg : r + w;
}

// The averaging ring of one substep: u and v on the faces i = ni+1 and
// j = nj+1 (unchanged by the step where the open-boundary masks are zero,
// advanced otherwise) and the w average there (w is not advanced at those
// points).  Point t walks the ring (i = 1..ni+1 at j = nj+1, then
// j = 1..nj at i = ni+1) at level k.  It reads uX/vX/ppdX and writes only
// ring points, which the column tiles neither read nor write, so the ring
// runs alongside them.
__device__ __forceinline__ void sound_ring_point(const KP &P, const SubArgs &a, int t, int k) {
  const int ni = P.ni, nj = P.nj;
  int i, j;
  if (t <= ni) {
    i = 1 + t;
    j = nj + 1;
  } else if (t <= ni + nj) {
    i = ni + 1;
    j = t - ni;
  } else {
    return;
  }
  const int c = IXD(P, i, j, k);
  put_u(a, c, sub_u(P, a, i, j, k));
  put_v(a, c, sub_v(P, a, i, j, k));
  put_w_avg(a, c, a.wR[c]);
}

// Operands of one level of the upward sweep of sound_sub (raw loads).
struct SubLoads {
  float uXc, uXe, utc, ute, vXc, vXn, vtc, vtn;
  float pdc, pdw, pde, pds, pdn;
  float thc, thw, the, ths, thn;
  float rruc, rrvc;
  float paw, pae, pas, pan, pa_p;
  float ppterm, ppten, wten, pp_c, w_p;
  float rw, re, rs, rn;
};

__device__ __forceinline__ void sub_load(const KP &P, const SubArgs &a, int c, int k, int i, int j, bool bw,
                                         bool be, bool bs, bool bn, SubLoads &L) {
  const int px = P.px, pxy = P.pxy;
  L.uXc = a.uX[c]; L.uXe = a.uX[c + 1]; L.utc = a.uten[c]; L.ute = a.uten[c + 1];
  L.vXc = a.vX[c]; L.vXn = a.vX[c + px]; L.vtc = a.vten[c]; L.vtn = a.vten[c + px];
  L.pdc = a.ppdX[c]; L.pdw = a.ppdX[c - 1]; L.pde = a.ppdX[c + 1]; L.pds = a.ppdX[c - px];
  L.pdn = a.ppdX[c + px];
  L.thc = a.thv[c]; L.thw = a.thv[c - 1]; L.the = a.thv[c + 1]; L.ths = a.thv[c - px]; L.thn = a.thv[c + px];
  L.rruc = a.first ? 0.0f : a.rru[c];
  L.rrvc = a.first ? 0.0f : a.rrv[c];
  L.paw = a.piadv[c - 1]; L.pae = a.piadv[c + 1]; L.pas = a.piadv[c - px]; L.pan = a.piadv[c + px];
  L.pa_p = a.piadv[c + pxy];
  L.ppterm = a.ppterm[c]; L.ppten = a.ppten[c]; L.wten = a.wten[c]; L.pp_c = a.ppR[c];
  L.w_p = (k == P.nk) ? a.w3d[c + pxy] : a.wR[c + pxy];
  L.rw = bw ? a.radbcw[(j + 2) + k * P.py] : 0.0f;
  L.re = be ? a.radbce[(j + 2) + k * P.py] : 0.0f;
  L.rs = bs ? a.radbcs[(i + 2) + k * px] : 0.0f;
  L.rn = bn ? a.radbcn[(i + 2) + k * px] : 0.0f;
}

// One acoustic substep for a column (i,j): u/v forward step (also the
// faces i+1 / j+1 the divergence needs, recomputed from the same inputs),
// explicit pi forcing fpk (src/sound.F:484-505), Thomas forward sweep of the
// implicit w-pi system (src/sound.F:659-717), back substitution (:720-733),
// pi update with divergence damping (:739-765) and the w average.  The
// solver coefficients qk, pk1, pk2, mm (src/sound.F:117-180) are evaluated
// on the fly with the same expressions.  u and v are read from uX/vX and
// written to uY/vY (pi'+damping likewise ppdX/ppdY) so that neighbouring
// columns see the old values.
//
// Both sweeps load the next level's operands before the current level is
// evaluated and stored, so two levels' loads are in flight per column.
//
// The a.nring workgroups after the column tiles do the averaging ring (one
// thread per ring point and level); they fill the CUs the last tiles leave
// idle.
__device__ __forceinline__ void sound_sub(const KP &P, const SubArgs &a, const ColGrid &cg, int bx) {
  const int nmain = cg.per * NXCD;
  if (bx >= nmain) {
    const int npts = P.ni + P.nj + 1;
    const int t = (bx - nmain) * (int)blockDim.x + (int)threadIdx.x;
    if (t < npts * P.nk) sound_ring_point(P, a, t % npts, 1 + t / npts);
    return;
  }
  int i, j;
  if (!col_tile_b(cg, bx, i, j)) return;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (i > ni || j > nj) return;
  const float dts = a.dts, rdz = P.rdz, cp = P.K.cp, smeps = P.K.smeps;
  const float rdx = P.rdx, rdy = P.rdy;
  const int px = P.px, pxy = P.pxy;
  const bool bw = a.opw && i == 1, be = a.ope && i == ni, bs = a.ops && j == 1, bn = a.opn && j == nj;
  // k-invariant factors
  const float dtu_c = dts * a.dtu0[IX2D(P, i, j)], dtu_e = dts * a.dtu0[IX2D(P, i + 1, j)];
  const float dtv_c = dts * a.dtv0[IX2D(P, i, j)], dtv_n = dts * a.dtv0[IX2D(P, i, j + 1)];
  const float uf_c = P.uf[i + 2], uf_e = P.uf[i + 3], vf_c = P.vf[j + 2], vf_n = P.vf[j + 3];
  const float uh_c = P.uh[i + 2], vh_c = P.vh[j + 2];
  const float tem1 = a.tem1, tem2 = a.tem2, temx = a.temx, temy = a.temy;
  const float alph = a.alph, beta = a.beta, asq = a.asq;

  // level k-1 quantities carried up the column
  float A_m = 0.0f, B_m = 0.0f, fpk_m = 0.0f, pp_m = 0.0f;
  float d1p = 0.0f, d2p = 0.0f, w_top = 0.0f, w_bot = 0.0f, w_nk1 = 0.0f;
  int c = IXD(P, i, j, 1);
  float pa_m = a.piadv[c - pxy], pa_c = a.piadv[c];
  float th_m = a.thv[c - pxy];
  float w_c = a.wR[c];
  SubLoads nx;
  sub_load(P, a, c, 1, i, j, bw, be, bs, bn, nx);
  for (int k = 1; k <= nk; k++, c += pxy) {
    // ---- this level's operands; the next level's are loaded before this
    // level is evaluated and stored
    const SubLoads L = nx;
    if (k < nk) sub_load(P, a, c + pxy, k + 1, i, j, bw, be, bs, bn, nx);
    const float uXc = L.uXc, uXe = L.uXe, utc = L.utc, ute = L.ute;
    const float vXc = L.vXc, vXn = L.vXn, vtc = L.vtc, vtn = L.vtn;
    const float pdc = L.pdc, pdw = L.pdw, pde = L.pde, pds = L.pds, pdn = L.pdn;
    const float thc = L.thc, thw = L.thw, the = L.the, ths = L.ths, thn = L.thn;
    const float rruc = L.rruc, rrvc = L.rrvc;
    const float paw = L.paw, pae = L.pae, pas = L.pas, pan = L.pan;
    const float pa_p = L.pa_p;
    const float ppterm = L.ppterm, ppten = L.ppten, wten = L.wten, pp_c = L.pp_c;
    const float w_p = L.w_p;
    const float rw = L.rw, re = L.re, rs = L.rs, rn = L.rn;
    // ---- u, v (ssopenbc* then src/sound.F:307-345)
    float uo0 = uXc, uo1 = uXe, vo0 = vXc, vo1 = vXn;
    if (bw) uo0 = uo0 + dts * (-rw * (uXe - uo0) * rdx * P.uh[1 + 2] + utc);
    if (be) uo1 = uo1 + dts * (-re * (uo1 - uXc) * rdx * P.uh[ni + 2] + ute);
    if (bs) vo0 = vo0 + dts * (-rs * (vXn - vo0) * rdy * P.vh[1 + 2] + vtc);
    if (bn) vo1 = vo1 + dts * (-rn * (vo1 - vXc) * rdy * P.vh[nj + 2] + vtn);
    const float un0 = uo0 + dtu_c * (utc - tem1 * (pdc - pdw) * uf_c * (thc + thw));
    const float un1 = uo1 + dtu_e * (ute - tem1 * (pde - pdc) * uf_e * (the + thc));
    const float vn0 = vo0 + dtv_c * (vtc - tem2 * (pdc - pds) * vf_c * (thc + ths));
    const float vn1 = vo1 + dtv_n * (vtn - tem2 * (pdn - pdc) * vf_n * (thn + thc));
    // ---- coefficients (src/sound.F:117-180)
    const float *cf = a.coef + 8 * k;
    const float tem = cf[0], r1 = cf[1], r2 = cf[2], temm = cf[3];
    const float c1a = cf[4], c1b = cf[5], c2a = cf[6], c2b = cf[7];
    const float qk = tem * ppterm;
    const float pk2 = (k == nk) ? 0.0f : r2 * (-c2b * pa_p + (1.0f - c1b) * pa_c);
    const float pk1 = (k == 1) ? 0.0f : r1 * (+c1a * pa_m + (c2a - 1.0f) * pa_c);
    const float mm = temm * (c2a * thc + c1a * th_m);
    const float A = pk1 + qk, B = pk2 - qk;
    // ---- explicit pi forcing
    float div = (un1 - un0) * rdx * uh_c + (vn1 - vn0) * rdy * vh_c;
    if (fabsf(div) < smeps) div = 0.0f;
    const float fpk = dts * (ppten - ppterm * div) +
                      (-(un0 * (pa_c - paw) + un1 * (pae - pa_c)) * temx * uh_c -
                       (vn0 * (pa_c - pas) + vn1 * (pan - pa_c)) * temy * vh_c) +
                      beta * (w_c * (pk1 + qk) + w_p * (pk2 - qk));
    // ---- Thomas forward sweep
    if (k == 1) {
      w_bot = w_c;
    } else {
      const float aa = -asq * (+mm * A_m);
      const float cc = -asq * (-mm * B);
      const float bb = 1.0f - asq * (+mm * (-A + B_m));
      const float fwk = dts * wten + beta * (-mm * (pp_c - pp_m));
      const float dd = w_c + fwk + alph * (-mm * ((pp_c + fpk) - (pp_m + fpk_m)));
      if (k == 2) {
        const float rr = 1.0f / bb;
        d1p = -cc * rr;
        d2p = dd * rr;
      } else {
        const float rr = 1.0f / (aa * d1p + bb);
        d1p = -cc * rr;
        d2p = (dd - aa * d2p) * rr;
      }
      if (k == nk) w_top = d2p;
    }
    // ---- stores
    a.uY[c] = un0;
    a.rru[c] = a.last ? (rruc + un0) * a.tavg : rruc + un0;
    a.vY[c] = vn0;
    a.rrv[c] = a.last ? (rrvc + vn0) * a.tavg : rrvc + vn0;
    a.fpk[c] = fpk;
    // d1 depends only on the stage's fixed coefficients (thv, piadv,
    // ppterm, dts): stored on the first substep, read on the others
    if (k >= 2 && k < nk) {
      if (a.first) a.dum1[c] = d1p;
      a.dum2[c] = d2p;
    }
    if (k == nk) w_nk1 = w_p;
    A_m = A;
    B_m = B;
    fpk_m = fpk;
    pp_m = pp_c;
    pa_m = pa_c;
    pa_c = pa_p;
    th_m = thc;
    w_c = w_p;
  }

  // back substitution, pi update, w average (top down)
  c = IXD(P, i, j, nk);
  float w_above = w_nk1;
  float pa_p = a.piadv[c + pxy];
  float pa_c2 = a.piadv[c];
  // operands of the current level
  float q_d1 = 0.0f, q_d2 = 0.0f;
  float q_pp = a.ppR[c], q_fpk = a.fpk[c], q_pm = a.piadv[c - pxy], q_term = a.ppterm[c];
  float q_rrw = a.first ? a.wa[c] * a.beta : a.rrw[c];
  for (int k = nk; k >= 1; k--, c -= pxy) {
    const float p_old = q_pp, fk = q_fpk, pa_m2 = q_pm, pterm = q_term, rrwc = q_rrw;
    float wk;
    if (k == nk) {
      wk = w_top;
    } else if (k >= 2) {
      wk = q_d1 * w_above + q_d2;
    } else {
      wk = w_bot;
    }
    // prefetch the level below
    if (k > 1) {
      const int cb = c - pxy;
      if (k - 1 >= 2) {
        q_d1 = a.dum1[cb];
        q_d2 = a.dum2[cb];
      }
      q_pp = a.ppR[cb];
      q_fpk = a.fpk[cb];
      q_pm = a.piadv[cb - pxy];
      q_term = a.ppterm[cb];
      q_rrw = a.first ? a.wa[cb] * a.beta : a.rrw[cb];
    }
    const float *cf = a.coef + 8 * k;
    const float tem = cf[0], r1 = cf[1], r2 = cf[2];
    const float c1a = cf[4], c1b = cf[5], c2a = cf[6], c2b = cf[7];
    const float qk = tem * pterm;
    const float pk2 = (k == nk) ? 0.0f : r2 * (-c2b * pa_p + (1.0f - c1b) * pa_c2);
    const float pk1 = (k == 1) ? 0.0f : r1 * (+c1a * pa_m2 + (c2a - 1.0f) * pa_c2);
    float pn = p_old + fk + alph * (wk * (pk1 + qk) + w_above * (pk2 - qk));
    if (fabsf(pn) < smeps) pn = 0.0f;
    const float d = a.kdiv * (pn - p_old);
    if (k >= 2 || a.wR != a.w3d) a.w3d[c] = wk;
    a.pp3d[c] = pn;
    if (a.last) a.dum1[c] = d;
    a.ppdY[c] = pn + d;
    a.rrw[c] = a.last ? (rrwc + wk * alph) * a.tavg : rrwc + wk;
    w_above = wk;
    pa_p = pa_c2;
    pa_c2 = pa_m2;
  }
}

void sound_run(cm1g_model *m, int nrk, bool fold1) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt, alph = m->c.alph, kdiv = m->c.kdiv;
  int nloop;
  float dts;
  sound_nloop(nrk, m->c.nsound, dt, &nloop, &dts);
  const float beta = 1.0f - alph;
  float *dum1 = m->f[CM1G_DUM1];
  float *ppd = m->f[CM1G_DUM7];
  float *u3d = m->f[CM1G_U3D], *v3d = m->f[CM1G_V3D], *w3d = m->f[CM1G_W3D];
  const float *ua = m->f[CM1G_UA], *va = m->f[CM1G_VA], *wa = m->f[CM1G_WA];
  float *rru = m->f[CM1G_RRU], *rrv = m->f[CM1G_RRV], *rrw = m->f[CM1G_RRW];
  const float *ppi = m->f[CM1G_PPI];
  float *pp3d = m->f[CM1G_PP3D], *ppx = m->f[CM1G_PPX];

  SubArgs a;
  a.dts = dts;
  a.alph = alph;
  a.beta = beta;
  a.asq = alph * alph;
  a.kdiv = kdiv;
  a.tavg = 1.0f / (float)nloop;
  a.tem1 = P.rdx * P.K.cp * 0.5f;
  a.tem2 = P.rdy * P.K.cp * 0.5f;
  a.temx = dts * 0.5f * P.rdx;
  a.temy = dts * 0.5f * P.rdy;
  a.opw = (P.wbc == 2 && P.ibw == 1);
  a.ope = (P.ebc == 2 && P.ibe == 1);
  a.ops = (P.sbc == 2 && P.ibs == 1);
  a.opn = (P.nbc == 2 && P.ibn == 1);
  a.uten = m->f[CM1G_UTEN];
  a.vten = m->f[CM1G_VTEN];
  a.wten = m->f[CM1G_WTEN];
  a.thv = m->f[CM1G_T11];
  a.piadv = m->f[CM1G_SADV];
  a.ppterm = m->f[CM1G_T22];
  a.ppten = m->f[CM1G_PPTEN];
  a.dtu0 = m->dtu0;
  a.dtv0 = m->dtv0;
  a.radbcw = m->d2[CM1G_D_RADBCW];
  a.radbce = m->d2[CM1G_D_RADBCE];
  a.radbcs = m->d2[CM1G_D_RADBCS];
  a.radbcn = m->d2[CM1G_D_RADBCN];
  a.w3d = w3d;
  a.pp3d = pp3d;
  a.fpk = m->f[CM1G_DUM8];
  a.dum1 = dum1;
  a.dum2 = m->f[CM1G_DUM2];
  a.rru = rru;
  a.rrv = rrv;
  a.rrw = rrw;
  // The copies the reference makes before the substeps (u3d=ua, ..., ppd=ppx,
  // src/sound.F:191-250) are folded into the first substep, which reads ua,
  // va, wa, ppi, ppx directly (after the first RK stage) and starts the
  // averages itself.  u, v and ppd alternate between the state arrays and
  // scratch arrays so that neighbouring columns read the old values; the
  // last substep writes u3d/v3d.
  // (at the first stage u3d..pp3d equal ua..ppi where the substep reads
  // them when the caller says so, fold1: the folded form then needs no
  // copy of the scratch pair back into u3d/v3d)
  const bool fold = (nrk >= 2) || fold1;
  a.wa = wa;
  const ColGrid cg = col_grid(ni, nj, m->tile_rows);
  const dim3 blk = col_threads(cg), grd = col_blocks(cg);
  {
    const int nthr = (int)blk.x, npts = (ni + nj + 1) * nk;
    a.nring = (npts + nthr - 1) / nthr;
  }
  // the per-level solver coefficients of this stage's dts, with the
  // reference's expressions (src/sound.F:117-180)
  {
    float *cf = m->scoef;
    const float rdz = P.rdz, cp = P.K.cp;
    launch<64>(m, dim3((nk + 63) / 64), dim3(64), [=] __device__(int bx, int, int) {
      const int k = 1 + bx * 64 + (int)threadIdx.x;
      if (k > nk) return;
      float *o = cf + 8 * k;
      o[0] = dts * rdz * P.mh[k];
      o[1] = dts * rdz * P.mh[k] * P.rr0[k] * P.rf0[k];
      o[2] = dts * rdz * P.mh[k] * P.rr0[k] * P.rf0[k + 1];
      o[3] = dts * rdz * cp * P.mf[k];
      o[4] = P.c1[k];
      o[5] = P.c1[k + 1];
      o[6] = P.c2[k];
      o[7] = P.c2[k + 1];
    });
    a.coef = cf;
  }
  float *ppdb[2] = {ppd, m->tx};
  float *uprev = nullptr, *vprev = nullptr;
  for (int n = 1; n <= nloop; n++) {
    // destination of this substep: u3d on the last one, alternating before
    float *udst, *vdst;
    if (fold) {
      const bool scratch = ((nloop - n) & 1) != 0;
      udst = scratch ? m->scr1 : u3d;
      vdst = scratch ? m->scr2 : v3d;
    } else {
      const bool scratch = (n & 1) != 0;
      udst = scratch ? m->scr1 : u3d;
      vdst = scratch ? m->scr2 : v3d;
    }
    a.uX = (n == 1) ? (fold ? ua : u3d) : uprev;
    a.vX = (n == 1) ? (fold ? va : v3d) : vprev;
    a.uY = udst;
    a.vY = vdst;
    a.ppdX = (n == 1) ? ppx : ppdb[(n - 2) & 1];
    a.ppdY = ppdb[(n - 1) & 1];
    a.wR = (n == 1 && fold) ? wa : w3d;
    a.ppR = (n == 1 && fold) ? ppi : pp3d;
    a.first = (n == 1);
    a.last = (n == nloop);
    ktimer_mark(m, CM1G_KT_SOUND_SUBSTEP);
    launch<256>(m, dim3(grd.x + a.nring), blk,
                [=] __device__(int bx, int, int) { sound_sub(P, a, cg, bx); });
    ktimer_mark(m, CM1G_KT_SOUND_SUBSTEP);
    if (n < nloop) bc_fill(m, BC_P, a.ppdY);
    uprev = udst;
    vprev = vdst;
  }
  if (!fold && (nloop & 1)) {
    // the last substep wrote the scratch pair: copy the updated range back
    const float *us = m->scr1, *vs = m->scr2;
    for3(m, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      u3d[c] = us[c];
      v3d[c] = vs[c];
    });
  }
  if (nrk == 3)
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      ppx[c] = dum1[c];
    });
}

// Explicit acoustic substeps, psolver=2 (sounde, src/sounde.F:10-1026;
// Cartesian, no terrain).  Per substep: the radiative open-boundary
// velocities, then one kernel that advances u, v, w from the previous
// substep's ppd (src/sounde.F:333-352, 548-563) and folds them into the
// averages, then one kernel for the pressure (src/sounde.F:770-795) whose
// pi0 vertical-advection coefficients pk1/pk2 (src/sounde.F:117-136,
// 160-170) are evaluated on the fly.
void sounde_run(cm1g_model *m, int nrk) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float dt = m->c.dt, kdiv = m->c.kdiv;
  int nloop;
  float dts;
  sound_nloop(nrk, m->c.nsound, dt, &nloop, &dts);
  float *dum1 = m->f[CM1G_DUM1], *ppd = m->f[CM1G_DUM7];
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

  // prep copies (src/sounde.F:181-233); the first stage skips u3d = ua ...
  // because they are equal, except after an exchanging solve3, when the
  // 3-D arrays still hold the previous step
  if (nrk == 1 && !m->swapped) {
    for3(m, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      ppd[c] = ppx[c];
      rru[c] = 0.0f;
      rrv[c] = 0.0f;
      rrw[c] = 0.0f;
    });
  } else {
    for3(m, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      u3d[c] = ua[c];
      rru[c] = 0.0f;
      v3d[c] = va[c];
      rrv[c] = 0.0f;
      w3d[c] = wa[c];
      rrw[c] = 0.0f;
      pp3d[c] = ppi[c];
      ppd[c] = ppx[c];
    });
  }

  const bool opw = (P.wbc == 2 && P.ibw == 1), ope = (P.ebc == 2 && P.ibe == 1);
  const bool ops = (P.sbc == 2 && P.ibs == 1), opn = (P.nbc == 2 && P.ibn == 1);
  const float tavg = 1.0f / (float)nloop;
  for (int n = 1; n <= nloop; n++) {
    const bool last = (n == nloop);
    if (opw || ope)
      for2(m, 1, nj, 1, nk, [=] __device__(int j, int k) {
        if (opw)
          u3d[IXD(P, 1, j, k)] = u3d[IXD(P, 1, j, k)] + dts * (-radbcw[(j + 2) + k * P.py] *
              (u3d[IXD(P, 2, j, k)] - u3d[IXD(P, 1, j, k)]) * rdx * P.uh[1 + 2] + uten[IXD(P, 1, j, k)]);
        if (ope)
          u3d[IXD(P, ni + 1, j, k)] = u3d[IXD(P, ni + 1, j, k)] + dts * (-radbce[(j + 2) + k * P.py] *
              (u3d[IXD(P, ni + 1, j, k)] - u3d[IXD(P, ni, j, k)]) * rdx * P.uh[ni + 2] + uten[IXD(P, ni + 1, j, k)]);
      });
    if (ops || opn)
      for2(m, 1, ni, 1, nk, [=] __device__(int i, int k) {
        if (ops)
          v3d[IXD(P, i, 1, k)] = v3d[IXD(P, i, 1, k)] + dts * (-radbcs[(i + 2) + k * P.px] *
              (v3d[IXD(P, i, 2, k)] - v3d[IXD(P, i, 1, k)]) * rdy * P.vh[1 + 2] + vten[IXD(P, i, 1, k)]);
        if (opn)
          v3d[IXD(P, i, nj + 1, k)] = v3d[IXD(P, i, nj + 1, k)] + dts * (-radbcn[(i + 2) + k * P.px] *
              (v3d[IXD(P, i, nj + 1, k)] - v3d[IXD(P, i, nj, k)]) * rdy * P.vh[nj + 2] + vten[IXD(P, i, nj + 1, k)]);
      });
    // u, v, w forward and the velocity averages
    {
      const float tem1 = rdx * cp * 0.5f, tem2 = rdy * cp * 0.5f;
      for3(m, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
        const int c = IXD(P, i, j, k);
        const float dtu = dts * dtu0[IX2D(P, i, j)], dtv = dts * dtv0[IX2D(P, i, j)];
        const float u = u3d[c] + dtu * (uten[c] - tem1 * P.uf[i + 2] * (ppd[c] - ppd[c - 1]) *
                                                      (thv[c] + thv[c - 1]));
        const float v = v3d[c] + dtv * (vten[c] - tem2 * P.vf[j + 2] * (ppd[c] - ppd[c - P.px]) *
                                                      (thv[c] + thv[c - P.px]));
        float w = w3d[c];
        if (k >= 2 && i <= ni && j <= nj) {
          const float temw = rdz * cp * P.mf[k];
          w = w + dts * (wten[c] - temw * (ppd[c] - ppd[c - P.pxy]) *
                                       (P.c2[k] * thv[c] + P.c1[k] * thv[c - P.pxy]));
          w3d[c] = w;
        }
        u3d[c] = u;
        v3d[c] = v;
        if (last) {
          rru[c] = (rru[c] + u) * tavg;
          rrv[c] = (rrv[c] + v) * tavg;
          rrw[c] = (rrw[c] + w) * tavg;
        } else {
          rru[c] = rru[c] + u;
          rrv[c] = rrv[c] + v;
          rrw[c] = rrw[c] + w;
        }
      });
    }
    // pressure
    {
      const float temx = dts * 0.5f * rdx, temy = dts * 0.5f * rdy;
      for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
        const int c = IXD(P, i, j, k);
        const float mh = P.mh[k], rr0 = P.rr0[k];
        float pk1 = 0.0f, pk2 = 0.0f;
        if (k < nk) {
          const float p1 = dts * rdz * mh * rr0 * P.rf0[k + 1] * (1.0f - P.c1[k + 1]);
          const float p2 = -(dts * rdz * mh * rr0 * P.rf0[k + 1] * P.c2[k + 1]);
          pk2 = p2 * piadv[c + P.pxy] + p1 * piadv[c];
        }
        if (k > 1) {
          const float p3 = dts * rdz * mh * rr0 * P.rf0[k] * P.c1[k];
          const float p4 = dts * rdz * mh * rr0 * P.rf0[k] * (P.c2[k] - 1.0f);
          pk1 = p3 * piadv[c - P.pxy] + p4 * piadv[c];
        }
        const float u0 = u3d[c], u1 = u3d[c + 1], v0 = v3d[c], v1 = v3d[c + P.px];
        const float w0 = w3d[c], w1 = w3d[c + P.pxy];
        const float pa = piadv[c];
        float div = ((u1 - u0) * rdx * P.uh[i + 2] + (v1 - v0) * rdy * P.vh[j + 2]) + (w1 - w0) * rdz * mh;
        if (fabsf(div) < smeps) div = 0.0f;
        const float delp = dts * (ppten[c] - ppterm[c] * div) +
                           (-(u0 * (pa - piadv[c - 1]) + u1 * (piadv[c + 1] - pa)) * temx * P.uh[i + 2] -
                            (v0 * (pa - piadv[c - P.px]) + v1 * (piadv[c + P.px] - pa)) * temy * P.vh[j + 2]) +
                           (w0 * pk1 + w1 * pk2);
        float p = pp3d[c] + delp;
        if (fabsf(p) < smeps) p = 0.0f;
        pp3d[c] = p;
        const float d = kdiv * delp;
        if (last) dum1[c] = d;
        ppd[c] = p + d;
      });
    }
    if (n < nloop) bc_fill(m, BC_P, ppd);
  }
  if (nrk == 3)
    for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
      int c = IXD(P, i, j, k);
      ppx[c] = dum1[c];
    });
}

// getdiv (t11 from the time-averaged velocities) and getdivx (divx from the
// mass fluxes scaled by rho0/rf0; inside the time step the advection
// kernels apply the same scaling on the fly instead of storing scaled
// copies).  want_divx = 0 computes only the divergence into `div`.
void getdiv_both(cm1g_model *m, int want_t11, int want_divx, float *div) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const float *rru = m->f[CM1G_RRU], *rrv = m->f[CM1G_RRV], *rrw = m->f[CM1G_RRW];
  float *t11 = div ? div : m->f[CM1G_T11], *divx = m->f[CM1G_DIVX];
  const float smeps = P.K.smeps;
  for3(m, 1, ni, 1, nj, 1, nk, [=] __device__(int i, int j, int k) {
    int c = IXD(P, i, j, k);
    if (want_t11) {
      float d = ((rru[IXD(P, i + 1, j, k)] - rru[c]) * P.rdx * P.uh[i + 2] +
                 (rrv[IXD(P, i, j + 1, k)] - rrv[c]) * P.rdy * P.vh[j + 2]) +
                (rrw[IXD(P, i, j, k + 1)] - rrw[c]) * P.rdz * P.mh[k];
      if (fabsf(d) < smeps) d = 0.0f;
      t11[c] = d;
    }
    if (!want_divx) return;
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

// The in-place scaling of getdivx (src/misclibs.F:246-273): rru, rrv by
// rho0, rrw by rf0 with rrw = 0 at k = 1 and nk+1, over the reference's
// ranges.  Only the routine-level entry stores the scaled fluxes.
void scale_mass_fluxes(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  float *rru = m->f[CM1G_RRU], *rrv = m->f[CM1G_RRV], *rrw = m->f[CM1G_RRW];
  for3(m, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k);
    if (j <= nj) rru[c] = rru[c] * P.rho0[k];
    if (i <= ni) rrv[c] = rrv[c] * P.rho0[k];
    if (i <= ni && j <= nj) {
      if (k == 1) {
        rrw[c] = 0.0f;
        rrw[IXD(P, i, j, nk + 1)] = 0.0f;
      } else {
        rrw[c] = rrw[c] * P.rf0[k];
      }
    }
  });
}

// calccflquick: per-level maximum of the Courant magnitude, then the max.
#define CFL_NB 32
__device__ __forceinline__ void cfl_levels(const KP &P, float dt, const float *ua, const float *va,
                                           const float *wa, float *out, int bx, int by) {
  __shared__ float red[256];
  const int k = 1 + bx, b = by;
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
  cfl = block_reduce256(cfl, red, OpMaxF());
  if (threadIdx.x == 0) out[1 + (k - 1) * CFL_NB + b] = cfl;
}

// maximum over the per-level partial maxima (exact in any order)
__device__ __forceinline__ void cfl_final(int n, float *out) {
  float fmax = -99999999.0f;
  for (int t = 1; t <= n; t++) fmax = fmaxf(fmax, out[t]);
  out[0] = fmax;
}

void cfl_run(cm1g_model *m, float dt, const float *u, const float *v, const float *w) {
  const KP P = m->P;
  float *out = m->fred;
  launch<256>(m, dim3(P.nk, CFL_NB), dim3(256),
              [=] __device__(int bx, int by, int) { cfl_levels(P, dt, u, v, w, out, bx, by); });
  const int n = P.nk * CFL_NB;
  launch<64>(m, dim3(1), dim3(1), [=] __device__(int, int, int) { cfl_final(n, out); });
}
