// Momentum advection as column kernels: advu (src/adv.F:911-1839), advv
// (:1847-2770) and advw (:2778-3610), each one pass over the columns of its
// staggered grid.  For every level a thread evaluates the two horizontal
// face fluxes on each side of its point (hadv_flx5 or hadv_weno5,
// src/adv_routines.F:1335-1476 / :204-430), the vertical face flux from a
// register window (vadv_flx5 / vadv_weno5 with their lower-order boundary
// fluxes), the open-boundary inflow fixes (src/adv.F:1091-1230,
// :2010-2150, :2890-3100) and adds the tendency to uten/vten/wten in place.
// Every flux is the reference's FP32 expression, so results are bitwise
// those of separate flux and divergence passes.  The operands of level k+1
// are loaded before level k is evaluated and stored.
#include "cm1g_internal.h"
#include "cm1g_adv_common.h"

struct AdvmArgs {
  const float *a;                    // advected velocity component
  const float *rru, *rrv, *rrw, *divx, *dum8;
  const float *ten_in;               // tendency the advection is added to
  float *ten;                        // uten / vten / wten
  double weps;
  int buoy;                          // w: add the buoyancy term (src/solve2.F:899-907)
  int i1, i2, j1, j2;                // point range of the staggered grid
};

// Operands of one level (raw loads).  With OWN the advected component's
// own mass flux (rru for u, rrv for v, rrw for w) is not loaded: inside
// the time step it is rho0*u, rho0*v, rf0*w of the same array
// (src/solve2.F:490-545), formed from the values already in registers.
struct AdvmLevel {
  float xm3, xm2, xm1, xp1, xp2, xp3;  // a(i-3..i+3), centre excluded
  float ym3, ym2, ym1, yp1, yp2, yp3;  // a(j-3..j+3)
  float r0, r1, r2, r3, q0, q1, b8;    // horizontal mass fluxes (per staggering, below)
  float wz0, wz1;                      // vertical face mass fluxes
  float dv0, dv1, ten;                 // divx, tendency (w: b8 is dum8)
  float wnew;                          // a(k+4), entering the vertical window
};

template <int STAG, bool OWN>
__device__ __forceinline__ void advm_load(const KP &P, const AdvmArgs &g, int i, int j, int k,
                                          AdvmLevel &L) {
  const int px = P.px, pxy = P.pxy, nk = P.nk;
  const int c = IXD(P, i, j, k);
  const float *a = g.a;
  L.xm3 = a[c - 3]; L.xm2 = a[c - 2]; L.xm1 = a[c - 1];
  L.xp1 = a[c + 1]; L.xp2 = a[c + 2]; L.xp3 = a[c + 3];
  L.ym3 = a[c - 3 * px]; L.ym2 = a[c - 2 * px]; L.ym1 = a[c - px];
  L.yp1 = a[c + px]; L.yp2 = a[c + 2 * px]; L.yp3 = a[c + 3 * px];
  L.ten = g.ten_in[c];
  if (STAG == 2) {
    // rru(i-1), rru(i), rru(i+1); rrv(i,j), rrv(i-1,j), rrv(i,j+1), rrv(i-1,j+1)
    if (!OWN) {
      L.r0 = g.rru[c - 1]; L.r1 = g.rru[c]; L.r2 = g.rru[c + 1];
    }
    L.r3 = g.rrv[c]; L.q0 = g.rrv[c - 1]; L.q1 = g.rrv[c + px];
    L.b8 = g.rrv[c + px - 1];
    // vertical face k+1: rrw(i,k+1), rrw(i-1,k+1)
    L.wz0 = (k < nk) ? g.rrw[c + pxy] : 0.0f;
    L.wz1 = (k < nk) ? g.rrw[c + pxy - 1] : 0.0f;
    L.dv0 = g.divx[c]; L.dv1 = g.divx[c - 1];
    L.wnew = (k + 4 <= nk) ? a[c + 4 * pxy] : 0.0f;
  } else if (STAG == 3) {
    // rrv(j-1), rrv(j), rrv(j+1); rru(i,j), rru(i,j-1), rru(i+1,j), rru(i+1,j-1)
    if (!OWN) {
      L.r0 = g.rrv[c - px]; L.r1 = g.rrv[c]; L.r2 = g.rrv[c + px];
    }
    L.r3 = g.rru[c]; L.q0 = g.rru[c - px]; L.q1 = g.rru[c + 1];
    L.b8 = g.rru[c + 1 - px];
    L.wz0 = (k < nk) ? g.rrw[c + pxy] : 0.0f;
    L.wz1 = (k < nk) ? g.rrw[c + pxy - px] : 0.0f;
    L.dv0 = g.divx[c]; L.dv1 = g.divx[c - px];
    L.wnew = (k + 4 <= nk) ? a[c + 4 * pxy] : 0.0f;
  } else {
    // rru(i), rru(i+1), rrv(j), rrv(j+1) at level k
    L.r0 = g.rru[c]; L.r1 = g.rru[c + 1]; L.r2 = g.rrv[c]; L.r3 = g.rrv[c + px];
    L.q0 = L.q1 = 0.0f;
    // vertical face k (between w(k) and w(k+1)): rrw(k), rrw(k+1)
    if (!OWN) {
      L.wz0 = g.rrw[c];
      L.wz1 = g.rrw[c + pxy];
    }
    L.dv0 = g.divx[c]; L.dv1 = 0.0f;
    L.b8 = g.buoy ? g.dum8[c] : 0.0f;
    L.wnew = (k + 4 <= nk + 1) ? a[c + 4 * pxy] : 0.0f;
  }
}

template <int STAG, bool WENO, bool OWN>
__device__ __forceinline__ void advm_col(const KP &P, const AdvmArgs &g, const ColGrid &cg, int bx) {
  int i, j;
  if (!col_tile_b(cg, bx, i, j)) return;
  if (i < g.i1 || i > g.i2 || j < g.j1 || j > g.j2) return;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const bool doitw = (P.wbc == 2 && P.ibw == 1), doite = (P.ebc == 2 && P.ibe == 1);
  const bool doits = (P.sbc == 2 && P.ibs == 1), doitn = (P.nbc == 2 && P.ibn == 1);
  const float *a = g.a;
  const double weps = g.weps;
  auto val = [&](int kk) -> float {  // column value at level kk (stencil range only)
    if (kk < 0 || kk > nk + 1) return 0.0f;
    return a[IXD(P, i, j, kk)];
  };
  const int k0 = (STAG == 4) ? 2 : 1;
  // vertical window: u, v use S(kf-3..kf+2) for face kf = k+1 (w0..w5 =
  // a(k-2..k+3)); w uses S(k-2..k+3) for face k (w0..w5 = a(k-2..k+3))
  float w0 = val(k0 - 2), w1 = val(k0 - 1), w2 = val(k0), w3 = val(k0 + 1), w4 = val(k0 + 2),
        w5 = val(k0 + 3);
  float fz_lo = 0.0f;
  // w: the level k-1 horizontal mass fluxes, divx, dum8 and face k-1 flux
  float pr0 = 0.0f, pr1 = 0.0f, pr2 = 0.0f, pr3 = 0.0f, pdv = 0.0f, pb8 = 0.0f;
  if (STAG == 4) {
    const int c1 = IXD(P, i, j, 1);
    pr0 = g.rru[c1]; pr1 = g.rru[c1 + 1]; pr2 = g.rrv[c1]; pr3 = g.rrv[c1 + P.px];
    pdv = g.divx[c1];
    pb8 = g.buoy ? g.dum8[c1] : 0.0f;
    // face 1
    auto S1 = [=](int kk) -> float {  // S(kk) = a(kk): the window holds a(0..5)
      const int o = kk - 2;
      return o == -2 ? w0 : o == -1 ? w1 : o == 0 ? w2 : o == 1 ? w3 : o == 2 ? w4 : w5;
    };
    const float rw1 = OWN ? 0.0f : g.rrw[c1];
    // the window holds a(0..5) here: w(2) = w2
    const float rw2 = OWN ? ((nk >= 2) ? P.rf0[2] * w2 : 0.0f) : g.rrw[c1 + P.pxy];
    const float wb = 0.5f * (rw1 + rw2);
    fz_lo = WENO ? vfluxw_of(P, 4, wb, S1, 1, 0, weps) : vflux5_of(P, 4, wb, S1, 1);
  }
  AdvmLevel nx;
  advm_load<STAG, OWN>(P, g, i, j, k0, nx);
  for (int k = k0; k <= nk; k++) {
    AdvmLevel L = nx;
    if (k < nk) advm_load<STAG, OWN>(P, g, i, j, k + 1, nx);
    const int c = IXD(P, i, j, k);
    if (OWN) {
      if (STAG == 2) {
        const float r = P.rho0[k];
        L.r0 = r * L.xm1; L.r1 = r * w2; L.r2 = r * L.xp1;
      } else if (STAG == 3) {
        const float r = P.rho0[k];
        L.r0 = r * L.ym1; L.r1 = r * w2; L.r2 = r * L.yp1;
      } else {
        // the window holds a(k-2..k+3): w(k) = w2, w(k+1) = w3
        L.wz0 = (k == 1) ? 0.0f : P.rf0[k] * w2;
        L.wz1 = (k == nk) ? 0.0f : P.rf0[k + 1] * w3;
      }
    }
    float ax, ay, az, t;
    if (STAG == 2) {
      const float sc = w2;
      const float ub1 = 0.5f * (L.r1 + L.r2), ub0 = 0.5f * (L.r0 + L.r1);
      const float fx1 = hface5<WENO>(ub1, L.xm2, L.xm1, sc, L.xp1, L.xp2, L.xp3, 0, weps);
      const float fx0 = hface5<WENO>(ub0, L.xm3, L.xm2, L.xm1, sc, L.xp1, L.xp2, 0, weps);
      ax = -(fx1 - fx0) * P.rdx * P.uf[i + 2];
      // rrv(i,j) = r3, rrv(i-1,j) = q0, rrv(i,j+1) = q1, rrv(i-1,j+1) = b8
      const float vb0 = 0.5f * (L.r3 + L.q0), vb1 = 0.5f * (L.q1 + L.b8);
      float fy0 = hface5<WENO>(vb0, L.ym3, L.ym2, L.ym1, sc, L.yp1, L.yp2, 0, weps);
      float fy1 = hface5<WENO>(vb1, L.ym2, L.ym1, sc, L.yp1, L.yp2, L.yp3, 0, weps);
      const float fy0_raw = fy0;
      const bool fs = doits && j == 1 && (L.r3 + L.q0) >= 0.0f;
      const bool fn = doitn && j == nj && (L.q1 + L.b8) <= 0.0f;
      if (fs) fy0 = fy1;
      if (fn) fy1 = fy0_raw;
      ay = -(fy1 - fy0) * P.rdy * P.vh[j + 2];
      if (fs) ay = ay - sc * 0.5f * ((L.b8 - L.q0) + (L.q1 - L.r3)) * P.rdy * P.vh[1 + 2];
      if (fn) ay = ay - sc * 0.5f * ((L.b8 - L.q0) + (L.q1 - L.r3)) * P.rdy * P.vh[nj + 2];
      float fz_hi = 0.0f;
      if (k < nk) {
        const int kf = k + 1;
        auto S = [=](int kk) -> float {
          const int o = kk - kf;
          return o == -3 ? w0 : o == -2 ? w1 : o == -1 ? w2 : o == 0 ? w3 : o == 1 ? w4 : w5;
        };
        const float wb = 0.5f * (L.wz0 + L.wz1);
        fz_hi = WENO ? vfluxw_of(P, 2, wb, S, kf, 0, weps) : vflux5_of(P, 2, wb, S, kf);
      }
      if (k == 1) az = -fz_hi * P.rdz * P.mh[1];
      else if (k == nk) az = +fz_lo * P.rdz * P.mh[nk];
      else az = -(fz_hi - fz_lo) * P.rdz * P.mh[k];
      fz_lo = fz_hi;
      t = L.ten + ((ax + ay) + az + sc * 0.5f * (L.dv0 + L.dv1)) * P.rr0[k];
    } else if (STAG == 3) {
      const float sc = w2;
      // rru(i,j) = r3, rru(i,j-1) = q0, rru(i+1,j) = q1, rru(i+1,j-1) = b8
      const float ub0 = 0.5f * (L.r3 + L.q0), ub1 = 0.5f * (L.q1 + L.b8);
      float fx0 = hface5<WENO>(ub0, L.xm3, L.xm2, L.xm1, sc, L.xp1, L.xp2, 0, weps);
      float fx1 = hface5<WENO>(ub1, L.xm2, L.xm1, sc, L.xp1, L.xp2, L.xp3, 0, weps);
      const float fx0_raw = fx0;
      const bool fw = doitw && i == 1 && (L.r3 + L.q0) >= 0.0f;
      const bool fe = doite && i == ni && (L.q1 + L.b8) <= 0.0f;
      if (fw) fx0 = fx1;
      if (fe) fx1 = fx0_raw;
      ax = -(fx1 - fx0) * P.rdx * P.uh[i + 2];
      if (fw) ax = ax - sc * 0.5f * ((L.b8 - L.q0) + (L.q1 - L.r3)) * P.rdx * P.uh[1 + 2];
      if (fe) ax = ax - sc * 0.5f * ((L.b8 - L.q0) + (L.q1 - L.r3)) * P.rdx * P.uh[ni + 2];
      const float vb1 = 0.5f * (L.r1 + L.r2), vb0 = 0.5f * (L.r0 + L.r1);
      const float fy1 = hface5<WENO>(vb1, L.ym2, L.ym1, sc, L.yp1, L.yp2, L.yp3, 0, weps);
      const float fy0 = hface5<WENO>(vb0, L.ym3, L.ym2, L.ym1, sc, L.yp1, L.yp2, 0, weps);
      ay = -(fy1 - fy0) * P.rdy * P.vf[j + 2];
      float fz_hi = 0.0f;
      if (k < nk) {
        const int kf = k + 1;
        auto S = [=](int kk) -> float {
          const int o = kk - kf;
          return o == -3 ? w0 : o == -2 ? w1 : o == -1 ? w2 : o == 0 ? w3 : o == 1 ? w4 : w5;
        };
        const float wb = 0.5f * (L.wz0 + L.wz1);
        fz_hi = WENO ? vfluxw_of(P, 3, wb, S, kf, 0, weps) : vflux5_of(P, 3, wb, S, kf);
      }
      if (k == 1) az = -fz_hi * P.rdz * P.mh[1];
      else if (k == nk) az = +fz_lo * P.rdz * P.mh[nk];
      else az = -(fz_hi - fz_lo) * P.rdz * P.mh[k];
      fz_lo = fz_hi;
      t = L.ten + ((ax + ay) + az + sc * 0.5f * (L.dv0 + L.dv1)) * P.rr0[k];
    } else {
      const float sc = w2;
      const float cc2 = 0.5f * (P.c2[k] + P.c2[k]), cc1 = 1.0f - cc2;
      // rru(i) = r0, rru(i+1) = r1, rrv(j) = r2, rrv(j+1) = r3 at k; pr* at k-1
      const float ub0 = cc2 * L.r0 + cc1 * pr0, ub1 = cc2 * L.r1 + cc1 * pr1;
      float fx0 = hface5<WENO>(ub0, L.xm3, L.xm2, L.xm1, sc, L.xp1, L.xp2, 0, weps);
      float fx1 = hface5<WENO>(ub1, L.xm2, L.xm1, sc, L.xp1, L.xp2, L.xp3, 0, weps);
      const float fx0_raw = fx0;
      const bool fw = doitw && i == 1 && ub0 >= 0.0f;
      const bool fe = doite && i == ni && ub1 <= 0.0f;
      if (fw) fx0 = fx1;
      if (fe) fx1 = fx0_raw;
      ax = -(fx1 - fx0) * P.rdx * P.uh[i + 2];
      if (fw) ax = ax - sc * (P.c1[k] * (pr1 - pr0) + P.c2[k] * (L.r1 - L.r0)) * P.rdx * P.uh[1 + 2];
      if (fe) ax = ax - sc * (P.c1[k] * (pr1 - pr0) + P.c2[k] * (L.r1 - L.r0)) * P.rdx * P.uh[ni + 2];
      const float vb0 = cc2 * L.r2 + cc1 * pr2, vb1 = cc2 * L.r3 + cc1 * pr3;
      float fy0 = hface5<WENO>(vb0, L.ym3, L.ym2, L.ym1, sc, L.yp1, L.yp2, 0, weps);
      float fy1 = hface5<WENO>(vb1, L.ym2, L.ym1, sc, L.yp1, L.yp2, L.yp3, 0, weps);
      const float fy0_raw = fy0;
      const bool fs = doits && j == 1 && vb0 >= 0.0f;
      const bool fn = doitn && j == nj && vb1 <= 0.0f;
      if (fs) fy0 = fy1;
      if (fn) fy1 = fy0_raw;
      ay = -(fy1 - fy0) * P.rdy * P.vh[j + 2];
      if (fs) ay = ay - sc * (P.c1[k] * (pr3 - pr2) + P.c2[k] * (L.r3 - L.r2)) * P.rdy * P.vh[1 + 2];
      if (fn) ay = ay - sc * (P.c1[k] * (pr3 - pr2) + P.c2[k] * (L.r3 - L.r2)) * P.rdy * P.vh[nj + 2];
      // vertical face k from a(k-2..k+3)
      auto S = [=](int kk) -> float {
        const int o = kk - k;
        return o == -2 ? w0 : o == -1 ? w1 : o == 0 ? w2 : o == 1 ? w3 : o == 2 ? w4 : w5;
      };
      const float wb = 0.5f * (L.wz0 + L.wz1);
      const float fz = WENO ? vfluxw_of(P, 4, wb, S, k, 0, weps) : vflux5_of(P, 4, wb, S, k);
      az = -(fz - fz_lo) * P.rdz * P.mf[k];
      fz_lo = fz;
      t = L.ten + ((ax + ay) + az + sc * (P.c2[k] * L.dv0 + P.c1[k] * pdv)) * P.rrf0[k];
      if (g.buoy) t = t + (P.c1[k] * pb8 + P.c2[k] * L.b8);
      pr0 = L.r0; pr1 = L.r1; pr2 = L.r2; pr3 = L.r3; pdv = L.dv0; pb8 = L.b8;
    }
    g.ten[c] = t;
    w0 = w1;
    w1 = w2;
    w2 = w3;
    w3 = w4;
    w4 = w5;
    w5 = L.wnew;
  }
}

// With from_ten1 the tendencies start from uten1/vten1/wten1 (the copy of
// src/solve2.F:459-471): the advection kernels read them, and this kernel
// copies the points they do not write (the faces i=ni+1 / j=nj+1, the
// open-boundary faces outside the advected range, and w at k=1).
__device__ __forceinline__ void ten_edges(const KP &P, const float *u1, const float *v1, const float *w1,
                                          float *u, float *v, float *w, int ui1, int ui2, int vj1,
                                          int vj2, int bx, int by) {
  const int i = 1 + bx * (int)blockDim.x + (int)threadIdx.x;
  const int j = 1 + by * (int)blockDim.y + (int)threadIdx.y;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (i > ni + 1 || j > nj + 1) return;
  const bool ucol = (j == nj + 1) || i < ui1 || i > ui2;
  const bool vcol = (i == ni + 1) || j < vj1 || j > vj2;
  const bool wcol = (i == ni + 1) || (j == nj + 1);
  for (int k = 1; k <= nk; k++) {
    const int c = IXD(P, i, j, k);
    if (ucol) u[c] = u1[c];
    if (vcol) v[c] = v1[c];
    if (wcol || k == 1) w[c] = w1[c];
    if (!(ucol || vcol || wcol)) break;
  }
}

// uten/vten/wten += advection (the momentum part of adv_uvw); mask bit 1 u,
// 2 v, 4 w; buoy adds the buoyancy term to wten; own forms each
// component's own mass flux from the velocity (valid while rru/rrv/rrw
// hold rho0*u, rho0*v, rf0*w, as between src/solve2.F:490-545 and sound).
template <int STAG, bool WENO, bool OWN>
static void advm_launch(cm1g_model *m, const KP &P, const AdvmArgs &g, const ColGrid &cg) {
  launch<256>(m, col_blocks(cg), col_threads(cg),
              [=] __device__(int bx, int, int) { advm_col<STAG, WENO, OWN>(P, g, cg, bx); });
}

void advm_run(cm1g_model *m, int nrk, int mask, int buoy, int from_ten1, int own, const float *const *src3) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj;
  const bool weno = (m->c.advwenov == 1) || (m->c.advwenov == 2 && nrk == 3);
  AdvmArgs g;
  g.rru = m->f[CM1G_RRU];
  g.rrv = m->f[CM1G_RRV];
  g.rrw = m->f[CM1G_RRW];
  g.divx = m->f[CM1G_DIVX];
  g.dum8 = m->f[CM1G_DUM8];
  g.weps = (double)(100.0f * P.K.epsilon);
  g.buoy = buoy;
  const int ui1 = P.ibw == 1 ? 2 : 1, ui2 = P.ibe == 1 ? ni : ni + 1;
  const int vj1 = P.ibs == 1 ? 2 : 1, vj2 = P.ibn == 1 ? nj : nj + 1;
  if (from_ten1) {
    // the launch covers all three components, so every mask bit must be set
    const float *u1 = m->f[CM1G_UTEN1], *v1 = m->f[CM1G_VTEN1], *w1 = m->f[CM1G_WTEN1];
    float *u = m->f[CM1G_UTEN], *v = m->f[CM1G_VTEN], *w = m->f[CM1G_WTEN];
    launch<256>(m, dim3((ni + 1 + 63) / 64, (nj + 1 + 3) / 4, 1), dim3(64, 4, 1),
                [=] __device__(int bx, int by, int) {
                  ten_edges(P, u1, v1, w1, u, v, w, ui1, ui2, vj1, vj2, bx, by);
                });
  }
  if (mask & 1) {
    g.a = src3 ? src3[0] : m->f[CM1G_U3D];
    g.ten = m->f[CM1G_UTEN];
    g.ten_in = from_ten1 ? m->f[CM1G_UTEN1] : g.ten;
    g.i1 = ui1;
    g.i2 = ui2;
    g.j1 = 1;
    g.j2 = nj;
    const ColGrid cg = col_grid(ni + 1, nj, m->tile_rows);
    auto *kern = own ? (weno ? advm_launch<2, true, true> : advm_launch<2, false, true>)
                     : (weno ? advm_launch<2, true, false> : advm_launch<2, false, false>);
    kern(m, P, g, cg);
  }
  if (mask & 2) {
    g.a = src3 ? src3[1] : m->f[CM1G_V3D];
    g.ten = m->f[CM1G_VTEN];
    g.ten_in = from_ten1 ? m->f[CM1G_VTEN1] : g.ten;
    g.i1 = 1;
    g.i2 = ni;
    g.j1 = vj1;
    g.j2 = vj2;
    const ColGrid cg = col_grid(ni, nj + 1, m->tile_rows);
    auto *kern = own ? (weno ? advm_launch<3, true, true> : advm_launch<3, false, true>)
                     : (weno ? advm_launch<3, true, false> : advm_launch<3, false, false>);
    kern(m, P, g, cg);
  }
  if (mask & 4) {
    g.a = src3 ? src3[2] : m->f[CM1G_W3D];
    g.ten = m->f[CM1G_WTEN];
    g.ten_in = from_ten1 ? m->f[CM1G_WTEN1] : g.ten;
    g.i1 = 1;
    g.i2 = ni;
    g.j1 = 1;
    g.j2 = nj;
    const ColGrid cg = col_grid(ni, nj, m->tile_rows);
    auto *kern = own ? (weno ? advm_launch<4, true, true> : advm_launch<4, false, true>)
                     : (weno ? advm_launch<4, true, false> : advm_launch<4, false, false>);
    kern(m, P, g, cg);
  }
}

void advw_tke(cm1g_model *m, int nrk, int advweno, const float *rru, const float *rrv, const float *rrw,
              const float *a, const float *ten_in, float *ten) {
  const KP P = m->P;
  const bool weno = (advweno == 1) || (advweno == 2 && nrk == 3);
  AdvmArgs g;
  g.rru = rru;
  g.rrv = rrv;
  g.rrw = rrw;
  g.divx = m->f[CM1G_DIVX];
  g.dum8 = m->f[CM1G_DUM8];
  g.weps = (double)(100.0f * P.K.epsilon);
  g.buoy = 0;
  g.a = a;
  g.ten_in = ten_in;
  g.ten = ten;
  g.i1 = 1;
  g.i2 = P.ni;
  g.j1 = 1;
  g.j2 = P.nj;
  const ColGrid cg = col_grid(P.ni, P.nj, m->tile_rows);
  auto *kern = weno ? advm_launch<4, true, false> : advm_launch<4, false, false>;
  kern(m, P, g, cg);
}
