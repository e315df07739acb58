// Flux-form advection on the device.
// flx3/flx5: src/adv_routines.F:5971-5984, :6006-6019.  hadv_flx5
// :1335-1476, vadv_flx5 :3269-3388 (+ vadv_lwr_flx_23 :4817-4995,
// vadv_lwr_flx_35 :5001-5169).  advs: src/adv.F:14-903, advu :911-1839,
// advv :1847-2770, advw :2778-3610; pdef: src/pdef.F:15-581.
//
// Every flux is evaluated with the reference's operation order in FP32
// without contraction, so results are bitwise those of the reference.
#include "cm1g_adv_common.h"

// ------------------------------------------------------------------ pdef
// The positive-definite correction (pdscheme=1, src/pdef.F:15-581), per
// direction: dum = rho0*s0 + dt*(adv + s*div) (pdefx1/pdefy1/pdefz), its
// halo fill (bcs, or zero below and above the column), and the mass each
// cell takes from or gives to its two neighbours (pdefx2/pdefy2/pdefz).
// The reference scatters inside a sequential loop; here each cell gathers
// the same terms in the same order (left neighbour, itself, right
// neighbour), which reproduces the sequential sums bit for bit.  dum is not
// stored: each cell evaluates it at the five points it needs, mapping halo
// points to the interior point the fill copies.

// mass flux of direction dir at the lower face of (c, level k) and at the
// upper one (c + stride): with scale, the getdivx scaling rho0(k) for u,v
// and rf0 for w (zero at k=1 and nk+1)
__device__ __forceinline__ void pdef_faces(const KP &P, int dir, const float *vel, int scale, int c,
                                           int k, float &lo, float &hi) {
  if (dir == 0) {
    lo = vel[c];
    hi = vel[c + 1];
    if (scale) { lo = lo * P.rho0[k]; hi = hi * P.rho0[k]; }
  } else if (dir == 1) {
    lo = vel[c];
    hi = vel[c + P.px];
    if (scale) { lo = lo * P.rho0[k]; hi = hi * P.rho0[k]; }
  } else {
    lo = vel[c];
    hi = vel[c + P.pxy];
    if (scale) {
      lo = (k == 1) ? 0.0f : lo * P.rf0[k];
      hi = (k + 1 == P.nk + 1) ? 0.0f : hi * P.rf0[k + 1];
    }
  }
}

// dum at (i,j,k) for direction dir (0 x, 1 y, 2 z), interior point
__device__ __forceinline__ float pdef_dum(const KP &P, int dir, const float *vel, int scale,
                                          const float *adv, const float *s0, const float *s, float dt,
                                          int i, int j, int k) {
  const int c = IXD(P, i, j, k);
  float lo, hi;
  pdef_faces(P, dir, vel, scale, c, k, lo, hi);
  // s*(flux difference)*metric, associated left to right as the reference
  // writes it (src/pdef.F:44-45, :256-257, :482-483)
  const float sc = s[c];
  float sdiv;
  if (dir == 0) sdiv = sc * (hi - lo) * P.rdx * P.uh[i + 2];
  else if (dir == 1) sdiv = sc * (hi - lo) * P.rdy * P.vh[j + 2];
  else sdiv = sc * (hi - lo) * P.rdz * P.mh[k];
  float d = P.rho0[k] * s0[c] + dt * (adv[c] + sdiv);
  if (dir == 0) {
    if (P.wbc == 2 && P.ibw == 1 && i == 1 && lo >= 0.0f) d = P.rho0[k] * s0[c];
    if (P.ebc == 2 && P.ibe == 1 && i == P.ni && hi <= 0.0f) d = P.rho0[k] * s0[c];
  } else if (dir == 1) {
    if (P.sbc == 2 && P.ibs == 1 && j == 1 && lo >= 0.0f) d = P.rho0[k] * s0[c];
    if (P.nbc == 2 && P.ibn == 1 && j == P.nj && hi <= 0.0f) d = P.rho0[k] * s0[c];
  }
  return d;
}

// halo point of the horizontal fill (bcs, src/bc.F:180-364) -> the
// interior point whose value it holds
__device__ __forceinline__ int pdef_wrap(int m, int n, bool periodic) {
  if (m < 1) return periodic ? m + n : 1;
  if (m > n) return periodic ? m - n : n;
  return m;
}

// the corrected tendency of one cell from dum at offsets -2..2 along the
// direction; vm/vp: the neighbours -1/+1 are cells the reference visits
__device__ __forceinline__ float pdef_gather(float dm2, float dm1, float d0, float dp1, float dp2,
                                             bool vm, bool vp, float smeps, bool &flag) {
  float mass = 0.0f;
  flag = false;
  if (vm && dm1 < 0.0f) {
    const float foo1 = fmaxf(0.0f, dm2), foo2 = fmaxf(0.0f, d0);
    if (foo1 + foo2 > smeps) {
      const float foo3 = fmaxf(dm1, -(foo1 + foo2)) / (foo1 + foo2);
      mass = mass + foo2 * foo3;
      if (d0 > smeps) flag = true;
    }
  }
  if (d0 < 0.0f) {
    const float foo1 = fmaxf(0.0f, dm1), foo2 = fmaxf(0.0f, dp1);
    if (foo1 + foo2 > smeps) {
      const float foo3 = fmaxf(d0, -(foo1 + foo2)) / (foo1 + foo2);
      mass = mass - (foo1 + foo2) * foo3;
      flag = true;
    }
  }
  if (vp && dp1 < 0.0f) {
    const float foo1 = fmaxf(0.0f, d0), foo2 = fmaxf(0.0f, dp2);
    if (foo1 + foo2 > smeps) {
      const float foo3 = fmaxf(dp1, -(foo1 + foo2)) / (foo1 + foo2);
      mass = mass + foo1 * foo3;
      if (d0 > smeps) flag = true;
    }
  }
  return mass;
}

// x and y directions (pdefx1/x2, pdefy1/y2): corrected advx, advy
struct PdefArgs {
  const float *rru, *rrv, *rrw, *advx, *advy, *advz, *s0, *s;
  float *ax_out, *ay_out;
  const float *sten0, *t11, *thterm, *divx;
  float *out_sten, *out_new;
  float dt;
  int smeps_flush, scale;
};

// One 64 x 4 block per tile of one level.  Each cell's dum in x and in y
// is evaluated once and shared through LDS: a cell needs dum at i-2..i+2
// and at j-2..j+2, so the block evaluates its own 64 x 4 points in each
// direction plus the two-point halos (4 x 4 in x, 64 x 4 in y), at the
// indices the halo fill maps them to.
__device__ __forceinline__ void pdef_xy(const KP &P, const PdefArgs &a, int bx, int by, int bz) {
  __shared__ float dxs[4][68];
  __shared__ float dys[8][64];
  const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
  const int i0 = 1 + bx * 64, j0 = 1 + by * 4;
  const int i = i0 + tx, j = j0 + ty;
  const int k = 1 + bz;
  const int ni = P.ni, nj = P.nj;
  const bool perx = (P.wbc == 1), pery = (P.sbc == 1);
  const bool in = (i <= ni && j <= nj);
  {
    // x: own point of this row, and the row's halo points i0-2, i0-1, i0+64, i0+65
    const int jj = pdef_wrap(j, nj, pery);
    dxs[ty][tx + 2] = pdef_dum(P, 0, a.rru, a.scale, a.advx, a.s0, a.s, a.dt, pdef_wrap(i, ni, perx), jj, k);
    if (tx < 4) {
      const int o = (tx < 2) ? tx - 2 : 62 + tx;
      dxs[ty][(tx < 2) ? tx : 64 + tx] =
          pdef_dum(P, 0, a.rru, a.scale, a.advx, a.s0, a.s, a.dt, pdef_wrap(i0 + o, ni, perx), jj, k);
    }
    // y: own point of this column, and one halo row (j0-2, j0-1, j0+4, j0+5)
    const int ii = (i <= ni) ? i : ni;
    dys[ty + 2][tx] = pdef_dum(P, 1, a.rrv, a.scale, a.advy, a.s0, a.s, a.dt, ii, pdef_wrap(j, nj, pery), k);
    const int hr = (ty < 2) ? ty : ty + 4;
    dys[hr][tx] = pdef_dum(P, 1, a.rrv, a.scale, a.advy, a.s0, a.s, a.dt, ii,
                           pdef_wrap(j0 - 2 + hr, nj, pery), k);
  }
  __syncthreads();
  if (!in) return;
  const int c = IXD(P, i, j, k);
  const float smeps = P.K.smeps, rdt = 1.0f / a.dt;
  const float r0s0 = P.rho0[k] * a.s0[c];
  {
    const float *d = &dxs[ty][tx];
    bool flag;
    const float mass = pdef_gather(d[0], d[1], d[2], d[3], d[4], i - 1 >= 0, i + 1 <= ni + 1, smeps, flag);
    float out = a.advx[c];
    if (flag) {
      float lo, hi;
      pdef_faces(P, 0, a.rru, a.scale, c, k, lo, hi);
      out = (d[2] + mass - r0s0) * rdt - a.s[c] * (hi - lo) * P.rdx * P.uh[i + 2];
    }
    a.ax_out[c] = out;
  }
  {
    bool flag;
    const float mass = pdef_gather(dys[ty][tx], dys[ty + 1][tx], dys[ty + 2][tx], dys[ty + 3][tx],
                                   dys[ty + 4][tx], j - 1 >= 0, j + 1 <= nj + 1, smeps, flag);
    float out = a.advy[c];
    if (flag) {
      float lo, hi;
      pdef_faces(P, 1, a.rrv, a.scale, c, k, lo, hi);
      out = (dys[ty + 2][tx] + mass - r0s0) * rdt - a.s[c] * (hi - lo) * P.rdy * P.vh[j + 2];
    }
    a.ay_out[c] = out;
  }
}

// z direction (pdefz) marched up each column with a window of dum, then the
// tendency with the corrected advx, advy, advz and the s0 + dt*sten update
__device__ __forceinline__ void pdef_z(const KP &P, const PdefArgs &a, const ColGrid &cg, int bx) {
  int i, j;
  if (!col_tile_b(cg, bx, i, j)) return;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (i > ni || j > nj) return;
  const float smeps = P.K.smeps, rdt = 1.0f / a.dt;
  auto dz = [&](int kk) -> float {  // dum of pdefz; zero outside 1..nk
    if (kk < 1 || kk > nk) return 0.0f;
    return pdef_dum(P, 2, a.rrw, a.scale, a.advz, a.s0, a.s, a.dt, i, j, kk);
  };
  float dm2 = 0.0f, dm1 = 0.0f, d0 = dz(1), dp1 = dz(2), dp2 = dz(3);
  for (int k = 1; k <= nk; k++) {
    const int c = IXD(P, i, j, k);
    bool flag;
    const float mass = pdef_gather(dm2, dm1, d0, dp1, dp2, k - 1 >= 1, k + 1 <= nk, smeps, flag);
    const float sc = a.s[c];
    float az = a.advz[c];
    if (flag) {
      float lo, hi;
      pdef_faces(P, 2, a.rrw, a.scale, c, k, lo, hi);
      az = (d0 + mass - P.rho0[k] * a.s0[c]) * rdt - sc * (hi - lo) * P.rdz * P.mh[k];
    }
    float st = a.sten0[c];
    if (a.t11) st = st - a.t11[c] * a.thterm[c];
    st = st + ((a.ax_out[c] + a.ay_out[c]) + az + sc * a.divx[c]) * P.rr0[k];
    if (a.out_sten) a.out_sten[c] = st;
    if (a.out_new) {
      float v = a.s0[c] + a.dt * st;
      if (a.smeps_flush && fabsf(v) < P.K.smeps) v = 0.0f;
      a.out_new[c] = v;
    }
    dm2 = dm1;
    dm1 = d0;
    d0 = dp1;
    dp1 = dp2;
    dp2 = dz(k + 3);
  }
}

// ------------------------------------------------------ fused advs column
struct AdvsArgs {
  const float *s;                 // advected scalar
  float sadd;                     // nonzero: advect (th0(k)-th0r) + s
  const float *s0, *sten0, *t11, *thterm, *rru, *rrv, *rrw, *divx;
  float *out_sten, *out_new;      // tendency and/or s0 + dt*tendency
  float *advx, *advy, *advz;      // mode 1: the three advective tendencies
  float dt;
  int smeps_flush, mode;          // mode 0: final tendency; 1: advx/advy/advz only
  int pdef, pdefweno;             // WENO5 magnitude tests (horizontal / vertical)
  double weps;
  int scale;                      // rru/rrv/rrw are the averaged velocities: scale by
                                  // rho0 / rf0 on the fly (getdivx, src/misclibs.F:222-370)
};

// Operands of one level of a advs_col column (raw loads; the theta
// offset th0-th0r is applied when they are used).
struct AdvLevel {
  float xm3, xm2, xm1, xp1, xp2, xp3;  // s(i-3..i+3), centre excluded
  float ym3, ym2, ym1, yp1, yp2, yp3;  // s(j-3..j+3)
  float ru0, ru1, rv0, rv1, rw1;       // face mass fluxes (rw1: face k+1)
  float dv, st0, sv0, t11v, thtv;      // divx, initial tendency, s0, t11, thterm
  float wnew;                          // s(k+4), entering the vertical window
};

template <bool WENO, bool TH>
__device__ __forceinline__ void advs_load(const KP &P, const AdvsArgs &a, int i, int j, int k,
                                          AdvLevel &L) {
  const int px = P.px, pxy = P.pxy, nk = P.nk;
  const int c = IXD(P, i, j, k);
  const float *s = a.s;
  L.xm3 = s[c - 3]; L.xm2 = s[c - 2]; L.xm1 = s[c - 1];
  L.xp1 = s[c + 1]; L.xp2 = s[c + 2]; L.xp3 = s[c + 3];
  L.ym3 = s[c - 3 * px]; L.ym2 = s[c - 2 * px]; L.ym1 = s[c - px];
  L.yp1 = s[c + px]; L.yp2 = s[c + 2 * px]; L.yp3 = s[c + 3 * px];
  L.ru0 = a.rru[c]; L.ru1 = a.rru[c + 1];
  L.rv0 = a.rrv[c]; L.rv1 = a.rrv[c + px];
  L.rw1 = (k < nk) ? a.rrw[c + pxy] : 0.0f;
  if (a.scale) {
    const float r0 = P.rho0[k];
    L.ru0 = L.ru0 * r0; L.ru1 = L.ru1 * r0;
    L.rv0 = L.rv0 * r0; L.rv1 = L.rv1 * r0;
    if (k < nk) L.rw1 = L.rw1 * P.rf0[k + 1];
  }
  L.wnew = (k + 4 <= nk) ? s[c + 4 * pxy] : 0.0f;
  L.dv = a.divx[c];
  L.st0 = a.sten0[c];
  L.sv0 = a.s0[c];
  L.t11v = TH ? a.t11[c] : 0.0f;
  L.thtv = TH ? a.thterm[c] : 0.0f;
}

// One column (i,j) of advs (src/adv.F:14-658 with hadv_flx5/vadv_flx5, or
// hadv_weno5/vadv_weno5 when WENO): face fluxes at i, i+1, j, j+1 with the
// open-boundary inflow fixes (src/adv.F:211-340), the vertical fluxes
// marched up the column with a six-level register window, and the tendency
// assembly.  Each flux is the same FP32 expression as the reference's, so
// results are bitwise equal to separate flux / divergence passes.  The
// operands of level k+1 are loaded before level k is evaluated and stored.
// The uniform switches are compile-time: SADD (theta's offset form), TH
// (theta's t11*thterm term), MODE1 (advx/advy/advz for the pdef pass).
template <bool WENO, bool SADD, bool TH, bool MODE1>
__device__ __forceinline__ void advs_col(const KP &P, const AdvsArgs &a, const ColGrid &cg, int bx) {
  int i, j;
  if (!col_tile_b(cg, bx, i, j)) return;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  if (i > ni || j > nj) return;
  const bool doitw = (P.wbc == 2 && P.ibw == 1), doite = (P.ebc == 2 && P.ibe == 1);
  const bool doits = (P.sbc == 2 && P.ibs == 1), doitn = (P.nbc == 2 && P.ibn == 1);
  const float *s = a.s;
  constexpr bool sadd = SADD;
  const float th0r = P.K.th0r;
  auto val = [&](int kk) -> float {  // column value at level kk (1..nk)
    if (kk < 1 || kk > nk) return 0.0f;
    const float v = s[IXD(P, i, j, kk)];
    return sadd ? (P.th0[kk] - th0r) + v : v;
  };
  // window: wv[0..5] = S(k-2 .. k+3)
  float wv0 = val(-1), wv1 = val(0), wv2 = val(1), wv3 = val(2), wv4 = val(3), wv5 = val(4);
  float fz_lo = 0.0f;
  AdvLevel nx;
  advs_load<WENO, TH>(P, a, i, j, 1, nx);
  for (int k = 1; k <= nk; k++) {
    const AdvLevel L = nx;
    if (k < nk) advs_load<WENO, TH>(P, a, i, j, k + 1, nx);
    const int c = IXD(P, i, j, k);
    const float tem = sadd ? (P.th0[k] - th0r) : 0.0f;
    auto H = [=](float v) -> float { return sadd ? tem + v : v; };
    const float sc = wv2;
    // x faces i and i+1
    const float xm3 = H(L.xm3), xm2 = H(L.xm2), xm1 = H(L.xm1), xp1 = H(L.xp1), xp2 = H(L.xp2),
                xp3 = H(L.xp3);
    const float ru0 = L.ru0, ru1 = L.ru1;
    float fx0 = hface5<WENO>(ru0, xm3, xm2, xm1, sc, xp1, xp2, a.pdef, a.weps);
    float fx1 = hface5<WENO>(ru1, xm2, xm1, sc, xp1, xp2, xp3, a.pdef, a.weps);
    const float fx0_raw = fx0;
    if (doitw && i == 1 && ru0 >= 0.0f) fx0 = fx1;
    if (doite && i == ni && ru1 <= 0.0f) fx1 = fx0_raw;
    // y faces j and j+1
    const float ym3 = H(L.ym3), ym2 = H(L.ym2), ym1 = H(L.ym1), yp1 = H(L.yp1), yp2 = H(L.yp2),
                yp3 = H(L.yp3);
    const float rv0 = L.rv0, rv1 = L.rv1;
    float fy0 = hface5<WENO>(rv0, ym3, ym2, ym1, sc, yp1, yp2, a.pdef, a.weps);
    float fy1 = hface5<WENO>(rv1, ym2, ym1, sc, yp1, yp2, yp3, a.pdef, a.weps);
    const float fy0_raw = fy0;
    if (doits && j == 1 && rv0 >= 0.0f) fy0 = fy1;
    if (doitn && j == nj && rv1 <= 0.0f) fy1 = fy0_raw;
    float ax = -(fx1 - fx0) * P.rdx * P.uh[i + 2];
    float ay = -(fy1 - fy0) * P.rdy * P.vh[j + 2];
    if (doitw && i == 1 && ru0 >= 0.0f) ax = ax - sc * (ru1 - ru0) * P.rdx * P.uh[1 + 2];
    if (doite && i == ni && ru1 <= 0.0f) ax = ax - sc * (ru1 - ru0) * P.rdx * P.uh[ni + 2];
    if (doits && j == 1 && rv0 >= 0.0f) ay = ay - sc * (rv1 - rv0) * P.rdy * P.vh[1 + 2];
    if (doitn && j == nj && rv1 <= 0.0f) ay = ay - sc * (rv1 - rv0) * P.rdy * P.vh[nj + 2];
    // vertical face k+1 from S(k-2..k+3)
    float fz_hi = 0.0f;
    if (k + 1 <= nk) {
      const int kf = k + 1;
      auto S = [=](int kk) -> float {
        const int o = kk - kf;  // -3..2
        return o == -3 ? wv0 : o == -2 ? wv1 : o == -1 ? wv2 : o == 0 ? wv3 : o == 1 ? wv4 : wv5;
      };
      fz_hi = WENO ? vfluxw_of(P, 1, L.rw1, S, kf, a.pdefweno, a.weps)
                   : vflux5_of(P, 1, L.rw1, S, kf);
    }
    float az;
    if (k == 1) az = -fz_hi * P.rdz * P.mh[1];
    else if (k == nk) az = +fz_lo * P.rdz * P.mh[nk];
    else az = -(fz_hi - fz_lo) * P.rdz * P.mh[k];
    if (MODE1) {
      a.advx[c] = ax;
      a.advy[c] = ay;
      a.advz[c] = az;
    } else {
      float st = L.st0;
      if (TH) st = st - L.t11v * L.thtv;
      st = st + ((ax + ay) + az + sc * L.dv) * P.rr0[k];
      if (a.out_sten) a.out_sten[c] = st;
      if (a.out_new) {
        float v = L.sv0 + a.dt * st;
        if (a.smeps_flush && fabsf(v) < P.K.smeps) v = 0.0f;
        a.out_new[c] = v;
      }
    }
    fz_lo = fz_hi;
    wv0 = wv1;
    wv1 = wv2;
    wv2 = wv3;
    wv3 = wv4;
    wv4 = wv5;
    wv5 = (k + 4 <= nk) ? (sadd ? (P.th0[k + 4] - th0r) + L.wnew : L.wnew) : 0.0f;
  }
}

// ------------------------------------------------------------------ advs
// sadd: when nonzero, s(i,j,k) = (th0(k) - th0r) + s_in(i,j,k) is formed on
// the fly (the theta equation's sadv, src/solve2.F:1354-1374).
// sten0/thterm_div: initial tendency; with thterm_div != nullptr the
// initial tendency is thten1 - t11*thterm (src/solve2.F:1320-1330), where
// sten0 = thten1, thterm_div[0] = t11 and thterm_div[1] = thterm.
// out_new, if given, receives s0 + dt*tendency with the smeps flush
// (src/solve2.F:1416-1426, :1518-1528).
template <bool WENO, bool SADD, bool TH, bool MODE1>
static void advs_launch(cm1g_model *m, const KP &P, const AdvsArgs &a, const ColGrid &cg) {
  // 8 waves per SIMD (64 registers) for the fifth-order scalar passes: the
  // 4096 workgroups of a 1024 x 1024 grid then run in exactly two rounds (7
  // waves per SIMD leave a third round 29% full).  Measured per launch: q
  // 849 against 943 us, the pdef pass 793 against 866 us (both without
  // spilling), theta 1158 against 1185 us (20 bytes per lane spilled)
  constexpr int W = (CM1G_COL_WAVES >= 1 && !WENO && (!TH || CM1G_COL_WAVES >= 2)) ? 8 : 0;
  launch<256, W>(m, col_blocks(cg), col_threads(cg),
                 [=] __device__(int bx, int, int) { advs_col<WENO, SADD, TH, MODE1>(P, a, cg, bx); });
}

void advs_run(cm1g_model *m, int nrk, const float *s_in, float sadd, const float *s0,
              const float *sten0, const float *thterm_div, float *out_sten, float *out_new,
              int pdef, int pdefweno, float dt, double weps, int smeps_flush) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  // WENO5 on every stage (advwenos=1) or the last one (2), src/adv.F:75-78
  const bool weno = (m->c.advwenos == 1) || (m->c.advwenos == 2 && nrk == 3);
  AdvsArgs g;
  g.s = s_in;
  g.sadd = sadd;
  g.s0 = s0;
  g.sten0 = sten0;
  g.t11 = thterm_div;
  g.thterm = thterm_div ? m->f[CM1G_THTERM] : nullptr;
  g.rru = m->f[CM1G_RRU];
  g.rrv = m->f[CM1G_RRV];
  g.rrw = m->f[CM1G_RRW];
  g.scale = m->advs_unscaled ? 0 : 1;
  g.divx = m->f[CM1G_DIVX];
  g.out_sten = out_sten;
  g.out_new = out_new;
  g.advx = m->tx;
  g.advy = m->ty;
  g.advz = m->tz;
  g.dt = dt;
  g.smeps_flush = smeps_flush;
  g.pdef = pdef;
  g.pdefweno = pdefweno;
  g.weps = weps;
  // compile-time variants: theta (offset form, with or without the t11
  // term), any other scalar in the plain or the pdef (advx/advy/advz) pass
  const bool sadd_on = (sadd != 0.0f), th_on = (thterm_div != nullptr);
  using Launcher = void (*)(cm1g_model *, const KP &, const AdvsArgs &, const ColGrid &);
  auto pick = [&](bool mode1) -> Launcher {
    if (sadd_on && th_on && !mode1) return weno ? advs_launch<true, true, true, false> : advs_launch<false, true, true, false>;
    if (sadd_on && !th_on && !mode1) return weno ? advs_launch<true, true, false, false> : advs_launch<false, true, false, false>;
    if (!sadd_on && !th_on && !mode1) return weno ? advs_launch<true, false, false, false> : advs_launch<false, false, false, false>;
    if (!sadd_on && th_on && !mode1) return weno ? advs_launch<true, false, true, false> : advs_launch<false, false, true, false>;
    // the pdef pass: advx/advy/advz only (the t11 term is added by pdef_z)
    return weno ? advs_launch<true, false, false, true> : advs_launch<false, false, false, true>;
  };
  const ColGrid cg = col_grid(ni, nj, m->tile_rows);
  // an in-place update (out_new aliasing the advected field) writes a
  // partner array, since neighbouring columns still read the old values;
  // the field and its partner then exchange buffers
  const bool inplace = (out_new != nullptr && out_new == s_in);
  if (inplace) g.out_new = m->thb;
  if (pdef != 1) {
    g.mode = 0;
    pick(false)(m, P, g, cg);
  } else {
    // the positive-definite pass reads the advected values themselves
    // (never the theta offset form: pdef is for moisture and tracers)
    if (sadd != 0.0f) {
      cm1g_set_error("advs: pdef with an offset field is not supported");
      return;
    }
    g.mode = 1;
    pick(true)(m, P, g, cg);
    PdefArgs pa;
    pa.rru = g.rru;
    pa.rrv = g.rrv;
    pa.rrw = g.rrw;
    pa.scale = g.scale;
    pa.advx = m->tx;
    pa.advy = m->ty;
    pa.advz = m->tz;
    pa.s0 = s0;
    pa.s = s_in;
    pa.ax_out = m->sx;
    pa.ay_out = m->sy;
    pa.sten0 = sten0;
    pa.t11 = thterm_div;
    pa.thterm = g.thterm;
    pa.divx = g.divx;
    pa.out_sten = out_sten;
    pa.out_new = g.out_new;
    pa.dt = dt;
    pa.smeps_flush = smeps_flush;
    launch<256>(m, dim3((ni + 63) / 64, (nj + 3) / 4, nk), dim3(64, 4, 1),
                [=] __device__(int bx, int by, int bz) { pdef_xy(P, pa, bx, by, bz); });
    launch<256>(m, col_blocks(cg), col_threads(cg), [=] __device__(int bx, int, int) { pdef_z(P, pa, cg, bx); });
  }
  if (inplace) {
    // the partner now holds the new field: exchange the two buffers
    for (int id = 0; id < CM1G_NFIELDS; id++)
      if (m->f[id] == out_new) {
        m->f[id] = m->thb;
        m->thb = out_new;
        break;
      }
  }
}

// ----------------------------------------------------------- advu/v/w
// The momentum tendencies use the pre-acoustic mass fluxes rru/rrv/rrw
// (model fields) and divx; uten/vten/wten are updated in place, or, with
// from_ten1, set to uten1/vten1/wten1 plus advection (column kernels in
// cm1g_advm.hip).
void adv_uvw(cm1g_model *m, int nrk, int mask, int buoy, int from_ten1, int own, const float *const *src3) {
  advm_run(m, nrk, mask, buoy, from_ten1, own, src3);
}
