// LES subgrid turbulence: the Deardorff TKE scheme with the semi-slip
// surface of the reference's dry convective boundary layer
// (cm1setup = 1, sgsmodel = 1, tconfig = 1, bcturbs = 1, bbc = 3 with
// sfcmodel = 1, set_znt = 1, set_flx = 1; doimpl 0 or 1; dry, doubly
// periodic, no terrain).
//
//   turb_sfc   sfc_and_turb (src/turb.F:12-1700): surface layer (getcecd,
//              sfcflux: src/sfcphys.F:18-110, :313-330), calcnm, calcdef,
//              tkekm, tkebc, gettau with sfcstress, as 7 launches
//   turb_tend  the turbulence tendencies of solve1 (src/solve1.F:315-380,
//              1118-1240): tketen with turbt, turbu, turbv, turbw, turbs
//              for theta and the tracers (src/turbtend.F), one column
//              kernel each (the implicit vertical solves march down and
//              back up the column)
//   tke_stage  the TKE's advection and update of one RK stage
//              (src/solve2.F:1709-1800)
//   tke_copy   tkea = tke3d (src/solve3.F:563-575)
//
// Every expression is the reference's FP32 expression in its order;
// alog and ** are the host library's logf and powf restated (cm1g_libm.h),
// so the fields equal the reference build's bit for bit.  The horizontal
// fluxes that the reference stores in dumx/dumy are recomputed by each
// point that needs them (the same operands in the same order, so the same
// bits), which turns each routine into one pass.
#include "cm1g_internal.h"

namespace {

constexpr float KARMAN = 0.40f, C_M = 0.10f, C_L = 0.8165f, RI_C = 0.25f;
constexpr float TKE_MIN = 1.0e-10f, NM_MIN = 1.0e-6f, SMALL_LEN_FRAC = 0.001f;
constexpr float VIALPHA = 1.0f, VIBETA = 1.0f - VIALPHA;

// the LES constants of every column (src/init3d.F:2672-2680 with
// cm0 = c_m): cme, ce1, ce2, evaluated as the reference evaluates them
struct TkeConsts {
  float cme, ce1, ce2;
};
TkeConsts tke_consts() {
  const float pi = 3.1415926535897932384626433f;
  TkeConsts t;
  t.cme = C_M;
  t.ce1 = t.cme * C_L * C_L * (1.0f / RI_C - 1.0f);
  t.ce2 = fmaxf(0.0f, t.cme * pi * pi - t.ce1);
  return t;
}

// the interior point a periodic halo point repeats (src/bc.F:12-172)
__device__ __forceinline__ int wrap(int i, int n) { return ((i - 1) % n + n) % n + 1; }

// scalar value relative to th0r for theta (src/solve1.F:1184-1206), the
// field itself for a tracer
struct ScalarIn {
  const float *s;
  const float *th0;
  float th0r;
  int theta;
  __device__ __forceinline__ float at(const KP &P, int i, int j, int k) const {
    const float v = s[IXD(P, i, j, k)];
    if (theta) {
      const float tem = th0[k] - th0r;
      return tem + v;
    }
    return v;
  }
};

}  // namespace

// ------------------------------------------------------------ sfc_and_turb
void turb_sfc(cm1g_model *m, int getsfc) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  float **f = m->f;
  const float *ua = f[CM1G_UA], *va = f[CM1G_VA], *wa = f[CM1G_WA], *tha = f[CM1G_THA];
  const float *rho = f[CM1G_RHO], *rf = f[CM1G_RF];
  float *t11 = f[CM1G_T11], *t12 = f[CM1G_T12], *t13 = f[CM1G_T13], *t22 = f[CM1G_T22],
        *t23 = f[CM1G_T23], *t33 = f[CM1G_T33];
  float *nm = f[CM1G_NM], *defv = f[CM1G_DEFV], *defh = f[CM1G_DEFH], *diss = f[CM1G_DISSTEN];
  float *kmh = f[CM1G_KMH], *kmv = f[CM1G_KMV], *khh = f[CM1G_KHH], *khv = f[CM1G_KHV];
  float *tkea = f[CM1G_TKEA], *tke3d = f[CM1G_TKE3D], *lenscl = f[CM1G_LENSCL];
  float *ust = m->d2[CM1G_D_UST], *u1 = m->d2[CM1G_D_U1], *v1 = m->d2[CM1G_D_V1], *s1 = m->d2[CM1G_D_S1],
        *thflux = m->d2[CM1G_D_THFLUX];
  const float *znt = m->d2[CM1G_D_ZNT];
  const float *tcol = m->tcol;
  const float ustlog = m->ustlog, shflx = m->turb.cnst_shflx;
  const TkeConsts T = tke_consts();

  // u1, v1, s1, ust (getcecd, set_znt = 1) and the specified heat flux on
  // the interior, and on the 3-wide periodic halo the values of the
  // interior points they repeat (the bc2d fills of src/turb.F:1138-1242);
  // getsfc = 0 (the first step after a restart, src/cm1.F:3726-3728) keeps
  // the stored ones
  if (getsfc) for2(m, -2, ni + 3, -2, nj + 3, [=] __device__(int i, int j) {
    const int ii = wrap(i, ni), jj = wrap(j, nj);
    const float a = 0.5f * (ua[IXD(P, ii, jj, 1)] + ua[IXD(P, ii + 1, jj, 1)]);
    const float b = 0.5f * (va[IXD(P, ii, jj, 1)] + va[IXD(P, ii, jj + 1, 1)]);
    const float s = sqrtf(a * a + b * b);
    const int c = IX2D(P, i, j);
    u1[c] = a;
    v1[c] = b;
    s1[c] = s;
    ust[c] = fmaxf(s * KARMAN / ustlog, 1.0e-6f);
    if (i >= 1 && i <= ni && j >= 1 && j <= nj) thflux[c] = shflx;
  });
  // the bc2d fills run with getsfc = 0 as well: the stored interior values
  // repeat on the halo
  else for2(m, -2, ni + 3, -2, nj + 3, [=] __device__(int i, int j) {
    if (i >= 1 && i <= ni && j >= 1 && j <= nj) return;
    const int s = IX2D(P, wrap(i, ni), wrap(j, nj)), c = IX2D(P, i, j);
    ust[c] = ust[s];
    u1[c] = u1[s];
    v1[c] = v1[s];
    s1[c] = s1[s];
  });

  // calcdef's rho*S_ij (src/turb.F:4750-4830, no terrain) and calcnm's
  // dry N^2 (:5450-5470); dissten = 0 on the boundary levels
  for3(m, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
    const Consts &K = P.K;
    const int c = IXD(P, i, j, k);
    const float r = rho[c];
    t11[c] = r * (ua[c + 1] - ua[c]) * P.rdx * P.uh[i + 2];
    t22[c] = r * (va[c + P.px] - va[c]) * P.rdy * P.vh[j + 2];
    const float temz = P.rdz * P.mh[k];
    t33[c] = r * (wa[c + P.pxy] - wa[c]) * temz;
    t12[c] = 0.5f * ((ua[c] - ua[c - P.px]) * P.rdy * P.vf[j + 2] + (va[c] - va[c - 1]) * P.rdx * P.uf[i + 2]) *
             0.25f * ((rho[c - 1 - P.px] + r) + (rho[c - 1] + rho[c - P.px]));
    if (k >= 2 && i >= 1 && j >= 1) {
      const float mf2 = P.mf[k] + P.mf[k];
      t13[c] = 0.5f * ((wa[c] - wa[c - 1]) * P.rdx * P.uf[i + 2] + (ua[c] - ua[c - P.pxy]) * P.rdz * 0.5f * mf2) *
               0.5f * (rf[c - 1] + rf[c]);
      t23[c] = 0.5f * ((wa[c] - wa[c - P.px]) * P.rdy * P.vf[j + 2] + (va[c] - va[c - P.pxy]) * P.rdz * 0.5f * mf2) *
               0.5f * (rf[c - P.px] + rf[c]);
    }
    if (i < 1 || i > ni || j < 1 || j > nj) return;
    if (k >= 2)
      nm[c] = cm1libm::logf_host((P.th0[k] + tha[c]) / (P.th0[k - 1] + tha[c - P.pxy])) * K.g * P.rdz * P.mf[k];
    else {
      nm[c] = 0.0f;
      diss[c] = 0.0f;
    }
    if (k == nk) {
      nm[c + P.pxy] = 0.0f;
      diss[c + P.pxy] = 0.0f;
    }
  });

  // deformation (src/turb.F:5340-5380) and tkekm (:3030-3090)
  for3(m, 1, ni, 1, nj, 2, nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k), d = c - P.pxy;
    const int px = P.px;
    auto sq = [](float x) { return x * x; };
    const float c1 = P.c1[k], c2 = P.c2[k];
    const float tmp11 = (c1 * sq(t11[d]) + c2 * sq(t11[c]));
    const float tmp22 = (c1 * sq(t22[d]) + c2 * sq(t22[c]));
    const float tmp33 = (c1 * sq(t33[d]) + c2 * sq(t33[c]));
    const float tmp12 = 0.25f * (c1 * ((sq(t12[d]) + sq(t12[d + 1 + px])) + (sq(t12[d + px]) + sq(t12[d + 1]))) +
                                 c2 * ((sq(t12[c]) + sq(t12[c + 1 + px])) + (sq(t12[c + px]) + sq(t12[c + 1]))));
    const float tmp13 = 0.5f * (sq(t13[c]) + sq(t13[c + 1]));
    const float tmp23 = 0.5f * (sq(t23[c]) + sq(t23[c + px]));
    const float rrf = 1.0f / sq(rf[c]);
    defv[c] = 4.0f * (tmp13 + tmp23) * rrf;
    defh[c] = (2.0f * ((tmp11 + tmp22) + tmp33) + 4.0f * tmp12) * rrf;
    // length scale, dissipation, km and kh
    const float grdscl = tcol[k];
    const float rgrdscl = 1.0f / grdscl;
    const float n2 = nm[c];
    float tk, len;
    if (tkea[c] <= TKE_MIN) {
      tk = TKE_MIN;
      len = SMALL_LEN_FRAC * grdscl;
    } else {
      tk = tkea[c];
      len = grdscl;
      if (n2 > NM_MIN) {
        len = C_L * sqrtf(tk / n2);
        len = fminf(len, grdscl);
        len = fmaxf(len, SMALL_LEN_FRAC * grdscl);
      }
    }
    lenscl[c] = len;
    diss[c] = 0.0f + (T.ce1 + T.ce2 * len * rgrdscl) * tk * sqrtf(tk) / len;
    const float km = T.cme * sqrtf(tk) * len;
    float prinv = 3.00f;
    if (n2 > NM_MIN) prinv = fminf(1.0f + 2.00f * len * rgrdscl, 3.00f);
    kmh[c] = km;
    kmv[c] = km;
    khh[c] = km * prinv;
    khv[c] = km * prinv;
  });
  {
    float *arr[4] = {kmh, kmv, khh, khv};
    bc_fill_n(m, BC_W1, arr, 4);
  }

  // tkebc (src/turb.F:3240-3275): surface tke and viscosity, the bc2d fill
  // of the k = 1 level, kh from the level above; tke3d at the two ends
  for2(m, -2, ni + 3, -2, nj + 3, [=] __device__(int i, int j) {
    const int cw = IX2D(P, wrap(i, ni), wrap(j, nj));
    const float km1 = KARMAN * znt[cw] * ust[cw];
    const int c = IXD(P, i, j, 1);
    tkea[c] = 0.0f;
    kmh[c] = km1;
    if (i < 0 || i > ni + 1 || j < 0 || j > nj + 1) return;
    kmv[c] = km1;
    khv[c] = km1 * (khv[c + P.pxy] / (1.0e-10f + kmv[c + P.pxy]));
    khh[c] = km1 * (khh[c + P.pxy] / (1.0e-10f + kmh[c + P.pxy]));
    tke3d[c] = 0.0f;
    tke3d[c + nk * P.pxy] = tkea[c + nk * P.pxy];
  });

  // gettau (src/turb.F:4090-4160): t_ij = 2 rho K S_ij
  for3(m, 0, ni + 1, 0, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k), u = c + P.pxy, px = P.px;
    const float tem = (kmh[c] + kmh[u]);
    t11[c] = t11[c] * tem;
    t22[c] = t22[c] * tem;
    t33[c] = t33[c] * tem;
    t12[c] = t12[c] * 0.25f *
             (((kmh[c - 1 - px] + kmh[c]) + (kmh[c - 1] + kmh[c - px])) +
              ((kmh[u - 1 - px] + kmh[u]) + (kmh[u - 1] + kmh[u - px])));
    if (k >= 2 && i >= 1 && j >= 1) {
      t13[c] = t13[c] * (kmv[c - 1] + kmv[c]);
      t23[c] = t23[c] * (kmv[c - px] + kmv[c]);
    }
  });
  // sfcstress with bbc = 3 (src/turb.F:4580-4590) and the free-slip top
  // (tbc = 1, :4425-4432)
  for2(m, 1, ni + 1, 1, nj + 1, [=] __device__(int i, int j) {
    const int c2 = IX2D(P, i, j), w2 = c2 - 1, s2 = c2 - P.px;
    const int c = IXD(P, i, j, 1);
    t13[c] = 0.25f * ((ust[w2] * ust[w2]) * (u1[w2] / fmaxf(s1[w2], 0.01f)) +
                      (ust[c2] * ust[c2]) * (u1[c2] / fmaxf(s1[c2], 0.01f))) *
             (rf[c - 1] + rf[c]);
    t23[c] = 0.25f * ((ust[s2] * ust[s2]) * (v1[s2] / fmaxf(s1[s2], 0.01f)) +
                      (ust[c2] * ust[c2]) * (v1[c2] / fmaxf(s1[c2], 0.01f))) *
             (rf[c - P.px] + rf[c]);
    const int top = IXD(P, i, j, nk + 1), below = top - P.pxy;
    t13[top] = t13[below];
    t23[top] = t23[below];
  });
}

// ---------------------------------------------------------- solve1 tendencies
namespace {

// turbt (src/turbtend.F:700-1050) with the TKE equation's own terms
// (src/solve1.F:320-332): tketen for one column, k = 2..nk
struct TurbtArgs {
  const float *t, *rho, *rf, *kmh, *kmv, *khv, *nm, *diss, *defv, *defh;
  float *tten, *cx, *cy;
  int doimpl;
  int rhs;     // 1: tten = the TKE source terms + transport; 0: tten += transport
  float dt;
};
__device__ __forceinline__ float tke_rhs(const TurbtArgs &a, int c) {
  if (!a.rhs) return a.tten[c];
  return -a.khv[c] * a.nm[c] - a.diss[c] + a.kmv[c] * fmaxf(0.0f, (a.defv[c] + a.defh[c]));
}
// the horizontal part at level k: turbx + turby
__device__ __forceinline__ float turbt_h(const KP &P, const TurbtArgs &a, int i, int j, int c) {
  const float *t = a.t, *rf = a.rf, *kmh = a.kmh;
  const int px = P.px;
  auto fx = [&](int cc, int ii) {
    return -0.25f * (rf[cc] + rf[cc - 1]) * 2.0f * (kmh[cc] + kmh[cc - 1]) * (t[cc] - t[cc - 1]) * P.rdx * P.uf[ii + 2];
  };
  auto fy = [&](int cc, int jj) {
    return -0.25f * (rf[cc] + rf[cc - px]) * 2.0f * (kmh[cc] + kmh[cc - px]) * (t[cc] - t[cc - px]) * P.rdy * P.vf[jj + 2];
  };
  const float tx = -(fx(c + 1, i + 1) - fx(c, i)) * P.rdx * P.uh[i + 2];
  const float ty = -(fy(c + px, j + 1) - fy(c, j)) * P.rdy * P.vh[j + 2];
  return tx + ty;
}
__device__ void turbt_col(const KP &P, const TurbtArgs &a, int i, int j) {
  const int nk = P.nk, pxy = P.pxy;
  const float *t = a.t, *rho = a.rho, *rf = a.rf, *kmv = a.kmv;
  const float rdz = P.rdz;
  if (!a.doimpl) {
    // dumz(k) = -(kmv(k)+kmv(k+1))*(t(k+1)-t(k))*rdz*mh(k)*rho(k); dumz(1) = 0 (bbc = 3)
    auto dz = [&](int k) {
      const int c = IXD(P, i, j, k);
      return -(kmv[c] + kmv[c + pxy]) * (t[c + pxy] - t[c]) * rdz * P.mh[k] * rho[c];
    };
    float lo = 0.0f;
    for (int k = 2; k <= nk; k++) {
      const int c = IXD(P, i, j, k);
      const float hi = dz(k);
      const float tz = -(hi - lo) * rdz * P.mf[k];
      a.tten[c] = tke_rhs(a, c) + (turbt_h(P, a, i, j, c) + tz) / rf[c];
      lo = hi;
    }
    return;
  }
  const float rdt = 1.0f / a.dt;
  const float tema = -1.0f * a.dt * VIALPHA * rdz * rdz;
  const float temb = a.dt * VIBETA * rdz * rdz;
  float *cx = a.cx, *cy = a.cy;
  // k = 2 .. nk-1 forward
  for (int k = 2; k <= nk - 1; k++) {
    const int c = IXD(P, i, j, k);
    const float rrf = P.mf[k] / rf[c];
    const float r2 = (kmv[c] + kmv[c + pxy]) * P.mh[k] * rho[c] * rrf;
    const float cfc = tema * r2;
    if (k == 2) {
      const float cfb = 1.0f - cfc;
      const float cfd = t[c] + temb * (r2 * t[c + pxy] - r2 * t[c]);
      const float tem = 1.0f / cfb;
      cx[c] = -cfc * tem;
      cy[c] = cfd * tem;
    } else {
      const float r1 = (kmv[c - pxy] + kmv[c]) * P.mh[k - 1] * rho[c - pxy] * rrf;
      const float cfa = tema * r1;
      const float cfb = 1.0f - cfa - cfc;
      const float cfd = t[c] + temb * (r2 * t[c + pxy] - (r1 + r2) * t[c] + r1 * t[c - pxy]);
      const float tem = 1.0f / (cfa * cx[c - pxy] + cfb);
      cx[c] = -cfc * tem;
      cy[c] = (cfd - cfa * cy[c - pxy]) * tem;
    }
  }
  float up;
  {
    const int k = nk, c = IXD(P, i, j, k);
    const float rrf = P.mf[k] / rf[c];
    const float r1 = (kmv[c - pxy] + kmv[c]) * P.mh[k - 1] * rho[c - pxy] * rrf;
    const float cfa = tema * r1;
    const float cfb = 1.0f - cfa;
    const float cfd = t[c] + temb * (-r1 * t[c] + r1 * t[c - pxy]);
    const float tem = 1.0f / (cfa * cx[c - pxy] + cfb);
    up = (cfd - cfa * cy[c - pxy]) * tem;
    const float tz = rf[c] * (up - t[c]) * rdt;
    a.tten[c] = tke_rhs(a, c) + (turbt_h(P, a, i, j, c) + tz) / rf[c];
  }
  for (int k = nk - 1; k >= 2; k--) {
    const int c = IXD(P, i, j, k);
    up = cx[c] * up + cy[c];
    const float tz = rf[c] * (up - t[c]) * rdt;
    a.tten[c] = tke_rhs(a, c) + (turbt_h(P, a, i, j, c) + tz) / rf[c];
  }
}

// turbu with turbuz (src/turbtend.F:1052-1816), one u column (i = 1..ni+1)
struct TurbuvArgs {
  const float *vel, *w, *rho, *rf, *kmv;
  const float *ta, *tb, *tz;   // u: t11, t12, t13; v: t12, t22, t23
  float *ten, *d1, *d2, *d7;
  int doimpl;
  float dt;
};
template <int V>
__device__ void turbuv_col(const KP &P, const TurbuvArgs &a, int i, int j) {
  const int nk = P.nk, pxy = P.pxy;
  // the neighbour across the staggered point: i-1 for u, j-1 for v
  const int o = V ? P.px : 1;
  const float *rho = a.rho, *rf = a.rf, *kmv = a.kmv, *u = a.vel, *w = a.w;
  const float rdz = P.rdz;
  auto horiz = [&](int c) {
    if (V == 0) {
      const float tx = (a.ta[c] - a.ta[c - 1]) * P.rdx * P.uf[i + 2];
      const float ty = (a.tb[c + P.px] - a.tb[c]) * P.rdy * P.vh[j + 2];
      return tx + ty;
    } else {
      const float tx = (a.ta[c + 1] - a.ta[c]) * P.rdx * P.uh[i + 2];
      const float ty = (a.tb[c] - a.tb[c - P.px]) * P.rdy * P.vf[j + 2];
      return tx + ty;
    }
  };
  auto add = [&](int c, float tz) {
    const float r0 = 1.0f / (0.5f * (rho[c - o] + rho[c]));
    a.ten[c] = a.ten[c] + (horiz(c) + tz) * r0;
  };
  if (!a.doimpl) {
    const float tem = rdz * 0.5f;
    for (int k = 1; k <= nk; k++) {
      const int c = IXD(P, i, j, k);
      add(c, (a.tz[c + pxy] - a.tz[c]) * tem * (P.mh[k] + P.mh[k]));
    }
    return;
  }
  const float rdt = 0.5f / a.dt;
  const float tema = -0.0625f * a.dt * VIALPHA * rdz * rdz;
  const float temb = 0.0625f * a.dt * VIBETA * rdz * rdz;
  const float temc = 0.5f * a.dt * rdz;
  // the explicit dw/dx (u) or dw/dy (v) piece at level k
  auto d7f = [&](int c) {
    const float g = V ? (w[c] - w[c - o]) * P.rdy * P.vf[j + 2] : (w[c] - w[c - o]) * P.rdx * P.uf[i + 2];
    return g * 0.25f * (kmv[c - o] + kmv[c]) * (rf[c - o] + rf[c]);
  };
  float *d1 = a.d1, *d2 = a.d2, *d7 = a.d7;
  for (int k = 1; k <= nk - 1; k++) {
    const int c = IXD(P, i, j, k);
    const float r0 = 1.0f / (0.5f * (rho[c - o] + rho[c]));
    float tem = (P.mh[k] + P.mh[k]) * r0;
    const float r2 = (kmv[c - o + pxy] + kmv[c + pxy]) * (P.mf[k + 1] + P.mf[k + 1]) *
                     (rf[c - o + pxy] + rf[c + pxy]) * tem;
    const float cfc = tema * r2;
    if (k == 1) {
      const float cfb = 1.0f - cfc;
      const float cfd = u[c] + temb * (r2 * u[c + pxy] - r2 * u[c]) - temc * a.tz[c] * (P.mh[1] + P.mh[1]) * r0;
      tem = 1.0f / cfb;
      d1[c] = -cfc * tem;
      d2[c] = cfd * tem;
      d7[c] = 0.0f;
    } else {
      const float r1 = (kmv[c - o] + kmv[c]) * (P.mf[k] + P.mf[k]) * (rf[c - o] + rf[c]) * tem;
      const float cfa = tema * r1;
      const float cfb = 1.0f - cfa - cfc;
      const float cfd = u[c] + temb * (r2 * u[c + pxy] - (r1 + r2) * u[c] + r1 * u[c - pxy]);
      tem = 1.0f / (cfa * d1[c - pxy] + cfb);
      d1[c] = -cfc * tem;
      d2[c] = (cfd - cfa * d2[c - pxy]) * tem;
      d7[c] = d7f(c);
    }
  }
  float up, d7up;
  {
    const int k = nk, c = IXD(P, i, j, k);
    const float r0 = 1.0f / (0.5f * (rho[c - o] + rho[c]));
    float tem = (P.mh[k] + P.mh[k]) * r0;
    const float r1 = (kmv[c - o] + kmv[c]) * (P.mf[k] + P.mf[k]) * (rf[c - o] + rf[c]) * tem;
    const float cfa = tema * r1;
    const float cfb = 1.0f - cfa;
    const float cfd = u[c] + temb * (-r1 * u[c] + r1 * u[c - pxy]) + temc * a.tz[c + pxy] * (P.mh[nk] + P.mh[nk]) * r0;
    tem = 1.0f / (cfa * d1[c - pxy] + cfb);
    up = (cfd - cfa * d2[c - pxy]) * tem;
    float tz = (rho[c - o] + rho[c]) * (up - u[c]) * rdt;
    d7up = d7f(c);
    tz = tz + (0.0f - d7up) * rdz * 0.5f * (P.mh[k] + P.mh[k]);
    add(c, tz);
  }
  for (int k = nk - 1; k >= 1; k--) {
    const int c = IXD(P, i, j, k);
    up = d1[c] * up + d2[c];
    float tz = (rho[c - o] + rho[c]) * (up - u[c]) * rdt;
    const float d7k = d7[c];
    tz = tz + (d7up - d7k) * rdz * 0.5f * (P.mh[k] + P.mh[k]);
    add(c, tz);
    d7up = d7k;
  }
}

// turbw (src/turbtend.F:2473-2892), one w column, k = 2..nk
struct TurbwArgs {
  const float *w, *rho, *rf, *kmh, *t13, *t23, *t33;
  float *ten, *d1, *d2;
  int doimpl;
  float dt;
};
__device__ void turbw_col(const KP &P, const TurbwArgs &a, int i, int j) {
  const int nk = P.nk, pxy = P.pxy, px = P.px;
  const float *w = a.w, *rho = a.rho, *rf = a.rf, *kmh = a.kmh, *t33 = a.t33;
  const float rdz = P.rdz;
  auto add = [&](int c, float tz) {
    const float tx = (a.t13[c + 1] - a.t13[c]) * P.rdx * P.uh[i + 2];
    const float ty = (a.t23[c + px] - a.t23[c]) * P.rdy * P.vh[j + 2];
    a.ten[c] = a.ten[c] + ((tx + ty) + tz) / rf[c];
  };
  if (!a.doimpl) {
    for (int k = 2; k <= nk; k++) {
      const int c = IXD(P, i, j, k);
      add(c, (t33[c] - t33[c - pxy]) * rdz * P.mf[k]);
    }
    return;
  }
  const float rdt = 1.0f / a.dt;
  const float tema = -1.0f * a.dt * VIALPHA * rdz * rdz;
  const float temb = a.dt * VIBETA * rdz * rdz;
  const float temc = a.dt * rdz;
  float *d1 = a.d1, *d2 = a.d2;
  for (int k = 2; k <= nk - 1; k++) {
    const int c = IXD(P, i, j, k);
    const float rrf = P.mf[k] / rf[c];
    const float r2 = (kmh[c] + kmh[c + pxy]) * P.mh[k] * rho[c] * rrf;
    const float cfc = tema * r2;
    if (k == 2) {
      const float cfb = 1.0f - cfc;
      const float cfd = w[c] + temb * (r2 * w[c + pxy] - r2 * w[c]) - temc * t33[c - pxy] * rrf;
      const float tem = 1.0f / cfb;
      d1[c] = -cfc * tem;
      d2[c] = cfd * tem;
    } else {
      const float r1 = (kmh[c - pxy] + kmh[c]) * P.mh[k - 1] * rho[c - pxy] * rrf;
      const float cfa = tema * r1;
      const float cfb = 1.0f - cfa - cfc;
      const float cfd = w[c] + temb * (r2 * w[c + pxy] - (r1 + r2) * w[c] + r1 * w[c - pxy]);
      const float tem = 1.0f / (cfa * d1[c - pxy] + cfb);
      d1[c] = -cfc * tem;
      d2[c] = (cfd - cfa * d2[c - pxy]) * tem;
    }
  }
  float up;
  {
    const int k = nk, c = IXD(P, i, j, k);
    const float rrf = P.mf[k] / rf[c];
    const float r1 = (kmh[c - pxy] + kmh[c]) * P.mh[k - 1] * rho[c - pxy] * rrf;
    const float cfa = tema * r1;
    const float cfb = 1.0f - cfa;
    const float cfd = w[c] + temb * (-r1 * w[c] + r1 * w[c - pxy]) + temc * t33[c] * rrf;
    const float tem = 1.0f / (cfa * d1[c - pxy] + cfb);
    up = (cfd - cfa * d2[c - pxy]) * tem;
    add(c, rf[c] * (up - w[c]) * rdt);
  }
  for (int k = nk - 1; k >= 2; k--) {
    const int c = IXD(P, i, j, k);
    up = d1[c] * up + d2[c];
    add(c, rf[c] * (up - w[c]) * rdt);
  }
}

// turbs with turbsz (src/turbtend.F:18-698), one scalar column
struct TurbsArgs {
  ScalarIn s;
  const float *rho, *rr, *rf, *khh, *khv, *sflux;
  float *sten, *d1, *d2;
  int iflux, dosfcflx, doimpl;
  float dt;
};
__device__ __forceinline__ float turbs_h(const KP &P, const TurbsArgs &a, int i, int j, int k) {
  const int c = IXD(P, i, j, k), px = P.px, pxy = P.pxy;
  const float *rho = a.rho, *khh = a.khh;
  auto fx = [&](int cc, int ii) {
    return -0.125f * (rho[cc] + rho[cc - 1]) * ((khh[cc] + khh[cc - 1]) + (khh[cc + pxy] + khh[cc - 1 + pxy])) *
           (a.s.at(P, ii, j, k) - a.s.at(P, ii - 1, j, k)) * P.rdx * P.uf[ii + 2];
  };
  auto fy = [&](int cc, int jj) {
    return -0.125f * (rho[cc] + rho[cc - px]) * ((khh[cc] + khh[cc - px]) + (khh[cc + pxy] + khh[cc - px + pxy])) *
           (a.s.at(P, i, jj, k) - a.s.at(P, i, jj - 1, k)) * P.rdy * P.vf[jj + 2];
  };
  const float tx = -(fx(c + 1, i + 1) - fx(c, i)) * P.rdx * P.uh[i + 2];
  const float ty = -(fy(c + px, j + 1) - fy(c, j)) * P.rdy * P.vh[j + 2];
  return tx + ty;
}
__device__ void turbs_col(const KP &P, const TurbsArgs &a, int i, int j) {
  const int nk = P.nk, pxy = P.pxy;
  const float *rho = a.rho, *rr = a.rr, *rf = a.rf, *khv = a.khv;
  const float rdz = P.rdz;
  const int c1 = IXD(P, i, j, 1);
  const float bot = (a.iflux == 1 && a.dosfcflx) ? a.sflux[IX2D(P, i, j)] * rf[c1] : 0.0f;
  auto add = [&](int k, float tz) {
    const int c = IXD(P, i, j, k);
    a.sten[c] = a.sten[c] + ((turbs_h(P, a, i, j, k)) + tz) * rr[c];
  };
  if (!a.doimpl) {
    // dumz(k) = -khv(k)*(s(k)-s(k-1))*rdz*mf(k)*rf(k), k = 2..nk; dumz(1)
    // the surface flux (or 0), dumz(nk+1) = 0
    float lo = bot;
    for (int k = 1; k <= nk; k++) {
      float hi = 0.0f;
      if (k < nk) {
        const int cu = IXD(P, i, j, k + 1);
        hi = -khv[cu] * (a.s.at(P, i, j, k + 1) - a.s.at(P, i, j, k)) * rdz * P.mf[k + 1] * rf[cu];
      }
      add(k, -(hi - lo) * rdz * P.mh[k]);
      lo = hi;
    }
    return;
  }
  const float rdt = 1.0f / a.dt;
  const float tema = -1.0f * a.dt * VIALPHA * rdz * rdz;
  const float temb = a.dt * VIBETA * rdz * rdz;
  const float temc = a.dt * rdz;
  float *d1 = a.d1, *d2 = a.d2;
  // dum7(k) = khv(k) mf(k) rf(k) mh(k) rr(k), dum8(k) = khv(k+1) mf(k+1) rf(k+1) mh(k) rr(k)
  auto r7 = [&](int k) {
    const int c = IXD(P, i, j, k);
    return khv[c] * P.mf[k] * rf[c] * P.mh[k] * rr[c];
  };
  auto r8 = [&](int k) {
    const int c = IXD(P, i, j, k);
    return khv[c + pxy] * P.mf[k + 1] * rf[c + pxy] * P.mh[k] * rr[c];
  };
  for (int k = 1; k <= nk - 1; k++) {
    const int c = IXD(P, i, j, k);
    const float sk = a.s.at(P, i, j, k), su = a.s.at(P, i, j, k + 1);
    const float r2 = r8(k);
    const float cfc = tema * r2;
    if (k == 1) {
      const float cfb = 1.0f - cfc;
      const float cfd = sk + temb * (-r2 * sk + r2 * su) + temc * bot * P.mh[1] * rr[c];
      const float tem = 1.0f / cfb;
      d1[c] = -cfc * tem;
      d2[c] = cfd * tem;
    } else {
      const float r1 = r7(k);
      const float cfa = tema * r1;
      const float cfb = 1.0f - cfa - cfc;
      const float cfd = sk + temb * (r1 * a.s.at(P, i, j, k - 1) - (r1 + r2) * sk + r2 * su);
      const float tem = 1.0f / (cfa * d1[c - pxy] + cfb);
      d1[c] = -cfc * tem;
      d2[c] = (cfd - cfa * d2[c - pxy]) * tem;
    }
  }
  float up;
  {
    const int k = nk, c = IXD(P, i, j, k);
    const float sk = a.s.at(P, i, j, k);
    const float r1 = r7(k);
    const float cfa = tema * r1;
    const float cfb = 1.0f - cfa;
    // dumz(nk+1) = 0 (bcturbs = 1)
    const float cfd = sk + temb * (r1 * a.s.at(P, i, j, k - 1) - r1 * sk) - temc * 0.0f * P.mh[nk] * rr[c];
    const float tem = 1.0f / (cfa * d1[c - pxy] + cfb);
    up = (cfd - cfa * d2[c - pxy]) * tem;
    add(k, rho[c] * (up - sk) * rdt);
  }
  for (int k = nk - 1; k >= 1; k--) {
    const int c = IXD(P, i, j, k);
    up = d1[c] * up + d2[c];
    add(k, rho[c] * (up - a.s.at(P, i, j, k)) * rdt);
  }
}

// turbs of up to 8 scalars of one column in one pass: the solver's
// forward coefficient d1 depends only on khv, rf and rr, so it is formed
// once; each scalar keeps its own recursion (d2, the back-substituted
// value) in registers and its own d2 column in scratch.  Every scalar's
// arithmetic is turbs_col's, in the same order.
constexpr int TURBS_MAX = 8;
struct TurbsMultiArgs {
  const float *s[TURBS_MAX];
  float *sten[TURBS_MAX], *d2[TURBS_MAX];
  const float *th0;
  float th0r;
  int theta0;   // scalar 0 is theta, taken relative to th0r
  int flux0;    // scalar 0 takes the surface flux
  const float *rho, *rr, *rf, *khh, *khv, *sflux;
  float *d1;
  int doimpl;
  float dt;
};
template <int NS>
__device__ void turbs_multi_col(const KP &P, const TurbsMultiArgs &a, int i, int j) {
  const int nk = P.nk, pxy = P.pxy, px = P.px;
  const float *rho = a.rho, *rr = a.rr, *rf = a.rf, *khh = a.khh, *khv = a.khv;
  const float rdz = P.rdz;
  auto sv = [&](int n, int c, int k) {
    const float v = a.s[n][c];
    if (n == 0 && a.theta0) {
      const float tem = a.th0[k] - a.th0r;
      return tem + v;
    }
    return v;
  };
  const int c1 = IXD(P, i, j, 1);
  const float bot0 = a.flux0 ? a.sflux[IX2D(P, i, j)] * rf[c1] : 0.0f;
  // the horizontal terms of every scalar at level k (turbs_h), their shared
  // coefficient parts formed once: fx = ((q * ds) * rdx) * uf
  auto horiz = [&](int k, float (&h)[NS]) {
    const int c = IXD(P, i, j, k);
    auto qx = [&](int cc) {
      return -0.125f * (rho[cc] + rho[cc - 1]) * ((khh[cc] + khh[cc - 1]) + (khh[cc + pxy] + khh[cc - 1 + pxy]));
    };
    auto qy = [&](int cc) {
      return -0.125f * (rho[cc] + rho[cc - px]) * ((khh[cc] + khh[cc - px]) + (khh[cc + pxy] + khh[cc - px + pxy]));
    };
    const float qxw = qx(c), qxe = qx(c + 1), qys = qy(c), qyn = qy(c + px);
#pragma unroll
    for (int n = 0; n < NS; n++) {
      const float sc = sv(n, c, k);
      const float fxw = qxw * (sc - sv(n, c - 1, k)) * P.rdx * P.uf[i + 2];
      const float fxe = qxe * (sv(n, c + 1, k) - sc) * P.rdx * P.uf[i + 3];
      const float fys = qys * (sc - sv(n, c - px, k)) * P.rdy * P.vf[j + 2];
      const float fyn = qyn * (sv(n, c + px, k) - sc) * P.rdy * P.vf[j + 3];
      const float tx = -(fxe - fxw) * P.rdx * P.uh[i + 2];
      const float ty = -(fyn - fys) * P.rdy * P.vh[j + 2];
      h[n] = tx + ty;
    }
  };
  auto add = [&](int k, const float (&tz)[NS]) {
    const int c = IXD(P, i, j, k);
    float h[NS];
    horiz(k, h);
    const float r = rr[c];
#pragma unroll
    for (int n = 0; n < NS; n++) a.sten[n][c] = a.sten[n][c] + (h[n] + tz[n]) * r;
  };
  if (!a.doimpl) {
    float lo[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) lo[n] = (n == 0) ? bot0 : 0.0f;
    for (int k = 1; k <= nk; k++) {
      float tz[NS];
      const int c = IXD(P, i, j, k), cu = c + pxy;
      const float kv = (k < nk) ? khv[cu] : 0.0f, rfu = (k < nk) ? rf[cu] : 0.0f;
#pragma unroll
      for (int n = 0; n < NS; n++) {
        float hi = 0.0f;
        if (k < nk) hi = -kv * (sv(n, cu, k + 1) - sv(n, c, k)) * rdz * P.mf[k + 1] * rfu;
        tz[n] = -(hi - lo[n]) * rdz * P.mh[k];
        lo[n] = hi;
      }
      add(k, tz);
    }
    return;
  }
  const float rdt = 1.0f / a.dt;
  const float tema = -1.0f * a.dt * VIALPHA * rdz * rdz;
  const float temb = a.dt * VIBETA * rdz * rdz;
  const float temc = a.dt * rdz;
  float *d1 = a.d1;
  float d1p = 0.0f, d2p[NS], sm[NS], s0[NS], sp[NS];
  // sm, s0, sp: the scalar at k-1, k, k+1
#pragma unroll
  for (int n = 0; n < NS; n++) {
    s0[n] = sv(n, c1, 1);
    sm[n] = 0.0f;
  }
  for (int k = 1; k <= nk - 1; k++) {
    const int c = IXD(P, i, j, k);
    const float rrc = rr[c];
    // r2 = khv(k+1) mf(k+1) rf(k+1) mh(k) rr(k), r1 = khv(k) mf(k) rf(k) mh(k) rr(k)
    const float r2 = khv[c + pxy] * P.mf[k + 1] * rf[c + pxy] * P.mh[k] * rrc;
    const float cfc = tema * r2;
#pragma unroll
    for (int n = 0; n < NS; n++) sp[n] = sv(n, c + pxy, k + 1);
    if (k == 1) {
      const float cfb = 1.0f - cfc;
      const float tem = 1.0f / cfb;
      d1p = -cfc * tem;
#pragma unroll
      for (int n = 0; n < NS; n++) {
        const float bot = (n == 0) ? bot0 : 0.0f;
        const float cfd = s0[n] + temb * (-r2 * s0[n] + r2 * sp[n]) + temc * bot * P.mh[1] * rrc;
        d2p[n] = cfd * tem;
      }
    } else {
      const float r1 = khv[c] * P.mf[k] * rf[c] * P.mh[k] * rrc;
      const float cfa = tema * r1;
      const float cfb = 1.0f - cfa - cfc;
      const float tem = 1.0f / (cfa * d1p + cfb);
      d1p = -cfc * tem;
#pragma unroll
      for (int n = 0; n < NS; n++) {
        const float cfd = s0[n] + temb * (r1 * sm[n] - (r1 + r2) * s0[n] + r2 * sp[n]);
        d2p[n] = (cfd - cfa * d2p[n]) * tem;
      }
    }
    d1[c] = d1p;
#pragma unroll
    for (int n = 0; n < NS; n++) {
      a.d2[n][c] = d2p[n];
      sm[n] = s0[n];
      s0[n] = sp[n];
    }
  }
  float up[NS];
  {
    const int k = nk, c = IXD(P, i, j, k);
    const float rrc = rr[c];
    const float r1 = khv[c] * P.mf[k] * rf[c] * P.mh[k] * rrc;
    const float cfa = tema * r1;
    const float cfb = 1.0f - cfa;
    const float tem = 1.0f / (cfa * d1p + cfb);
    const float rh = rho[c];
    float tz[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
      // the top flux is 0 (bcturbs = 1)
      const float cfd = s0[n] + temb * (r1 * sm[n] - r1 * s0[n]) - temc * 0.0f * P.mh[nk] * rrc;
      up[n] = (cfd - cfa * d2p[n]) * tem;
      tz[n] = rh * (up[n] - s0[n]) * rdt;
    }
    add(k, tz);
  }
  for (int k = nk - 1; k >= 1; k--) {
    const int c = IXD(P, i, j, k);
    const float dd = d1[c], rh = rho[c];
    float tz[NS];
#pragma unroll
    for (int n = 0; n < NS; n++) {
      up[n] = dd * up[n] + a.d2[n][c];
      tz[n] = rh * (up[n] - sv(n, c, k)) * rdt;
    }
    add(k, tz);
  }
}

}  // namespace

namespace {

// turbt; rhs = 1 also forms the TKE's source terms (src/solve1.F:320-332)
void run_turbt(cm1g_model *m, int rhs) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj;
  float **f = m->f;
  const int doimpl = m->turb.doimpl;
  const float dt = m->c.dt;
  {
    TurbtArgs a;
    a.t = f[CM1G_TKEA];
    a.rho = f[CM1G_RHO];
    a.rf = f[CM1G_RF];
    a.kmh = f[CM1G_KMH];
    a.kmv = f[CM1G_KMV];
    a.khv = f[CM1G_KHV];
    a.nm = f[CM1G_NM];
    a.diss = f[CM1G_DISSTEN];
    a.defv = f[CM1G_DEFV];
    a.defh = f[CM1G_DEFH];
    a.tten = f[CM1G_TKETEN];
    a.cx = f[CM1G_DUM1];
    a.cy = f[CM1G_DUM2];
    a.doimpl = doimpl;
    a.rhs = rhs;
    a.dt = dt;
    for2(m, 1, ni, 1, nj, [=] __device__(int i, int j) { turbt_col(P, a, i, j); });
  }
}

// turbu (mask 1), turbv (2), turbw (4) added to uten1, vten1, wten1
void run_turbuvw(cm1g_model *m, int mask) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj;
  float **f = m->f;
  const int doimpl = m->turb.doimpl;
  const float dt = m->c.dt;
  {
    TurbuvArgs a;
    a.w = f[CM1G_WA];
    a.rho = f[CM1G_RHO];
    a.rf = f[CM1G_RF];
    a.kmv = f[CM1G_KMV];
    a.d1 = f[CM1G_DUM1];
    a.d2 = f[CM1G_DUM2];
    a.d7 = f[CM1G_DUM7];
    a.doimpl = doimpl;
    a.dt = dt;
    a.vel = f[CM1G_UA];
    a.ta = f[CM1G_T11];
    a.tb = f[CM1G_T12];
    a.tz = f[CM1G_T13];
    a.ten = f[CM1G_UTEN1];
    if (mask & 1) for2(m, 1, ni + 1, 1, nj, [=] __device__(int i, int j) { turbuv_col<0>(P, a, i, j); });
    a.vel = f[CM1G_VA];
    a.ta = f[CM1G_T12];
    a.tb = f[CM1G_T22];
    a.tz = f[CM1G_T23];
    a.ten = f[CM1G_VTEN1];
    if (mask & 2) for2(m, 1, ni, 1, nj + 1, [=] __device__(int i, int j) { turbuv_col<1>(P, a, i, j); });
  }
  if (mask & 4) {
    TurbwArgs a;
    a.w = f[CM1G_WA];
    a.rho = f[CM1G_RHO];
    a.rf = f[CM1G_RF];
    a.kmh = f[CM1G_KMH];
    a.t13 = f[CM1G_T13];
    a.t23 = f[CM1G_T23];
    a.t33 = f[CM1G_T33];
    a.ten = f[CM1G_WTEN1];
    a.d1 = f[CM1G_DUM1];
    a.d2 = f[CM1G_DUM2];
    a.doimpl = doimpl;
    a.dt = dt;
    for2(m, 1, ni, 1, nj, [=] __device__(int i, int j) { turbw_col(P, a, i, j); });
  }
}

// turbs of s into sten; theta: s is tha, taken relative to th0r as solve1's
// sadv (src/solve1.F:1184-1206)
void run_turbs(cm1g_model *m, const float *s, int theta, float *sten, int iflux, int dosfcflx) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj;
  float **f = m->f;
  {
    TurbsArgs a;
    a.s.s = s;
    a.s.th0 = P.th0;
    a.s.th0r = P.K.th0r;
    a.s.theta = theta;
    a.rho = f[CM1G_RHO];
    a.rr = f[CM1G_RR];
    a.rf = f[CM1G_RF];
    a.khh = f[CM1G_KHH];
    a.khv = f[CM1G_KHV];
    a.sflux = m->d2[CM1G_D_THFLUX];
    a.sten = sten;
    a.d1 = f[CM1G_DUM1];
    a.d2 = f[CM1G_DUM2];
    a.iflux = iflux;
    a.dosfcflx = dosfcflx;
    a.doimpl = m->turb.doimpl;
    a.dt = m->c.dt;
    for2(m, 1, ni, 1, nj, [=] __device__(int i, int j) { turbs_col(P, a, i, j); });
  }
}

template <int NS>
void turbs_multi_launch(cm1g_model *m, const TurbsMultiArgs &a) {
  const KP P = m->P;
  for2(m, 1, P.ni, 1, P.nj, [=] __device__(int i, int j) { turbs_multi_col<NS>(P, a, i, j); });
}

// turbs of ns scalars (ns <= TURBS_MAX) in one launch: s[n] into sten[n];
// scalar 0 is theta (relative to th0r, with the surface flux when
// dosfcflx) when theta0; d1 in dum1, the scalars' d2 in dum2..dum9
void run_turbs_multi(cm1g_model *m, int ns, const float *const *s, float *const *sten, int theta0,
                     int dosfcflx) {
  const KP P = m->P;
  float **f = m->f;
  TurbsMultiArgs a = {};
  for (int n = 0; n < ns; n++) {
    a.s[n] = s[n];
    a.sten[n] = sten[n];
    a.d2[n] = f[CM1G_DUM2 + n];
  }
  a.th0 = P.th0;
  a.th0r = P.K.th0r;
  a.theta0 = theta0;
  a.flux0 = theta0 && dosfcflx;
  a.rho = f[CM1G_RHO];
  a.rr = f[CM1G_RR];
  a.rf = f[CM1G_RF];
  a.khh = f[CM1G_KHH];
  a.khv = f[CM1G_KHV];
  a.sflux = m->d2[CM1G_D_THFLUX];
  a.d1 = f[CM1G_DUM1];
  a.doimpl = m->turb.doimpl;
  a.dt = m->c.dt;
  switch (ns) {
    case 1: turbs_multi_launch<1>(m, a); break;
    case 2: turbs_multi_launch<2>(m, a); break;
    case 3: turbs_multi_launch<3>(m, a); break;
    case 4: turbs_multi_launch<4>(m, a); break;
    case 5: turbs_multi_launch<5>(m, a); break;
    case 6: turbs_multi_launch<6>(m, a); break;
    case 7: turbs_multi_launch<7>(m, a); break;
    case 8: turbs_multi_launch<8>(m, a); break;
    default: break;
  }
}

}  // namespace

void turb_tend(cm1g_model *m) {
  float **f = m->f;
  run_turbt(m, 1);
  run_turbuvw(m, 7);
  // theta (iflux = 1: the surface flux) and the tracers (iflux = 0), up
  // to eight scalars a launch
  const int dosfc = (m->turb.isfcflx == 1);
  const float *s[1 + 8];
  float *sten[1 + 8];
  int ns = 0;
  s[ns] = f[CM1G_THA];
  sten[ns++] = f[CM1G_THTEN1];
  for (int n = 0; n < m->c.npt; n++) {
    s[ns] = f[CM1G_PTA0 + n];
    sten[ns++] = f[CM1G_PTTEN0 + n];
  }
  for (int n0 = 0; n0 < ns; n0 += TURBS_MAX) {
    const int cnt = (ns - n0 < TURBS_MAX) ? ns - n0 : TURBS_MAX;
    run_turbs_multi(m, cnt, s + n0, sten + n0, n0 == 0, dosfc);
  }
}

// --------------------------------------------------------- TKE of a stage
void tke_stage(cm1g_model *m, int nrk, float dttmp) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  float **f = m->f;
  float *wten = f[CM1G_WTEN], *tke3d = f[CM1G_TKE3D];
  const float *tkea = f[CM1G_TKEA];
  // the reference advects the TKE with the mass fluxes getdivx has scaled in
  // place (src/misclibs.F:246-273); the stage keeps them unscaled (the
  // scalar advection scales as it reads), so the scaled values go to
  // dum3..dum5 over the same ranges
  const float *rru = f[CM1G_RRU], *rrv = f[CM1G_RRV], *rrw = f[CM1G_RRW];
  float *sru = f[CM1G_DUM3], *srv = f[CM1G_DUM4], *srw = f[CM1G_DUM5];
  for3(m, 1, ni + 1, 1, nj + 1, 1, nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k);
    if (j <= nj) sru[c] = rru[c] * P.rho0[k];
    if (i <= ni) srv[c] = rrv[c] * P.rho0[k];
    if (i <= ni && j <= nj) {
      if (k == 1) {
        srw[c] = 0.0f;
        srw[IXD(P, i, j, nk + 1)] = 0.0f;
      } else {
        srw[c] = rrw[c] * P.rf0[k];
      }
    }
  });
  // wten = tketen + advection of tke3d (advw with the scalars' WENO switch)
  advw_tke(m, nrk, m->c.advwenos, sru, srv, srw, tke3d, f[CM1G_TKETEN], wten);
  for3(m, 1, ni, 1, nj, 2, nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k);
    float t = tkea[c] + dttmp * wten[c];
    if (t < 1.0e-10f) t = 0.0f;
    tke3d[c] = t;
  });
  bc_fill(m, BC_W1, tke3d);
}

void tke_copy(cm1g_model *m) {
  const KP P = m->P;
  const float *tke3d = m->f[CM1G_TKE3D];
  float *tkea = m->f[CM1G_TKEA];
  for3(m, 0, P.ni + 1, 0, P.nj + 1, 1, P.nk, [=] __device__(int i, int j, int k) {
    const int c = IXD(P, i, j, k);
    tkea[c] = tke3d[c];
  });
}

// getset's TKE (src/init3d.F:2160, 2250-2266, 2322-2331)
int tke_init(cm1g_model *m) {
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj;
  float *tkea = m->f[CM1G_TKEA];
  bc_fill(m, BC_W1, tkea);
  for2(m, -2, ni + 3, -2, nj + 3, [=] __device__(int i, int j) {
    tkea[IXD(P, i, j, 1)] = tkea[IXD(P, wrap(i, ni), wrap(j, nj), 2)];
  });
  HIPCHK(hipMemcpyAsync(m->f[CM1G_TKE3D], tkea, m->ntot * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
  return check_launch();
}

// --------------------------------------------------------------- entries
extern "C" int cm1g_set_turbulence(cm1g_model *m, const cm1g_turb_config *t) {
  if (!m || !t) {
    cm1g_set_error("cm1g_set_turbulence: null argument");
    return CM1G_E_ARG;
  }
  if (t->cm1setup == 0) {
    m->turb = *t;
    graphs_clear(m);
    return CM1G_OK;
  }
  const bool periodic = m->c.wbc == 1 && m->c.ebc == 1 && m->c.sbc == 1 && m->c.nbc == 1;
  if (t->cm1setup != 1 || (t->doimpl != 0 && t->doimpl != 1) || t->bbc != 3 || t->isfcflx != 1 ||
      !(t->cnst_znt > 0.0f) || m->c.imoist != 0 || !periodic || m->c.nk < 4) {
    cm1g_set_error("cm1g_set_turbulence: supported: cm1setup 1, doimpl 0/1, bbc 3, isfcflx 1, "
                   "cnst_znt > 0, on a dry doubly periodic grid with nk >= 4");
    return CM1G_E_ARG;
  }
  for (int id = CM1G_TKEA; id <= CM1G_LENSCL; id++)
    if (!m->f[id]) {
      const int rc = dalloc(&m->f[id], m->ntot);
      if (rc) return rc;
    }
  const int nk = m->c.nk, pz = nk + 3;
  if (!m->tcol) {
    const int rc = dalloc(&m->tcol, (size_t)pz);
    if (rc) return rc;
  }
  m->turb = *t;
  // per-level grid scale of tkekm, tconfig = 1 (src/turb.F:3040): uh = vh
  // = 1 on these grids
  {
    std::vector<float> g(pz, 0.0f);
    const float *rmf = m->hk[CM1G_K_RMF];
    for (int k = 2; k <= nk; k++)
      g[k] = cm1libm::powf_host(((m->c.dx * 1.0f) * (m->c.dy * 1.0f)) * (m->c.dz * rmf[k]), 0.33333333f);
    HIPCHK(hipMemcpy(m->tcol, g.data(), pz * sizeof(float), hipMemcpyHostToDevice));
  }
  // roughness length (src/init_surface.F:251) and alog(za*rznt) of the
  // log-law friction velocity (src/sfcphys.F:91-98); za = zh(1), no terrain
  {
    const float znt = fmaxf(1.0e-12f, t->cnst_znt);
    const float za = m->hk[CM1G_K_ZH][1];
    const float rznt = 1.0f / znt;
    m->ustlog = cm1libm::logf_host(za * rznt);
    std::vector<float> h((size_t)m->P.pxy, znt);
    HIPCHK(hipMemcpy(m->d2[CM1G_D_ZNT], h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  graphs_clear(m);
  return CM1G_OK;
}

extern "C" int cm1g_sfc_and_turb(cm1g_model *m, int getsfc) {
  if (!m || m->turb.cm1setup != 1) {
    cm1g_set_error("cm1g_sfc_and_turb: turbulence not set (cm1g_set_turbulence)");
    return CM1G_E_ARG;
  }
  int rc;
  if ((rc = unswap(m))) return rc;
  turb_sfc(m, getsfc != 0);
  return check_launch();
}

// ------------------------------------------------- routine-level entries
static int turb_ready(cm1g_model *m, const char *who) {
  if (!m || m->turb.cm1setup != 1) {
    cm1g_set_error(std::string(who) + ": turbulence not set (cm1g_set_turbulence)");
    return CM1G_E_ARG;
  }
  return CM1G_OK;
}

extern "C" int cm1g_turbs(cm1g_model *m, int iflux, int dosfcflx, int s_id, int sten_id) {
  int rc = turb_ready(m, "cm1g_turbs");
  if (rc) return rc;
  if (!field_ok(m, s_id) || !field_ok(m, sten_id) || (iflux != 0 && iflux != 1)) {
    cm1g_set_error("cm1g_turbs: bad field id or iflux");
    return CM1G_E_ARG;
  }
  if ((rc = unswap(m))) return rc;
  run_turbs(m, m->f[s_id], 0, m->f[sten_id], iflux, dosfcflx != 0);
  return check_launch();
}

extern "C" int cm1g_turbt(cm1g_model *m) {
  int rc = turb_ready(m, "cm1g_turbt");
  if (rc) return rc;
  run_turbt(m, 0);
  return check_launch();
}

extern "C" int cm1g_turbu(cm1g_model *m) {
  int rc = turb_ready(m, "cm1g_turbu");
  if (rc) return rc;
  if ((rc = unswap(m))) return rc;
  run_turbuvw(m, 1);
  return check_launch();
}

extern "C" int cm1g_turbv(cm1g_model *m) {
  int rc = turb_ready(m, "cm1g_turbv");
  if (rc) return rc;
  if ((rc = unswap(m))) return rc;
  run_turbuvw(m, 2);
  return check_launch();
}

extern "C" int cm1g_turbw(cm1g_model *m) {
  int rc = turb_ready(m, "cm1g_turbw");
  if (rc) return rc;
  if ((rc = unswap(m))) return rc;
  run_turbuvw(m, 4);
  return check_launch();
}

extern "C" int cm1g_advw_scalar(cm1g_model *m, int nrk, int advweno, int a_id, int ten_id) {
  int rc;
  if (!m || nrk < 1 || nrk > 3 || advweno < 0 || advweno > 2 || !field_ok(m, a_id) || !field_ok(m, ten_id)) {
    cm1g_set_error("cm1g_advw_scalar: bad argument");
    return CM1G_E_ARG;
  }
  if ((rc = unswap(m))) return rc;
  float **f = m->f;
  advw_tke(m, nrk, advweno, f[CM1G_RRU], f[CM1G_RRV], f[CM1G_RRW], f[a_id], f[ten_id], f[ten_id]);
  return check_launch();
}
