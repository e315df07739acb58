/*
 * CPU restatement (oracle) of the CM1 r21.1 dynamical-core time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the HIP path is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it.
 *
 * Each routine follows the reference loop nests statement by statement
 * (operation order, single precision where the reference uses REAL,
 * double where it uses DOUBLE PRECISION, quad where it uses real(kind=qp))
 * so that, built with -O2 -ffp-contract=off, it reproduces the reference
 * FP32 build bit for bit wherever no libm transcendental is involved.
 *
 * Layout: every 3-D array (scalar, u, v and w staggering alike) is stored
 * in one padded shape covering i=-2..ni+4, j=-2..nj+4, k=0..nk+2, i fastest
 * (a superset of the reference bounds ib:ie+1, jb:je+1, kb:ke+1 with
 * ngxy=3, ngz=1; src/cm1.F:747-777).
 */
#ifndef CM1O_H
#define CM1O_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Configuration: the namelist subset the hot path reads (src/input.F). */
typedef struct cm1o_config {
  int32_t ni, nj, nk;
  int32_t wbc, ebc, sbc, nbc;     /* 1 = periodic, 2 = open-radiative */
  int32_t imoist;                 /* 0 dry, 1 moist (qv,qc,qr) */
  int32_t ptype;                  /* 0 none, 1 Kessler */
  int32_t psolver;                /* 3 = vertically implicit (sound.F), 2 = explicit (sounde.F) */
  int32_t advwenos;               /* 0 5th-order flux, 1 WENO5 every stage, 2 WENO5 last stage */
  int32_t advwenov;               /* same for velocities */
  int32_t apmasscon;              /* 1 = dry-mass pressure fix (src/solve2.F:1617-1656) */
  int32_t irbc;                   /* 4 = radbcew4/radbcns4 open-boundary speeds */
  int32_t nsound;                 /* acoustic steps per large step */
  int32_t npt;                    /* passive tracers (0 = none) */
  int32_t pdtra;                  /* positive-definite tracers */
  float dx, dy, dz, dt;
  float alph, kdiv;
  float csmax;                    /* 350 (src/constants.F) */
  int32_t irdamp;                 /* 1 = Rayleigh damping above zd (src/solve1.F:414-426) */
  float rdalpha, zd;              /* damping rate (1/s) and height (m) */
  /* LES subgrid turbulence (src/turb.F, src/turbtend.F): cm1setup 0 none,
   * 1 LES with sgsmodel 1 (TKE), tconfig 1, bcturbs 1, horizturb 0; doimpl
   * 0/1; bbc 3 with sfcmodel 1, set_znt 1, set_flx 1; tbc 1 */
  int32_t cm1setup, doimpl, bbc, isfcflx;
  float cnst_shflx, cnst_znt;
} cm1o_config;

/* Field identifiers for cm1o_field(). */
enum {
  F_UA, F_U3D, F_UTEN, F_UTEN1, F_VA, F_V3D, F_VTEN, F_VTEN1,
  F_WA, F_W3D, F_WTEN, F_WTEN1, F_RRU, F_RRV, F_RRW,
  F_PPI, F_PP3D, F_PPTEN, F_STEN, F_SADV, F_PPX,
  F_THA, F_TH3D, F_THTEN, F_THTEN1, F_THTERM,
  F_QPTEN, F_QTTEN, F_QVTEN, F_QCTEN,
  F_RHO, F_RR, F_RF, F_PRS, F_DIVX,
  F_DUM1, F_DUM2, F_DUM3, F_DUM4, F_DUM5, F_DUM6, F_DUM7, F_DUM8, F_DUM9,
  F_T11, F_T12, F_T13, F_T22, F_T23, F_T33,
  F_QA0, F_QA1, F_QA2, F_Q3D0, F_Q3D1, F_Q3D2, F_QTEN0, F_QTEN1, F_QTEN2,
  F_PTA0, F_PT3D0 = F_PTA0 + 8, F_PTTEN0 = F_PT3D0 + 8,
  F_TKEA = F_PTTEN0 + 8, F_TKE3D, F_TKETEN, F_KMH, F_KMV, F_KHH, F_KHV, F_NM, F_DEFV, F_DEFH,
  F_DISSTEN, F_LENSCL,
  F_NFIELDS
};

/* 1-D (k) metric/base-state identifiers for cm1o_kfield(). */
enum {
  K_MH, K_RMH, K_MF, K_RMF, K_C1, K_C2, K_ZH, K_ZF,
  K_PI0, K_TH0, K_PRS0, K_QV0, K_RHO0, K_RR0, K_RF0, K_RRF0, K_THV0, K_RTH0,
  K_QC0, K_QI0, K_U0, K_V0, K_TAUH, K_TAUF, K_NKFIELDS
};

/* 2-D fields */
enum { D_RAIN, D_PRATE, D_RADBCW, D_RADBCE, D_RADBCS, D_RADBCN,
       D_UST, D_U1, D_V1, D_S1, D_THFLUX, D_ZNT, D_N2D };

typedef struct cm1o_model cm1o_model;

cm1o_model *cm1o_create(const cm1o_config *cfg);
void   cm1o_destroy(cm1o_model *m);
float *cm1o_field(cm1o_model *m, int id);
float *cm1o_kfield(cm1o_model *m, int id);
float *cm1o_2dfield(cm1o_model *m, int id);
float *cm1o_hfield(cm1o_model *m, int which); /* 0 uh,1 ruh,2 uf,3 vh,4 rvh,5 vf */
void   cm1o_set_extrap(cm1o_model *m, const float *cgs6);
void   cm1o_dims(const cm1o_model *m, int64_t *px, int64_t *py, int64_t *pz);
double *cm1o_scalars(cm1o_model *m);  /* [0]=mass1 [1]=train [2]=cflmax [3]=mass2 */

/* derived base state (rho0, rr0, rf0, rrf0, thv0, rth0, prs0) from
 * pi0, th0, qv0: src/base.F:2085-2198. */
void cm1o_base_derive(cm1o_model *m);

/* getset-equivalent preparation of the t=0 state (src/init3d.F:2032-2872). */
void cm1o_getset(cm1o_model *m, int fresh_start);
/* One full large step: solve1, solve2, mp_driver, solve3 (src/cm1.F:3313-3476). */
void cm1o_step(cm1o_model *m);
/* Its four parts (for phase-by-phase comparisons). */
void cm1o_solve1(cm1o_model *m);
void cm1o_solve2(cm1o_model *m);
void cm1o_mp_driver(cm1o_model *m);
void cm1o_solve3(cm1o_model *m);
/* LES subgrid turbulence set-up of a step (src/turb.F:12, sfc_and_turb):
 * surface layer, N^2, deformation, eddy viscosities, stresses */
void cm1o_sfc_and_turb(cm1o_model *m);

/* Individual routines on the model's own arrays (for unit parity tests). */
void cm1o_bc(cm1o_model *m, int id, int kind);  /* kind 0 s,1 p,2 u,3 v,4 w(flag1),5 w(flag0) */
void cm1o_advs(cm1o_model *m, int nrk, int s_id, int s0_id, int sten_id,
               int pdef, int pdefweno, double weps);
void cm1o_advu(cm1o_model *m, int nrk);
void cm1o_advv(cm1o_model *m, int nrk);
void cm1o_advw(cm1o_model *m, int nrk);
void cm1o_sound(cm1o_model *m, int nrk, float dttmp);
/* exp/pow evaluation: 0 host expf/powf (default, the reference), 1 double
 * precision rounded once (the HIP kernels' evaluation). Process-wide. */
void cm1o_set_libm(int mode);
void cm1o_kessler_branch(cm1o_model *m);
void cm1o_getdiv(cm1o_model *m, int div_id);      /* getdiv(rru,rrv,rrw -> div) */
void cm1o_getdivx(cm1o_model *m);                 /* scales rru/rrv/rrw in place, divx */
void cm1o_calccflquick(cm1o_model *m, float dt, int u_id, int v_id, int w_id); /* -> scalars[2] */
void cm1o_pdefq(cm1o_model *m, float rmax, int q_id);

#ifdef __cplusplus
}
#endif
#endif
