/*
 * cm1gpu: MI355X (gfx950) implementation of the CM1 r21.1 dynamical-core
 * time step behind a plain C ABI.
 *
 * Every prognostic and scratch array of the step lives in HBM for the life
 * of a model handle.  The host hands arrays over once (cm1g_put_field),
 * runs steps or individual routines, and reads arrays back only when a
 * host-side routine needs them (cm1g_get_field), which is the residency
 * protocol a Fortran host uses around its non-ported code (statistics,
 * output, restart).
 *
 * Array layout on the device: every 3-D array (scalar, u, v and w
 * staggerings alike) is one padded box i = -2..ni+4, j = -2..nj+4,
 * k = 0..nk+2, i fastest, element (i,j,k) at
 *     (i + CM1G_IOFF) + (j + 2) * px + k * px * py,
 * with px = ni + 36 rounded up to a multiple of 64 (i = 1 starts a 128-byte
 * line) and py = nj + 7.  This is
 * a superset of the reference bounds ib:ie(+1), jb:je(+1), kb:ke(+1) with
 * ngxy = 3 and ngz = 1 (src/cm1.F:747-777); cm1g_put_field/cm1g_get_field
 * convert from/to the caller's own Fortran-ordered box.
 *
 * All entry points return 0 on success and a nonzero CM1G_E_* code on
 * failure; cm1g_last_error() returns the message.  The reference has no
 * return codes: its fatal paths call stopcm1 (src/stopcm1.F), which a
 * Fortran wrapper does on a nonzero return.  Calls on one handle must not
 * overlap; the library is single-stream per handle.
 */
#ifndef CM1GPU_H
#define CM1GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  CM1G_OK = 0,
  CM1G_E_ARG = 1,      /* invalid argument or configuration */
  CM1G_E_HIP = 2,      /* HIP runtime error */
  CM1G_E_NOMEM = 3,    /* device allocation failed */
  CM1G_E_UNSUPPORTED = 4,
  CM1G_E_SATADJ = 5,   /* saturation adjustment did not converge (src/kessler.F:253-257) */
  CM1G_E_CFL = 6,      /* cflmax >= 1.5 (src/cm1.F:3020-3024) */
  CM1G_E_IO = 7        /* an output file could not be written */
};

/* Namelist subset read by the hot path (module input, src/input.F:261-351). */
typedef struct cm1g_config {
  int32_t ni, nj, nk;             /* nx, ny, nz of one rank (single GPU: the whole domain) */
  int32_t wbc, ebc, sbc, nbc;     /* 1 periodic, 2 open-radiative (src/bc.F) */
  int32_t imoist;                 /* 0 dry, 1 moist */
  int32_t ptype;                  /* 0 none, 1 Kessler */
  int32_t psolver;                /* 3 vertically implicit (sound.F); 2 explicit (sounde.F) */
  int32_t advwenos, advwenov;     /* 0 flux-form 5th order; 1 WENO5 every stage; 2 last stage */
  int32_t apmasscon;              /* 1: dry-mass conserving pi' adjustment */
  int32_t irbc;                   /* 4: radbcew4/radbcns4 */
  int32_t nsound;                 /* acoustic steps per large step (see cm1g_dtsmall) */
  int32_t npt;                    /* passive tracers */
  int32_t pdtra;                  /* positive-definite tracers */
  float dx, dy, dz, dt;
  float alph, kdiv;               /* param3 */
  float csmax;                    /* 350 (src/constants.F) */
  int32_t irdamp;                 /* 0 none; 1 Rayleigh damping above zd toward the base
                                     state (src/solve1.F:414-426, src/param.F:8633-8668) */
  float rdalpha, zd;              /* damping rate (1/s) and height (m), param3 */
} cm1g_config;

/* 3-D field ids (names follow the reference's arrays). */
enum {
  CM1G_UA, CM1G_U3D, CM1G_UTEN, CM1G_UTEN1, CM1G_VA, CM1G_V3D, CM1G_VTEN, CM1G_VTEN1,
  CM1G_WA, CM1G_W3D, CM1G_WTEN, CM1G_WTEN1, CM1G_RRU, CM1G_RRV, CM1G_RRW,
  CM1G_PPI, CM1G_PP3D, CM1G_PPTEN, CM1G_STEN, CM1G_SADV, CM1G_PPX,
  CM1G_THA, CM1G_TH3D, CM1G_THTEN, CM1G_THTEN1, CM1G_THTERM,
  CM1G_QPTEN, CM1G_QTTEN, CM1G_QVTEN, CM1G_QCTEN,
  CM1G_RHO, CM1G_RR, CM1G_RF, CM1G_PRS, CM1G_DIVX,
  CM1G_DUM1, CM1G_DUM2, CM1G_DUM3, CM1G_DUM4, CM1G_DUM5, CM1G_DUM6, CM1G_DUM7,
  CM1G_DUM8, CM1G_DUM9, CM1G_T11, CM1G_T12, CM1G_T13, CM1G_T22, CM1G_T23, CM1G_T33,
  CM1G_QA0, CM1G_QA1, CM1G_QA2, CM1G_Q3D0, CM1G_Q3D1, CM1G_Q3D2,
  CM1G_QTEN0, CM1G_QTEN1, CM1G_QTEN2,
  CM1G_PTA0, CM1G_PT3D0 = CM1G_PTA0 + 8, CM1G_PTTEN0 = CM1G_PT3D0 + 8,
  /* LES subgrid turbulence (allocated by cm1g_set_turbulence): the TKE
   * (tkea, tke3d, tketen), eddy viscosities and diffusivities, N^2,
   * deformation, dissipation and length scale, all on w levels */
  CM1G_TKEA = CM1G_PTTEN0 + 8, CM1G_TKE3D, CM1G_TKETEN, CM1G_KMH, CM1G_KMV, CM1G_KHH, CM1G_KHV,
  CM1G_NM, CM1G_DEFV, CM1G_DEFH, CM1G_DISSTEN, CM1G_LENSCL,
  CM1G_NFIELDS
};

/* 1-D vertical metric and base-state columns, index k = 0..nk+2. */
enum {
  CM1G_K_MH, CM1G_K_RMH, CM1G_K_MF, CM1G_K_RMF, CM1G_K_C1, CM1G_K_C2, CM1G_K_ZH, CM1G_K_ZF,
  CM1G_K_PI0, CM1G_K_TH0, CM1G_K_PRS0, CM1G_K_QV0, CM1G_K_RHO0, CM1G_K_RR0, CM1G_K_RF0,
  CM1G_K_RRF0, CM1G_K_THV0, CM1G_K_RTH0, CM1G_K_QC0, CM1G_K_QI0, CM1G_K_U0, CM1G_K_V0,
  CM1G_K_TAUH, CM1G_K_TAUF,       /* Rayleigh-damping coefficients (set by cm1g_metrics) */
  CM1G_K_NFIELDS
};

/* 2-D fields: rain, prate are (i,j); radbcw/e are (j,k); radbcs/n are (i,k). */
enum { CM1G_D_RAIN, CM1G_D_PRATE, CM1G_D_RADBCW, CM1G_D_RADBCE, CM1G_D_RADBCS,
       CM1G_D_RADBCN,
       /* surface layer of the turbulence (i,j): friction velocity, the
        * lowest-level wind at scalar points and its speed, surface heat
        * flux, roughness length */
       CM1G_D_UST, CM1G_D_U1, CM1G_D_V1, CM1G_D_S1, CM1G_D_THFLUX, CM1G_D_ZNT,
       CM1G_D_NFIELDS };

typedef struct cm1g_model cm1g_model;

/* Device layout of a 3-D field (cm1g_field_ptr): element (i,j,k) at
 * (i + CM1G_IOFF) + (j+2)*px + k*px*py floats, i = -2..ni+4,
 * j = -2..nj+4, k = 0..nk+2 (px, py from cm1g_dims). */
#define CM1G_IOFF 31

/* --- set-up, residency ------------------------------------------------ */

/* Replaces the module-scalar hand-over of param()/set_constants()
 * (src/param.F:189-277, src/constants.F:69-170) and the device allocation
 * of every array of src/cm1.F:779-1700.  Vertical metrics default to an
 * unstretched grid (src/param.F:7790-7810); override them with
 * cm1g_put_column and call cm1g_metrics. */
int cm1g_create(const cm1g_config *cfg, cm1g_model **out);
int cm1g_destroy(cm1g_model *m);
const char *cm1g_last_error(void);
int cm1g_dims(const cm1g_model *m, int64_t *px, int64_t *py, int64_t *pz);

/* Copy a host column (k = 0..nk+2, nk+3 values) into the model. */
int cm1g_put_column(cm1g_model *m, int id, const float *host);
int cm1g_get_column(cm1g_model *m, int id, float *host);
/* Recompute rmh, rmf, c1, c2 and the extrapolation coefficients from
 * zh, zf, mh, mf (src/param.F:8590-8622, :8980-9060). */
int cm1g_metrics(cm1g_model *m);
/* rho0, rr0, rf0, rrf0, thv0, rth0, prs0 from pi0, th0, qv0
 * (src/base.F:2085-2198). */
int cm1g_base_derive(cm1g_model *m);

/* Copy a whole host box in/out: (nk+3)*(nj+7)*(ni+7) floats, i = -2..ni+4
 * fastest, then j = -2..nj+4, then k = 0..nk+2. */
int cm1g_put_field(cm1g_model *m, int id, const float *host);
int cm1g_get_field(cm1g_model *m, int id, float *host);
/* The same for a host array in the reference's own shape (what a Fortran
 * host holds, src/cm1.F:747-777): stag CM1G_STAG_S (ib:ie,jb:je,kb:ke),
 * _U (ib:ie+1,...), _V (...,jb:je+1,...), _W (...,kb:ke+1), ngxy=3,
 * ngz=1, i fastest.  Replaces the host<->device moves around non-ported
 * code (statpack, writeout, restart). */
enum { CM1G_STAG_S = 0, CM1G_STAG_U = 1, CM1G_STAG_V = 2, CM1G_STAG_W = 3 };
int cm1g_put_array(cm1g_model *m, int id, int stag, const float *host);
int cm1g_get_array(cm1g_model *m, int id, int stag, float *host);
int cm1g_put_2d(cm1g_model *m, int id, const float *host);
int cm1g_get_2d(cm1g_model *m, int id, float *host);
/* The same for a 2-D host array in the reference's own shape: rain and
 * prate (ib:ie,jb:je) (one slice of rain(ib:ie,jb:je,nrain)), radbcw and
 * radbce (jb:je,kb:ke), radbcs and radbcn (ib:ie,kb:ke) (src/cm1.F:
 * 747-777, src/bc.F:2054-2172), and the surface fields CM1G_D_UST ..
 * CM1G_D_ZNT (ib:ie,jb:je) (src/cm1.F:747-777), first index fastest. */
int cm1g_put_2d_array(cm1g_model *m, int id, const float *host);
int cm1g_get_2d_array(cm1g_model *m, int id, float *host);
/* Device address of a field (for zero-copy use by a caller that already
 * holds device memory, e.g. a Fortran host with its own device mirrors).
 * The address is valid until the next call that runs kernels on the handle:
 * the scalar updates exchange th3d, q3d and pt3d with partner buffers, so
 * query it again after each step or routine call (a step also ends by
 * exchanging the buffers of ua..qa, pta with those of u3d..q3d, pt3d instead
 * of copying them, src/solve3.F's ua = u3d ...).  Once the address of
 * ua, va, wa, ppi, u3d, v3d, w3d or pp3d has been handed out, the caller may
 * write those arrays, and every later step copies instead of exchanging and
 * reads u3d..pp3d as the reference does instead of assuming they equal
 * ua..ppi after solve3 (correct for any writes, at the cost of the copies). */
int cm1g_field_ptr(cm1g_model *m, int id, float **dptr);
/* Wait for all queued work of this handle; CM1G_E_SATADJ if a saturation
 * adjustment of any step so far failed to converge (src/kessler.F:343-356). */
int cm1g_sync(cm1g_model *m);

/* --- LES subgrid turbulence ------------------------------------------- */

/* The reference's switches of its dry convective boundary-layer LES
 * (run/config_files/les_ConvBoundLayer) as far as the device path carries
 * them: cm1setup = 1 with sgsmodel = 1 (Deardorff TKE), tconfig = 1,
 * bcturbs = 1, horizturb = 0 and ramp_sgs = 0 are implied; doimpl
 * chooses explicit (0) or implicit (1) vertical diffusion; bbc = 3 is the
 * semi-slip surface with sfcmodel = 1 and set_znt = 1 (the neutral log-law
 * friction velocity over the roughness length cnst_znt); isfcflx = 1 with
 * set_flx = 1 is the specified surface kinematic heat flux cnst_shflx
 * (K m/s); tbc = 1.  Dry (imoist = 0), doubly periodic grids only. */
typedef struct cm1g_turb_config {
  int32_t cm1setup;               /* 1: LES with the TKE scheme; 0: off */
  int32_t doimpl;                 /* 0 explicit, 1 implicit vertical diffusion */
  int32_t bbc;                    /* 3: semi-slip surface (src/turb.F:522-1124) */
  int32_t isfcflx;                /* 1: surface heat flux */
  float cnst_shflx;               /* surface kinematic heat flux (K m/s) */
  float cnst_znt;                 /* roughness length (m) */
} cm1g_turb_config;
/* Turns the subgrid turbulence on for a created model (before cm1g_getset):
 * allocates the TKE, viscosity and diagnostic fields, sets the roughness
 * length (src/init_surface.F:251) and the per-level grid scale of tkekm
 * (src/turb.F:3040-3046).  Replaces the namelist hand-over of param2 and
 * param12 (src/param.F:189-277) for these switches.  CM1G_E_ARG for a
 * combination outside the ones above. */
int cm1g_set_turbulence(cm1g_model *m, const cm1g_turb_config *t);
/* sfc_and_turb(getsfc=.true., ...), src/turb.F:12-1700, for the switches
 * above: u1, v1, s1 and ust (getcecd, src/sfcphys.F:18-110), the
 * specified heat flux (sfcflux, :313-330), N^2 (calcnm), rho S_ij in
 * t11..t33 and the deformation (calcdef), km, kh, dissten, lenscl (tkekm),
 * the surface values (tkebc) and the stresses (gettau, sfcstress);
 * getsfc = 0 keeps the stored interior u1, v1, s1, ust and heat flux (the
 * first call after a restart, src/cm1.F:3726-3728) and refills the periodic
 * halos of u1, v1, s1, ust from them (bc2d, src/turb.F:1134-1160).  cm1g_step runs it (getsfc = 1) at
 * the start of every large step, where the reference runs it at the end of
 * the one before (src/cm1.F:3737). */
int cm1g_sfc_and_turb(cm1g_model *m, int getsfc);
/* The turbulence tendencies solve1 calls (src/solve1.F:1118-1240), with the
 * model's arrays as solve1 passes them (src/turbtend.F, the branch set
 * above): turbs(iflux,...,sflux,...,s,sten,khh,khv,...) adds the subgrid
 * flux divergence of field s_id to field sten_id, sflux being the 2-D
 * field CM1G_D_THFLUX (src/turbtend.F:18-698); turbt adds the TKE's own
 * transport to tketen from tkea (:700-1050); turbu, turbv, turbw add the
 * stress divergence to uten1, vten1, wten1 from ua, va, wa, t11..t33
 * (:1052-2892). */
int cm1g_turbs(cm1g_model *m, int iflux, int dosfcflx, int s_id, int sten_id);
int cm1g_turbt(cm1g_model *m);
int cm1g_turbu(cm1g_model *m);
int cm1g_turbv(cm1g_model *m);
int cm1g_turbw(cm1g_model *m);
/* advw(...,w3d=a,wten=ten,...,advweno,...,iamw=.false.) as solve2 calls it
 * for the TKE (src/solve2.F:1746-1749, src/adv.F:2778-3610): ten += the
 * advection of the w-level field a_id with the model's rru, rrv, rrw and
 * divx as stored, WENO5 when advweno is 1 (or 2 on the last stage). */
int cm1g_advw_scalar(cm1g_model *m, int nrk, int advweno, int a_id, int ten_id);

/* --- the time step ---------------------------------------------------- */

/* getset: halo fills, 3d = a copies, rho/prs, rr/rf, mass1, ppx
 * (src/init3d.F:2032-2872).  fresh_start=0 is the restart path. */
int cm1g_getset(cm1g_model *m, int fresh_start);
/* One large step: solve1 + solve2 + mp_driver + solve3
 * (src/cm1.F:3313-3476).  With cflmax non-NULL the step is waited for, the
 * Courant maximum is returned in *cflmax, and the step returns
 * CM1G_E_SATADJ if the saturation adjustment has failed to converge
 * (src/kessler.F:343-356) or CM1G_E_CFL if cflmax >= 1.5, where the
 * reference's MPI build stops (src/solve2.F:2082, src/cm1.F:4385-4397; its
 * serial build does not check, src/misclibs.F:2810); the step's results are
 * in place either way.  With cflmax NULL nothing is read back and the step
 * is only queued; cm1g_sync then reports a failed saturation adjustment.
 * A model whose adjustment failed keeps reporting it (the reference stops
 * the run there). */
int cm1g_step(cm1g_model *m, float *cflmax);
/* No counterpart in the reference.  on=1 (the default): cm1g_step replays
 * each large step as one captured HIP graph (captured on first use, one
 * per arrangement of the exchanged scalar buffers), which removes the
 * per-kernel launch cost that dominates small grids; on=0: kernel-by-kernel
 * launches.  Results are identical either way. */
int cm1g_set_graphs(cm1g_model *m, int on);
/* No counterpart in the reference.  Rows of 64 columns per workgroup of the
 * column kernels (advection, pdef, mass fluxes, acoustic substep): 1, 2 or
 * 4 (the default).  Changes which workgroup computes which column, never a
 * result. */
int cm1g_set_tiles(cm1g_model *m, int ty);

/* --- ensembles (no counterpart in the reference) ------------------------
 * Members of any grid size and sounding, created and set up as single
 * models, advanced together: each kernel of the large step runs once for
 * all members (member-indexed workgroups), and the whole merged step is
 * replayed as one captured HIP graph.  The members must share the switches
 * that decide which kernels a step launches (psolver, imoist, ptype,
 * advwenos/v, npt, apmasscon, open or periodic sides); sizes, soundings,
 * dt, nsound-independent scalars may differ, nsound may not.  Results equal
 * those of the members stepped one by one, bit for bit.  Work is queued on
 * the ensemble's own stream: call cm1g_ensemble_sync before using a member
 * through any other entry point, and do not step a member on its own while
 * it belongs to a live ensemble. */
typedef struct cm1g_ensemble cm1g_ensemble;
int cm1g_ensemble_create(cm1g_model *const *members, int n, cm1g_ensemble **out);
int cm1g_ensemble_destroy(cm1g_ensemble *e);   /* the members stay */
int cm1g_ensemble_step(cm1g_ensemble *e);      /* one large step of every member */
int cm1g_ensemble_sync(cm1g_ensemble *e);     /* wait; CM1G_E_SATADJ if a member's adjustment failed */
int cm1g_ensemble_set_tiles(cm1g_ensemble *e, int ty);
/* Autotuner: times `steps` ensemble steps at each tile height (1, 2, 4),
 * keeps the fastest in *ty_best and, if ms_per_step is non-NULL, stores the
 * three timings.  The 4*(steps+1) steps it takes are ordinary steps of the
 * run. */
int cm1g_ensemble_tune(cm1g_ensemble *e, int steps, int *ty_best, float *ms_per_step);
/* The order of the members' workgroups in every merged launch (dispatched
 * in that order): 0 as given, 1 deepest columns (largest nk) first, then
 * the larger members, 2 the larger members first.  Never changes a result. */
int cm1g_ensemble_set_order(cm1g_ensemble *e, int mode);
/* Autotuner over both schedule knobs: every tile height (1, 2, 4) with
 * every member order (0, 1, 2), `steps` timed steps each after one warm-up
 * schedule; keeps the fastest in *ty_best / *order_best and, if ms9 is
 * non-NULL, stores the nine timings as ms9[3*order + k] (k: 0 for height
 * 1, 1 for 2, 2 for 4).  Its 10*(steps+1) steps are ordinary steps. */
int cm1g_ensemble_tune_schedule(cm1g_ensemble *e, int steps, int *ty_best, int *order_best, float *ms9);

/* --- learned Kessler surrogate (no counterpart in the reference) --------
 * A per-cell MLP, 8 -> 128 -> 128 -> 128 -> 4 with SiLU hidden layers, BF16
 * operands and FP32 accumulation on the matrix cores, standing in for
 * kessler (src/kessler.F:8-138: autoconversion, accretion, rain
 * evaporation); sedimentation before it and satadj after it stay exact.
 * Inputs (raw): T, p, rho, qv, qc, qr, dt, z/ztop of a cell entering
 * kessler; outputs: the changes of theta, qv, qc, qr over kessler.
 * Parameters are FP32, weights row-major [in][out] in the order
 * W1 b1 W2 b2 W3 b3 W4 b4; normalisation norm[40] = xmean[8] xstd[8]
 * ymean[4] ystd[4] xlo[8] xhi[8] (x = (raw - xmean)/xstd,
 * change = y*ystd + ymean; [xlo, xhi] the range of normalised inputs
 * trusted, -8..8 by default).  Device arrays below are device pointers. */
typedef struct cm1g_mlp cm1g_mlp;
int cm1g_mlp_nparams(int64_t *n);
int cm1g_mlp_create(uint64_t seed, cm1g_mlp **out);   /* He-uniform weights, zero biases */
int cm1g_mlp_destroy(cm1g_mlp *net);
int cm1g_mlp_get_params(cm1g_mlp *net, float *host);
int cm1g_mlp_set_params(cm1g_mlp *net, const float *host);
int cm1g_mlp_set_norm(cm1g_mlp *net, const float *norm40);
/* y[rows x 4] = net(x[rows x 8]) on normalised values (fused inference) */
int cm1g_mlp_forward(cm1g_mlp *net, const float *x, int64_t rows, float *y);
/* Mean squared error of net(x) against t[rows x 4] (normalised) and its
 * gradient (kept for cm1g_mlp_adam); *loss and grad[nparams] (host) are
 * filled when non-NULL. */
int cm1g_mlp_loss_grad(cm1g_mlp *net, const float *x, const float *t, int64_t rows, float *loss, float *grad);
int cm1g_mlp_adam(cm1g_mlp *net, float lr);          /* one Adam step (0.9, 0.999, 1e-8) */
int cm1g_mlp_sync(cm1g_mlp *net);
/* Replace kessler inside cm1g_step's Kessler branch with the network (NULL:
 * plain kessler).  Cells where none of kessler's processes can act stay
 * unchanged, as kessler leaves them.  Elsewhere the proposal is projected
 * onto what kessler can do there: cloud only lost (at most all of it),
 * vapour only gained where rain evaporates, rain taking the difference (so
 * qv + qc + qr is kept by construction, no species negative); theta, pi'
 * and p then follow from the evaporation as in kessler.  A column with a
 * cell whose inputs leave [xlo, xhi] or whose proposal is not finite runs
 * plain kessler instead. */
int cm1g_set_surrogate(cm1g_model *m, cm1g_mlp *net);
/* While buf is set, every Kessler branch of the plain model appends
 * (inputs[8], changes over kessler[4]) float pairs to buf (device, capacity
 * pairs): all cells with cloud, rain or any change, and one in keep_one_in
 * of the others.  buf = NULL stops. */
int cm1g_set_harvest(cm1g_model *m, float *buf, int64_t capacity, int keep_one_in);
/* out6 = pairs harvested (may exceed the capacity), surrogate local steps,
 * cells that fell back to plain Kessler, harvest calls, changing cells
 * outside the trusted input range, unusable proposals */
int cm1g_surrogate_counts(cm1g_model *m, int64_t *out6);
int cm1g_solve1(cm1g_model *m);   /* src/solve1.F:77-1791 (ten1, ppten) */
int cm1g_solve2(cm1g_model *m);   /* src/solve2.F:77-2094 (RK3 loop) */
int cm1g_mp_driver(cm1g_model *m);/* src/mp_driver.F:17-1706 (ptype 0/1) */
int cm1g_solve3(cm1g_model *m);   /* src/solve3.F:77-1143 */

/* --- individual routines (same argument meaning as the reference) ------ */

/* advs(nrk,...,s0,s,sten,pdef,pdefweno,dt,weps,...), src/adv.F:14-903.
 * Uses the model's rru/rrv/rrw/divx; the s, s0 and sten fields are given
 * by id. */
int cm1g_advs(cm1g_model *m, int nrk, int s_id, int s0_id, int sten_id,
              int pdef, int pdefweno, float dt, double weps);
int cm1g_advu(cm1g_model *m, int nrk);   /* src/adv.F:911-1839  */
int cm1g_advv(cm1g_model *m, int nrk);   /* src/adv.F:1847-2770 */
int cm1g_advw(cm1g_model *m, int nrk);   /* src/adv.F:2778-3610 */
/* sound(dt,...,nrk,dttmp,...), src/sound.F:10-859 (psolver=3), or
 * sounde(...), src/sounde.F:10-1026 (psolver=2), with the solve2 calls'
 * arrays: thv=t11, ppterm=t22, piadv=sadv (src/solve2.F:1100-1136). */
int cm1g_sound(cm1g_model *m, int nrk, float dttmp);
/* Kessler branch of mp_driver: pdefq x3, k_fallout, fallout, kessler,
 * satadj (src/mp_driver.F:190-232, src/kessler.F). */
int cm1g_kessler_branch(cm1g_model *m);
/* The routines of kessler_module one by one, on the model's th3d, pp3d,
 * prs, rho and q3d(nqv..nqr) fields, with qten(nqr) (CM1G_QTEN2) as the
 * fall speed, in the order mp_driver calls them (src/mp_driver.F:195-232):
 * k_fallout(rho,qr3d,vr), src/kessler.F:368-396: vr into CM1G_QTEN2.
 * fallout(dt,train,...,rain,prate,rr,rho,q3d,vq), src/kessler.F:404-477:
 *   sedimentation of qr, rain/prate, and *train (host in/out) += the
 *   surface rain budget.
 * geterain(dt,cpx,lx1,erain,ruh,rvh,t,rr,q3d,vr), src/kessler.F:530-562:
 *   *erain (host in/out) += the rain-energy budget of fields t_id, rr_id,
 *   q_id with the fall speed CM1G_QTEN2.
 * kessler(dt,...), src/kessler.F:8-138 (the process budgets tauto, taccr,
 *   tevar are not formed).
 * satadj(4,dt,...), src/kessler.F:146-360 (tcond, tevac not formed);
 *   CM1G_E_SATADJ where the adjustment does not converge. */
int cm1g_k_fallout(cm1g_model *m);
int cm1g_fallout(cm1g_model *m, double *train);
int cm1g_geterain(cm1g_model *m, float dt, float cpx, float lx1, int t_id, int rr_id, int q_id, double *erain);
int cm1g_kessler(cm1g_model *m);
int cm1g_satadj(cm1g_model *m);
/* pdefq(rmax,asq,...,rho,q3d), src/misclibs.F:1037-1117 */
int cm1g_pdefq(cm1g_model *m, float rmax, int q_id);
/* Boundary fills, src/bc.F: kind 0 bcs :180, 1 bcp :372, 2 bcu :528,
 * 3 bcv :722, 4 bcw(flag=1) :942, 5 bcw(flag=0). */
int cm1g_bc(cm1g_model *m, int id, int kind);
/* calccflquick(dt,uh,vh,mh,ua,va,wa,...), src/misclibs.F:2726-2814 */
int cm1g_calccflquick(cm1g_model *m, float dt, int u_id, int v_id, int w_id, float *cflmax);
/* getdiv (src/misclibs.F:108-214): divergence of the model's rru/rrv/rrw
 * into field div_id (t11 at src/solve2.F:1280).  getdivx
 * (src/misclibs.F:222-370): scales rru/rrv by rho0 and rrw by rf0 in place
 * (rrw = 0 at k = 1 and nk+1) and computes divx from them. */
int cm1g_getdiv(cm1g_model *m, int div_id);
int cm1g_getdivx(cm1g_model *m);
/* maxmin(izz,jzz,kzz,f,...), src/maxmin.F:5-119: the largest and smallest
 * interior value of field id (izz, jzz, kzz = ni, nj, nk plus one along the
 * staggered direction of stag, as statpack passes them, src/statpack.F:138,
 * 214) and their (i,j,k): among equal extremes the smallest k, then j, then
 * i, as maxmin's strict comparisons pick them.  Reduced on the device. */
int cm1g_maxmin(cm1g_model *m, int id, int stag, float *vmax, int32_t *locmax, float *vmin, int32_t *locmin);
/* No counterpart in the reference.  cm1g_maxmin for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): the same result for a device
 * field `field` in this library's layout (as cm1g_field_ptr returns it) of
 * an ni x nj x nk grid, with no model handle and no allocation.  Its
 * partial results live in a file-scope device array of 64 KiB per GPU,
 * held until the process ends; ni, nj and nk must each lie in
 * 1..CM1G_NOALLOC_MAXN and the interior below 2^32-1 points, else
 * CM1G_E_ARG.  Runs on `stream` (a hipStream_t, NULL for the default
 * stream) and waits for it.  Every call on one GPU shares that array:
 * calls that overlap on one GPU race. */
#define CM1G_NOALLOC_MAXN 16384
int cm1g_maxmin_noalloc(const float *field, int ni, int nj, int nk, int stag, void *stream,
                        float *vmax, int32_t *locmax, float *vmin, int32_t *locmin);
/* dtsmall: nsound for dt (src/misclibs.F:818-915).  min_dz is the smallest
 * level thickness, min over k of zf(k+1)-zf(k) (src/param.F:8522-8541),
 * which psolver 2 uses; min_dz <= 0 takes cfg->dz (an unstretched grid). */
int cm1g_dtsmall(const cm1g_config *cfg, float min_dz, int32_t *nsound);
/* Domain budgets in double precision: dry-air mass (calcmass,
 * src/misclibs.F:3134-3191) and total water qv+qc+qr (totmois,
 * src/misclibs.F:3199-3264). */
int cm1g_budgets(cm1g_model *m, double *tmass, double *tmois);
/* Device-state checkpoint (the restart write/read of
 * src/restart_write.F:173-306 and src/restart_read.F, in this library's own
 * layout): every array and scalar a step carries into the next, taken
 * between steps.  cm1g_state_bytes gives the size; cm1g_state_save copies
 * the state into a host buffer of that size; cm1g_state_load restores it
 * into a model created with the same configuration (checked; a model with
 * the subgrid turbulence also carries its TKE and surface fields, so load
 * into a model with the same cm1g_set_turbulence), after which stepping
 * continues bit for bit as if uninterrupted.  Layout: an 88-byte
 * header, then 3-D fields as logical boxes (cm1g_put_field order), the box
 * of the scalar updates' partner buffer, then the 2-D fields in device
 * order. */
int cm1g_state_bytes(cm1g_model *m, int64_t *nbytes);
int cm1g_state_save(cm1g_model *m, void *host, int64_t nbytes);
int cm1g_state_load(cm1g_model *m, const void *host, int64_t nbytes);

/* GrADS output (writeout, src/writeout.F:20-9530, with output_format = 1,
 * output_filetype = 1, outunits = 1: x, y, z in km).  Appends output time
 * nwrite (1, 2, ...; 1 starts the files afresh) of the variables the mask
 * selects to cm1out_s.dat, cm1out_u.dat, cm1out_v.dat and cm1out_w.dat in
 * the existing directory dir, in the reference's file order, one record of
 * ni*nj ((ni+1)*nj for u, ni*(nj+1) for v) REAL*4 values per level, i
 * fastest; rewrites the descriptors cm1out_{s,u,v,w}.ctl for nwrite times
 * (src/writeout.F:9278-9530); and writes record nwrite of
 * cm1out_metadata.dat/.ctl (mtime, nstep, dt, nwrite; src/writeout.F:
 * 9536-9757).  Mask bits are the reference's output_* switches (param9);
 * passive tracers pt1.. are written whenever npt > 0 (iptra = 1), and
 * CM1G_OUT_Q writes qc and qr of a moist model.  iorigin (param2: 1 or 2)
 * and stretch_z (param6: 0, or 1 for a descriptor that lists the levels)
 * only shape the descriptors.  Every value is formed on the device with the
 * reference's expression (th0+th', prs-prs0, the buoyancy, the averages
 * to scalar points) and streamed to the files one variable at a time; the
 * call waits for the model's queued work and returns when the files are
 * closed.  The staging (two device and two pinned host buffers of the
 * largest variable) is allocated on first use and kept with the handle. */
enum {
  CM1G_OUT_RAIN = 1u << 0,       /* rain, prate (2-D) */
  CM1G_OUT_ZH = 1u << 1,         /* zhval */
  CM1G_OUT_TH = 1u << 2,         /* th = th0 + th' */
  CM1G_OUT_THPERT = 1u << 3,     /* thpert */
  CM1G_OUT_PRS = 1u << 4,        /* prs */
  CM1G_OUT_PRSPERT = 1u << 5,    /* prspert = prs - prs0 */
  CM1G_OUT_PI = 1u << 6,         /* pi = pi0 + pi' */
  CM1G_OUT_PIPERT = 1u << 7,     /* pipert */
  CM1G_OUT_RHO = 1u << 8,        /* rho */
  CM1G_OUT_RHOPERT = 1u << 9,    /* rhopert = rho - rho0 */
  CM1G_OUT_QV = 1u << 10,        /* qv */
  CM1G_OUT_QVPERT = 1u << 11,    /* qvpert = qv - qv0 */
  CM1G_OUT_Q = 1u << 12,         /* qc, qr */
  CM1G_OUT_BUOYANCY = 1u << 13,  /* g th'/th0 + g repsm1 (qv-qv0) - g qc - g qr */
  CM1G_OUT_UINTERP = 1u << 14,   /* u, v, w averaged to scalar points */
  CM1G_OUT_VINTERP = 1u << 15,
  CM1G_OUT_WINTERP = 1u << 16,
  CM1G_OUT_BASESTATE = 1u << 17, /* pi0, th0, prs0, qv0; u0; v0 */
  CM1G_OUT_U = 1u << 18,
  CM1G_OUT_UPERT = 1u << 19,
  CM1G_OUT_V = 1u << 20,
  CM1G_OUT_VPERT = 1u << 21,
  CM1G_OUT_W = 1u << 22,
  CM1G_OUT_ALL = (1u << 23) - 1u,  /* every switch above (no subgrid turbulence needed) */
  /* the subgrid turbulence's fields, after w (src/writeout.F:7788-8200);
   * they need cm1g_set_turbulence, and hold what cm1g_sfc_and_turb formed
   * last (the reference writes after its end-of-step sfc_and_turb) */
  CM1G_OUT_TKE = 1u << 23,         /* tke */
  CM1G_OUT_KM = 1u << 24,          /* kmh, kmv */
  CM1G_OUT_KH = 1u << 25,          /* khh, khv */
  CM1G_OUT_DISSTEN = 1u << 26,     /* lenscl, dissten */
  CM1G_OUT_NM = 1u << 27,          /* nm */
  CM1G_OUT_DEF = 1u << 28,         /* defv, defh, epst (0: the reference forms epst
                                      only for its domain diagnostics) */
  CM1G_OUT_TURB = CM1G_OUT_TKE | CM1G_OUT_KM | CM1G_OUT_KH | CM1G_OUT_DISSTEN | CM1G_OUT_NM | CM1G_OUT_DEF
};
int cm1g_write_grads(cm1g_model *m, const char *dir, int nwrite, uint32_t mask, int iorigin,
                     int stretch_z, double mtime, int nstep);

/* statpack (src/statpack.F:22-660, with output_format = 1): one record of
 * domain statistics of the resident state, in the reference's order,
 * written as record nrec (1, 2, ...; 1 starts the file afresh) of
 * cm1out_stats.dat in the existing directory dir, with the descriptor
 * cm1out_stats.ctl for nrec records (src/statpack.F:1460-1530).  Mask bits
 * are the reference's stat_* switches (param10); the record always starts
 * with mtime, and tracer maxima/minima follow whenever npt > 0 (iptra = 1).
 * Every maximum and minimum is maxmin's (src/maxmin.F): the value found
 * first in the reference's scan order, the level of w's and the wind
 * speed's extremes from its position; the mass totals (calcmass, totmois,
 * totq, src/misclibs.F:3134-3330) add the FP32 cell terms in double in the
 * reference's order.  All reductions run on the device; only the record
 * comes back.  rstat (may be NULL) receives the record (at most
 * CM1G_STAT_MAXN values), nstat (may be NULL) its length. */
enum {
  CM1G_STAT_W = 1u << 0,        /* wmax, wmin, zwmax, zwmin */
  CM1G_STAT_WLEVS = 1u << 1,    /* w extremes at 0.5, 1, 2.5, 5, 10 km (model top >= 10 km) */
  CM1G_STAT_U = 1u << 2,        /* umax, umin, sumax, sumin */
  CM1G_STAT_V = 1u << 3,        /* vmax, vmin, svmax, svmin */
  CM1G_STAT_PIPERT = 1u << 4,   /* ppimax, ppimin */
  CM1G_STAT_PRSPERT = 1u << 5,  /* ppmax, ppmin (prs - p00 pi0^(cp/rd)) */
  CM1G_STAT_THPERT = 1u << 6,   /* thpmax, thpmin, sthpmax, sthpmin */
  CM1G_STAT_Q = 1u << 7,        /* max/min of qv, qc, qr; pratemax, pratemin */
  CM1G_STAT_SFCPRS = 1u << 8,   /* sprsmax, sprsmin, psfcmax, psfcmin */
  CM1G_STAT_WSP = 1u << 9,      /* wspmax, wspmin, zwspmax, zwspmin, swspmax, swspmin */
  CM1G_STAT_CFL = 1u << 10,     /* cflmax (calccfl, src/misclibs.F:2584-2718) */
  CM1G_STAT_TMASS = 1u << 11,   /* tmass */
  CM1G_STAT_TMOIS = 1u << 12,   /* tmois */
  CM1G_STAT_QMASS = 1u << 13,   /* massqv, massqc, massqr */
  CM1G_STAT_ALL = (1u << 14) - 1u,  /* every switch above (no subgrid turbulence needed) */
  /* the subgrid turbulence's, after the q block (src/statpack.F:258-278);
   * they need cm1g_set_turbulence */
  CM1G_STAT_TKE = 1u << 14,     /* tkemax, tkemin */
  CM1G_STAT_KM = 1u << 15,      /* kmhmax, kmhmin, kmvmax, kmvmin */
  CM1G_STAT_KH = 1u << 16,      /* khhmax, khhmin, khvmax, khvmin */
  CM1G_STAT_TURB = CM1G_STAT_TKE | CM1G_STAT_KM | CM1G_STAT_KH
};
#define CM1G_STAT_MAXPAIRS 48
#define CM1G_STAT_MAXN 128
int cm1g_statpack(cm1g_model *m, const char *dir, int nrec, uint32_t mask, double mtime, float *rstat,
                  int32_t *nstat);

/* Scalars carried between steps: [0] mass1, [1] train (domain rain
 * accumulation, src/kessler.F:472), [2] cflmax, [3] mass2. */
int cm1g_scalars(cm1g_model *m, double *out4);

/* Kernel timer (measurement hook, no counterpart in the reference): while
 * started, every launch of the selected kernel is bracketed by HIP events
 * on the model's stream.  cm1g_ktimer_stop synchronizes the stream and
 * returns the summed kernel time, the launch count and the algorithmic
 * bytes of one launch (the distinct full-size fields the reference routine
 * reads plus writes, 4 B per interior cell each, scratch excluded: 27 for
 * an acoustic substep, src/sound.F:307-835). */
enum { CM1G_KT_NONE = 0, CM1G_KT_SOUND_SUBSTEP = 1 };
int cm1g_ktimer_start(cm1g_model *m, int which);
int cm1g_ktimer_stop(cm1g_model *m, double *total_ms, int64_t *launches,
                     double *bytes_per_launch, char *name, int name_len);

#ifdef __cplusplus
}
#endif
#endif
