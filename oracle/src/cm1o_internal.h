/* Internal declarations of the CPU oracle (test infrastructure only). */
#ifndef CM1O_INTERNAL_H
#define CM1O_INTERNAL_H

/* CM1O_FP64: the same restatement carried out in double precision, the
 * analogue of the reference's -DDP build (every REAL becomes REAL(8),
 * src/Makefile:73-76: -r8 with -DDP).  The Makefile builds it from
 * copies of the sources with the 'f' suffix taken off every literal, so
 * constants are double too.  System headers are read before 'float' is
 * redefined. */
#ifdef CM1O_FP64
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static inline double o64_fabs(double x) { return fabs(x); }
static inline double o64_fmax(double x, double y) { return fmax(x, y); }
static inline double o64_fmin(double x, double y) { return fmin(x, y); }
static inline double o64_sqrt(double x) { return sqrt(x); }
static inline double o64_exp(double x) { return exp(x); }
static inline double o64_pow(double x, double y) { return pow(x, y); }
static inline double o64_copysign(double x, double y) { return copysign(x, y); }
static inline long o64_lround(double x) { return lround(x); }
static inline double o64_cos(double x) { return cos(x); }
static inline double o64_log(double x) { return log(x); }
#define float double
#define fabsf o64_fabs
#define fmaxf o64_fmax
#define fminf o64_fmin
#define sqrtf o64_sqrt
#define expf o64_exp
#define powf o64_pow
#define copysignf o64_copysign
#define lroundf o64_lround
#define cosf o64_cos
#define logf o64_log
#endif

#include "cm1o.h"
#include <math.h>

typedef __float128 quad;

/* Single-precision exp and pow.  Mode 0 (default): the host C library's
 * expf/powf, which is what the reference's REAL exp() and ** lower to;
 * the oracle then reproduces the reference bit for bit.  Mode 1: the
 * value evaluated in double precision and rounded once, which is how the
 * HIP kernels evaluate them; used to separate the effect of libm rounding
 * from everything else when the GPU is compared with the reference. */
extern int cm1o_libm_mode;
static inline float o_expf(float x) { return cm1o_libm_mode ? (float)exp((double)x) : expf(x); }
static inline float o_powf(float x, float y) {
  return cm1o_libm_mode ? (float)pow((double)x, (double)y) : powf(x, y);
}

/* Physical constants, set exactly as set_constants() does for testcase=0
 * (src/constants.F:69-170): single-precision arithmetic at run time. */
typedef struct {
  float g, rd, rv, cp, cv, cpv, cvv, cpdcv, rovcp, rddcv, cpdrd, eps, reps,
        repsm1, cpl, cpi, lv1, lv2, xlv;
  float p00, rp00, th0r, smeps, tsmall, qsmall;
  float epsilon;
} cm1o_const;

struct cm1o_model {
  cm1o_config c;
  cm1o_const k;
  int ni, nj, nk;
  long px, py, pz, pxy, ntot;
  int ibw, ibe, ibs, ibn;
  float rdx, rdy, rdz;
  float cgs1, cgs2, cgs3, cgt1, cgt2, cgt3;
  float *f[F_NFIELDS];
  float *kf[K_NKFIELDS];
  float *h[6];           /* uh, ruh, uf, vh, rvh, vf */
  float *d2[D_N2D];
  float *dtu, *dtv, *dtu0, *dtv0;  /* 2-D (px*py) */
  /* scratch */
  float *s1, *s2, *s3, *s4, *s5, *s6, *s7, *s8;
  unsigned char *flag;
  double mass1, mass2, train, cflmax;
  float qmag[3];
};

#define IX(m, i, j, k) ((size_t)((i) + 2) + (size_t)((j) + 2) * (size_t)(m)->px + (size_t)(k) * (size_t)(m)->pxy)
#define IX2(m, i, j) ((size_t)((i) + 2) + (size_t)((j) + 2) * (size_t)(m)->px)

/* accessor helpers used by the routine files */
#define A3(arr, i, j, k) (arr)[IX(M, i, j, k)]
#define UH(i)  M->h[0][(i) + 2]
#define RUH(i) M->h[1][(i) + 2]
#define UF(i)  M->h[2][(i) + 2]
#define VH(j)  M->h[3][(j) + 2]
#define RVH(j) M->h[4][(j) + 2]
#define VF(j)  M->h[5][(j) + 2]

/* bc.F */
void o_bcs(cm1o_model *M, float *s);
void o_bcp(cm1o_model *M, float *p);
void o_bcu(cm1o_model *M, float *u);
void o_bcv(cm1o_model *M, float *v);
void o_bcw(cm1o_model *M, float *w, int flag);
void o_radbcew4(cm1o_model *M, float *radbcw, float *radbce, const float *u1, const float *u2, float dt);
void o_radbcns4(cm1o_model *M, float *radbcs, float *radbcn, const float *v1, const float *v2, float dt);

/* adv.F / adv_routines.F / pdef.F */
float o_flx3(float s1, float s2, float s3);
float o_flx5(float s1, float s2, float s3, float s4, float s5);
float o_weno5(float s1, float s2, float s3, float s4, float s5, double weps);
float o_upstrpd(float s1, float s2, float s3, double weps);
void o_advs(cm1o_model *M, int nrk, const float *rru, const float *rrv, const float *rrw,
            const float *divx, const float *s0, const float *s, float *sten,
            int pdef, int pdefweno, float dt, double weps);
void o_advu(cm1o_model *M, int nrk, const float *rru, const float *u3d, float *uten,
            const float *rrv, const float *rrw, const float *divx);
void o_advv(cm1o_model *M, int nrk, const float *rru, const float *rrv, const float *v3d,
            float *vten, const float *rrw, const float *divx);
void o_advw(cm1o_model *M, int nrk, const float *rru, const float *rrv, const float *rrw,
            const float *w3d, float *wten, const float *divx);
void o_advw_w(cm1o_model *M, int nrk, int doweno, const float *rru, const float *rrv, const float *rrw,
              const float *w3d, float *wten, const float *divx);

/* sound.F */
void o_sound(cm1o_model *M, int nrk, float dttmp);
void o_sounde(cm1o_model *M, int nrk, float dttmp);

/* misclibs.F */
void o_getdiv(cm1o_model *M, const float *u, const float *v, const float *w, float *div);
void o_getdivx(cm1o_model *M, float *rru, float *rrv, float *rrw, float *divx);
void o_calccflquick(cm1o_model *M, float dt, const float *ua, const float *va, const float *wa);
void o_pdefq(cm1o_model *M, float rmax, const float *rho, float *q3d);

/* kessler.F */
void o_k_fallout(cm1o_model *M, const float *rho, const float *qr3d, float *vr);
void o_fallout(cm1o_model *M, float dt, const float *rho, float *q3d, float *vq);
void o_kessler(cm1o_model *M, float dt, float *tmp, float *rho, float *pp3d, float *th3d,
               float *prs, float *qv3d, float *qc3d, float *qr3d);
int  o_satadj(cm1o_model *M, int nrk, float dt, float *rho, float *pp3d, float *prs,
              float *th3d, float *qv, float *qc, float *qr);

/* turb.F, turbtend.F, sfcphys.F (LES subgrid turbulence) */
void o_bc2d(cm1o_model *M, float *s);   /* k = 1 slice or a 2-D array, periodic */
void o_sfc_and_turb(cm1o_model *M);
void o_turb_tend(cm1o_model *M);
void o_tke_step(cm1o_model *M, int nrk, float dttmp);
void o_tke_copy(cm1o_model *M);

/* solve*.F */
void o_solve1(cm1o_model *M);
void o_solve2(cm1o_model *M);
void o_mp_driver(cm1o_model *M);
void o_solve3(cm1o_model *M);

#endif
