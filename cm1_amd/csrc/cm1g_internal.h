// Internal declarations of the cm1gpu HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>
#include "../../include/cm1gpu.h"
#include "cm1g_libm.h"

// Physical constants, single precision as set_constants() computes them
// (src/constants.F:69-170, testcase=0).
struct Consts {
  float g, rd, rv, cp, cv, cpv, cvv, cpdcv, rovcp, rddcv, cpdrd, eps, reps, repsm1,
      cpl, cpi, lv1, lv2, xlv;
  float p00, rp00, th0r, smeps, tsmall, qsmall, epsilon;
};

// Everything a kernel needs about the grid: sizes, boundary switches,
// horizontal metrics and the vertical (k) metric/base-state columns.
struct KP {
  int ni, nj, nk;
  int px, py;
  int pxy;
  int wbc, ebc, sbc, nbc;
  int ibw, ibe, ibs, ibn;
  float rdx, rdy, rdz, dx, dy, dz;
  float cgs1, cgs2, cgs3, cgt1, cgt2, cgt3;
  Consts K;
  const float *mh, *rmh, *mf, *c1, *c2, *pi0, *th0, *qv0, *rho0, *rr0, *rf0, *rrf0,
      *thv0, *rth0, *qc0, *qi0, *u0, *v0, *tauh, *tauf;
  int irdamp;
  float rdalpha;
  const float *uh, *ruh, *uf, *vh, *rvh, *vf;
};

// Element (i,j,k) of a 3-D array, i = -2..ni+4, j = -2..nj+4, k = 0..nk+2.
// i = 1 sits at a multiple of 32 floats (CM1G_IOFF), so a wavefront's 64
// consecutive interior i fill exactly two 128-B cache lines.
__device__ __host__ __forceinline__ int IXD(const KP &P, int i, int j, int k) {
  return (i + CM1G_IOFF) + (j + 2) * P.px + k * P.pxy;
}
__device__ __host__ __forceinline__ int IX2D(const KP &P, int i, int j) {
  return (i + CM1G_IOFF) + (j + 2) * P.px;
}

// bc kinds
enum { BC_S = 0, BC_P = 1, BC_U = 2, BC_V = 3, BC_W1 = 4, BC_W0 = 5 };

struct cm1g_model {
  cm1g_config c;
  KP P;                         // kernel parameters (device pointers inside)
  float *f[CM1G_NFIELDS];       // 3-D device fields
  float *kf[CM1G_K_NFIELDS];    // 1-D device columns
  float *hk[CM1G_K_NFIELDS];    // host copies of the columns
  float *h[6];                  // device uh, ruh, uf, vh, rvh, vf
  float *d2[CM1G_D_NFIELDS];    // 2-D device fields
  float *dtu0, *dtv0;           // 2-D open-boundary masks (src/param.F:9182-9218)
  // scratch (device)
  float *sx, *sy, *sz, *tx, *ty, *tz;   // advection fluxes / pdef work
  float *thb;                   // ping-pong partner for th3d / q3d updates
  float *scr1, *scr2;           // u/v partner arrays of the acoustic substeps
  double *dred;                 // reduction workspace (per-k partial sums etc.)
  float *fred;                  // float reduction workspace
  double *dscal;                // device scalars: [0] mass1 [1] mass2 [2] tem(pp shift) [3] train
  int *iflag;                   // device flags: [0] satadj failure
  hipStream_t stream;
  size_t ntot;
  double mass1, train;
  float cflmax;
  int advs_unscaled;            // cm1g_advs: use rru/rrv/rrw as given (no rho0/rf0 scaling)
  // kernel timer: selected kernel id and the event pairs recorded so far
  int kt_which;
  std::vector<hipEvent_t> kt_ev;
  size_t kt_used;
};

// bracket a launch of kernel `which` with timer events (no-op unless selected)
void ktimer_mark(cm1g_model *m, int which);

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) {                                                        \
      cm1g_set_error(std::string(#x) + ": " + hipGetErrorString(e_));              \
      return CM1G_E_HIP;                                                           \
    }                                                                              \
  } while (0)

void cm1g_set_error(const std::string &s);

// ---------------------------------------------------------------- launching
// Element-wise launches.  Point i = 1 sits at the start of a 128-byte line
// (CM1G_IOFF), so the launch origin in i is rounded down to 1 - 64m: every
// wavefront then covers whole lines of one row, also for ranges that start
// in the halo (i = 0, -1, -2), whose points go to one mostly idle wavefront
// instead of shifting every wavefront across a line boundary (which costs
// up to twice the time of the aligned launch).
inline int i_origin(int ilo) { return ilo >= 1 ? 1 + ((ilo - 1) / 64) * 64 : 1 - ((64 - ilo) / 64) * 64; }

template <class F>
__global__ void __launch_bounds__(256) k_for3(F f, int i0, int ilo, int ihi, int j0, int nj, int k0) {
  int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  int k = k0 + (int)blockIdx.z;
  if (i >= ilo && i <= ihi && j < j0 + nj) f(i, j, k);
}

// run f(i,j,k) over the inclusive box [ilo,ihi]x[jlo,jhi]x[klo,khi]
template <class F>
inline void for3(hipStream_t s, int ilo, int ihi, int jlo, int jhi, int klo, int khi, F f) {
  int nj = jhi - jlo + 1, nk = khi - klo + 1;
  if (ihi < ilo || nj <= 0 || nk <= 0) return;
  const int i0 = i_origin(ilo), ni = ihi - i0 + 1;
  dim3 b(64, 4, 1);
  dim3 g((ni + 63) / 64, (nj + 3) / 4, nk);
  hipLaunchKernelGGL(k_for3<F>, g, b, 0, s, f, i0, ilo, ihi, jlo, nj, klo);
}

template <class F>
__global__ void __launch_bounds__(256) k_for2(F f, int i0, int ilo, int ihi, int j0, int nj) {
  int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i >= ilo && i <= ihi && j < j0 + nj) f(i, j);
}

// run f(i,j) over [ilo,ihi]x[jlo,jhi] (column kernels loop k inside)
template <class F>
inline void for2(hipStream_t s, int ilo, int ihi, int jlo, int jhi, F f) {
  int nj = jhi - jlo + 1;
  if (ihi < ilo || nj <= 0) return;
  const int i0 = i_origin(ilo), ni = ihi - i0 + 1;
  dim3 b(64, 4, 1);
  dim3 g((ni + 63) / 64, (nj + 3) / 4, 1);
  hipLaunchKernelGGL(k_for2<F>, g, b, 0, s, f, i0, ilo, ihi, jlo, nj);
}

// Column kernels run on a 1-D grid of 64 x ty column tiles (one wavefront
// per tile row).  The hardware hands consecutive workgroups to the 8 XCDs in
// turn (each with its own L2); this maps the workgroup id to a tile so that
// every XCD owns one contiguous band of tile rows, walked in order, and the
// tiles that share halo rows run on the same XCD at about the same time.
// The grid is padded to a multiple of 8 workgroups; tiles past the end
// return at once.
constexpr int NXCD = 8;
struct ColGrid {
  int gx, gy, ty, nb, per;  // tiles in i and j, tile rows, tiles, tiles per XCD
  int i0, j0;               // first column of tile 0 (i0 on a line start)
};
// columns i = 1..ni, j = 1..nj
inline ColGrid col_grid(int ni, int nj, int ty = 4) {
  ColGrid g;
  g.ty = ty;
  g.gx = (ni + 63) / 64;
  g.gy = (nj + ty - 1) / ty;
  g.nb = g.gx * g.gy;
  g.per = (g.nb + NXCD - 1) / NXCD;
  g.i0 = 1;
  g.j0 = 1;
  return g;
}
// columns covering [ilo,ihi] x [jlo,jhi], tiles aligned as for3 aligns them
// (the kernel checks the range itself)
inline ColGrid col_grid_box(int ilo, int ihi, int jlo, int jhi, int ty = 4) {
  const int i0 = i_origin(ilo);
  ColGrid g = col_grid(ihi - i0 + 1, jhi - jlo + 1, ty);
  g.i0 = i0;
  g.j0 = jlo;
  return g;
}
inline dim3 col_blocks(const ColGrid &g) { return dim3(g.per * NXCD, 1, 1); }
inline dim3 col_threads(const ColGrid &g) { return dim3(64 * g.ty, 1, 1); }
// returns false for padding workgroups; b is the workgroup's index in the
// column grid
__device__ __forceinline__ bool col_tile_b(const ColGrid &g, int b, int &i, int &j) {
  const int t = (b % NXCD) * g.per + b / NXCD;
  if (t >= g.nb) return false;
  i = g.i0 + (t % g.gx) * 64 + (int)(threadIdx.x & 63);
  j = g.j0 + (t / g.gx) * g.ty + (int)(threadIdx.x >> 6);
  return true;
}
__device__ __forceinline__ bool col_tile(const ColGrid &g, int &i, int &j) {
  return col_tile_b(g, (int)blockIdx.x, i, j);
}

// ---------------------------------------------------------------- modules
// bc fills (cm1g_bc.hip)
void bc_fill(cm1g_model *m, int kind, float *a);
void bc_fill_n(cm1g_model *m, int kind, float *const *arrs, int n);
void radbc4(cm1g_model *m, const float *ua, const float *u3d, const float *va,
            const float *v3d, float dt);
// advection (cm1g_adv.hip)
void adv_uvw(cm1g_model *m, int nrk, int mask, int buoy, int from_ten1 = 0, int own = 0);
void advm_run(cm1g_model *m, int nrk, int mask, int buoy, int from_ten1, int own);  // cm1g_advm.hip
void advs_run(cm1g_model *m, int nrk, const float *s_in, float sadd, const float *s0,
              const float *sten0, const float *thterm_div, float *out_sten, float *out_new,
              int pdef, int pdefweno, float dt, double weps, int smeps_flush);
// acoustic (cm1g_sound.hip)
void sound_run(cm1g_model *m, int nrk);
void sounde_run(cm1g_model *m, int nrk);
void getdiv_both(cm1g_model *m, int want_t11, int want_divx = 1, float *div = nullptr);
void scale_mass_fluxes(cm1g_model *m);
void cfl_run(cm1g_model *m, float dt, const float *u, const float *v, const float *w);
// microphysics (cm1g_micro.hip)
void pdefq_run(cm1g_model *m, float rmax, const float *rho, float *q);
void kessler_branch_run(cm1g_model *m);
// reductions
void mass_sums(cm1g_model *m);           // dumk3/dumk4 -> dscal
void budgets_run(cm1g_model *m, double *tmass, double *tmois);

// FP32 exp and pow with the host C library's rounding (what the
// reference's REAL exp() and ** lower to), restated in cm1g_libm.h.
__device__ __forceinline__ float exp_f(float x) { return cm1libm::expf_host(x); }
__device__ __forceinline__ float pow_f(float x, float y) { return cm1libm::powf_host(x, y); }
