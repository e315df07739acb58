// Internal declarations of the cm1gpu HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <type_traits>
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
  float *scoef;                 // per-level acoustic solver coefficients, 8 x (nk+3)
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
  // Whole large steps captured as HIP graphs.  The scalar updates exchange
  // field buffers with the partner array, so the buffers a step reads depend
  // on the arrangement it starts from: one graph per arrangement (the
  // exchanges cycle through a few), keyed by the field pointers before the
  // step, with the arrangement the step leaves behind.
  struct StepGraph {
    std::vector<float *> before, after;
    hipGraphExec_t exec;
  };
  std::vector<StepGraph> graphs;
  int no_graphs;                // 1: cm1g_step launches kernel by kernel
  int tile_rows;                // rows of 64 columns per workgroup of the column kernels (1, 2 or 4)
  // 1 while ua, va, wa, ppi equal u3d, v3d, w3d, pp3d wherever the first
  // acoustic substep reads them (set by solve3 and getset, cleared by
  // anything that may write either): the first RK stage's acoustic
  // substeps may then read the a-arrays directly
  int synced;
  // 1 once cm1g_field_ptr has handed out the address of one of those eight
  // arrays: the caller may then write them where the library cannot see
  // it, so synced is never set again on this handle
  int ext_writers;
  // 1 after a step that ended by exchanging the a-arrays with the 3-D
  // arrays (solve3) instead of copying: the a-arrays hold the state, the
  // 3-D arrays are stale until the first RK stage rewrites them (it reads
  // the a-arrays in their place), or until unswap() copies them back for a
  // caller that reads them
  int swapped;
  // non-null while an ensemble records this member's step: launches are
  // appended here instead of run
  std::vector<struct LaunchRec> *rec;
  // learned surrogate of the local Kessler step (cm1g_set_surrogate), and the
  // harvest of (state, change) pairs from the plain one (cm1g_set_harvest)
  struct cm1g_mlp *sur;
  float *harv;
  int64_t harv_cap;
  int harv_keep;
  unsigned long long *dcount;   // [0] pairs harvested [1] surrogate local steps [2] fallback cells
                                // [3] harvest calls [4] cells out of the trained range [5] unusable proposals
  int *colflag;                 // per column: the surrogate fell back to plain Kessler
  int *actlist;                 // the surrogate's list of cells kessler can act on (ni*nj*nk)
  // GrADS output staging (cm1g_write_grads): two device / pinned host
  // buffer pairs of io_cap floats, and the event of each buffer's copy
  float *io_dev[2], *io_host[2];
  hipEvent_t io_ev[2];
  size_t io_cap;
  // statpack (cm1g_statpack): max/min keys, picked values and budget sums
  void *stat_buf;
  size_t stat_cap;
  // LES subgrid turbulence (cm1g_set_turbulence; cm1setup 0: off)
  cm1g_turb_config turb;
  float *tcol;                  // per level: tkekm's grid scale (k = 2..nk)
  float ustlog;                 // alog(za/znt) of the log-law friction velocity
  // pinned host staging of cm1g_put_array / cm1g_get_array
  float *pin;
  size_t pin_cap;
};

// zeroed device allocation of n floats
int dalloc(float **p, size_t n);
// LES subgrid turbulence (cm1g_turb.hip): sfc_and_turb, the solve1
// tendencies, the TKE of one RK stage, solve3's tkea = tke3d, getset's TKE
void turb_sfc(cm1g_model *m, int getsfc);
void turb_tend(cm1g_model *m);
void tke_stage(cm1g_model *m, int nrk, float dttmp);
void tke_copy(cm1g_model *m);
int tke_init(cm1g_model *m);
// ten = ten_in + advection of a on w levels with the scalars' WENO switch
// and the mass fluxes given (advw as solve2 calls it for the TKE,
// src/solve2.F:1746-1749)
void advw_tke(cm1g_model *m, int nrk, int advweno, const float *rru, const float *rrv, const float *rrw,
              const float *a, const float *ten_in, float *ten);

// release the output staging and the statpack buffer (cm1g_io.hip)
void io_free(cm1g_model *m);
// make the 3-D arrays equal the a-arrays again after an exchanging solve3
// (entry points that read or write fields by id call this first)
int unswap(cm1g_model *m);
// drop the captured steps (after anything that changes launch arguments)
void graphs_clear(cm1g_model *m);
// the host side of one large step (solve1, solve2, mp_driver, solve3)
int step_body(cm1g_model *m);
int satadj_check(cm1g_model *m);
// a valid 3-D field id of the model
bool field_ok(const cm1g_model *m, int id);
// CM1G_E_HIP if a launch since the last check failed
int check_launch();
// the rain budget of the fallout pass's columns added to train (src/kessler.F:470-473)
void train_accumulate(cm1g_model *m);
// every field pointer and the scalar updates' partner, and back
std::vector<float *> field_map(const cm1g_model *m);
void set_field_map(cm1g_model *m, const std::vector<float *> &v);

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
// Every kernel of the time step is a device functor f(bx, by, bz), run by
// each workgroup (bx, by, bz) of its grid and launched through launch<NT>()
// (NT = the workgroup size it is compiled for).  A model runs it as a kernel
// of its own (k_one).  An ensemble instead records the launches of every
// member's step and runs the matching launches of all members as one
// kernel (k_many): the members' grids are laid end to end, each starting on
// a multiple of 8 workgroups, and each workgroup finds its member in a
// table of start offsets and runs that member's functor.
struct BatchSlot {
  int start, n, gx, gy;  // first workgroup, workgroups, grid x and y of one member
};

template <int NT, class F>
__global__ void __launch_bounds__(NT) k_one(F f) {
  f((int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}

// The member search and functor call of k_many (below).
template <class F>
__device__ __forceinline__ void many_body(const F *__restrict__ fs, const BatchSlot *__restrict__ slots, int nmem) {
  // the search is uniform across the workgroup: keep it in scalar registers
  // so the member's functor is read with scalar loads
  const int b = (int)blockIdx.x;
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (__builtin_amdgcn_readfirstlane(slots[mid].start) <= b) lo = mid;
    else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const int start = __builtin_amdgcn_readfirstlane(slots[lo].start);
  const int n = __builtin_amdgcn_readfirstlane(slots[lo].n);
  const int gx = __builtin_amdgcn_readfirstlane(slots[lo].gx);
  const int gy = __builtin_amdgcn_readfirstlane(slots[lo].gy);
  const int l = b - start;
  if (l >= n) return;  // padding up to the next member's start
  const F f = fs[lo];    // read before the functor stores anything
  f(l % gx, (l / gx) % gy, l / (gx * gy));
}

template <int NT, class F>
__global__ void __launch_bounds__(NT) k_many(const F *__restrict__ fs, const BatchSlot *__restrict__ slots,
                                             int nmem) {
  many_body(fs, slots, nmem);
}

// The same two launchers compiled for at least W waves per SIMD (the
// register allocation is capped to fit W): used where the grid's wavefront
// count is a multiple of what W waves per SIMD hold on the chip, so no
// partly filled last round is left.
template <int NT, int W, class F>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W))) k_one_w(F f) {
  f((int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z);
}
template <int NT, int W, class F>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W)))
k_many_w(const F *__restrict__ fs, const BatchSlot *__restrict__ slots, int nmem) {
  many_body(fs, slots, nmem);
}

// One recorded launch: the functor's bytes, its grid and workgroup, and the
// launcher of the k_many instance for its type (which also identifies it).
struct LaunchRec {
  void (*many)(hipStream_t, const void *fs, const BatchSlot *slots, int nmem, int nblocks, dim3 block);
  dim3 grid, block;
  std::vector<unsigned char> fn;
};

template <int NT, int W, class F>
void launch_many(hipStream_t s, const void *fs, const BatchSlot *slots, int nmem, int nblocks, dim3 block) {
  if constexpr (W > 0)
    hipLaunchKernelGGL((k_many_w<NT, W, F>), dim3(nblocks), block, 0, s, (const F *)fs, slots, nmem);
  else
    hipLaunchKernelGGL((k_many<NT, F>), dim3(nblocks), block, 0, s, (const F *)fs, slots, nmem);
}

// Scalar advection column kernels compiled for 8 waves per SIMD: 0 none
// (the compiler's choice), 1 the fifth-order passes of q and the pdef pass,
// 2 also theta's.
#ifndef CM1G_COL_WAVES
#define CM1G_COL_WAVES 2
#endif

// W > 0: the kernel is compiled for at least W waves per SIMD
template <int NT, int W = 0, class F>
inline void launch(cm1g_model *m, dim3 g, dim3 b, const F &f) {
  static_assert(std::is_trivially_copyable<F>::value, "launch: functor must be trivially copyable");
  if (m->rec) {
    LaunchRec r;
    r.many = &launch_many<NT, W, F>;
    r.grid = g;
    r.block = b;
    r.fn.resize(sizeof(F));
    std::memcpy(r.fn.data(), (const void *)&f, sizeof(F));
    m->rec->push_back(std::move(r));
    return;
  }
  if constexpr (W > 0)
    hipLaunchKernelGGL((k_one_w<NT, W, F>), g, b, 0, m->stream, f);
  else
    hipLaunchKernelGGL((k_one<NT, F>), g, b, 0, m->stream, f);
}

// Element-wise launches.  Point i = 1 sits at the start of a 128-byte line
// (CM1G_IOFF), so the launch origin in i is rounded down to 1 - 64m: every
// wavefront then covers whole lines of one row, also for ranges that start
// in the halo (i = 0, -1, -2), whose points go to one mostly idle wavefront
// instead of shifting every wavefront across a line boundary (which costs
// up to twice the time of the aligned launch).
inline int i_origin(int ilo) { return ilo >= 1 ? 1 + ((ilo - 1) / 64) * 64 : 1 - ((64 - ilo) / 64) * 64; }

// run f(i,j,k) over the inclusive box [ilo,ihi]x[jlo,jhi]x[klo,khi]
template <class F>
inline void for3(cm1g_model *m, int ilo, int ihi, int jlo, int jhi, int klo, int khi, F f) {
  const int nj = jhi - jlo + 1, nk = khi - klo + 1;
  if (ihi < ilo || nj <= 0 || nk <= 0) return;
  const int i0 = i_origin(ilo), ni = ihi - i0 + 1;
  launch<256>(m, dim3((ni + 63) / 64, (nj + 3) / 4, nk), dim3(64, 4, 1),
              [=] __device__(int bx, int by, int bz) {
                const int i = i0 + bx * 64 + (int)threadIdx.x;
                const int j = jlo + by * 4 + (int)threadIdx.y;
                const int k = klo + bz;
                if (i >= ilo && i <= ihi && j <= jhi) f(i, j, k);
              });
}

// run f(i,j) over [ilo,ihi]x[jlo,jhi] (column kernels loop k inside)
template <class F>
inline void for2(cm1g_model *m, int ilo, int ihi, int jlo, int jhi, F f) {
  const int nj = jhi - jlo + 1;
  if (ihi < ilo || nj <= 0) return;
  const int i0 = i_origin(ilo), ni = ihi - i0 + 1;
  launch<256>(m, dim3((ni + 63) / 64, (nj + 3) / 4, 1), dim3(64, 4, 1),
              [=] __device__(int bx, int by, int) {
                const int i = i0 + bx * 64 + (int)threadIdx.x;
                const int j = jlo + by * 4 + (int)threadIdx.y;
                if (i >= ilo && i <= ihi && j <= jhi) f(i, j);
              });
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

// ---------------------------------------------------------- reductions
// Reduction over a 256-thread workgroup in the pairing order of a halving
// tree, v(t) op v(t+128), then op v(t+64), then op v(t+32) ... v(t+1): the
// two steps that cross wavefronts go through LDS, the six inside the first
// wavefront are shuffles.  The order is fixed, so sums are deterministic;
// the result is valid in thread 0.  lds: 256 elements of T.
template <class T, class OP>
__device__ __forceinline__ T block_reduce256(T x, T *lds, OP op) {
  const int t = (int)threadIdx.x;
  lds[t] = x;
  __syncthreads();
  if (t < 128) lds[t] = op(lds[t], lds[t + 128]);
  __syncthreads();
  T v = x;
  if (t < 64) {
    v = op(lds[t], lds[t + 64]);
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) v = op(v, __shfl_down(v, s, 64));
  }
  return v;
}
struct OpAdd {
  template <class T> __device__ __forceinline__ T operator()(T a, T b) const { return a + b; }
};
struct OpMaxF {
  __device__ __forceinline__ float operator()(float a, float b) const { return fmaxf(a, b); }
};

// ---------------------------------------------------------------- modules
// bc fills (cm1g_bc.hip)
void bc_fill(cm1g_model *m, int kind, float *a);
void bc_fill_n(cm1g_model *m, int kind, float *const *arrs, int n);
void radbc4(cm1g_model *m, const float *ua, const float *u3d, const float *va,
            const float *v3d, float dt);
// advection (cm1g_adv.hip)
// src3 (optional): the u, v, w arrays to advect instead of u3d, v3d, w3d
void adv_uvw(cm1g_model *m, int nrk, int mask, int buoy, int from_ten1 = 0, int own = 0,
             const float *const *src3 = nullptr);
void advm_run(cm1g_model *m, int nrk, int mask, int buoy, int from_ten1, int own,
              const float *const *src3 = nullptr);  // cm1g_advm.hip
void advs_run(cm1g_model *m, int nrk, const float *s_in, float sadd, const float *s0,
              const float *sten0, const float *thterm_div, float *out_sten, float *out_new,
              int pdef, int pdefweno, float dt, double weps, int smeps_flush);
// acoustic (cm1g_sound.hip)
void sound_run(cm1g_model *m, int nrk, bool fold1 = false);
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
