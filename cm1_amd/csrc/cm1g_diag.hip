// Device-side diagnostics: maxmin (src/maxmin.F) over the interior of a
// field, so statistics need no host copy of the array.
#include "cm1g_internal.h"
#include <mutex>

#define ARGCHK(cond, msg)          \
  do {                             \
    if (!(cond)) {                 \
      cm1g_set_error(msg);         \
      return CM1G_E_ARG;           \
    }                              \
  } while (0)

// A float mapped to an unsigned integer of the same order.
__device__ __forceinline__ uint32_t ordered(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// maxmin takes, level by level, the first point (in j, then i) that beats
// the running extreme, then the first level (in k) that beats the others:
// among equal extremes the one with the smallest k, then j, then i.  Keys
// (ordered value, position) with the position inverted make that a plain
// maximum of 64-bit keys for the maximum and for the minimum; NaNs never win
// a comparison there and are skipped here.
__device__ __forceinline__ void mm_cell(float v, uint32_t pos, unsigned long long &kmax,
                                        unsigned long long &kmin) {
  if (v != v) return;
  v = v + 0.0f;  // -0 and +0 compare equal: one key for both
  const unsigned long long a = ((unsigned long long)ordered(v) << 32) | (0xffffffffu - pos);
  const unsigned long long b = ((unsigned long long)(0xffffffffu - ordered(v)) << 32) | (0xffffffffu - pos);
  kmax = a > kmax ? a : kmax;
  kmin = b > kmin ? b : kmin;
}

// maximum of a 64-bit key over the 64 lanes of a wavefront (butterfly)
__device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
  for (int d = 32; d >= 1; d >>= 1) {
    const unsigned long long y = __shfl_xor(x, d, 64);
    x = y > x ? y : x;
  }
  return x;
}

// The two reduced keys back to values and (i,j,k) positions; with no
// comparable value (every point NaN) maxmin's initial values.
static void mm_unpack(const unsigned long long h[2], int izz, int jzz, float *vmax, int32_t *locmax, float *vmin,
                      int32_t *locmin) {
  if (h[0] == 0ULL) {
    *vmax = -1.e30f;
    *vmin = 1.e30f;
    locmax[0] = locmax[1] = locmax[2] = 1;
    locmin[0] = locmin[1] = locmin[2] = 1;
    return;
  }
  auto unpack = [&](unsigned long long kk, bool is_max, float *v, int32_t *loc) {
    uint32_t o = (uint32_t)(kk >> 32);
    if (!is_max) o = 0xffffffffu - o;
    const uint32_t u = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
    std::memcpy(v, &u, sizeof(float));
    const uint32_t pos = 0xffffffffu - (uint32_t)(kk & 0xffffffffu);
    loc[0] = 1 + (int32_t)(pos % (uint32_t)izz);
    loc[1] = 1 + (int32_t)((pos / (uint32_t)izz) % (uint32_t)jzz);
    loc[2] = 1 + (int32_t)(pos / ((uint32_t)izz * (uint32_t)jzz));
  };
  unpack(h[0], true, vmax, locmax);
  unpack(h[1], false, vmin, locmin);
}

extern "C" int cm1g_maxmin(cm1g_model *m, int id, int stag, float *vmax, int32_t *locmax, float *vmin,
                           int32_t *locmin) {
  ARGCHK(m && id >= 0 && id < CM1G_NFIELDS && m->f[id], "cm1g_maxmin: bad field id");
  ARGCHK(vmax && locmax && vmin && locmin, "cm1g_maxmin: null argument");
  ARGCHK(stag >= CM1G_STAG_S && stag <= CM1G_STAG_W, "cm1g_maxmin: bad staggering");
  const KP P = m->P;
  const int izz = P.ni + (stag == CM1G_STAG_U), jzz = P.nj + (stag == CM1G_STAG_V),
            kzz = P.nk + (stag == CM1G_STAG_W);
  ARGCHK((int64_t)izz * jzz * kzz < 0xffffffffLL, "cm1g_maxmin: field too large");
  const float *f = m->f[id];
  unsigned long long *key = (unsigned long long *)m->dred;  // two keys
  launch<64>(m, dim3(1), dim3(1), [=] __device__(int, int, int) {
    key[0] = 0ULL;
    key[1] = 0ULL;
  });
  // each workgroup walks rows (j, k) of the interior, lanes along i; every
  // lane keeps its own two keys, a wavefront combines its lanes' keys with
  // shuffles and issues one atomic maximum per key
  const int rows = jzz * kzz;
  const int nblk = rows < 4096 ? rows : 4096;
  launch<256>(m, dim3(nblk), dim3(256), [=] __device__(int bx, int, int) {
    unsigned long long kmax = 0ULL, kmin = 0ULL;
    for (int r = bx; r < rows; r += nblk) {
      const int j = 1 + r % jzz, k = 1 + r / jzz;
      const uint32_t row0 = (uint32_t)r * (uint32_t)izz;
      for (int i = 1 + (int)threadIdx.x; i <= izz; i += 256)
        mm_cell(f[IXD(P, i, j, k)], row0 + (uint32_t)(i - 1), kmax, kmin);
    }
    kmax = wave_max(kmax);
    kmin = wave_max(kmin);
    if ((threadIdx.x & 63) == 0) {
      if (kmax) atomicMax(&key[0], kmax);
      if (kmin) atomicMax(&key[1], kmin);
    }
  });
  unsigned long long h[2];
  HIPCHK(hipMemcpyAsync(h, key, sizeof(h), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));
  mm_unpack(h, izz, jzz, vmax, locmax, vmin, locmin);
  return 0;
}

// ---------------------------------------------------------------------------
// maxmin without allocation (no counterpart in the reference): the same
// reduction over a caller's device field in this library's layout, with its
// per-workgroup partial keys in a file-scope device array instead of a
// model's workspace.  Two kernels: every workgroup writes its own pair of
// keys (so nothing needs clearing between calls), then one workgroup reduces
// the pairs in a fixed order and writes the final pair after them.
#define MM_NOALLOC_BLOCKS 4096
static __device__ __align__(256) unsigned long long scratch[2 * MM_NOALLOC_BLOCKS + 2];

// per device: the address of that device's copy of scratch (set once, with
// the device current, by hipGetSymbolAddress; never the array's host name)
static unsigned long long *scratch_dev[64];
static std::once_flag scratch_once[64];
static hipError_t scratch_err[64];

__global__ void __launch_bounds__(256) k_mm_partial(const float *__restrict__ f, int px, int pxy, int izz,
                                                    int jzz, int rows, int nblk,
                                                    unsigned long long *__restrict__ part) {
  __shared__ unsigned long long lmax[4], lmin[4];
  unsigned long long kmax = 0ULL, kmin = 0ULL;
  for (int r = blockIdx.x; r < rows; r += nblk) {
    const int j = 1 + r % jzz, k = 1 + r / jzz;
    const uint32_t row0 = (uint32_t)r * (uint32_t)izz;
    const float *row = f + (CM1G_IOFF + (j + 2) * px + (size_t)k * pxy);
    for (int i = 1 + (int)threadIdx.x; i <= izz; i += 256) mm_cell(row[i], row0 + (uint32_t)(i - 1), kmax, kmin);
  }
  kmax = wave_max(kmax);
  kmin = wave_max(kmin);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    lmax[w] = kmax;
    lmin[w] = kmin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = lmax[0], b = lmin[0];
    for (int q = 1; q < 4; ++q) {
      a = lmax[q] > a ? lmax[q] : a;
      b = lmin[q] > b ? lmin[q] : b;
    }
    part[2 * blockIdx.x] = a;
    part[2 * blockIdx.x + 1] = b;
  }
}

__global__ void __launch_bounds__(256) k_mm_final(int nblk, unsigned long long *__restrict__ part) {
  __shared__ unsigned long long lmax[4], lmin[4];
  unsigned long long kmax = 0ULL, kmin = 0ULL;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    const unsigned long long a = part[2 * b], c = part[2 * b + 1];
    kmax = a > kmax ? a : kmax;
    kmin = c > kmin ? c : kmin;
  }
  kmax = wave_max(kmax);
  kmin = wave_max(kmin);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    lmax[w] = kmax;
    lmin[w] = kmin;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = lmax[0], c = lmin[0];
    for (int q = 1; q < 4; ++q) {
      a = lmax[q] > a ? lmax[q] : a;
      c = lmin[q] > c ? lmin[q] : c;
    }
    part[2 * MM_NOALLOC_BLOCKS] = a;
    part[2 * MM_NOALLOC_BLOCKS + 1] = c;
  }
}

extern "C" int cm1g_maxmin_noalloc(const float *field, int ni, int nj, int nk, int stag, void *stream,
                                   float *vmax, int32_t *locmax, float *vmin, int32_t *locmin) {
  ARGCHK(field && vmax && locmax && vmin && locmin, "cm1g_maxmin_noalloc: null argument");
  ARGCHK(stag >= CM1G_STAG_S && stag <= CM1G_STAG_W, "cm1g_maxmin_noalloc: bad staggering");
  ARGCHK(ni >= 1 && nj >= 1 && nk >= 1 && ni <= CM1G_NOALLOC_MAXN && nj <= CM1G_NOALLOC_MAXN &&
             nk <= CM1G_NOALLOC_MAXN,
         "cm1g_maxmin_noalloc: ni, nj and nk must lie in 1..CM1G_NOALLOC_MAXN (16384)");
  const int izz = ni + (stag == CM1G_STAG_U), jzz = nj + (stag == CM1G_STAG_V), kzz = nk + (stag == CM1G_STAG_W);
  ARGCHK((int64_t)izz * jzz * kzz < 0xffffffffLL,
         "cm1g_maxmin_noalloc: more than 2^32-2 interior points (the position key's limit)");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  ARGCHK(dev >= 0 && dev < 64, "cm1g_maxmin_noalloc: device index beyond 63");
  std::call_once(scratch_once[dev], [dev] {
    void *p = nullptr;
    scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    scratch_dev[dev] = (unsigned long long *)p;
  });
  if (scratch_err[dev] != hipSuccess || !scratch_dev[dev]) {
    cm1g_set_error(std::string("cm1g_maxmin_noalloc: hipGetSymbolAddress(scratch): ") +
                   hipGetErrorString(scratch_err[dev]));
    return CM1G_E_HIP;
  }
  unsigned long long *part = scratch_dev[dev];
  const int px = ((ni + CM1G_IOFF + 5) + 63) / 64 * 64, py = nj + 7;  // as cm1g_create lays fields out
  const int rows = jzz * kzz;
  const int nblk = rows < MM_NOALLOC_BLOCKS ? rows : MM_NOALLOC_BLOCKS;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mm_partial, dim3(nblk), dim3(256), 0, s, field, px, px * py, izz, jzz, rows, nblk, part);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_mm_final, dim3(1), dim3(256), 0, s, nblk, part);
  HIPCHK(hipGetLastError());
  unsigned long long h[2];
  HIPCHK(hipMemcpyAsync(h, part + 2 * MM_NOALLOC_BLOCKS, sizeof(h), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  mm_unpack(h, izz, jzz, vmax, locmax, vmin, locmin);
  return 0;
}
