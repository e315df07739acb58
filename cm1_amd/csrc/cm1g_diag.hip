// Device-side diagnostics: maxmin (src/maxmin.F) over the interior of a
// field, so statistics need no host copy of the array.
#include "cm1g_internal.h"
#include <algorithm>
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

// maxmin reports the value it found at the winning position: read it back
// there, so a zero extreme keeps the sign of the point the scan met first
// (the keys hold -0 and +0 as one value).  f: the field's origin, px/pxy
// its strides.
static int mm_value_at(const float *f, int px, int pxy, const int32_t *loc, float *v, hipStream_t s) {
  const size_t c = (size_t)(loc[0] + CM1G_IOFF) + (size_t)(loc[1] + 2) * px + (size_t)loc[2] * pxy;
  HIPCHK(hipMemcpyAsync(v, f + c, sizeof(float), hipMemcpyDeviceToHost, s));
  return 0;
}

extern "C" int cm1g_maxmin(cm1g_model *m, int id, int stag, float *vmax, int32_t *locmax, float *vmin,
                           int32_t *locmin) {
  ARGCHK(m && id >= 0 && id < CM1G_NFIELDS && m->f[id], "cm1g_maxmin: bad field id");
  { const int rc_ = unswap(m); if (rc_) return rc_; }
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
  if (h[0] != 0ULL) {
    int rc = mm_value_at(f, P.px, P.pxy, locmax, vmax, m->stream);
    if (!rc) rc = mm_value_at(f, P.px, P.pxy, locmin, vmin, m->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(m->stream));
  }
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
  if (h[0] != 0ULL) {
    int rc = mm_value_at(field, px, px * py, locmax, vmax, s);
    if (!rc) rc = mm_value_at(field, px, px * py, locmin, vmin, s);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(s));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// statpack (src/statpack.F:22-660): one record of domain statistics of the
// resident state, formed on the device, appended to cm1out_stats.dat with
// its descriptor.  Each max/min pair is a maxmin (src/maxmin.F) of a field
// or of an expression evaluated on the fly, reduced to two 64-bit keys as
// cm1g_maxmin does; the mass budgets (calcmass, totmois, totq,
// src/misclibs.F:3134-3330) add their FP32 cell terms in double precision in
// the reference's order (i fastest, then j, per level; then the levels in
// k order), so they round to the same REAL.

namespace {

// max/min keys of val(i,j,k) over i = 1..izz, j = 1..jzz, k = 1..kzz into
// key[0], key[1] (zeroed beforehand)
template <class V>
void mm_reduce(cm1g_model *m, int izz, int jzz, int kzz, unsigned long long *key, V val) {
  const int rows = jzz * kzz;
  const int nblk = rows < 4096 ? rows : 4096;
  launch<256>(m, dim3(nblk), dim3(256), [=] __device__(int bx, int, int) {
    unsigned long long kmax = 0ULL, kmin = 0ULL;
    for (int r = bx; r < rows; r += nblk) {
      const int j = 1 + r % jzz, k = 1 + r / jzz;
      const uint32_t row0 = (uint32_t)r * (uint32_t)izz;
      for (int i = 1 + (int)threadIdx.x; i <= izz; i += 256)
        mm_cell(val(i, j, k), row0 + (uint32_t)(i - 1), kmax, kmin);
    }
    kmax = wave_max(kmax);
    kmin = wave_max(kmin);
    if ((threadIdx.x & 63) == 0) {
      if (kmax) atomicMax(&key[0], kmax);
      if (kmin) atomicMax(&key[1], kmin);
    }
  });
}

// The value at the winning position of each key (maxmin keeps the value
// found there, so a zero extreme keeps its sign as the reference's scan
// finds it): out[0] for the maximum, out[1] for the minimum.
template <class V>
void mm_pick(cm1g_model *m, int izz, int jzz, const unsigned long long *key, float *out, V val) {
  launch<64>(m, dim3(1), dim3(64), [=] __device__(int, int, int) {
    const int t = (int)threadIdx.x;
    if (t < 2) {
      const unsigned long long kk = key[t];
      if (kk == 0ULL) {
        out[t] = t == 0 ? -1.e30f : 1.e30f;
      } else {
        const uint32_t pos = 0xffffffffu - (uint32_t)(kk & 0xffffffffu);
        const int i = 1 + (int)(pos % (uint32_t)izz), j = 1 + (int)((pos / (uint32_t)izz) % (uint32_t)jzz),
                  k = 1 + (int)(pos / ((uint32_t)izz * (uint32_t)jzz));
        out[t] = val(i, j, k);
      }
    }
  });
}

// Per level k and budget term q: the sum of the term's FP32 cell values in
// double, added in the reference's order (i fastest, then j).  One
// wavefront per (k, q): its lanes form 64 consecutive terms, lane 0 adds
// them in order.
enum { BT_TMASS, BT_TMOIS, BT_Q0 };
__global__ void __launch_bounds__(64) k_budget_ordered(KP P, const float *rho, const float *qv, const float *qc,
                                                        const float *qr, double *sums) {
  __shared__ float t[64];
  const int k = 1 + (int)blockIdx.x, term = (int)blockIdx.y;
  const int ni = P.ni, n = P.ni * P.nj;
  const float *q = term == BT_Q0 ? qv : term == BT_Q0 + 1 ? qc : qr;
  double s = 0.0;
  for (int t0 = 0; t0 < n; t0 += 64) {
    const int l = t0 + (int)threadIdx.x;
    float v = 0.0f;
    if (l < n) {
      const int i = 1 + l % ni, j = 1 + l / ni;
      const int c = IXD(P, i, j, k);
      float w;
      if (term == BT_TMASS) {
        w = rho[c];
      } else if (term == BT_TMOIS) {
        float ql = 0.0f;  // getqli (src/misclibs.F:923-967): liquid qc + qr, no ice
        ql = ql + qc[c];
        ql = ql + qr[c];
        const float qi = 0.0f;
        w = rho[c] * (qv[c] + ql + qi);
      } else {
        w = rho[c] * q[c];
      }
      v = w * P.ruh[i + 2] * P.rvh[j + 2] * P.rmh[k];
    }
    t[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      const int e = n - t0 < 64 ? n - t0 : 64;
      for (int a = 0; a < e; a++) s = s + (double)t[a];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) sums[(size_t)term * (P.nk + 1) + k] = s;
}

// One statistic of the record: its descriptor entry and where its value
// comes from (a key's value, the level of a key's position, a budget sum,
// the model time).
struct StatEntry {
  std::string name, desc, unit;
  int src;     // 0 mtime, 1 key value, 2 zf at the key's level, 3 zh at the key's level, 4 budget
  int key;     // key index (max = 2p, min = 2p+1) or budget term
};

std::string fpad(const std::string &s, size_t w) {
  std::string r = s;
  r.resize(w, ' ');
  return r;
}

}  // namespace

extern "C" int cm1g_statpack(cm1g_model *m, const char *dir, int nrec, uint32_t mask, double mtime,
                             float *rstat, int32_t *nstat) {
  ARGCHK(m && dir && nrec >= 1, "cm1g_statpack: null argument or nrec < 1");
  ARGCHK((mask & ~(uint32_t)(CM1G_STAT_ALL | CM1G_STAT_TURB)) == 0,
         "cm1g_statpack: unknown statistics switch in mask");
  ARGCHK(m->turb.cm1setup == 1 || (mask & CM1G_STAT_TURB) == 0,
         "cm1g_statpack: turbulence statistics need the subgrid turbulence (cm1g_set_turbulence)");
  // with the semi-slip surface the reference's stat_wsp adds the 10 m wind
  // and its stat_cfl the diffusion numbers (src/statpack.F:453-461, calccfl),
  // which are not formed here
  ARGCHK(m->turb.cm1setup != 1 || (mask & (CM1G_STAT_WSP | CM1G_STAT_CFL)) == 0,
         "cm1g_statpack: stat_wsp and stat_cfl with the subgrid turbulence are not supported");
  const KP P = m->P;
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  ARGCHK((int64_t)(ni + 1) * (nj + 1) * (nk + 1) < 0xffffffffLL, "cm1g_statpack: grid too large");
  const bool moist = m->c.imoist == 1;
  auto on = [&](uint32_t b) { return (mask & b) != 0; };
  const int nterm = BT_Q0 + 3;
  const size_t nkeys = 2 * CM1G_STAT_MAXPAIRS, nsum = (size_t)nterm * (nk + 1);
  const size_t bytes = nkeys * sizeof(unsigned long long) + nkeys * sizeof(float) + nsum * sizeof(double);
  if (bytes > m->stat_cap) {
    if (m->stat_buf) hipFree(m->stat_buf);
    m->stat_buf = nullptr;
    m->stat_cap = 0;
    if (hipMalloc(&m->stat_buf, bytes) != hipSuccess) {
      cm1g_set_error("cm1g_statpack: cannot allocate the reduction buffer");
      return CM1G_E_NOMEM;
    }
    m->stat_cap = bytes;
  }
  unsigned long long *keys = (unsigned long long *)m->stat_buf;
  float *picks = (float *)(keys + nkeys);
  double *sums = (double *)(picks + nkeys);  // nkeys is even: 8-byte aligned
  launch<256>(m, dim3(1), dim3(256), [=] __device__(int, int, int) {
    for (size_t t = threadIdx.x; t < nkeys; t += 256) keys[t] = 0ULL;
  });

  std::vector<StatEntry> E;
  std::vector<int> box;  // izz, jzz of each max/min pair (for the key positions)
  E.push_back({"mtime", "model time (seconds since beginning of simulation)", "s", 0, 0});
  // a max/min pair of val over the box, with its two descriptor entries
  auto pair = [&](int izz, int jzz, int kzz, const std::string &nmax, const std::string &dmax,
                  const std::string &nmin, const std::string &dmin, const std::string &unit, auto val) {
    const int p = (int)box.size() / 2;
    if (p < CM1G_STAT_MAXPAIRS) {
      mm_reduce(m, izz, jzz, kzz, keys + 2 * p, val);
      mm_pick(m, izz, jzz, keys + 2 * p, picks + 2 * p, val);
    }
    box.push_back(izz);
    box.push_back(jzz);
    E.push_back({nmax, dmax, unit, 1, 2 * p});
    E.push_back({nmin, dmin, unit, 1, 2 * p + 1});
    return p;
  };
  const float *wa = m->f[CM1G_WA], *ua = m->f[CM1G_UA], *va = m->f[CM1G_VA], *tha = m->f[CM1G_THA],
              *ppi = m->f[CM1G_PPI], *prs = m->f[CM1G_PRS];
  const float *prate = m->d2[CM1G_D_PRATE];
  const float *zf = m->hk[CM1G_K_ZF], *zh = m->hk[CM1G_K_ZH];
  if (on(CM1G_STAT_W)) {
    const int p = pair(ni, nj, nk + 1, "wmax", "max vertical velocity", "wmin", "min vertical velocity", "m/s",
                       [=] __device__(int i, int j, int k) { return wa[IXD(P, i, j, k)]; });
    E.push_back({"zwmax", "level of max vertical velocity", "m AGL", 2, 2 * p});
    E.push_back({"zwmin", "level of min vertical velocity", "m AGL", 2, 2 * p + 1});
  }
  if (on(CM1G_STAT_WLEVS) && zf[nk + 1] >= 10000.0f) {
    // w interpolated to 500 m ... 10 km AGL (src/statpack.F:147-210)
    const float zl[5] = {500.0f, 1000.0f, 2500.0f, 5000.0f, 10000.0f};
    const char *tag[5] = {"500", "1000", "2500", "5000", "10k"};
    const char *dtag[5] = {"500 m", "1000 m", "2500 m", "5000 m", "10 km"};
    for (int q = 0; q < 5; q++) {
      const float zlev = zl[q];
      int nkm = nk + 1;
      float r1 = 0.0f, r2 = 0.0f;
      const bool inside = zlev < zf[nk + 1];
      if (inside) {
        while (zf[nkm] > zlev && nkm > 1) nkm--;
        r2 = (zlev - zf[nkm]) / (zf[nkm + 1] - zf[nkm]);
        r1 = 1.0f - r2;
      }
      const int kk = nkm;
      pair(ni, nj, 1, std::string("wmax") + tag[q], std::string("max vertical velocity at ") + dtag[q] + " AGL",
           std::string("wmin") + tag[q], std::string("min vertical velocity at ") + dtag[q] + " AGL", "m/s",
           [=] __device__(int i, int j, int) {
             return inside ? r1 * wa[IXD(P, i, j, kk)] + r2 * wa[IXD(P, i, j, kk + 1)] : 0.0f;
           });
    }
  }
  if (on(CM1G_STAT_U)) {
    pair(ni + 1, nj, nk, "umax", "max u velocity (grid-rel)", "umin", "min u velocity (grid-rel)", "m/s",
         [=] __device__(int i, int j, int k) { return ua[IXD(P, i, j, k)]; });
    pair(ni + 1, nj, 1, "sumax", "max u velocity at lowst mod lev (grid-rel)", "sumin",
         "min u velocity at lowst mod lev (grid-rel)", "m/s",
         [=] __device__(int i, int j, int) { return ua[IXD(P, i, j, 1)]; });
  }
  if (on(CM1G_STAT_V)) {
    pair(ni, nj + 1, nk, "vmax", "max v velocity (grid-rel)", "vmin", "min v velocity (grid-rel)", "m/s",
         [=] __device__(int i, int j, int k) { return va[IXD(P, i, j, k)]; });
    pair(ni, nj + 1, 1, "svmax", "max v velocity at lowst mod lev (grid-rel)", "svmin",
         "min v velocity at lowst mod lev (grid-rel)", "m/s",
         [=] __device__(int i, int j, int) { return va[IXD(P, i, j, 1)]; });
  }
  if (on(CM1G_STAT_PIPERT))
    pair(ni, nj, nk, "ppimax", "max nondimensional pressure perturbation", "ppimin",
         "min nondimensional pressure perturbation", "nondimensional",
         [=] __device__(int i, int j, int k) { return ppi[IXD(P, i, j, k)]; });
  if (on(CM1G_STAT_PRSPERT)) {
    // prs - p00*pi0**cpdrd: prs0 of cm1g_base_derive is that expression
    // with the host library's powf, as the reference evaluates it
    const float *prs0 = m->kf[CM1G_K_PRS0];
    pair(ni, nj, nk, "ppmax", "max pressure perturbation", "ppmin", "min pressure perturbation", "Pa",
         [=] __device__(int i, int j, int k) { return prs[IXD(P, i, j, k)] - prs0[k]; });
  }
  if (on(CM1G_STAT_THPERT)) {
    pair(ni, nj, nk, "thpmax", "max potential temperature perturbation", "thpmin",
         "min potential temperature perturbation", "K",
         [=] __device__(int i, int j, int k) { return tha[IXD(P, i, j, k)]; });
    pair(ni, nj, 1, "sthpmax", "max pot. temp. pert. at lowest model level", "sthpmin",
         "min pot. temp. pert. at lowest model level", "K",
         [=] __device__(int i, int j, int) { return tha[IXD(P, i, j, 1)]; });
  }
  if (on(CM1G_STAT_Q) && moist) {
    const char *qn[3] = {"qv", "qc", "qr"};
    for (int n = 0; n < 3; n++) {
      const float *q = m->f[CM1G_QA0 + n];
      pair(ni, nj, nk, std::string("max") + qn[n], std::string("max ") + qn[n], std::string("min") + qn[n],
           std::string("min ") + qn[n], "kg/kg", [=] __device__(int i, int j, int k) { return q[IXD(P, i, j, k)]; });
    }
    pair(ni, nj, 1, "pratemax", "maximum surface precipitation rate", "pratemin",
         "minimum surface precipitation rate", "kg/m2/s",
         [=] __device__(int i, int j, int) { return prate[IX2D(P, i, j)]; });
  }
  // the subgrid turbulence (src/statpack.F:258-278): w-level fields, nk+1 levels
  auto wpair = [&](int id, const char *nmax, const char *dmax, const char *nmin, const char *dmin,
                   const char *unit) {
    const float *a = m->f[id];
    pair(ni, nj, nk + 1, nmax, dmax, nmin, dmin, unit,
         [=] __device__(int i, int j, int k) { return a[IXD(P, i, j, k)]; });
  };
  if (on(CM1G_STAT_TKE))
    wpair(CM1G_TKEA, "tkemax", "max subgrid tke", "tkemin", "min subgrid tke", "m^2/s^2");
  if (on(CM1G_STAT_KM)) {
    wpair(CM1G_KMH, "kmhmax", "max horiz eddy viscosity for momentum", "kmhmin",
          "min horiz eddy viscosity for momentum", "m^2/s");
    wpair(CM1G_KMV, "kmvmax", "max vert eddy viscosity for momentum", "kmvmin",
          "min vert eddy viscosity for momentum", "m^2/s");
  }
  if (on(CM1G_STAT_KH)) {
    wpair(CM1G_KHH, "khhmax", "max horiz eddy diffusivity for scalars", "khhmin",
          "min horiz eddy diffusivity for scalars", "m^2/s");
    wpair(CM1G_KHV, "khvmax", "max vert eddy diffusivity for scalars", "khvmin",
          "min vert eddy diffusivity for scalars", "m^2/s");
  }
  for (int n = 0; n < m->c.npt; n++) {  // passive tracers (iptra = 1)
    const float *pt = m->f[CM1G_PTA0 + n];
    const std::string s = std::to_string(n + 1);
    pair(ni, nj, nk, "maxpt" + s, "max pt" + s, "minpt" + s, "min pt" + s, "kg/kg",
         [=] __device__(int i, int j, int k) { return pt[IXD(P, i, j, k)]; });
  }
  if (on(CM1G_STAT_SFCPRS)) {
    pair(ni, nj, 1, "sprsmax", "max pressure at lowest model level", "sprsmin", "min pressure at lowest model level",
         "Pa", [=] __device__(int i, int j, int) { return prs[IXD(P, i, j, 1)]; });
    pair(ni, nj, 1, "psfcmax", "max pressure at surface", "psfcmin", "min pressure at surface", "Pa",
         [=] __device__(int i, int j, int) {
           return P.cgs1 * prs[IXD(P, i, j, 1)] + P.cgs2 * prs[IXD(P, i, j, 2)] + P.cgs3 * prs[IXD(P, i, j, 3)];
         });
  }
  if (on(CM1G_STAT_WSP)) {
    // grid-relative horizontal speed at scalar points (umove = vmove = 0)
    auto wsp = [=] __device__(int i, int j, int k) {
      const int c = IXD(P, i, j, k);
      const float a = 0.0f + 0.5f * (ua[c] + ua[c + 1]);
      const float b = 0.0f + 0.5f * (va[c] + va[c + P.px]);
      return sqrtf(a * a + b * b);
    };
    const int p = pair(ni, nj, nk, "wspmax", "max grid-rel horiz wind speed", "wspmin",
                       "min grid-rel horiz wind speed", "m/s", wsp);
    E.push_back({"zwspmax", "level of max grid-rel horiz wind speed", "m AGL", 3, 2 * p});
    E.push_back({"zwspmin", "level of min grid-rel horiz wind speed", "m AGL", 3, 2 * p + 1});
    pair(ni, nj, 1, "swspmax", "max grid-rel horiz wind speed at l.m.l.", "swspmin",
         "min grid-rel horiz wind speed at l.m.l.", "m/s", [=] __device__(int i, int j, int) { return wsp(i, j, 1); });
  }
  if (m->turb.cm1setup == 1 && m->turb.bbc == 3) {
    // with a surface model (bbc = 3, sfcmodel = 1) the record always holds
    // the extremes of the diagnosed PBL depth (src/statpack.F:465-467);
    // without a PBL scheme and outside the reference's LES test cases
    // (testcase 1-7, src/turb.F:563-583) hpbl keeps the 100 m it is
    // allocated with (src/cm1.F:2004)
    pair(ni, nj, 1, "hpblmax", "max diagnosed pbl depth", "hpblmin", "min diagnosed pbl depth", "m",
         [=] __device__(int, int, int) { return 100.0f; });
  }
  if (on(CM1G_STAT_CFL)) {
    // calccfl (src/misclibs.F:2584-2718) for nx > 1 and ny > 1: only the
    // maximum is recorded
    const float dt = m->c.dt;
    const float dtdx = 0.5f * dt * P.rdx, dtdy = 0.5f * dt * P.rdy, dtdz = 0.5f * dt * P.rdz;
    pair(ni, nj, nk, "cflmax", "max Courant number", "", "", "nondimensional", [=] __device__(int i, int j, int k) {
      const int c = IXD(P, i, j, k);
      const float a = (ua[c] + ua[c + 1]) * dtdx * P.uh[i + 2];
      const float b = (va[c] + va[c + P.px]) * dtdy * P.vh[j + 2];
      const float d = (wa[c] + wa[c + P.pxy]) * dtdz * P.mh[k];
      return sqrtf((a * a + b * b) + d * d);
    });
    E.pop_back();  // no minimum
  }
  const int npair = (int)box.size() / 2;
  if (npair > CM1G_STAT_MAXPAIRS) {
    cm1g_set_error("cm1g_statpack: more than CM1G_STAT_MAXPAIRS max/min pairs");
    return CM1G_E_ARG;
  }
  const bool want_budget = on(CM1G_STAT_TMASS) || (moist && (on(CM1G_STAT_TMOIS) || on(CM1G_STAT_QMASS)));
  if (on(CM1G_STAT_TMASS)) E.push_back({"tmass", "total mass of dry air", "kg", 4, BT_TMASS});
  if (moist && on(CM1G_STAT_TMOIS)) E.push_back({"tmois", "total mass of moisture", "kg", 4, BT_TMOIS});
  if (moist && on(CM1G_STAT_QMASS)) {
    const char *qn[3] = {"qv", "qc", "qr"};
    for (int n = 0; n < 3; n++)
      E.push_back({std::string("mass") + qn[n], std::string("total mass of ") + qn[n], "kg", 4, BT_Q0 + n});
  }
  if (want_budget)
    hipLaunchKernelGGL(k_budget_ordered, dim3(nk, moist ? nterm : 1), dim3(64), 0, m->stream, P,
                       m->f[CM1G_RHO], m->f[CM1G_QA0], m->f[CM1G_QA1], m->f[CM1G_QA2], sums);
  int rc = check_launch();
  if (rc) return rc;
  std::vector<unsigned long long> hk(2 * (size_t)npair + 2);
  std::vector<float> hp(2 * (size_t)npair + 2);
  std::vector<double> hs(nsum);
  if (npair) {
    HIPCHK(hipMemcpyAsync(hk.data(), keys, 2 * (size_t)npair * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          m->stream));
    HIPCHK(hipMemcpyAsync(hp.data(), picks, 2 * (size_t)npair * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  }
  if (want_budget) HIPCHK(hipMemcpyAsync(hs.data(), sums, nsum * sizeof(double), hipMemcpyDeviceToHost, m->stream));
  HIPCHK(hipStreamSynchronize(m->stream));

  // the record (src/statpack.F:123-640) and the totals' final scaling
  const float vol = m->c.dx * m->c.dy * m->c.dz;  // dx*dy*dz in REAL
  std::vector<float> rec;
  for (const StatEntry &e : E) {
    float v = 0.0f;
    if (e.src == 0) {
      v = (float)mtime;
    } else if (e.src == 4) {
      double t = 0.0;
      for (int k = 1; k <= nk; k++) t = t + hs[(size_t)e.key * (nk + 1) + k];
      t = t * (double)vol;
      v = (float)t;
    } else {
      const int p = e.key / 2;
      float vmax, vmin;
      int32_t lmax[3], lmin[3];
      const unsigned long long h[2] = {hk[2 * p], hk[2 * p + 1]};
      mm_unpack(h, box[2 * p], box[2 * p + 1], &vmax, lmax, &vmin, lmin);
      const bool is_max = (e.key % 2) == 0;
      if (e.src == 1) v = hp[e.key];
      else if (e.src == 2) v = zf[is_max ? lmax[2] : lmin[2]];
      else v = zh[is_max ? lmax[2] : lmin[2]];
    }
    rec.push_back(v);
  }
  const int n = (int)rec.size();
  if (rstat) std::memcpy(rstat, rec.data(), std::min(n, CM1G_STAT_MAXN) * sizeof(float));
  if (nstat) *nstat = n;

  // cm1out_stats.dat: record nrec of n REAL*4 values; the descriptor
  // (write_statsctl, src/statpack.F:1460-1530) for nrec times
  const std::string d(dir);
  const std::string dat = d + "/cm1out_stats.dat";
  FILE *f = fopen(dat.c_str(), nrec == 1 ? "wb" : "r+b");
  if (!f) f = fopen(dat.c_str(), "w+b");
  bool ok = f && fseeko(f, (off_t)(nrec - 1) * (off_t)n * 4, SEEK_SET) == 0 &&
            fwrite(rec.data(), sizeof(float), (size_t)n, f) == (size_t)n;
  if (f) ok = (fclose(f) == 0) && ok;
  if (!ok) {
    cm1g_set_error("cm1g_statpack: cannot write " + dat);
    return CM1G_E_IO;
  }
  FILE *c = fopen((d + "/cm1out_stats.ctl").c_str(), "w");
  if (!c) {
    cm1g_set_error("cm1g_statpack: cannot write cm1out_stats.ctl");
    return CM1G_E_IO;
  }
  fprintf(c, "dset ^%s\n", fpad("cm1out_stats.dat", 120).c_str());
  fprintf(c, "title CM1 stats output, using version cm1r21.1; time is generic, see variable mtime for actual times\n");
  fprintf(c, "undef %10.1f\n", -999999.9);
  fprintf(c, "xdef 1 linear 0 1\nydef 1 linear 0 1\nzdef 1 linear 0 1\n");
  fprintf(c, "tdef %10d linear 00:00Z01JAN0001 1YR\n", nrec);
  fprintf(c, "vars %6d\n", n);
  for (const StatEntry &e : E) {
    std::string a16(16, ' ');  // '(' unit(1:14) ')' (src/statpack.F:1510-1515)
    a16.replace(1, 14, fpad(e.unit, 14));
    a16[0] = '(';
    const size_t nn = e.unit.size();
    if (nn + 1 < 16) a16[nn + 1] = ')';
    fprintf(c, "%s 1 99 %s %s\n", fpad(e.name, 12).c_str(), fpad(e.desc, 40).c_str(), a16.c_str());
  }
  fprintf(c, "endvars\n");
  if (fclose(c) != 0) {
    cm1g_set_error("cm1g_statpack: error closing cm1out_stats.ctl");
    return CM1G_E_IO;
  }
  return 0;
}
