// Device-side diagnostics: maxmin (src/maxmin.F) over the interior of a
// field, so statistics need no host copy of the array.
#include "cm1g_internal.h"

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
  if (h[0] == 0ULL) {  // no comparable value (all NaN): maxmin's initial values
    *vmax = -1.e30f;
    *vmin = 1.e30f;
    locmax[0] = locmax[1] = locmax[2] = 1;
    locmin[0] = locmin[1] = locmin[2] = 1;
    return 0;
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
  return 0;
}
