// The per-cell MLP of the learned Kessler surrogate: sizes, parameter
// layout and the fused inference routine (cm1g_mlp.hip has the rest).
#pragma once
#include "cm1g_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int MLP_IN = 8, MLP_H = 128, MLP_OUT = 4;
// flat FP32 parameters, weights row-major [in][out]
constexpr int OW1 = 0, OB1 = OW1 + MLP_IN * MLP_H, OW2 = OB1 + MLP_H, OB2 = OW2 + MLP_H * MLP_H,
              OW3 = OB2 + MLP_H, OB3 = OW3 + MLP_H * MLP_H, OW4 = OB3 + MLP_H,
              OB4 = OW4 + MLP_H * MLP_OUT, MLP_NPARAM = OB4 + MLP_OUT;
// normalisation: x = (raw - xmean) / xstd, raw output = y * ystd + ymean;
// then the range of normalised inputs the network was trained on
// (xmean[8] xstd[8] ymean[4] ystd[4] xlo[8] xhi[8])
constexpr int MLP_NNORM = 4 * MLP_IN + 2 * MLP_OUT;
constexpr int NM_XLO = 2 * MLP_IN + 2 * MLP_OUT, NM_XHI = NM_XLO + MLP_IN;

struct cm1g_mlp;
const float *mlp_params(const cm1g_mlp *n);
const float *mlp_norm(const cm1g_mlp *n);

__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float silu_f(float z) { return z / (1.0f + expf(-z)); }

// LDS of the fused inference: BF16 weights transposed to [out][in] (the four
// inputs a lane feeds one matrix instruction are then contiguous), rows
// padded by 4 against bank conflicts; per wavefront a 16-sample activation
// block and the 16 x 4 outputs.
constexpr int MLP_P1 = 20, MLP_PH = 132;
constexpr int MLP_WAVES = 8;  // wavefronts per workgroup (512 threads)
struct MlpLds {
  unsigned short w1[MLP_H * MLP_P1];
  unsigned short w2[MLP_H * MLP_PH];
  unsigned short w3[MLP_H * MLP_PH];
  unsigned short w4[16 * MLP_PH];
  float b[3 * MLP_H + 16];
  unsigned short act[MLP_WAVES][16 * MLP_PH];
  float out[MLP_WAVES][16 * MLP_OUT];
};

// one hidden or output layer of a 16-sample block: act (16 x K, BF16) times
// W^T rows of `w` (pitch pitch), nt output tiles of 16; SiLU on hidden layers
// writes back into act, the last layer leaves its 16 x 4 block in out
template <int KSTEPS, int NT, bool LAST>
__device__ __forceinline__ void mlp_layer(unsigned short *act, const unsigned short *w, int pitch, const float *bias,
                                          float *out, int l) {
  bf16x4 a[KSTEPS];
  for (int ks = 0; ks < KSTEPS; ks++)
    for (int t = 0; t < 4; t++) a[ks][t] = (short)act[(l & 15) * MLP_PH + 16 * ks + 4 * (l >> 4) + t];
  __syncthreads();
  for (int nt = 0; nt < NT; nt++) {
    floatx4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ks = 0; ks < KSTEPS; ks++) {
      bf16x4 b;
      const unsigned short *pb = &w[(16 * nt + (l & 15)) * pitch + 16 * ks + 4 * (l >> 4)];
      for (int t = 0; t < 4; t++) b[t] = (short)pb[t];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[ks], b, acc, 0, 0, 0);
    }
    const int col = 16 * nt + (l & 15);
    for (int t = 0; t < 4; t++) {
      const int row = 4 * (l >> 4) + t;
      const float z = acc[t] + bias[col];
      if (LAST) {
        if (col < MLP_OUT) out[row * MLP_OUT + col] = z;
      } else {
        act[row * MLP_PH + col] = bf16_rne(silu_f(z));
      }
    }
  }
  __syncthreads();
}

// Run the network over samples 0..n-1: in(s, x) gives the 8 normalised
// inputs of sample s, out(s, y) takes its 4 normalised outputs.  Workgroup
// bx of nbx (512 threads) stages the weights once, then its wavefronts take
// 16-sample blocks bx*8 + w, bx*8 + w + nbx*8, ...
template <class IN, class OUT>
__device__ void mlp_tiles(const float *__restrict__ prm, int64_t n, const IN &in, const OUT &out, int bx, int nbx) {
  __shared__ MlpLds S;
  const int tid = (int)threadIdx.x, l = tid & 63, w = tid >> 6;
  for (int e = tid; e < MLP_H * 16; e += 512) {  // e = k*128 + n over the padded 16 inputs
    const int k = e / MLP_H, nn = e % MLP_H;
    S.w1[nn * MLP_P1 + k] = k < MLP_IN ? bf16_rne(prm[OW1 + k * MLP_H + nn]) : (unsigned short)0;
  }
  for (int e = tid; e < MLP_H * MLP_H; e += 512) {
    const int k = e / MLP_H, nn = e % MLP_H;
    S.w2[nn * MLP_PH + k] = bf16_rne(prm[OW2 + e]);
    S.w3[nn * MLP_PH + k] = bf16_rne(prm[OW3 + e]);
  }
  for (int e = tid; e < MLP_H * 16; e += 512) {  // e = k*16 + n over 16 padded outputs
    const int k = e / 16, nn = e % 16;
    S.w4[nn * MLP_PH + k] = nn < MLP_OUT ? bf16_rne(prm[OW4 + k * MLP_OUT + nn]) : (unsigned short)0;
  }
  for (int e = tid; e < 3 * MLP_H + 16; e += 512) {
    float v = 0.0f;
    if (e < MLP_H) v = prm[OB1 + e];
    else if (e < 2 * MLP_H) v = prm[OB2 + e - MLP_H];
    else if (e < 3 * MLP_H) v = prm[OB3 + e - 2 * MLP_H];
    else if (e < 3 * MLP_H + MLP_OUT) v = prm[OB4 + e - 3 * MLP_H];
    S.b[e] = v;
  }
  __syncthreads();
  unsigned short *act = S.act[w];
  float *o = S.out[w];
  const int64_t nblk = (n + 15) / 16;
  // every wavefront runs the same number of rounds (the barriers inside)
  for (int64_t base = (int64_t)bx * MLP_WAVES; base < nblk; base += (int64_t)nbx * MLP_WAVES) {
    const int64_t blk = base + w;
    const int64_t s = blk * 16 + l;
    const bool live = (blk < nblk) && (l < 16) && (s < n);
    if (l < 16) {
      float x[MLP_IN];
      for (int k = 0; k < MLP_IN; k++) x[k] = 0.0f;
      if (live) in(s, x);
      for (int k = 0; k < MLP_IN; k++) act[l * MLP_PH + k] = bf16_rne(x[k]);
      for (int k = MLP_IN; k < 16; k++) act[l * MLP_PH + k] = 0;
    }
    __syncthreads();
    mlp_layer<1, 8, false>(act, S.w1, MLP_P1, S.b, o, l);
    mlp_layer<8, 8, false>(act, S.w2, MLP_PH, S.b + MLP_H, o, l);
    mlp_layer<8, 8, false>(act, S.w3, MLP_PH, S.b + 2 * MLP_H, o, l);
    mlp_layer<8, 1, true>(act, S.w4, MLP_PH, S.b + 3 * MLP_H, o, l);
    if (live) out(s, &o[l * MLP_OUT]);
    __syncthreads();
  }
}
