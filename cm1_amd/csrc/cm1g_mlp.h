// The per-cell MLP of the learned Kessler surrogate: sizes, parameter
// layout and the fused inference routine (cm1g_mlp.hip has the rest).
#pragma once
#include "cm1g_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8s __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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
// SiLU with the hardware exponential and reciprocal (training and inference
// use the same form)
__device__ __forceinline__ float silu_f(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// Fused inference (the first layer with v_mfma_f32_16x16x16_bf16, the
// others with v_mfma_f32_16x16x32_bf16).  Each layer is computed
// transposed, Y^T = W^T X^T: the
// weight block is the first matrix operand and the 16-sample activation
// block the second, so a lane's result, Y(sample l%16, outputs 16nt +
// 4(l/16) + t), is exactly the second operand the next layer needs for
// k-step nt.  Activations therefore stay in registers from layer to layer;
// only the weights sit in LDS, as BF16 in fragment order (for output tile
// nt and k-step ks the 64 lanes' 4-element fragments lie consecutively,
// lane l at ((nt*KS + ks)*64 + l)*4, one 8-byte load per lane).
constexpr int MLP_WAVES = 8;          // wavefronts per workgroup (512 threads)
constexpr int MLP_G = 1;              // 16-sample groups per wavefront and round
constexpr int MLP_F1 = 8 * 1 * 256;   // layer 1: 8 output tiles x 1 k-step of 16 x 64 lanes x 4
constexpr int MLP_FH = 8 * 4 * 512;   // hidden layers: 8 x 4 k-steps of 32 x 64 lanes x 8
constexpr int MLP_F4 = 1 * 4 * 512;   // output layer: 1 (4 of 16 outputs used) x 4
struct MlpLds {
  bf16x4 w1[MLP_F1 / 4];
  bf16x8s w2[MLP_FH / 8];
  bf16x8s w3[MLP_FH / 8];
  bf16x8s w4[MLP_F4 / 8];
  float b[3 * MLP_H + 16];
};

// fragment-order index of element (k, n) of a weight with KS k-steps
__device__ __forceinline__ int mlp_wfrag(int k, int n, int ks_count) {
  return (((n >> 4) * ks_count + (k >> 4)) * 64 + (n & 15) + 16 * ((k & 15) >> 2)) * 4 + (k & 3);
}

// The 32-deep instruction (v_mfma_f32_16x16x32_bf16) reads lane l's 8
// second-operand values as k = 8(l/16) + p, p = 0..7.  A lane holds the
// previous layer's outputs 16nt + 4(l/16) + t for two consecutive tiles,
// i.e. inputs 4(l/16) + t and 16 + 4(l/16) + t of a 32-input block: so the
// weights are staged with the same permutation of k, and position p of
// lane l's fragment is input 4(l/16) + p (p < 4) or 16 + 4(l/16) + p - 4.
__device__ __forceinline__ int mlp_wfrag32(int k, int n, int kb_count) {
  const int kl = k & 31;
  const int g = (kl & 15) >> 2, p = (kl & 3) + (kl >= 16 ? 4 : 0);
  return (((n >> 4) * kb_count + (k >> 5)) * 64 + (n & 15) + 16 * g) * 8 + p;
}

// four floats to BF16 (round to nearest, ties to even: the hardware's
// packed conversion, the same rounding as bf16_rne for finite values)
__device__ __forceinline__ bf16x4 mlp_pack(float a, float b, float c, float d) {
  bf16x4 v;
  v[0] = __builtin_bit_cast(short, (__bf16)a);
  v[1] = __builtin_bit_cast(short, (__bf16)b);
  v[2] = __builtin_bit_cast(short, (__bf16)c);
  v[3] = __builtin_bit_cast(short, (__bf16)d);
  return v;
}

// accumulators start from the bias of the outputs each lane receives
__device__ __forceinline__ void mlp_bias_init(floatx4 (&acc)[MLP_G][8], const float *bias, int l) {
#pragma unroll
  for (int nt = 0; nt < 8; nt++) {
    const int c0 = 16 * nt + 4 * (l >> 4);
    const floatx4 b = {bias[c0], bias[c0 + 1], bias[c0 + 2], bias[c0 + 3]};
#pragma unroll
    for (int g = 0; g < MLP_G; g++) acc[g][nt] = b;
  }
}
__device__ __forceinline__ void mlp_silu_out(const floatx4 (&acc)[MLP_G][8], bf16x4 (&x)[MLP_G][8]) {
#pragma unroll
  for (int nt = 0; nt < 8; nt++)
#pragma unroll
    for (int g = 0; g < MLP_G; g++)
      x[g][nt] = mlp_pack(silu_f(acc[g][nt][0]), silu_f(acc[g][nt][1]), silu_f(acc[g][nt][2]),
                          silu_f(acc[g][nt][3]));
}
// lane l's 8 inputs of 32-input block kb: outputs of tiles 2kb and 2kb+1
__device__ __forceinline__ bf16x8 mlp_join(const bf16x4 &lo, const bf16x4 &hi) {
  const bf16x8s v = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  return __builtin_bit_cast(bf16x8, v);
}

// first layer (8 inputs padded to one 16-deep step): x[g][0] -> x[g][nt]
__device__ __forceinline__ void mlp_first(bf16x4 (&x)[MLP_G][8], const bf16x4 *wf, const float *bias, int l) {
  floatx4 acc[MLP_G][8];
  mlp_bias_init(acc, bias, l);
#pragma unroll
  for (int nt = 0; nt < 8; nt++) {
    const bf16x4 w = wf[nt * 64 + l];
#pragma unroll
    for (int g = 0; g < MLP_G; g++)
      acc[g][nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w, x[g][0], acc[g][nt], 0, 0, 0);
  }
  mlp_silu_out(acc, x);
}

// a hidden layer (128 -> 128) in four 32-deep steps.  The next step's
// weight fragments load while this one's matrix instructions run; the
// scheduling barrier keeps the compiler from hoisting every step's loads at
// once (which would exceed the registers).
__device__ __forceinline__ void mlp_hidden(bf16x4 (&x)[MLP_G][8], const bf16x8s *wf, const float *bias, int l) {
  floatx4 acc[MLP_G][8];
  mlp_bias_init(acc, bias, l);
  bf16x8s wc[8], wn[8];
#pragma unroll
  for (int nt = 0; nt < 8; nt++) wc[nt] = wf[(nt * 4) * 64 + l];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    if (kb + 1 < 4) {
#pragma unroll
      for (int nt = 0; nt < 8; nt++) wn[nt] = wf[(nt * 4 + kb + 1) * 64 + l];
    }
    bf16x8 xb[MLP_G];
#pragma unroll
    for (int g = 0; g < MLP_G; g++) xb[g] = mlp_join(x[g][2 * kb], x[g][2 * kb + 1]);
#pragma unroll
    for (int nt = 0; nt < 8; nt++)
#pragma unroll
      for (int g = 0; g < MLP_G; g++)
        acc[g][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, wc[nt]), xb[g],
                                                              acc[g][nt], 0, 0, 0);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int nt = 0; nt < 8; nt++) wc[nt] = wn[nt];
  }
  mlp_silu_out(acc, x);
}

// Run the network over samples 0..n-1.  load(s, r) fills r[MLP_KEEP] with
// the raw values sample s needs (loads only), prep(s, r, x) turns them into
// the 8 normalised inputs, out(s, y, r) takes the 4 normalised outputs.
// In a wavefront, lanes l and l+16 (l < 16) of group g work on sample
// base + 16g + l: both load and prepare it (each keeps its half of the
// inputs), lane l takes the outputs.  The loads of the next round are
// issued before the current round's layers, so their latency overlaps the
// matrix work.  Workgroup bx of nbx stages the weights once; its
// wavefronts then run independently.
constexpr int MLP_KEEP = 8;
// n_dev, if not null, holds a count (in device memory, written by an
// earlier launch) that caps n: the samples are then 0..min(n, *n_dev)-1.
template <class LOAD, class PREP, class OUT>
__device__ void mlp_tiles(const float *__restrict__ prm, int64_t n, const unsigned long long *n_dev,
                          const LOAD &load, const PREP &prep, const OUT &out, int bx, int nbx) {
  __shared__ MlpLds S;
  if (n_dev) {
    const int64_t nd = (int64_t)*n_dev;
    n = nd < n ? nd : n;
  }
  const int tid = (int)threadIdx.x, l = tid & 63, w = tid >> 6;
  unsigned short *w1 = (unsigned short *)S.w1, *w2 = (unsigned short *)S.w2, *w3 = (unsigned short *)S.w3,
                 *w4 = (unsigned short *)S.w4;
  for (int e = tid; e < 16 * MLP_H; e += 512) {  // layer 1, inputs padded to 16
    const int k = e / MLP_H, nn = e % MLP_H;
    w1[mlp_wfrag(k, nn, 1)] = k < MLP_IN ? bf16_rne(prm[OW1 + k * MLP_H + nn]) : (unsigned short)0;
  }
  for (int e = tid; e < MLP_H * MLP_H; e += 512) {
    const int k = e / MLP_H, nn = e % MLP_H;
    w2[mlp_wfrag32(k, nn, 4)] = bf16_rne(prm[OW2 + e]);
    w3[mlp_wfrag32(k, nn, 4)] = bf16_rne(prm[OW3 + e]);
  }
  for (int e = tid; e < MLP_H * 16; e += 512) {  // output layer, outputs padded to 16
    const int k = e / 16, nn = e % 16;
    w4[mlp_wfrag32(k, nn, 4)] = nn < MLP_OUT ? bf16_rne(prm[OW4 + k * MLP_OUT + nn]) : (unsigned short)0;
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
  constexpr int PER = 16 * MLP_G;  // samples per wavefront and round
  const int64_t stride = (int64_t)nbx * MLP_WAVES * PER;
  const bool loader = l < 32;
  int64_t base = ((int64_t)bx * MLP_WAVES + w) * PER;
  float r[MLP_G][MLP_KEEP];
  auto fetch = [&](int64_t b0, float (&rr)[MLP_G][MLP_KEEP]) {
#pragma unroll
    for (int g = 0; g < MLP_G; g++) {
      for (int q = 0; q < MLP_KEEP; q++) rr[g][q] = 0.0f;
      const int64_t s = b0 + 16 * g + (l & 15);
      if (loader && s < n) load(s, rr[g]);
    }
  };
  fetch(base, r);
  for (; base < n; base += stride) {
    bf16x4 x[MLP_G][8];
#pragma unroll
    for (int g = 0; g < MLP_G; g++) {
      const int64_t s = base + 16 * g + (l & 15);
      float xin[MLP_IN];
      for (int q = 0; q < MLP_IN; q++) xin[q] = 0.0f;
      if (loader && s < n) prep(s, r[g], xin);
      const int h = (l >> 4) & 1;  // lanes 0-15 take inputs 0-3, lanes 16-31 inputs 4-7
      x[g][0] = loader ? mlp_pack(xin[4 * h], xin[4 * h + 1], xin[4 * h + 2], xin[4 * h + 3])
                       : mlp_pack(0.0f, 0.0f, 0.0f, 0.0f);
    }
    float rn[MLP_G][MLP_KEEP];
    fetch(base + stride, rn);
    mlp_first(x, S.w1, S.b, l);
    mlp_hidden(x, S.w2, S.b + MLP_H, l);
    mlp_hidden(x, S.w3, S.b + 2 * MLP_H, l);
    // output layer: one tile of 16 outputs (4 used); lane l < 16 ends up
    // with outputs 0..3 of its sample
    floatx4 acc[MLP_G];
#pragma unroll
    for (int g = 0; g < MLP_G; g++) acc[g] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {
      const bf16x8 wv = __builtin_bit_cast(bf16x8, S.w4[kb * 64 + l]);
#pragma unroll
      for (int g = 0; g < MLP_G; g++)
        acc[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wv, mlp_join(x[g][2 * kb], x[g][2 * kb + 1]), acc[g],
                                                         0, 0, 0);
    }
#pragma unroll
    for (int g = 0; g < MLP_G; g++) {
      const int64_t s = base + 16 * g + l;
      if (l < 16 && s < n) {
        const float y[MLP_OUT] = {acc[g][0] + S.b[3 * MLP_H], acc[g][1] + S.b[3 * MLP_H + 1],
                                  acc[g][2] + S.b[3 * MLP_H + 2], acc[g][3] + S.b[3 * MLP_H + 3]};
        out(s, y, r[g]);
      }
    }
#pragma unroll
    for (int g = 0; g < MLP_G; g++)
      for (int q = 0; q < MLP_KEEP; q++) r[g][q] = rn[g][q];
  }
}
