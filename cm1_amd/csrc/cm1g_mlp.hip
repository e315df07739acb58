// A small per-cell MLP (8 -> 128 -> 128 -> 128 -> 4, SiLU hidden layers)
// with BF16 operands and FP32 accumulation on the CDNA matrix cores, used as
// a learned surrogate for the local part of Kessler microphysics.  No
// counterpart in the reference.
//
// Matrix products use v_mfma_f32_16x16x16_bf16: a wavefront multiplies a
// 16x16 block of A by a 16x16 block of B into a 16x16 FP32 block, where
//   lane l holds A(l%16, 4*(l/16) + t) and B(4*(l/16) + t, l%16), t = 0..3,
//   and receives D(4*(l/16) + t, l%16).
// Operands are rounded from FP32 to BF16 to nearest, ties to even.
//
// Training (forward with saved activations, backward, Adam) runs through one
// tiled matrix-product kernel (k_gemm: 64x64 output tiles, K in steps of 16
// staged through LDS, split over K for the weight gradients) and a few
// element-wise kernels; the loss is the mean squared error of the
// normalised outputs.  Inference runs fused (mlp_tiles in cm1g_mlp.h).
#include "cm1g_mlp.h"

#include <random>

#define ARGCHK(cond, msg)          \
  do {                             \
    if (!(cond)) {                 \
      cm1g_set_error(msg);         \
      return CM1G_E_ARG;           \
    }                              \
  } while (0)

struct cm1g_mlp {
  float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;  // params, grads, Adam moments
  float *dnorm = nullptr;                                          // normalisation (device)
  int64_t adam_t = 0;
  hipStream_t stream = nullptr;
  // training workspace for `cap` rows
  int64_t cap = 0;
  float *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *z3 = nullptr, *h3 = nullptr;
  float *y = nullptr, *dy = nullptr, *da = nullptr, *dz = nullptr, *part = nullptr, *red = nullptr;
};

const float *mlp_params(const cm1g_mlp *n) { return n->p; }
const float *mlp_norm(const cm1g_mlp *n) { return n->dnorm; }

// ------------------------------------------------------------ k_gemm
// C(m,n) = sum_k A(m,k) B(k,n) over this split's K range, with
// A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn]; split s writes
// C + s*cs.  64x64 output tile per workgroup, four wavefronts of 32x32.
#define GP 20  // LDS row pitch of the staged tiles (16 + 4, in BF16 elements)
__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, int kchunk, const float *__restrict__ A,
                                              long sam, long sak, const float *__restrict__ B, long sbk,
                                              long sbn, float *__restrict__ C, long ldc, long cs) {
  __shared__ unsigned short As[64 * GP];  // [m][k]
  __shared__ unsigned short Bs[64 * GP];  // [n][k]
  const int tid = (int)threadIdx.x, l = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1;
  const int m0 = (int)blockIdx.x * 64, n0 = (int)blockIdx.y * 64, s = (int)blockIdx.z;
  const int kb = s * kchunk, ke = min(K, kb + kchunk);
  const bool akf = (sak == 1), bkf = (sbk == 1);
  floatx4 acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = kb; k0 < ke; k0 += 16) {
    for (int q = 0; q < 4; q++) {
      const int e = tid + 256 * q;  // 0..1023 over the 64 x 16 tile
      // the fast index follows the unit stride of each operand
      const int ra = akf ? e / 16 : e % 64, ka = akf ? e % 16 : e / 64;
      const int m = m0 + ra, kk = k0 + ka;
      As[ra * GP + ka] = (m < M && kk < ke) ? bf16_rne(A[(long)m * sam + (long)kk * sak]) : (unsigned short)0;
      const int cb = bkf ? e / 16 : e % 64, kb2 = bkf ? e % 16 : e / 64;
      const int n = n0 + cb, kk2 = k0 + kb2;
      Bs[cb * GP + kb2] = (n < N && kk2 < ke) ? bf16_rne(B[(long)kk2 * sbk + (long)n * sbn]) : (unsigned short)0;
    }
    __syncthreads();
    bf16x4 af[2], bq[2];
    for (int t = 0; t < 2; t++) {
      const unsigned short *pa = &As[(32 * wm + 16 * t + (l & 15)) * GP + 4 * (l >> 4)];
      const unsigned short *pb = &Bs[(32 * wn + 16 * t + (l & 15)) * GP + 4 * (l >> 4)];
      for (int u = 0; u < 4; u++) {
        af[t][u] = (short)pa[u];
        bq[t][u] = (short)pb[u];
      }
    }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[a], bq[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int t = 0; t < 4; t++) {
        const int m = m0 + 32 * wm + 16 * a + 4 * (l >> 4) + t, n = n0 + 32 * wn + 16 * b + (l & 15);
        if (m < M && n < N) C[(long)s * cs + (long)m * ldc + n] = acc[a][b][t];
      }
}

// C(m,n) = sum over splits of the partial products, in split order
__global__ void __launch_bounds__(256) k_sum_splits(int M, int N, long ldc, int splits, long cs,
                                                    const float *__restrict__ part, float *__restrict__ C) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= M * N) return;
  float s = 0.0f;
  for (int q = 0; q < splits; q++) s += part[(long)q * cs + i];
  C[(long)(i / N) * ldc + (i % N)] = s;
}

static void gemm(hipStream_t st, int M, int N, int K, const float *A, long sam, long sak, const float *B,
                 long sbk, long sbn, float *C, long ldc, float *part, int splits) {
  if (splits <= 1 || K <= 1024) {
    hipLaunchKernelGGL(k_gemm, dim3((M + 63) / 64, (N + 63) / 64, 1), dim3(256), 0, st, M, N, K, K, A, sam,
                       sak, B, sbk, sbn, C, ldc, 0L);
    return;
  }
  int kchunk = (K + splits - 1) / splits;
  kchunk = (kchunk + 15) / 16 * 16;
  splits = (K + kchunk - 1) / kchunk;
  const long cs = (long)M * N;
  hipLaunchKernelGGL(k_gemm, dim3((M + 63) / 64, (N + 63) / 64, splits), dim3(256), 0, st, M, N, K, kchunk, A,
                     sam, sak, B, sbk, sbn, part, (long)N, cs);
  hipLaunchKernelGGL(k_sum_splits, dim3((M * N + 255) / 256), dim3(256), 0, st, M, N, ldc, splits, cs, part, C);
}

// ------------------------------------------------------------ element-wise
// Z += bias (per column); H = SiLU(Z) if H is given
__global__ void __launch_bounds__(256) k_bias_act(float *__restrict__ Z, int64_t rows, int width,
                                                  const float *__restrict__ bias, float *__restrict__ H) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows * width) return;
  const float z = Z[i] + bias[i % width];
  Z[i] = z;
  if (H) H[i] = silu_f(z);
}

// dZ = dH * SiLU'(Z)
__global__ void __launch_bounds__(256) k_silu_back(const float *__restrict__ dH, const float *__restrict__ Z,
                                                   int64_t count, float *__restrict__ dZ) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float z = Z[i];
  const float sg = 1.0f / (1.0f + expf(-z));
  dZ[i] = dH[i] * (sg * (1.0f + z * (1.0f - sg)));
}

// dY = 2 (Y - T) / (rows * 4); per-block partial sums of (Y - T)^2
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             int64_t count, float *__restrict__ dY, float *__restrict__ part) {
  __shared__ float r[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (i < count) {
    const float d = Y[i] - T[i];
    dY[i] = 2.0f * d / (float)count;
    e2 = d * d;
  }
  e2 = block_reduce256(e2, r, OpAdd());
  if (threadIdx.x == 0) part[blockIdx.x] = e2;
}

// out[0] = sum(part[0..n)) / count, in a fixed order
__global__ void __launch_bounds__(256) k_loss_final(const float *__restrict__ part, int n, int64_t count,
                                                    float *__restrict__ out) {
  __shared__ double r[256];
  double s = 0.0;
  for (int i = (int)threadIdx.x; i < n; i += 256) s += (double)part[i];
  s = block_reduce256(s, r, OpAdd());
  if (threadIdx.x == 0) out[0] = (float)(s / (double)count);
}

// out[c] = sum over rows of D[r*width + c]: one workgroup per column
__global__ void __launch_bounds__(256) k_colsum(const float *__restrict__ D, int64_t rows, int width,
                                                float *__restrict__ out) {
  __shared__ float r[256];
  const int c = (int)blockIdx.x;
  float s = 0.0f;
  for (int64_t q = threadIdx.x; q < rows; q += 256) s += D[q * width + c];
  s = block_reduce256(s, r, OpAdd());
  if (threadIdx.x == 0) out[c] = s;
}

__global__ void __launch_bounds__(256) k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m1, float *__restrict__ m2, int n, float lr,
                                              float bc1, float bc2) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1.0e-8f;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / bc1) / (sqrtf(b / bc2) + eps);
}

// fused inference over a row-major [n x 8] input, [n x 4] output
template <class LOAD, class PREP, class OUT>
__global__ void __launch_bounds__(512) k_mlp(const float *prm, int64_t n, LOAD load, PREP prep, OUT out) {
  mlp_tiles(prm, n, nullptr, load, prep, out, (int)blockIdx.x, (int)gridDim.x);
}

// ------------------------------------------------------------ C ABI
static inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

extern "C" int cm1g_mlp_nparams(int64_t *n) {
  ARGCHK(n, "cm1g_mlp_nparams: null argument");
  *n = MLP_NPARAM;
  return 0;
}

extern "C" int cm1g_mlp_destroy(cm1g_mlp *n) {
  if (!n) return 0;
  if (n->stream) hipStreamSynchronize(n->stream);
  float *bufs[] = {n->p, n->g, n->m1, n->m2, n->dnorm, n->z1, n->h1, n->z2, n->h2, n->z3, n->h3,
                   n->y, n->dy, n->da, n->dz, n->part, n->red};
  for (float *b : bufs)
    if (b) hipFree(b);
  if (n->stream) hipStreamDestroy(n->stream);
  delete n;
  return 0;
}

extern "C" int cm1g_mlp_create(uint64_t seed, cm1g_mlp **out) {
  ARGCHK(out, "cm1g_mlp_create: null argument");
  cm1g_mlp *n = new cm1g_mlp();
  bool ok = hipStreamCreate(&n->stream) == hipSuccess;
  ok = ok && hipMalloc((void **)&n->p, MLP_NPARAM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&n->g, MLP_NPARAM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&n->m1, MLP_NPARAM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&n->m2, MLP_NPARAM * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&n->dnorm, MLP_NNORM * sizeof(float)) == hipSuccess;
  if (!ok) {
    cm1g_mlp_destroy(n);
    cm1g_set_error("cm1g_mlp_create: allocation failed");
    return CM1G_E_HIP;
  }
  // He-uniform weights (fan-in), zero biases; identity normalisation
  std::vector<float> p(MLP_NPARAM, 0.0f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fin, int fout) {
    const float a = std::sqrt(6.0f / (float)fin);
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < fin * fout; i++) p[off + i] = u(rng);
  };
  fill(OW1, MLP_IN, MLP_H);
  fill(OW2, MLP_H, MLP_H);
  fill(OW3, MLP_H, MLP_H);
  fill(OW4, MLP_H, MLP_OUT);
  std::vector<float> nm(MLP_NNORM, 0.0f);
  for (int k = 0; k < MLP_IN; k++) nm[MLP_IN + k] = 1.0f;
  for (int k = 0; k < MLP_OUT; k++) nm[2 * MLP_IN + MLP_OUT + k] = 1.0f;
  for (int k = 0; k < MLP_IN; k++) {
    nm[NM_XLO + k] = -8.0f;
    nm[NM_XHI + k] = 8.0f;
  }
  std::vector<float> z(MLP_NPARAM, 0.0f);
  ok = hipMemcpy(n->p, p.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(n->m1, z.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(n->m2, z.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(n->g, z.data(), MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(n->dnorm, nm.data(), MLP_NNORM * sizeof(float), hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    cm1g_mlp_destroy(n);
    cm1g_set_error("cm1g_mlp_create: upload failed");
    return CM1G_E_HIP;
  }
  *out = n;
  return 0;
}

extern "C" int cm1g_mlp_get_params(cm1g_mlp *n, float *host) {
  ARGCHK(n && host, "cm1g_mlp_get_params: null argument");
  HIPCHK(hipStreamSynchronize(n->stream));
  HIPCHK(hipMemcpy(host, n->p, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  return 0;
}

extern "C" int cm1g_mlp_set_params(cm1g_mlp *n, const float *host) {
  ARGCHK(n && host, "cm1g_mlp_set_params: null argument");
  HIPCHK(hipStreamSynchronize(n->stream));
  HIPCHK(hipMemcpy(n->p, host, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_mlp_set_norm(cm1g_mlp *n, const float *norm) {
  ARGCHK(n && norm, "cm1g_mlp_set_norm: null argument");
  for (int k = 0; k < MLP_IN; k++) ARGCHK(norm[MLP_IN + k] > 0.0f, "cm1g_mlp_set_norm: xstd must be > 0");
  for (int k = 0; k < MLP_OUT; k++)
    ARGCHK(norm[2 * MLP_IN + MLP_OUT + k] > 0.0f, "cm1g_mlp_set_norm: ystd must be > 0");
  for (int k = 0; k < MLP_IN; k++)
    ARGCHK(norm[NM_XLO + k] < norm[NM_XHI + k], "cm1g_mlp_set_norm: xlo must be < xhi");
  HIPCHK(hipStreamSynchronize(n->stream));
  HIPCHK(hipMemcpy(n->dnorm, norm, MLP_NNORM * sizeof(float), hipMemcpyHostToDevice));
  return 0;
}

extern "C" int cm1g_mlp_forward(cm1g_mlp *n, const float *x, int64_t rows, float *y) {
  ARGCHK(n && x && y && rows >= 0, "cm1g_mlp_forward: bad argument");
  if (rows == 0) return 0;
  auto load = [=] __device__(int64_t s, float *r) {
    for (int k = 0; k < MLP_IN; k++) r[k] = x[s * MLP_IN + k];
  };
  auto prep = [=] __device__(int64_t, const float *r, float *v) {
    for (int k = 0; k < MLP_IN; k++) v[k] = r[k];
  };
  auto out = [=] __device__(int64_t s, const float *v, const float *) {
    for (int k = 0; k < MLP_OUT; k++) y[s * MLP_OUT + k] = v[k];
  };
  const int64_t groups = (rows + 16 * MLP_G * MLP_WAVES - 1) / (16 * MLP_G * MLP_WAVES);
  const int nbx = (int)(groups < 1024 ? groups : 1024);
  hipLaunchKernelGGL((k_mlp<decltype(load), decltype(prep), decltype(out)>), dim3(nbx), dim3(512), 0, n->stream,
                     n->p, rows, load, prep, out);
  HIPCHK(hipGetLastError());
  return 0;
}

static int ensure_workspace(cm1g_mlp *n, int64_t rows) {
  if (rows <= n->cap) return 0;
  HIPCHK(hipStreamSynchronize(n->stream));
  float **bufs[] = {&n->z1, &n->h1, &n->z2, &n->h2, &n->z3, &n->h3, &n->da, &n->dz};
  for (float **b : bufs) {
    if (*b) hipFree(*b);
    *b = nullptr;
    HIPCHK(hipMalloc((void **)b, (size_t)rows * MLP_H * sizeof(float)));
  }
  for (float **b : {&n->y, &n->dy}) {
    if (*b) hipFree(*b);
    *b = nullptr;
    HIPCHK(hipMalloc((void **)b, (size_t)rows * MLP_OUT * sizeof(float)));
  }
  if (n->part) hipFree(n->part);
  if (n->red) hipFree(n->red);
  n->part = n->red = nullptr;
  HIPCHK(hipMalloc((void **)&n->part, (size_t)64 * MLP_H * MLP_H * sizeof(float)));
  HIPCHK(hipMalloc((void **)&n->red, (size_t)(nb256(rows * MLP_OUT) + 1) * sizeof(float)));
  n->cap = rows;
  return 0;
}

// forward with saved activations, MSE loss, backward into n->g
extern "C" int cm1g_mlp_loss_grad(cm1g_mlp *n, const float *x, const float *t, int64_t rows, float *loss,
                                  float *grad) {
  ARGCHK(n && x && t && rows >= 1 && rows <= (1 << 24), "cm1g_mlp_loss_grad: bad argument");
  int rc = ensure_workspace(n, rows);
  if (rc) return rc;
  hipStream_t st = n->stream;
  const float *P = n->p;
  float *G = n->g;
  const int R = (int)rows, H = MLP_H;
  const int splits = 64;
  const unsigned eh = nb256(rows * H);
  // forward
  gemm(st, R, H, MLP_IN, x, MLP_IN, 1, P + OW1, H, 1, n->z1, H, n->part, 1);
  hipLaunchKernelGGL(k_bias_act, dim3(eh), dim3(256), 0, st, n->z1, rows, H, P + OB1, n->h1);
  gemm(st, R, H, H, n->h1, H, 1, P + OW2, H, 1, n->z2, H, n->part, 1);
  hipLaunchKernelGGL(k_bias_act, dim3(eh), dim3(256), 0, st, n->z2, rows, H, P + OB2, n->h2);
  gemm(st, R, H, H, n->h2, H, 1, P + OW3, H, 1, n->z3, H, n->part, 1);
  hipLaunchKernelGGL(k_bias_act, dim3(eh), dim3(256), 0, st, n->z3, rows, H, P + OB3, n->h3);
  gemm(st, R, MLP_OUT, H, n->h3, H, 1, P + OW4, MLP_OUT, 1, n->y, MLP_OUT, n->part, 1);
  hipLaunchKernelGGL(k_bias_act, dim3(nb256(rows * MLP_OUT)), dim3(256), 0, st, n->y, rows, MLP_OUT, P + OB4,
                     (float *)nullptr);
  const unsigned nbl = nb256(rows * MLP_OUT);
  hipLaunchKernelGGL(k_mse, dim3(nbl), dim3(256), 0, st, n->y, t, rows * MLP_OUT, n->dy, n->red + 1);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, st, n->red + 1, (int)nbl, rows * MLP_OUT, n->red);
  // backward
  gemm(st, H, MLP_OUT, R, n->h3, 1, H, n->dy, MLP_OUT, 1, G + OW4, MLP_OUT, n->part, splits);
  hipLaunchKernelGGL(k_colsum, dim3(MLP_OUT), dim3(256), 0, st, n->dy, rows, MLP_OUT, G + OB4);
  gemm(st, R, H, MLP_OUT, n->dy, MLP_OUT, 1, P + OW4, 1, MLP_OUT, n->da, H, n->part, 1);
  hipLaunchKernelGGL(k_silu_back, dim3(eh), dim3(256), 0, st, n->da, n->z3, rows * H, n->dz);
  gemm(st, H, H, R, n->h2, 1, H, n->dz, H, 1, G + OW3, H, n->part, splits);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, st, n->dz, rows, H, G + OB3);
  gemm(st, R, H, H, n->dz, H, 1, P + OW3, 1, H, n->da, H, n->part, 1);
  hipLaunchKernelGGL(k_silu_back, dim3(eh), dim3(256), 0, st, n->da, n->z2, rows * H, n->dz);
  gemm(st, H, H, R, n->h1, 1, H, n->dz, H, 1, G + OW2, H, n->part, splits);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, st, n->dz, rows, H, G + OB2);
  gemm(st, R, H, H, n->dz, H, 1, P + OW2, 1, H, n->da, H, n->part, 1);
  hipLaunchKernelGGL(k_silu_back, dim3(eh), dim3(256), 0, st, n->da, n->z1, rows * H, n->dz);
  gemm(st, MLP_IN, H, R, x, 1, MLP_IN, n->dz, H, 1, G + OW1, H, n->part, splits);
  hipLaunchKernelGGL(k_colsum, dim3(H), dim3(256), 0, st, n->dz, rows, H, G + OB1);
  HIPCHK(hipGetLastError());
  if (loss) {
    HIPCHK(hipMemcpyAsync(loss, n->red, sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  if (grad) {
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(grad, G, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost));
  }
  return 0;
}

extern "C" int cm1g_mlp_adam(cm1g_mlp *n, float lr) {
  ARGCHK(n && lr > 0.0f, "cm1g_mlp_adam: bad argument");
  n->adam_t++;
  const float bc1 = 1.0f - std::pow(0.9f, (float)n->adam_t), bc2 = 1.0f - std::pow(0.999f, (float)n->adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nb256(MLP_NPARAM)), dim3(256), 0, n->stream, n->p, n->g, n->m1, n->m2,
                     MLP_NPARAM, lr, bc1, bc2);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int cm1g_mlp_sync(cm1g_mlp *n) {
  ARGCHK(n, "cm1g_mlp_sync: null argument");
  HIPCHK(hipStreamSynchronize(n->stream));
  return 0;
}
