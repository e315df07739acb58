// Boundary / halo fills on the device (single GPU: periodic and open
// boundaries become on-device copies instead of comm.F MPI exchanges).
// Semantics follow src/bc.F (non-MPI build): bcs :180-364, bcp :372-520,
// bcu :528-714, bcv :722-934, bcw :942-1143.  Each fill runs as two
// phases in the reference's order (west/east then south/north; v: south/
// north then west/east) because the corner halos of the second phase read
// values the first phase wrote.  Within a phase the two sides write
// disjoint cells and read only interior cells, so they run concurrently.
#include "cm1g_internal.h"

struct BcArrs {
  float *a[8];
  int n;
};

// phase A / B for one array at (line index t, level k)
__device__ __forceinline__ void bc_cell(const KP &P, int kind, int phase, float *a, int t, int k) {
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  auto X = [&](int i, int j) -> float & { return a[IXD(P, i, j, k)]; };
  if (kind == BC_P) {
    if (k < 1 || k > nk) return;
    if (phase == 0) {
      int j = t;
      if (j < 0 || j > nj + 1) return;
      X(0, j) = (P.wbc == 1) ? X(ni, j) : X(1, j);
      X(ni + 1, j) = (P.ebc == 1) ? X(1, j) : X(ni, j);
    } else {
      int i = t;
      if (i < 0 || i > ni + 1) return;
      X(i, 0) = (P.sbc == 1) ? X(i, nj) : X(i, 1);
      X(i, nj + 1) = (P.nbc == 1) ? X(i, 1) : X(i, nj);
    }
    return;
  }
  if (kind == BC_S || kind == BC_W1 || kind == BC_W0) {
    int klo = (kind == BC_S) ? 1 : 2;
    if (phase == 2) {  // bcw top/bottom, flag=1
      int i = t % (ni + 2), j = t / (ni + 2);
      if (j > nj + 1) return;
      a[IXD(P, i, j, 1)] = 0.0f;
      a[IXD(P, i, j, nk + 1)] = 0.0f;
      return;
    }
    if (k < klo || k > nk) return;
    if (phase == 0) {
      int j = t;
      if (P.wbc == 1) {
        if (j >= -2 && j <= nj + 3)
          for (int i = 1; i <= 3; i++) X(1 - i, j) = X(ni + 1 - i, j);
      } else {
        if (j >= 0 && j <= nj + 1) {
          float v = X(1, j);
          for (int i = -2; i <= 0; i++) X(i, j) = v;
        }
      }
      if (P.ebc == 1) {
        if (j >= -2 && j <= nj + 3)
          for (int i = 1; i <= 3; i++) X(ni + i, j) = X(i, j);
      } else {
        if (j >= 0 && j <= nj + 1) {
          float v = X(ni, j);
          for (int i = ni + 1; i <= ni + 3; i++) X(i, j) = v;
        }
      }
    } else {
      int i = t;
      if (P.sbc == 1) {
        if (i >= -2 && i <= ni + 3)
          for (int j = 1; j <= 3; j++) X(i, 1 - j) = X(i, nj + 1 - j);
      } else {
        if (i >= 0 && i <= ni + 1) {
          float v = X(i, 1);
          for (int j = -2; j <= 0; j++) X(i, j) = v;
        }
      }
      if (P.nbc == 1) {
        if (i >= -2 && i <= ni + 3)
          for (int j = 1; j <= 3; j++) X(i, nj + j) = X(i, j);
      } else {
        if (i >= 0 && i <= ni + 1) {
          float v = X(i, nj);
          for (int j = nj + 1; j <= nj + 3; j++) X(i, j) = v;
        }
      }
    }
    return;
  }
  if (kind == BC_U) {
    if (k < 1 || k > nk) return;
    if (phase == 0) {
      int j = t;
      if (P.wbc == 1) {
        if (j >= -2 && j <= nj + 3)
          for (int i = 1; i <= 3; i++) X(1 - i, j) = X(ni + 1 - i, j);
      } else if (j >= 0 && j <= nj + 1) {
        float v = X(1, j);
        for (int i = -2; i <= 0; i++) X(i, j) = v;
      }
      if (P.ebc == 1) {
        if (j >= -2 && j <= nj + 3)
          for (int i = 1; i <= 3; i++) X(ni + 1 + i, j) = X(1 + i, j);
      } else if (j >= 0 && j <= nj + 1) {
        float v = X(ni + 1, j);
        for (int i = ni + 2; i <= ni + 4; i++) X(i, j) = v;
      }
    } else {
      int i = t;
      if (P.sbc == 1) {
        if (i >= -2 && i <= ni + 4)
          for (int j = 1; j <= 3; j++) X(i, 1 - j) = X(i, nj + 1 - j);
      } else if (i >= 0 && i <= ni + 2) {
        float v = X(i, 1);
        for (int j = -2; j <= 0; j++) X(i, j) = v;
      }
      if (P.nbc == 1) {
        if (i >= -2 && i <= ni + 4)
          for (int j = 1; j <= 3; j++) X(i, nj + j) = X(i, j);
      } else if (i >= 0 && i <= ni + 2) {
        float v = X(i, nj);
        for (int j = nj + 1; j <= nj + 3; j++) X(i, j) = v;
      }
    }
    return;
  }
  if (kind == BC_V) {
    if (k < 1 || k > nk) return;
    if (phase == 0) {  // south / north
      int i = t;
      if (P.sbc == 1) {
        if (i >= -2 && i <= ni + 3)
          for (int j = 1; j <= 3; j++) X(i, 1 - j) = X(i, nj + 1 - j);
      } else if (i >= 0 && i <= ni + 1) {
        float v = X(i, 1);
        for (int j = -2; j <= 0; j++) X(i, j) = v;
      }
      if (P.nbc == 1) {
        if (i >= -2 && i <= ni + 3)
          for (int j = 1; j <= 3; j++) X(i, nj + 1 + j) = X(i, 1 + j);
      } else if (i >= 0 && i <= ni + 1) {
        float v = X(i, nj + 1);
        for (int j = nj + 2; j <= nj + 4; j++) X(i, j) = v;
      }
    } else {  // west / east
      int j = t;
      if (P.wbc == 1) {
        if (j >= -2 && j <= nj + 4)
          for (int i = 1; i <= 3; i++) X(1 - i, j) = X(ni + 1 - i, j);
      } else if (j >= 0 && j <= nj + 2) {
        float v = X(1, j);
        for (int i = -2; i <= 0; i++) X(i, j) = v;
      }
      if (P.ebc == 1) {
        if (j >= -2 && j <= nj + 4)
          for (int i = 1; i <= 3; i++) X(ni + i, j) = X(i, j);
      } else if (j >= 0 && j <= nj + 2) {
        float v = X(ni, j);
        for (int i = ni + 1; i <= ni + 3; i++) X(i, j) = v;
      }
    }
  }
}

__device__ __forceinline__ void bc_run(const KP &P, int kind, int phase, const BcArrs &A, int tmin, int nt,
                                       int bx, int by) {
  const int t = tmin + bx * (int)blockDim.x + (int)threadIdx.x;
  const int k = by;
  if (t >= tmin + nt) return;
  for (int n = 0; n < A.n; n++) bc_cell(P, kind, phase, A.a[n], t, k);
}

void bc_fill_n(cm1g_model *m, int kind, float *const *arrs, int n) {
  const KP P = m->P;
  BcArrs A;
  A.n = n;
  for (int q = 0; q < n && q < 8; q++) A.a[q] = arrs[q];
  const int tmin = -2;
  const int nt = (P.ni > P.nj ? P.ni : P.nj) + 7;
  dim3 b(256), g((nt + 255) / 256, P.nk + 2);
  for (int phase = 0; phase < 2; phase++)
    launch<256>(m, g, b, [=] __device__(int bx, int by, int) { bc_run(P, kind, phase, A, tmin, nt, bx, by); });
  if (kind == BC_W1) {
    const int ntt = (P.ni + 2) * (P.nj + 2);
    launch<256>(m, dim3((ntt + 255) / 256, 1), b,
                [=] __device__(int bx, int by, int) { bc_run(P, kind, 2, A, 0, ntt, bx, by); });
  }
}

void bc_fill(cm1g_model *m, int kind, float *a) { bc_fill_n(m, kind, &a, 1); }

// radbcew4 / radbcns4 (src/bc.F:2054-2172): per boundary column, the
// vertical mean of the clipped phase speed.  One thread per column; the k
// sum runs sequentially in the reference's order.
__device__ __forceinline__ void radbc_cols(const KP &P, const float *u1, const float *u2, const float *v1,
                                           const float *v2, float dt, float csmax, float *radbcw,
                                           float *radbce, float *radbcs, float *radbcn, int bx, int by) {
  const int ni = P.ni, nj = P.nj, nk = P.nk;
  const int t = 1 + bx * (int)blockDim.x + (int)threadIdx.x;
  const int side = by;
  const float umax = csmax;
  if (side == 0 && P.ibw == 1 && P.wbc == 2 && t <= nj) {
    int j = t;
    float avgw = 0.0f;
    for (int k = 1; k <= nk; k++) {
      float foo = u1[IXD(P, 3, j, k)] - u1[IXD(P, 2, j, k)];
      float cbcw = P.dx * (1.0f / P.uf[2 + 2]) * (u1[IXD(P, 2, j, k)] - u2[IXD(P, 2, j, k)]) /
                   (dt * (copysignf(1.e-10f, foo) + foo));
      cbcw = fmaxf(fminf(cbcw, 0.0f), -umax);
      avgw = avgw + cbcw;
    }
    avgw = avgw / (float)nk;
    for (int k = 1; k <= nk; k++) radbcw[(j + 2) + k * P.py] = avgw;
  }
  if (side == 1 && P.ibe == 1 && P.ebc == 2 && t <= nj) {
    int j = t;
    float avge = 0.0f;
    for (int k = 1; k <= nk; k++) {
      float foo = u1[IXD(P, ni, j, k)] - u1[IXD(P, ni - 1, j, k)];
      float cbce = P.dx * (1.0f / P.uf[ni + 2]) * (u1[IXD(P, ni, j, k)] - u2[IXD(P, ni, j, k)]) /
                   (dt * (copysignf(1.e-10f, foo) + foo));
      cbce = fminf(fmaxf(cbce, 0.0f), umax);
      avge = avge + cbce;
    }
    avge = avge / (float)nk;
    for (int k = 1; k <= nk; k++) radbce[(j + 2) + k * P.py] = avge;
  }
  if (side == 2 && P.ibs == 1 && P.sbc == 2 && t <= ni) {
    int i = t;
    float avgs = 0.0f;
    for (int k = 1; k <= nk; k++) {
      float foo = v1[IXD(P, i, 3, k)] - v1[IXD(P, i, 2, k)];
      float cbcs = P.dy * (1.0f / P.vf[2 + 2]) * (v1[IXD(P, i, 2, k)] - v2[IXD(P, i, 2, k)]) /
                   (dt * (copysignf(1.e-10f, foo) + foo));
      cbcs = fmaxf(fminf(cbcs, 0.0f), -umax);
      avgs = avgs + cbcs;
    }
    avgs = avgs / (float)nk;
    for (int k = 1; k <= nk; k++) radbcs[(i + 2) + k * P.px] = avgs;
  }
  if (side == 3 && P.ibn == 1 && P.nbc == 2 && t <= ni) {
    int i = t;
    float avgn = 0.0f;
    for (int k = 1; k <= nk; k++) {
      float foo = v1[IXD(P, i, nj, k)] - v1[IXD(P, i, nj - 1, k)];
      float cbcn = P.dy * (1.0f / P.vf[nj + 2]) * (v1[IXD(P, i, nj, k)] - v2[IXD(P, i, nj, k)]) /
                   (dt * (copysignf(1.e-10f, foo) + foo));
      cbcn = fminf(fmaxf(cbcn, 0.0f), umax);
      avgn = avgn + cbcn;
    }
    avgn = avgn / (float)nk;
    for (int k = 1; k <= nk; k++) radbcn[(i + 2) + k * P.px] = avgn;
  }
}

void radbc4(cm1g_model *m, const float *ua, const float *u3d, const float *va, const float *v3d,
            float dt) {
  const KP P = m->P;
  const int n = (P.ni > P.nj ? P.ni : P.nj);
  const float csmax = m->c.csmax;
  float *rw = m->d2[CM1G_D_RADBCW], *re = m->d2[CM1G_D_RADBCE], *rs = m->d2[CM1G_D_RADBCS],
        *rn = m->d2[CM1G_D_RADBCN];
  launch<64>(m, dim3((n + 63) / 64, 4), dim3(64), [=] __device__(int bx, int by, int) {
    radbc_cols(P, ua, u3d, va, v3d, dt, csmax, rw, re, rs, rn, bx, by);
  });
}
