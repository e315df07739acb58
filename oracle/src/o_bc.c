/* Oracle: boundary/halo fills (test infrastructure only).
 * Follows src/bc.F (non-MPI build): bcs :180-364, bcp :372-520,
 * bcu :528-714, bcv :722-934, bcw :942-1143, radbcew4 :2054-2109,
 * radbcns4 :2117-2172.  Fill order (west, east, south, north; v: south,
 * north, west, east) is kept because corner halos read values written by
 * the earlier fills. */
#include "cm1o_internal.h"

void o_bcs(cm1o_model *M, float *s) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 1; k <= nk; k++) {
    if (M->c.wbc == 1) {
      for (j = -2; j <= nj + 3; j++)
        for (i = 1; i <= 3; i++) A3(s, 1 - i, j, k) = A3(s, ni + 1 - i, j, k);
    } else if (M->c.wbc == 2) {
      for (j = 0; j <= nj + 1; j++)
        for (i = -2; i <= 0; i++) A3(s, i, j, k) = A3(s, 1, j, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.ebc == 1) {
      for (j = -2; j <= nj + 3; j++)
        for (i = 1; i <= 3; i++) A3(s, ni + i, j, k) = A3(s, i, j, k);
    } else if (M->c.ebc == 2) {
      for (j = 0; j <= nj + 1; j++)
        for (i = ni + 1; i <= ni + 3; i++) A3(s, i, j, k) = A3(s, ni, j, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.sbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(s, i, 1 - j, k) = A3(s, i, nj + 1 - j, k);
    } else if (M->c.sbc == 2) {
      for (j = -2; j <= 0; j++)
        for (i = 0; i <= ni + 1; i++) A3(s, i, j, k) = A3(s, i, 1, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.nbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(s, i, nj + j, k) = A3(s, i, j, k);
    } else if (M->c.nbc == 2) {
      for (j = nj + 1; j <= nj + 3; j++)
        for (i = 0; i <= ni + 1; i++) A3(s, i, j, k) = A3(s, i, nj, k);
    }
  }
}

void o_bcp(cm1o_model *M, float *p) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 1; k <= nk; k++)
    for (j = 0; j <= nj + 1; j++)
      A3(p, 0, j, k) = (M->c.wbc == 1) ? A3(p, ni, j, k) : A3(p, 1, j, k);
  for (k = 1; k <= nk; k++)
    for (j = 0; j <= nj + 1; j++)
      A3(p, ni + 1, j, k) = (M->c.ebc == 1) ? A3(p, 1, j, k) : A3(p, ni, j, k);
  for (k = 1; k <= nk; k++)
    for (i = 0; i <= ni + 1; i++)
      A3(p, i, 0, k) = (M->c.sbc == 1) ? A3(p, i, nj, k) : A3(p, i, 1, k);
  for (k = 1; k <= nk; k++)
    for (i = 0; i <= ni + 1; i++)
      A3(p, i, nj + 1, k) = (M->c.nbc == 1) ? A3(p, i, 1, k) : A3(p, i, nj, k);
}

void o_bcu(cm1o_model *M, float *u) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 1; k <= nk; k++) {
    if (M->c.wbc == 1) {
      for (j = -2; j <= nj + 3; j++)
        for (i = 1; i <= 3; i++) A3(u, 1 - i, j, k) = A3(u, ni + 1 - i, j, k);
    } else {
      for (j = 0; j <= nj + 1; j++)
        for (i = -2; i <= 0; i++) A3(u, i, j, k) = A3(u, 1, j, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.ebc == 1) {
      for (j = -2; j <= nj + 3; j++)
        for (i = 1; i <= 3; i++) A3(u, ni + 1 + i, j, k) = A3(u, 1 + i, j, k);
    } else {
      for (j = 0; j <= nj + 1; j++)
        for (i = ni + 2; i <= ni + 4; i++) A3(u, i, j, k) = A3(u, ni + 1, j, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.sbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 4; i++) A3(u, i, 1 - j, k) = A3(u, i, nj + 1 - j, k);
    } else {
      for (j = -2; j <= 0; j++)
        for (i = 0; i <= ni + 2; i++) A3(u, i, j, k) = A3(u, i, 1, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.nbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 4; i++) A3(u, i, nj + j, k) = A3(u, i, j, k);
    } else {
      for (j = nj + 1; j <= nj + 3; j++)
        for (i = 0; i <= ni + 2; i++) A3(u, i, j, k) = A3(u, i, nj, k);
    }
  }
}

void o_bcv(cm1o_model *M, float *v) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 1; k <= nk; k++) {
    if (M->c.sbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(v, i, 1 - j, k) = A3(v, i, nj + 1 - j, k);
    } else {
      for (j = -2; j <= 0; j++)
        for (i = 0; i <= ni + 1; i++) A3(v, i, j, k) = A3(v, i, 1, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.nbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(v, i, nj + 1 + j, k) = A3(v, i, 1 + j, k);
    } else {
      for (j = nj + 2; j <= nj + 4; j++)
        for (i = 0; i <= ni + 1; i++) A3(v, i, j, k) = A3(v, i, nj + 1, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.wbc == 1) {
      for (j = -2; j <= nj + 4; j++)
        for (i = 1; i <= 3; i++) A3(v, 1 - i, j, k) = A3(v, ni + 1 - i, j, k);
    } else {
      for (j = 0; j <= nj + 2; j++)
        for (i = -2; i <= 0; i++) A3(v, i, j, k) = A3(v, 1, j, k);
    }
  }
  for (k = 1; k <= nk; k++) {
    if (M->c.ebc == 1) {
      for (j = -2; j <= nj + 4; j++)
        for (i = 1; i <= 3; i++) A3(v, ni + i, j, k) = A3(v, i, j, k);
    } else {
      for (j = 0; j <= nj + 2; j++)
        for (i = ni + 1; i <= ni + 3; i++) A3(v, i, j, k) = A3(v, ni, j, k);
    }
  }
}

void o_bcw(cm1o_model *M, float *w, int flag) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  int i, j, k;
  for (k = 2; k <= nk; k++) {
    if (M->c.wbc == 1) {
      for (j = -2; j <= nj + 3; j++)
        for (i = 1; i <= 3; i++) A3(w, 1 - i, j, k) = A3(w, ni + 1 - i, j, k);
    } else {
      for (j = 0; j <= nj + 1; j++)
        for (i = -2; i <= 0; i++) A3(w, i, j, k) = A3(w, 1, j, k);
    }
  }
  for (k = 2; k <= nk; k++) {
    if (M->c.ebc == 1) {
      for (j = -2; j <= nj + 3; j++)
        for (i = 1; i <= 3; i++) A3(w, ni + i, j, k) = A3(w, i, j, k);
    } else {
      for (j = 0; j <= nj + 1; j++)
        for (i = ni + 1; i <= ni + 3; i++) A3(w, i, j, k) = A3(w, ni, j, k);
    }
  }
  for (k = 2; k <= nk; k++) {
    if (M->c.sbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(w, i, 1 - j, k) = A3(w, i, nj + 1 - j, k);
    } else {
      for (j = -2; j <= 0; j++)
        for (i = 0; i <= ni + 1; i++) A3(w, i, j, k) = A3(w, i, 1, k);
    }
  }
  for (k = 2; k <= nk; k++) {
    if (M->c.nbc == 1) {
      for (j = 1; j <= 3; j++)
        for (i = -2; i <= ni + 3; i++) A3(w, i, nj + j, k) = A3(w, i, j, k);
    } else {
      for (j = nj + 1; j <= nj + 3; j++)
        for (i = 0; i <= ni + 1; i++) A3(w, i, j, k) = A3(w, i, nj, k);
    }
  }
  if (flag == 1) {
    for (j = 0; j <= nj + 1; j++)
      for (i = 0; i <= ni + 1; i++) {
        A3(w, i, j, 1) = 0.0f;
        A3(w, i, j, nk + 1) = 0.0f;
      }
  }
}

/* radbcew4: src/bc.F:2054-2109.  radbcw/radbce are (j,k) arrays stored as
 * [j+2 + k*py]. */
void o_radbcew4(cm1o_model *M, float *radbcw, float *radbce, const float *u1,
                const float *u2, float dt) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float umax = M->c.csmax;
  const float dx = M->c.dx;
  int j, k;
  if (M->ibw == 1 && M->c.wbc == 2) {
    for (j = 1; j <= nj; j++) {
      float avgw = 0.0f;
      for (k = 1; k <= nk; k++) {
        float foo = A3(u1, 3, j, k) - A3(u1, 2, j, k);
        float cbcw = dx * (1.0f / UF(2)) * (A3(u1, 2, j, k) - A3(u2, 2, j, k)) /
                     (dt * (copysignf(1.e-10f, foo) + foo));
        cbcw = fmaxf(fminf(cbcw, 0.0f), -umax);
        avgw = avgw + cbcw;
      }
      avgw = avgw / (float)nk;
      for (k = 1; k <= nk; k++) radbcw[(j + 2) + (size_t)k * M->py] = avgw;
    }
  }
  if (M->ibe == 1 && M->c.ebc == 2) {
    for (j = 1; j <= nj; j++) {
      float avge = 0.0f;
      for (k = 1; k <= nk; k++) {
        float foo = A3(u1, ni + 1 - 1, j, k) - A3(u1, ni + 1 - 2, j, k);
        float cbce = dx * (1.0f / UF(ni + 1 - 1)) * (A3(u1, ni + 1 - 1, j, k) - A3(u2, ni + 1 - 1, j, k)) /
                     (dt * (copysignf(1.e-10f, foo) + foo));
        cbce = fminf(fmaxf(cbce, 0.0f), umax);
        avge = avge + cbce;
      }
      avge = avge / (float)nk;
      for (k = 1; k <= nk; k++) radbce[(j + 2) + (size_t)k * M->py] = avge;
    }
  }
}

/* radbcns4: src/bc.F:2117-2172.  radbcs/radbcn are (i,k): [i+2 + k*px]. */
void o_radbcns4(cm1o_model *M, float *radbcs, float *radbcn, const float *v1,
                const float *v2, float dt) {
  const int ni = M->ni, nj = M->nj, nk = M->nk;
  const float umax = M->c.csmax;
  const float dy = M->c.dy;
  int i, k;
  if (M->ibs == 1 && M->c.sbc == 2) {
    for (i = 1; i <= ni; i++) {
      float avgs = 0.0f;
      for (k = 1; k <= nk; k++) {
        float foo = A3(v1, i, 3, k) - A3(v1, i, 2, k);
        float cbcs = dy * (1.0f / VF(2)) * (A3(v1, i, 2, k) - A3(v2, i, 2, k)) /
                     (dt * (copysignf(1.e-10f, foo) + foo));
        cbcs = fmaxf(fminf(cbcs, 0.0f), -umax);
        avgs = avgs + cbcs;
      }
      avgs = avgs / (float)nk;
      for (k = 1; k <= nk; k++) radbcs[(i + 2) + (size_t)k * M->px] = avgs;
    }
  }
  if (M->ibn == 1 && M->c.nbc == 2) {
    for (i = 1; i <= ni; i++) {
      float avgn = 0.0f;
      for (k = 1; k <= nk; k++) {
        float foo = A3(v1, i, nj + 1 - 1, k) - A3(v1, i, nj + 1 - 2, k);
        float cbcn = dy * (1.0f / VF(nj + 1 - 1)) * (A3(v1, i, nj + 1 - 1, k) - A3(v2, i, nj + 1 - 1, k)) /
                     (dt * (copysignf(1.e-10f, foo) + foo));
        cbcn = fminf(fmaxf(cbcn, 0.0f), umax);
        avgn = avgn + cbcn;
      }
      avgn = avgn / (float)nk;
      for (k = 1; k <= nk; k++) radbcn[(i + 2) + (size_t)k * M->px] = avgn;
    }
  }
}
