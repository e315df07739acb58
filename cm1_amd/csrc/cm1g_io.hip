// GrADS output of the resident state (src/writeout.F, output_format = 1,
// output_filetype = 1, outunits = 1): the variables the reference's output_*
// switches select, in the reference's file order, packed on the device one
// variable at a time and streamed to cm1out_{s,u,v,w}.dat while the next
// variable is packed, the four descriptors and the metadata file.
#include "cm1g_internal.h"

#include <algorithm>
#include <cerrno>
#include <sys/stat.h>

#define ARGCHK(cond, msg)  \
  do {                     \
    if (!(cond)) {         \
      cm1g_set_error(msg); \
      return CM1G_E_ARG;   \
    }                      \
  } while (0)

namespace {

// How one output value is formed from the resident arrays (the expressions
// of src/writeout.F; a: a 3-D or 2-D array, col: a base-state or grid column).
enum OutKind {
  OK_FIELD,      // a(i,j,k)
  OK_COL_PLUS,   // col(k) + a(i,j,k)            th, pi
  OK_MINUS_COL,  // a(i,j,k) - col(k)            prspert, rhopert, qvpert, upert, vpert
  OK_COL,        // col(k)                       zhval, pi0, th0, prs0, qv0, u0, v0
  OK_AVG_I,      // 0.5*(a(i,j,k)+a(i+1,j,k))   uinterp
  OK_AVG_J,      // 0.5*(a(i,j,k)+a(i,j+1,k))   vinterp
  OK_AVG_K,      // 0.5*(a(i,j,k)+a(i,j,k+1))   winterp
  OK_BUOY,       // g*th'/th0 + g*repsm1*(qv-qv0) - g*qc - g*qr
  OK_2D,         // a 2-D field (rain, prate)
  OK_ZERO        // 0 (epst outside the reference's domain diagnostics)
};

struct OutVar {
  const char *name, *desc, *unit;
  char grid;      // '2', 's', 'u', 'v', 'w' (the reference's grid_output)
  int kind;
  int fld, col;   // 3-D field id / column id / 2-D field id
};

// Levels and row lengths of one record set.
struct Shape {
  int nx, ny, nlev;
};

Shape shape_of(const cm1g_model *m, char grid) {
  const int ni = m->P.ni, nj = m->P.nj, nk = m->P.nk;
  switch (grid) {
    case '2': return {ni, nj, 1};
    case 'u': return {ni + 1, nj, nk};
    case 'v': return {ni, nj + 1, nk};
    case 'w': return {ni, nj, nk + 1};
    default: return {ni, nj, nk};
  }
}

const char *TRACER_NAMES[8] = {"pt1", "pt2", "pt3", "pt4", "pt5", "pt6", "pt7", "pt8"};

// The variables of one output time in file order (src/writeout.F:361-403,
// 3170-3600, 4108-4212, 4430-4500, 6825-6890, 7286-7355, 7748-7760).
std::vector<OutVar> out_list(const cm1g_model *m, uint32_t mask) {
  std::vector<OutVar> v;
  auto on = [&](uint32_t b) { return (mask & b) != 0; };
  const bool moist = m->c.imoist == 1;
  if (on(CM1G_OUT_RAIN)) {
    v.push_back({"rain", "accumulated surface rainfall", "cm", '2', OK_2D, CM1G_D_RAIN, 0});
    v.push_back({"prate", "surface precipitation rate", "kg/m2/s", '2', OK_2D, CM1G_D_PRATE, 0});
  }
  if (on(CM1G_OUT_ZH)) v.push_back({"zhval", "height on model levels", "m", 's', OK_COL, 0, CM1G_K_ZH});
  if (on(CM1G_OUT_TH))
    v.push_back({"th", "potential temperature", "K", 's', OK_COL_PLUS, CM1G_THA, CM1G_K_TH0});
  if (on(CM1G_OUT_THPERT))
    v.push_back({"thpert", "potential temperature perturbation", "K", 's', OK_FIELD, CM1G_THA, 0});
  if (on(CM1G_OUT_PRS)) v.push_back({"prs", "pressure", "Pa", 's', OK_FIELD, CM1G_PRS, 0});
  if (on(CM1G_OUT_PRSPERT))
    v.push_back({"prspert", "pressure perturbation", "Pa", 's', OK_MINUS_COL, CM1G_PRS, CM1G_K_PRS0});
  if (on(CM1G_OUT_PI))
    v.push_back({"pi", "nondimensional pressure", "nondimensional", 's', OK_COL_PLUS, CM1G_PPI,
                 CM1G_K_PI0});
  if (on(CM1G_OUT_PIPERT))
    v.push_back({"pipert", "nondimensional pressure perturbation", "nondimensional", 's', OK_FIELD,
                 CM1G_PPI, 0});
  if (on(CM1G_OUT_RHO)) v.push_back({"rho", "dry-air density", "kg/m^3", 's', OK_FIELD, CM1G_RHO, 0});
  if (on(CM1G_OUT_RHOPERT))
    v.push_back({"rhopert", "dry-air density perturbation", "kg/m^3", 's', OK_MINUS_COL, CM1G_RHO,
                 CM1G_K_RHO0});
  for (int n = 0; n < m->c.npt; n++)
    v.push_back({TRACER_NAMES[n], "passive tracer mixing ratio", "kg/kg", 's', OK_FIELD, CM1G_PTA0 + n, 0});
  if (on(CM1G_OUT_QV))
    v.push_back({"qv", "water vapor mixing ratio", "kg/kg", 's', OK_FIELD, CM1G_QA0, 0});
  if (on(CM1G_OUT_QVPERT))
    v.push_back({"qvpert", "water vapor mixing ratio perturbation", "kg/kg", 's', OK_MINUS_COL, CM1G_QA0,
                 CM1G_K_QV0});
  if (on(CM1G_OUT_Q) && moist) {
    v.push_back({"qc", "qc", "kg/kg", 's', OK_FIELD, CM1G_QA1, 0});
    v.push_back({"qr", "qr", "kg/kg", 's', OK_FIELD, CM1G_QA2, 0});
  }
  if (on(CM1G_OUT_BUOYANCY)) v.push_back({"buoyancy", "buoyancy", "m/s^2", 's', OK_BUOY, CM1G_THA, 0});
  if (on(CM1G_OUT_UINTERP))
    v.push_back({"uinterp", "u interpolated to scalar points (grid-relative)", "m/s", 's', OK_AVG_I,
                 CM1G_UA, 0});
  if (on(CM1G_OUT_VINTERP))
    v.push_back({"vinterp", "v interpolated to scalar points (grid-relative)", "m/s", 's', OK_AVG_J,
                 CM1G_VA, 0});
  if (on(CM1G_OUT_WINTERP))
    v.push_back({"winterp", "w interpolated to scalar points", "m/s", 's', OK_AVG_K, CM1G_WA, 0});
  if (on(CM1G_OUT_BASESTATE)) {
    v.push_back({"pi0", "base-state nondimensional pressure", "nondimensional", 's', OK_COL, 0, CM1G_K_PI0});
    v.push_back({"th0", "base-state potential temperature", "K", 's', OK_COL, 0, CM1G_K_TH0});
    v.push_back({"prs0", "base-state pressure", "Pa", 's', OK_COL, 0, CM1G_K_PRS0});
    v.push_back({"qv0", "base-state water vapor mixing ratio", "kg/kg", 's', OK_COL, 0, CM1G_K_QV0});
  }
  if (on(CM1G_OUT_U)) v.push_back({"u", "E-W (x) velocity (grid-relative)", "m/s", 'u', OK_FIELD, CM1G_UA, 0});
  if (on(CM1G_OUT_UPERT))
    v.push_back({"upert", "u perturbation (grid-relative)", "m/s", 'u', OK_MINUS_COL, CM1G_UA, CM1G_K_U0});
  if (on(CM1G_OUT_BASESTATE))
    v.push_back({"u0", "base-state u (grid-relative)", "m/s", 'u', OK_COL, 0, CM1G_K_U0});
  if (on(CM1G_OUT_V)) v.push_back({"v", "N-S (y) velocity (grid-relative)", "m/s", 'v', OK_FIELD, CM1G_VA, 0});
  if (on(CM1G_OUT_VPERT))
    v.push_back({"vpert", "v perturbation (grid-relative)", "m/s", 'v', OK_MINUS_COL, CM1G_VA, CM1G_K_V0});
  if (on(CM1G_OUT_BASESTATE))
    v.push_back({"v0", "base-state v (grid-relative)", "m/s", 'v', OK_COL, 0, CM1G_K_V0});
  if (on(CM1G_OUT_W)) v.push_back({"w", "vertical velocity", "m/s", 'w', OK_FIELD, CM1G_WA, 0});
  // the subgrid turbulence (src/writeout.F:7788-7890, 8125-8205)
  if (on(CM1G_OUT_TKE))
    v.push_back({"tke", "subgrid turbulence kinetic energy", "m^2/s^2", 'w', OK_FIELD, CM1G_TKEA, 0});
  if (on(CM1G_OUT_KM)) {
    v.push_back({"kmh", "horizontal eddy viscosity for momentum", "m^2/s", 'w', OK_FIELD, CM1G_KMH, 0});
    v.push_back({"kmv", "vertical eddy viscosity for momentum", "m^2/s", 'w', OK_FIELD, CM1G_KMV, 0});
  }
  if (on(CM1G_OUT_KH)) {
    v.push_back({"khh", "horizontal eddy diffusivity for scalars", "m^2/s", 'w', OK_FIELD, CM1G_KHH, 0});
    v.push_back({"khv", "vertical eddy diffusivity for scalars", "m^2/s", 'w', OK_FIELD, CM1G_KHV, 0});
  }
  if (on(CM1G_OUT_DISSTEN)) {
    v.push_back({"lenscl", "lenscl", "m", 'w', OK_FIELD, CM1G_LENSCL, 0});
    v.push_back({"dissten", "dissipation rate", "m^2/s^3", 'w', OK_FIELD, CM1G_DISSTEN, 0});
  }
  if (on(CM1G_OUT_NM)) v.push_back({"nm", "squared Brunt-Vaisala frequency", "1/s^2", 'w', OK_FIELD, CM1G_NM, 0});
  if (on(CM1G_OUT_DEF)) {
    v.push_back({"defv", "vertical deformation", "1/s^2", 'w', OK_FIELD, CM1G_DEFV, 0});
    v.push_back({"defh", "horizontal deformation", "1/s^2", 'w', OK_FIELD, CM1G_DEFH, 0});
    v.push_back({"epst", "epst", "1/s^2", 'w', OK_ZERO, 0, 0});
  }
  return v;
}

// One variable into out[(k-1)*nx*ny + (j-1)*nx + (i-1)], k = 1..nlev.
void pack(cm1g_model *m, const OutVar &ov, float *out) {
  const KP P = m->P;
  const Shape s = shape_of(m, ov.grid);
  const int kind = ov.kind, nx = s.nx, ny = s.ny;
  const float *a = ov.kind == OK_2D ? m->d2[ov.fld]
                                    : ((ov.kind == OK_COL || ov.kind == OK_ZERO) ? nullptr : m->f[ov.fld]);
  const float *col = m->kf[ov.col];
  const float *qv = m->f[CM1G_QA0], *qc = m->f[CM1G_QA1], *qr = m->f[CM1G_QA2];
  const bool moist = m->c.imoist == 1;
  const float gr = P.K.g * P.K.repsm1;  // g*repsm1, evaluated first as written
  for3(m, 1, nx, 1, ny, 1, s.nlev, [=] __device__(int i, int j, int k) {
    const size_t o = ((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1);
    const int c = IXD(P, i, j, k);
    float r;
    switch (kind) {
      case OK_FIELD: r = a[c]; break;
      case OK_COL_PLUS: r = col[k] + a[c]; break;
      case OK_MINUS_COL: r = a[c] - col[k]; break;
      case OK_COL: r = col[k]; break;
      case OK_AVG_I: r = 0.5f * (a[c] + a[c + 1]); break;
      case OK_AVG_J: r = 0.5f * (a[c] + a[c + P.px]); break;
      case OK_AVG_K: r = 0.5f * (a[c] + a[c + P.pxy]); break;
      case OK_BUOY: {
        r = P.K.g * a[c] / P.th0[k];
        if (moist) {
          r = r + gr * (qv[c] - P.qv0[k]);
          r = r - P.K.g * qc[c];
          r = r - P.K.g * qr[c];
        }
        break;
      }
      case OK_ZERO: r = 0.0f; break;
      default: r = a[IX2D(P, i, j)]; break;
    }
    out[o] = r;
  });
}

// Fortran edit descriptors of the descriptor files.
std::string padr(const char *s, size_t w) {  // A edit: left-justified, blank-filled, truncated
  std::string r(s);
  r.resize(w, ' ');
  return r;
}
std::string es14(float v) {  // ES14.6E2
  char b[64];
  snprintf(b, sizeof b, "%14.6E", (double)v);
  return b;
}
std::string unit12(const char *u) {  // '(' unit(1:10) ')' in 12 characters (src/writeout.F:9486-9491)
  std::string a(12, ' ');
  const size_t nn = std::min<size_t>(strlen(u), 10);
  const std::string u10 = padr(u, 10);
  a.replace(1, 10, u10);
  a[0] = '(';
  a[nn + 1] = ')';
  return a;
}

const char *VERSION = "cm1r21.1";

// cm1out_{s,u,v,w}.ctl (src/writeout.F:9278-9530), output_filetype = 1.
int write_ctl(const cm1g_model *m, const std::string &dir, char grid, const std::vector<OutVar> &vars,
              int nwrite, int iorigin, int stretch_z) {
  int n = 0;
  for (const OutVar &v : vars)
    if (v.grid == grid || (grid == 's' && v.grid == '2')) n++;
  if (n == 0) return 0;
  const std::string base = std::string("cm1out_") + grid;
  FILE *f = fopen((dir + "/" + base + ".ctl").c_str(), "w");
  if (!f) {
    cm1g_set_error("cm1g_write_grads: cannot write " + dir + "/" + base + ".ctl: " + strerror(errno));
    return CM1G_E_IO;
  }
  const float conv = 0.001f;  // outunitconv for outunits = 1 (km)
  const int nx = m->P.ni, ny = m->P.nj, nk = m->P.nk;
  const double dx = m->c.dx, dy = m->c.dy;
  // xf(1), xh(1) as src/param.F:6646-6687 forms them (double, then REAL)
  const double xf1 = iorigin == 2 ? -0.5 * dx * nx : 0.0, yf1 = iorigin == 2 ? -0.5 * dy * ny : 0.0;
  const float xf = (float)xf1, yf = (float)yf1;
  const float xh = (float)(0.5 * (xf1 + (dx + xf1))), yh = (float)(0.5 * (yf1 + (dy + yf1)));
  fprintf(f, "dset ^%s\n", padr((base + ".dat").c_str(), 120).c_str());
  fprintf(f, "title CM1 output, using version %s; units of x,y,z are km; time is generic, see "
             "cm1out_metadata for actual times\n", VERSION);
  fprintf(f, "undef %10.1f\n", -999999.9);
  if (grid == 'u') fprintf(f, "xdef %6d linear %s %s\n", nx + 1, es14(conv * xf).c_str(), es14(conv * m->c.dx).c_str());
  else fprintf(f, "xdef %6d linear %s %s\n", nx, es14(conv * xh).c_str(), es14(conv * m->c.dx).c_str());
  if (grid == 'v') fprintf(f, "ydef %6d linear %s %s\n", ny + 1, es14(conv * yf).c_str(), es14(conv * m->c.dy).c_str());
  else fprintf(f, "ydef %6d linear %s %s\n", ny, es14(conv * yh).c_str(), es14(conv * m->c.dy).c_str());
  const float *zh = m->hk[CM1G_K_ZH], *zf = m->hk[CM1G_K_ZF];
  if (grid == 'w') {
    if (!stretch_z) {
      fprintf(f, "zdef %6d linear %s %s\n", nk + 1, es14(0.0f).c_str(), es14(conv * m->c.dz).c_str());
    } else {
      fprintf(f, "zdef %6d levels \n", nk + 1);
      for (int k = 1; k <= nk + 1; k++) fprintf(f, "  %s\n", es14(conv * zf[k]).c_str());
    }
  } else {
    if (!stretch_z) {
      fprintf(f, "zdef %6d linear %s %s\n", nk, es14(conv * zh[1]).c_str(), es14(conv * m->c.dz).c_str());
    } else {
      fprintf(f, "zdef %6d levels \n", nk);
      for (int k = 1; k <= nk; k++) fprintf(f, "  %s\n", es14(conv * zh[k]).c_str());
    }
  }
  fprintf(f, "tdef %10d linear 00:00Z01JAN0001 1YR\n", nwrite);
  fprintf(f, "vars %6d\n", n);
  for (const OutVar &v : vars) {
    if (!(v.grid == grid || (grid == 's' && v.grid == '2'))) continue;
    const int levs = v.grid == '2' ? 0 : shape_of(m, v.grid).nlev;
    fprintf(f, "%s %6d 99 %s %s\n", padr(v.name, 12).c_str(), levs, padr(v.desc, 60).c_str(),
            unit12(v.unit).c_str());
  }
  fprintf(f, "endvars\n");
  if (fclose(f) != 0) {
    cm1g_set_error("cm1g_write_grads: error closing " + base + ".ctl");
    return CM1G_E_IO;
  }
  return 0;
}

// cm1out_metadata.dat/.ctl (src/writeout.F:9536-9757), adapt_dt = 0.
int write_metadata(const cm1g_model *m, const std::string &dir, int nwrite, double mtime, int nstep) {
  const std::string dat = dir + "/cm1out_metadata.dat";
  FILE *f = fopen(dat.c_str(), nwrite == 1 ? "wb" : "r+b");
  if (!f) f = fopen(dat.c_str(), "w+b");
  if (!f) {
    cm1g_set_error("cm1g_write_grads: cannot write " + dat + ": " + strerror(errno));
    return CM1G_E_IO;
  }
  const float rec[4] = {(float)mtime, (float)nstep, m->c.dt, (float)nwrite};
  bool ok = fseek(f, (long)(nwrite - 1) * (long)sizeof rec, SEEK_SET) == 0 &&
            fwrite(rec, sizeof rec, 1, f) == 1;
  ok = (fclose(f) == 0) && ok;
  if (!ok) {
    cm1g_set_error("cm1g_write_grads: error writing " + dat);
    return CM1G_E_IO;
  }
  FILE *c = fopen((dir + "/cm1out_metadata.ctl").c_str(), "w");
  if (!c) {
    cm1g_set_error("cm1g_write_grads: cannot write cm1out_metadata.ctl");
    return CM1G_E_IO;
  }
  fprintf(c, "dset ^cm1out_metadata.dat\n");
  fprintf(c, "title CM1 metadata for 3d output files; using version %s\n", VERSION);
  fprintf(c, "undef %10.1f\n", -999999.9);
  fprintf(c, "xdef 1 linear 0 1\nydef 1 linear 0 1\nzdef 1 linear 0 1\n");
  fprintf(c, "tdef %10d linear 00:00Z01JAN0001 1YR\n", nwrite);
  fprintf(c, "vars %6d\n", 4);
  const char *names[4] = {"mtime", "nstep", "dt", "nwrite"};
  const char *descs[4] = {"model time (seconds since beginning of simulation)",
                          "time steps (since beginning of simulation)", "current time step (seconds)",
                          "writeout number"};
  for (int i = 0; i < 4; i++) fprintf(c, "%s %6d 99 %s\n", padr(names[i], 12).c_str(), 0, padr(descs[i], 60).c_str());
  fprintf(c, "endvars\n");
  if (fclose(c) != 0) {
    cm1g_set_error("cm1g_write_grads: error closing cm1out_metadata.ctl");
    return CM1G_E_IO;
  }
  return 0;
}

}  // namespace

static void staging_free(cm1g_model *m) {
  for (int b = 0; b < 2; b++) {
    if (m->io_dev[b]) hipFree(m->io_dev[b]);
    if (m->io_host[b]) hipHostFree(m->io_host[b]);
    if (m->io_ev[b]) hipEventDestroy(m->io_ev[b]);
    m->io_dev[b] = nullptr;
    m->io_host[b] = nullptr;
    m->io_ev[b] = nullptr;
  }
  m->io_cap = 0;
}

void io_free(cm1g_model *m) {
  staging_free(m);
  if (m->stat_buf) hipFree(m->stat_buf);
  m->stat_buf = nullptr;
  m->stat_cap = 0;
}

extern "C" int cm1g_write_grads(cm1g_model *m, const char *dir, int nwrite, uint32_t mask, int iorigin,
                                int stretch_z, double mtime, int nstep) {
  ARGCHK(m && dir && nwrite >= 1, "cm1g_write_grads: null argument or nwrite < 1");
  ARGCHK((mask & ~(uint32_t)(CM1G_OUT_ALL | CM1G_OUT_TURB)) == 0,
         "cm1g_write_grads: unknown output switch in mask");
  ARGCHK(m->turb.cm1setup == 1 || (mask & CM1G_OUT_TURB) == 0,
         "cm1g_write_grads: turbulence output needs the subgrid turbulence (cm1g_set_turbulence)");
  ARGCHK(iorigin == 1 || iorigin == 2, "cm1g_write_grads: iorigin must be 1 or 2");
  ARGCHK(m->c.imoist == 1 || (mask & (CM1G_OUT_QV | CM1G_OUT_QVPERT)) == 0,
         "cm1g_write_grads: qv output needs imoist = 1");
  struct stat st;
  ARGCHK(stat(dir, &st) == 0 && S_ISDIR(st.st_mode), "cm1g_write_grads: output directory does not exist");
  const std::string d(dir);
  const std::vector<OutVar> vars = out_list(m, mask);

  // staging: two device/pinned-host buffer pairs of the largest variable
  size_t cap = 0;
  for (const OutVar &v : vars) {
    const Shape s = shape_of(m, v.grid);
    cap = std::max(cap, (size_t)s.nx * s.ny * s.nlev);
  }
  if (cap > m->io_cap) {
    staging_free(m);
    for (int b = 0; b < 2; b++) {
      if (hipMalloc(&m->io_dev[b], cap * sizeof(float)) != hipSuccess ||
          hipHostMalloc(&m->io_host[b], cap * sizeof(float), hipHostMallocDefault) != hipSuccess) {
        staging_free(m);
        cm1g_set_error("cm1g_write_grads: cannot allocate the staging buffers");
        return CM1G_E_NOMEM;
      }
      HIPCHK(hipEventCreateWithFlags(&m->io_ev[b], hipEventDisableTiming));
    }
    m->io_cap = cap;
  }

  // one file per staggering; record offsets of output time nwrite
  const char grids[4] = {'s', 'u', 'v', 'w'};
  FILE *fp[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t per_time[4] = {0, 0, 0, 0};
  for (const OutVar &v : vars) {
    const int g = v.grid == '2' ? 0 : (int)(strchr(grids, v.grid) - grids);
    const Shape s = shape_of(m, v.grid);
    per_time[g] += (size_t)s.nx * s.ny * s.nlev * sizeof(float);
  }
  int rc = 0;
  for (int g = 0; g < 4 && rc == 0; g++) {
    if (!per_time[g]) continue;
    const std::string path = d + "/cm1out_" + grids[g] + ".dat";
    fp[g] = fopen(path.c_str(), nwrite == 1 ? "wb" : "r+b");
    if (!fp[g]) fp[g] = fopen(path.c_str(), "w+b");
    if (!fp[g] || fseeko(fp[g], (off_t)(nwrite - 1) * (off_t)per_time[g], SEEK_SET) != 0) {
      cm1g_set_error("cm1g_write_grads: cannot write " + path + ": " + strerror(errno));
      rc = CM1G_E_IO;
    }
  }
  // pack variable n+1 on the device while variable n is written
  const size_t nv = vars.size();
  auto issue = [&](size_t n) -> int {
    const int b = (int)(n & 1);
    const Shape s = shape_of(m, vars[n].grid);
    pack(m, vars[n], m->io_dev[b]);
    HIPCHK(hipMemcpyAsync(m->io_host[b], m->io_dev[b], (size_t)s.nx * s.ny * s.nlev * sizeof(float),
                          hipMemcpyDeviceToHost, m->stream));
    HIPCHK(hipEventRecord(m->io_ev[b], m->stream));
    return 0;
  };
  if (rc == 0 && nv > 0) rc = issue(0);
  for (size_t n = 0; n < nv && rc == 0; n++) {
    const int b = (int)(n & 1);
    if (n + 1 < nv && (rc = issue(n + 1)) != 0) break;
    if (hipEventSynchronize(m->io_ev[b]) != hipSuccess) {
      cm1g_set_error("cm1g_write_grads: device error while packing " + std::string(vars[n].name));
      rc = CM1G_E_HIP;
      break;
    }
    const Shape s = shape_of(m, vars[n].grid);
    const int g = vars[n].grid == '2' ? 0 : (int)(strchr(grids, vars[n].grid) - grids);
    const size_t cnt = (size_t)s.nx * s.ny * s.nlev;
    if (fwrite(m->io_host[b], sizeof(float), cnt, fp[g]) != cnt) {
      cm1g_set_error("cm1g_write_grads: error writing " + std::string(vars[n].name));
      rc = CM1G_E_IO;
    }
  }
  if (rc != 0) hipStreamSynchronize(m->stream);
  for (int g = 0; g < 4; g++)
    if (fp[g] && fclose(fp[g]) != 0 && rc == 0) {
      cm1g_set_error("cm1g_write_grads: error closing an output file");
      rc = CM1G_E_IO;
    }
  if (rc != 0) return rc;
  if ((rc = check_launch()) != 0) return rc;
  for (char g : grids)
    if ((rc = write_ctl(m, d, g, vars, nwrite, iorigin, stretch_z)) != 0) return rc;
  return write_metadata(m, d, nwrite, mtime, nstep);
}
