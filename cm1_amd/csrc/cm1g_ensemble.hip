// Ensembles: members of different grid sizes and soundings advanced together,
// one kernel launch per kernel of the large step for all members.
//
// Every kernel of the step is a device functor launched through launch<NT>()
// (cm1g_internal.h).  To build an ensemble step, the host side of each
// member's step runs with its launches recorded instead of issued; the t-th
// launches of all members (the same kernel, each with its member's grid,
// arrays and switches) are merged into one member-indexed launch (k_many),
// and the merged launches are captured as one HIP graph that every later
// step replays.  A member's workgroups do the same work as in its own
// launch, so the results equal those of the members stepped one by one, bit
// for bit.  As with one model, the scalar updates exchange buffers, so the
// ensemble keeps one graph per arrangement of the members' buffers.
//
// Two schedule knobs are tuned: the column kernels' tile height (rows of 64
// columns per workgroup; small members fill the chip better with short
// tiles) and the order in which the members' workgroups are laid out in a
// merged launch (workgroups are dispatched in that order, so laying out the
// deepest columns first leaves short workgroups for the tail of each
// launch).  Neither changes any result, only which workgroup computes which
// column and when.
#include "cm1g_internal.h"

#include <algorithm>
#include <chrono>

#define ARGCHK(cond, msg)          \
  do {                             \
    if (!(cond)) {                 \
      cm1g_set_error(msg);         \
      return CM1G_E_ARG;           \
    }                              \
  } while (0)

struct cm1g_ensemble {
  std::vector<cm1g_model *> mem;
  std::vector<int> order;  // layout position -> member
  hipStream_t stream = nullptr;
  struct Graph {
    std::vector<std::vector<float *>> before, after;
    hipGraphExec_t exec = nullptr;
    void *dbuf = nullptr;  // the members' functors and slot tables of every launch
  };
  std::vector<Graph> graphs;
};

static void ens_graphs_clear(cm1g_ensemble *e) {
  if (e->stream) hipStreamSynchronize(e->stream);
  for (auto &g : e->graphs) {
    if (g.exec) hipGraphExecDestroy(g.exec);
    if (g.dbuf) hipFree(g.dbuf);
  }
  e->graphs.clear();
}

extern "C" int cm1g_set_tiles(cm1g_model *m, int ty) {
  ARGCHK(m, "cm1g_set_tiles: null model");
  ARGCHK(ty == 1 || ty == 2 || ty == 4, "cm1g_set_tiles: rows per tile must be 1, 2 or 4");
  if (m->tile_rows != ty) {
    m->tile_rows = ty;
    graphs_clear(m);
  }
  return 0;
}

extern "C" int cm1g_ensemble_create(cm1g_model *const *members, int n, cm1g_ensemble **out) {
  ARGCHK(members && out && n >= 1 && n <= 4096, "cm1g_ensemble_create: bad argument");
  for (int i = 0; i < n; i++) ARGCHK(members[i], "cm1g_ensemble_create: null member");
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) ARGCHK(members[i] != members[j], "cm1g_ensemble_create: a member appears twice");
  cm1g_ensemble *e = new cm1g_ensemble();
  e->mem.assign(members, members + n);
  for (int i = 0; i < n; i++) e->order.push_back(i);
  // members' own queued work (set-up) finishes before the ensemble's
  for (auto *m : e->mem) HIPCHK(hipStreamSynchronize(m->stream));
  if (hipStreamCreate(&e->stream) != hipSuccess) {
    delete e;
    cm1g_set_error("cm1g_ensemble_create: hipStreamCreate failed");
    return CM1G_E_HIP;
  }
  *out = e;
  return 0;
}

extern "C" int cm1g_ensemble_destroy(cm1g_ensemble *e) {
  if (!e) return 0;
  ens_graphs_clear(e);
  if (e->stream) hipStreamDestroy(e->stream);
  delete e;
  return 0;
}

extern "C" int cm1g_ensemble_sync(cm1g_ensemble *e) {
  ARGCHK(e, "cm1g_ensemble_sync: null ensemble");
  HIPCHK(hipStreamSynchronize(e->stream));
  for (auto *m : e->mem) {
    int rc = satadj_check(m);
    if (rc) return rc;
  }
  return 0;
}

static std::vector<std::vector<float *>> ens_key(const cm1g_ensemble *e) {
  std::vector<std::vector<float *>> k;
  for (auto *m : e->mem) k.push_back(field_map(m));
  return k;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// record every member's step, merge, capture; on failure the members keep
// the arrangement they had (nothing ran)
static int ens_build(cm1g_ensemble *e, cm1g_ensemble::Graph &g) {
  const int n = (int)e->mem.size();
  g.before = ens_key(e);
  std::vector<std::vector<LaunchRec>> recs(n);
  auto restore = [&]() {
    for (int i = 0; i < n; i++) set_field_map(e->mem[i], g.before[i]);
  };
  for (int i = 0; i < n; i++) {
    cm1g_model *m = e->mem[i];
    m->rec = &recs[i];
    const int rc = step_body(m);
    m->rec = nullptr;
    if (rc) {
      restore();
      return rc;
    }
  }
  const size_t nl = recs[0].size();
  for (int i = 1; i < n; i++) {
    bool same = recs[i].size() == nl;
    for (size_t t = 0; same && t < nl; t++)
      same = recs[i][t].many == recs[0][t].many && recs[i][t].block.x == recs[0][t].block.x &&
             recs[i][t].block.y == recs[0][t].block.y && recs[i][t].block.z == recs[0][t].block.z;
    if (!same) {
      restore();
      cm1g_set_error("cm1g_ensemble_step: members' steps launch different kernels "
                     "(their switches differ: psolver, imoist, ptype, advweno, npt, apmasscon, ...)");
      return CM1G_E_ARG;
    }
  }
  // device layout: per launch, the n functors, then the n slots
  std::vector<size_t> fnoff(nl), sloff(nl);
  std::vector<int> nblk(nl);
  size_t off = 0;
  for (size_t t = 0; t < nl; t++) {
    fnoff[t] = off;
    off = align256(off + (size_t)n * recs[0][t].fn.size());
    sloff[t] = off;
    off = align256(off + (size_t)n * sizeof(BatchSlot));
  }
  std::vector<unsigned char> hb(off);
  for (size_t t = 0; t < nl; t++) {
    const size_t fs = recs[0][t].fn.size();
    BatchSlot *sl = (BatchSlot *)(hb.data() + sloff[t]);
    long long cur = 0;
    for (int p = 0; p < n; p++) {
      const LaunchRec &r = recs[e->order[p]][t];
      std::memcpy(hb.data() + fnoff[t] + (size_t)p * fs, r.fn.data(), fs);
      // each member starts on a multiple of 8 workgroups, so the column
      // kernels' workgroup-to-XCD mapping holds within the member
      const long long start = (cur + NXCD - 1) / NXCD * NXCD;
      const long long nb = (long long)r.grid.x * r.grid.y * r.grid.z;
      sl[p].start = (int)start;
      sl[p].n = (int)nb;
      sl[p].gx = (int)r.grid.x;
      sl[p].gy = (int)r.grid.y;
      cur = start + nb;
    }
    if (cur >= (1LL << 31)) {
      restore();
      cm1g_set_error("cm1g_ensemble_step: too many workgroups in one merged launch");
      return CM1G_E_ARG;
    }
    nblk[t] = (int)cur;
  }
  unsigned char *d = nullptr;
  if (hipMalloc((void **)&d, off) != hipSuccess) {
    restore();
    cm1g_set_error("cm1g_ensemble_step: hipMalloc of the launch tables failed");
    return CM1G_E_HIP;
  }
  g.dbuf = d;
  hipError_t err = hipMemcpy(d, hb.data(), off, hipMemcpyHostToDevice);
  if (err == hipSuccess) err = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
  if (err == hipSuccess) {
    for (size_t t = 0; t < nl; t++)
      recs[0][t].many(e->stream, d + fnoff[t], (const BatchSlot *)(d + sloff[t]), n, nblk[t], recs[0][t].block);
    hipGraph_t graph = nullptr;
    err = hipStreamEndCapture(e->stream, &graph);
    if (err == hipSuccess) err = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    if (graph) hipGraphDestroy(graph);
  }
  if (err != hipSuccess) {
    hipFree(d);
    g.dbuf = nullptr;
    restore();
    cm1g_set_error(std::string("cm1g_ensemble_step: ") + hipGetErrorString(err));
    return CM1G_E_HIP;
  }
  g.after = ens_key(e);
  return 0;
}

extern "C" int cm1g_ensemble_step(cm1g_ensemble *e) {
  ARGCHK(e, "cm1g_ensemble_step: null ensemble");
  const auto key = ens_key(e);
  for (auto &g : e->graphs)
    if (g.before == key) {
      HIPCHK(hipGraphLaunch(g.exec, e->stream));
      for (size_t i = 0; i < e->mem.size(); i++) set_field_map(e->mem[i], g.after[i]);
      return 0;
    }
  if (e->graphs.size() >= 16) ens_graphs_clear(e);
  cm1g_ensemble::Graph g;
  const int rc = ens_build(e, g);
  if (rc) return rc;
  e->graphs.push_back(g);
  HIPCHK(hipGraphLaunch(g.exec, e->stream));
  return 0;
}

extern "C" int cm1g_ensemble_set_tiles(cm1g_ensemble *e, int ty) {
  ARGCHK(e, "cm1g_ensemble_set_tiles: null ensemble");
  ARGCHK(ty == 1 || ty == 2 || ty == 4, "cm1g_ensemble_set_tiles: rows per tile must be 1, 2 or 4");
  ens_graphs_clear(e);
  for (auto *m : e->mem) {
    const int rc = cm1g_set_tiles(m, ty);
    if (rc) return rc;
  }
  return 0;
}

extern "C" int cm1g_ensemble_set_order(cm1g_ensemble *e, int mode) {
  ARGCHK(e && mode >= 0 && mode <= 2, "cm1g_ensemble_set_order: mode must be 0, 1 or 2");
  const int n = (int)e->mem.size();
  std::vector<int> o(n);
  for (int i = 0; i < n; i++) o[i] = i;
  auto cells = [&](int i) { return (long long)e->mem[i]->c.ni * e->mem[i]->c.nj * e->mem[i]->c.nk; };
  if (mode == 1)  // deepest columns first, then the larger members
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) {
      if (e->mem[a]->c.nk != e->mem[b]->c.nk) return e->mem[a]->c.nk > e->mem[b]->c.nk;
      return cells(a) > cells(b);
    });
  else if (mode == 2)  // the larger members first
    std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return cells(a) > cells(b); });
  if (o != e->order) {
    ens_graphs_clear(e);
    e->order = o;
  }
  return 0;
}

// Time `steps` ensemble steps of the current schedule (after one untimed
// step that captures its graph).
static int ens_time(cm1g_ensemble *e, int steps, float *ms) {
  int rc = cm1g_ensemble_step(e);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(e->stream));
  const auto t0 = std::chrono::steady_clock::now();
  for (int s = 0; s < steps; s++) {
    rc = cm1g_ensemble_step(e);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  *ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
  return 0;
}

// Every tile height (1, 2, 4) with every member order (0 as given, 1 deepest
// columns first, 2 largest first): `steps` timed steps each, after one
// warm-up schedule; keeps the fastest.  ms9[3*order + height index].
extern "C" int cm1g_ensemble_tune_schedule(cm1g_ensemble *e, int steps, int *ty_best, int *order_best,
                                           float *ms9) {
  ARGCHK(e && ty_best && order_best && steps >= 1, "cm1g_ensemble_tune_schedule: bad argument");
  static const int cand[3] = {1, 2, 4};
  float best = 0.0f, ms = 0.0f;
  int bty = 4, bo = 0;
  int rc = cm1g_ensemble_set_tiles(e, cand[0]);
  if (!rc) rc = cm1g_ensemble_set_order(e, 0);
  if (!rc) rc = ens_time(e, steps, &ms);  // the device's warm-up
  if (rc) return rc;
  for (int o = 0; o < 3; o++)
    for (int c = 0; c < 3; c++) {
      rc = cm1g_ensemble_set_order(e, o);
      if (!rc) rc = cm1g_ensemble_set_tiles(e, cand[c]);
      if (!rc) rc = ens_time(e, steps, &ms);
      if (rc) return rc;
      if (ms9) ms9[3 * o + c] = ms;
      if ((o == 0 && c == 0) || ms < best) {
        best = ms;
        bty = cand[c];
        bo = o;
      }
    }
  *ty_best = bty;
  *order_best = bo;
  rc = cm1g_ensemble_set_order(e, bo);
  if (!rc) rc = cm1g_ensemble_set_tiles(e, bty);
  return rc;
}

// Try each tile height for `steps` timed ensemble steps (after one untimed
// step that captures its graph) and keep the fastest.  The steps taken are
// ordinary steps of the run: 4 * (steps + 1) of them in all.
extern "C" int cm1g_ensemble_tune(cm1g_ensemble *e, int steps, int *ty_best, float *ms_per_step) {
  ARGCHK(e && ty_best && steps >= 1, "cm1g_ensemble_tune: bad argument");
  static const int cand[3] = {1, 2, 4};
  float best = 0.0f;
  int bty = 4;
  // the first candidate is timed twice: its first timing includes the
  // device's warm-up
  for (int c = -1; c < 3; c++) {
    const int ty = cand[c < 0 ? 0 : c];
    int rc = cm1g_ensemble_set_tiles(e, ty);
    if (!rc) rc = cm1g_ensemble_step(e);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(e->stream));
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < steps; s++) {
      rc = cm1g_ensemble_step(e);
      if (rc) return rc;
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    const float ms =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
    if (c < 0) continue;
    if (ms_per_step) ms_per_step[c] = ms;
    if (c == 0 || ms < best) {
      best = ms;
      bty = ty;
    }
  }
  *ty_best = bty;
  return cm1g_ensemble_set_tiles(e, bty);
}
