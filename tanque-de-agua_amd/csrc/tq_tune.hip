// tq_tune.hip -- the solver's launch autotuner (tq_autotune): the project's own tuner over the
// launch shapes and the kernel structure of the pressure solve (pressao-cpu.js:296-402 on the
// device), timed on the tank's current pressure system.  The reference has no counterpart (its
// WebGPU engine fixes its dispatch shapes, js/gpu/multigrid.js:283-287).
//
// Knobs, in the order they are tuned (coordinate descent, one pass):
//   mg_dense        the first one-workgroup level's V-cycle as a dense matrix, or in one workgroup
//   split_cells     levels above this many cells run each multigrid pass as its own kernel
//                   (candidates: one per distinct set of split levels)
//   spmv_grid       most blocks of spmv (candidates that change the launched grid only)
//   update_grid     most blocks of the PCG update / level-0 input kernel
//   pass_fold_grid  most blocks of a split level 0's folding pass-B kernel
// Each trial runs warm-up and timed pressure solves of a fixed iteration count (no stop test, so
// every trial does the same arithmetic) with HIP events around each solve, restoring the state a
// solve changes (q, the bubbles' y, the device scalars) after every one, so tuning leaves the
// simulation where it was.  A change is kept only if it is faster by more than 1 %.  Results are
// cached per (nx, ny, precision) for the process; tq_create applies a cached result.
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "tq_internal.hpp"

namespace tq {

struct TuneSet {
  int spmv_grid, update_grid, split_cells, pass_fold_grid, mg_dense;
};

static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int>, TuneSet>& tune_cache() {
  static std::map<std::tuple<int, int, int>, TuneSet> m;
  return m;
}

static TuneSet current(const tq_sim* s) {
  return {s->launch_spmv_grid, s->launch_update_grid, s->launch_split_cells, s->launch_pass_fold_grid, s->d.mg_dense};
}

static void apply(tq_sim* s, const TuneSet& t) {
  const bool changed = s->launch_spmv_grid != t.spmv_grid || s->launch_update_grid != t.update_grid ||
                       s->launch_split_cells != t.split_cells || s->launch_pass_fold_grid != t.pass_fold_grid ||
                       s->d.mg_dense != t.mg_dense;
  s->launch_spmv_grid = t.spmv_grid;
  s->launch_update_grid = t.update_grid;
  s->launch_split_cells = t.split_cells;
  s->launch_pass_fold_grid = t.pass_fold_grid;
  s->d.mg_dense = t.mg_dense;
  if (changed) solver_launch_changed(s);
}

// a cached setting for this grid, if any (called by tq_create)
void tune_apply_cached(tq_sim* s) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  auto it = tune_cache().find(std::make_tuple(s->nx, s->ny, s->d.precision));
  if (it != tune_cache().end()) apply(s, it->second);
}

namespace {
// the state a pressure solve changes, saved once and put back after every trial solve
struct SolveState {
  float* q = nullptr;
  float* yB = nullptr;
  DevScalars* sc = nullptr;
  ~SolveState() {
    if (q) (void)hipFree(q);
    if (yB) (void)hipFree(yB);
    if (sc) (void)hipFree(sc);
  }
};
}  // namespace

static int save_state(tq_sim* s, SolveState& st) {
  TQ_HIP(hipMalloc((void**)&st.q, sizeof(float) * (size_t)s->C));
  TQ_HIP(hipMalloc((void**)&st.yB, sizeof(float) * MAX_BOLHAS));
  TQ_HIP(hipMalloc((void**)&st.sc, sizeof(DevScalars)));
  TQ_HIP(hipMemcpyAsync(st.q, s->q, sizeof(float) * (size_t)s->C, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(st.yB, s->yB32, sizeof(float) * MAX_BOLHAS, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(st.sc, s->sc, sizeof(DevScalars), hipMemcpyDeviceToDevice, s->st));
  return 0;
}

static int restore_state(tq_sim* s, const SolveState& st) {
  TQ_HIP(hipMemcpyAsync(s->q, st.q, sizeof(float) * (size_t)s->C, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(s->yB32, st.yB, sizeof(float) * MAX_BOLHAS, hipMemcpyDeviceToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(s->sc, st.sc, sizeof(DevScalars), hipMemcpyDeviceToDevice, s->st));
  return 0;
}

// mean device time of `reps` solves (microseconds), after one warm-up solve
static int time_solves(tq_sim* s, const SolveState& st, hipEvent_t e0, hipEvent_t e1, int reps, double* us) {
  double total = 0;
  for (int k = 0; k <= reps; k++) {
    TQ_HIP(hipEventRecord(e0, s->st));
    if (int e = pressure_solve(s)) return e;
    TQ_HIP(hipEventRecord(e1, s->st));
    if (int e = restore_state(s, st)) return e;
    TQ_HIP(hipEventSynchronize(e1));
    float ms = 0;
    TQ_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (k > 0) total += ms;
  }
  *us = 1e3 * total / reps;
  return 0;
}

}  // namespace tq

using namespace tq;

int tq_autotune(tq_sim* s, int iterations, tq_tune_result* out) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (s->tape) { set_error("tq_autotune: not allowed while frames are recorded"); return TQ_E_STATE; }
  if (iterations <= 0 || iterations > 1000) { set_error("tq_autotune: iterations out of range"); return TQ_E_ARG; }
  const tq_desc saved = s->d;
  const TuneSet start = current(s);
  s->d.mg_dense = dense_wanted(s) ? 1 : 0;  // the caller's choice, explicit (the trials change tol)
  s->d.tol = 0;  // fixed iteration count: every trial does the same arithmetic
  s->d.iteracoes = iterations;
  SolveState st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = save_state(s, st);
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = hip_fail(hipErrorOutOfMemory, "hipEventCreate", __FILE__, __LINE__);
  const int reps = 5;
  int trials = 0;
  TuneSet best = {0, 0, 0, 0, s->d.mg_dense};
  double best_us = 0, default_us = 0;
  auto trial = [&](const TuneSet& t, double* us) -> int {
    apply(s, t);
    trials++;
    return time_solves(s, st, e0, e1, reps, us);
  };
  if (!rc) rc = trial(best, &best_us);
  default_us = best_us;
  auto consider = [&](TuneSet cand) -> int {
    double us = 0;
    if (int e = trial(cand, &us)) return e;
    if (us < 0.99 * best_us) {
      best = cand;
      best_us = us;
    }
    return 0;
  };
  // candidate values that change what is launched on this grid
  const int L = (int)s->lv.size();
  if (!rc && s->vop != nullptr) {
    TuneSet t = best;
    t.mg_dense = 1 - best.mg_dense;
    rc = consider(t);
  }
  if (!rc) {
    // one threshold per set of split levels: none, levels 0..k for k = 0..3
    auto split_set = [&](int thr) {
      const int eff = thr > 0 ? thr : (1 << 20);
      unsigned m = 0;
      for (int l = 0; l < L; l++)
        if (s->lv[l].n > eff) m |= 1u << l;
      return m;
    };
    const TuneSet base = best;
    std::vector<unsigned> seen = {split_set(base.split_cells)};
    std::vector<int> thrs = {s->lv[0].n};
    for (int k = 0; k < 4 && k < L; k++) thrs.push_back(std::max(1, s->lv[k].n - 1));
    for (int thr : thrs) {
      if (rc) break;
      const unsigned m = split_set(thr);
      bool dup = false;
      for (unsigned v : seen) dup = dup || v == m;
      if (dup) continue;
      seen.push_back(m);
      TuneSet t = base;
      t.split_cells = thr;
      rc = consider(t);
    }
  }
  if (!rc) {
    const int spmv_blocks = ((s->nx + 63) / 64) * ((s->ny + 15) / 16);  // spmv's 64x16 tiles
    const TuneSet base = best;
    for (int cap : {256, 512, 2048, 4096}) {
      if (rc || cap >= spmv_blocks) continue;
      TuneSet t = base;
      t.spmv_grid = cap;
      rc = consider(t);
    }
  }
  if (!rc) {
    const int groups = ((s->nx + 15) / 16 + 1) / 2 * (((s->ny + 15) / 16 + 1) / 2);
    const TuneSet base = best;
    for (int cap : {512, 768, 2048, 4096}) {
      if (rc || cap >= groups) continue;
      TuneSet t = base;
      t.update_grid = cap;
      rc = consider(t);
    }
  }
  if (!rc) {
    const int split0 = s->lv[0].n > (best.split_cells > 0 ? best.split_cells : (1 << 20));
    const int groups8 = ((s->nx + 8 + 15) / 16 + 1) / 2 * (((s->ny + 8 + 15) / 16 + 1) / 2);
    const TuneSet base = best;
    for (int cap : {512, 1024, 4096}) {
      if (rc || !split0 || cap >= groups8) continue;
      TuneSet t = base;
      t.pass_fold_grid = cap;
      rc = consider(t);
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  // put the caller's solve options back, with the chosen shapes (or the starting ones on failure)
  const TuneSet keep = rc ? start : best;
  s->d = saved;
  apply(s, keep);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_cache()[std::make_tuple(s->nx, s->ny, s->d.precision)] = best;
  }
  if (out) {
    out->spmv_grid = best.spmv_grid;
    out->update_grid = best.update_grid;
    out->split_cells = best.split_cells;
    out->pass_fold_grid = best.pass_fold_grid;
    out->mg_dense = best.mg_dense;
    out->default_us = default_us;
    out->tuned_us = best_us;
    out->trials = trials;
  }
  return 0;
}
