// tq_api.hip -- the C ABI (include/tanque.h): state allocation, the substep sequence
// (SimulacaoCPU.subpasso, js/fisica-cpu.js:176-189) and the state accessors.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "tq_device.hpp"

namespace tq {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_err = buf;
  return TQ_E_HIP;
}

// ------------------------------------------------------------------ timing

// waits for the stream and adds the pending event pairs to their kernels' totals
static void timing_collect(tq_sim* s) {
  (void)hipStreamSynchronize(s->st);
  for (size_t k = 0; k < s->ev_pending.size(); k++) {
    const auto& pe = s->ev_pending[k];
    float ms = 0;
    if (hipEventElapsedTime(&ms, s->ev_pool[pe.second].first, s->ev_pool[pe.second].second) == hipSuccess) {
      s->ktimes[pe.first].total_ms += ms;
      s->ktimes[pe.first].launches += 1;
      s->ktimes[pe.first].total_bytes += s->ev_pending_bytes[k];
    }
  }
  s->ev_pending.clear();
  s->ev_pending_bytes.clear();
  s->ev_used = 0;
}

int time_begin(tq_sim* s, const char* name, double bytes) {
  if (!s->timing || s->capturing || s->lrec) return -1;
  if (!s->timing_only.empty() && s->timing_only != name) return -1;
  int ki = -1;
  for (size_t i = 0; i < s->ktimes.size(); i++)
    if (s->ktimes[i].name == name) { ki = (int)i; break; }
  if (ki < 0) {
    KTime k;
    k.name = name;
    s->ktimes.push_back(k);
    ki = (int)s->ktimes.size() - 1;
  }
  if (s->ev_used >= (int)s->ev_pool.size()) {
    if (s->ev_pool.size() >= 8192) {
      // flush: wait and accumulate what is pending, then reuse the pool
      timing_collect(s);
    } else {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess) return -1;
      if (hipEventCreate(&b) != hipSuccess) return -1;
      s->ev_pool.emplace_back(a, b);
    }
  }
  const int slot = s->ev_used++;
  (void)hipEventRecord(s->ev_pool[slot].first, s->st);
  s->ev_pending.emplace_back(ki, slot);
  s->ev_pending_bytes.push_back(bytes);
  return slot;
}

void time_end(tq_sim* s, int slot) {
  if (slot < 0) return;
  (void)hipEventRecord(s->ev_pool[slot].second, s->st);
}

static void timing_flush(tq_sim* s) { timing_collect(s); }

// fills n 32-bit words with zero: the form of hipMemsetAsync that a launch log can record
__device__ __forceinline__ void k_zero_words(Blk B_, uint32_t* __restrict__ p, int n) {
  for (int i = B_.x * 256 + (int)threadIdx.x; i < n; i += B_.gx * 256) p[i] = 0u;
}

int zero_async(tq_sim* s, hipStream_t st, void* p, size_t bytes) {
  if (!s->lrec) {
    TQ_HIP(hipMemsetAsync(p, 0, bytes, st));
    return 0;
  }
  if (bytes % 4 || (uintptr_t)p % 4 || bytes / 4 > (size_t)INT32_MAX) {
    set_error("zero_async: the range is not whole aligned words");
    return TQ_E_ARG;
  }
  const int n = (int)(bytes / 4);
  launch<256, 0, k_zero_words>(s, st, dim3(std::max(1, std::min((n + 255) / 256, 256))), dim3(256), 0, (uint32_t*)p, n);
  return 0;
}

// ------------------------------------------------------------------ allocation
template <class T>
static int dalloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  TQ_HIP(hipMalloc((void**)p, n * sizeof(T)));
  return 0;
}
#define DALLOC(p, n)                                                                   \
  do {                                                                                 \
    if (int e_ = dalloc(&(p), (n))) return e_;                                         \
    s->bufs.emplace_back((void*)(p), sizeof(*(p)) * std::max<size_t>((size_t)(n), 1)); \
  } while (0)

static void frame_sdf(std::vector<float>& sdf, int nx, int ny) { tq_scene_frame(sdf.data(), nx, ny); }

static int alloc_all(tq_sim* s) {
  const size_t C = s->C, nu = s->nu, nv = s->nv, cap = s->cap;
  for (int b = 0; b < 2; b++) {
    DALLOC(s->pos[b], cap); DALLOC(s->vel[b], cap); DALLOC(s->cm[b], cap); DALLOC(s->ext[b], cap);
    DALLOC(s->seq[b], cap);
  }
  DALLOC(s->pkey, cap + 1); DALLOC(s->prank, cap + 1); DALLOC(s->addflag, cap + 1);
  DALLOC(s->stage, cap * 12);
  DALLOC(s->sdf, 4 * C); DALLOC(s->solid, C);
  DALLOC(s->countKey, (size_t)s->K); DALLOC(s->startKey, (size_t)s->K + 1); DALLOC(s->countCell, C);
  DALLOC(s->u, nu); DALLOC(s->v, nv); DALLOC(s->uw, nu); DALLOC(s->vw, nv); DALLOC(s->uTmp, nu); DALLOC(s->vTmp, nv);
  DALLOC(s->uVal, nu); DALLOC(s->vVal, nv); DALLOC(s->uValTmp, nu); DALLOC(s->vValTmp, nv);
  DALLOC(s->dens, C); DALLOC(s->dye, 3 * C); DALLOC(s->tipo, C);
  DALLOC(s->fase, C); DALLOC(s->nX, C); DALLOC(s->nY, C); DALLOC(s->kappa, C);
  DALLOC(s->qAr, C); DALLOC(s->div, C); DALLOC(s->q, C);
  DALLOC(s->rot, C); DALLOC(s->cflag, C); DALLOC(s->cand, C); DALLOC(s->pai, C); DALLOC(s->comp, C);
  DALLOC(s->rootFlag, C + 1); DALLOC(s->compIdx, C + 1);
  DALLOC(s->compCells, C); DALLOC(s->compAtm, C); DALLOC(s->compTopo, C); DALLOC(s->compAlvo, C);
  DALLOC(s->compDono, C); DALLOC(s->compLabel, C);
  s->hsize = 1;
  while (s->hsize < (int)(2 * C)) s->hsize <<= 1;
  DALLOC(s->hkeys, (size_t)s->hsize); DALLOC(s->hcount, (size_t)s->hsize); DALLOC(s->hfirst, (size_t)s->hsize);
  DALLOC(s->hlist, (size_t)s->hsize);
  DALLOC(s->ordBuf, 3 * (size_t)s->hsize);
  DALLOC(s->pBolha, MAX_BOLHAS); DALLOC(s->volB, MAX_BOLHAS); DALLOC(s->facesB, MAX_BOLHAS); DALLOC(s->cellsB, MAX_BOLHAS);
  DALLOC(s->rhsB, MAX_BOLHAS); DALLOC(s->rhsB32, MAX_BOLHAS); DALLOC(s->yB32, MAX_BOLHAS); DALLOC(s->volAlvo, MAX_BOLHAS);
  DALLOC(s->vivo, MAX_BOLHAS); DALLOC(s->vivoNovo, MAX_BOLHAS);
  DALLOC(s->bvec, BV_COUNT * (size_t)MAX_BOLHAS); DALLOC(s->activeList, MAX_BOLHAS);
  DALLOC(s->opMask, C); DALLOC(s->r, C); DALLOC(s->ap, C); DALLOC(s->qlo, C);
  for (int b = 0; b < 2; b++) { DALLOC(s->zbuf[b], C); DALLOC(s->pbuf[b], C); }
  DALLOC(s->partials, RED_SLOTS * (size_t)MAX_RED_BLOCKS);
  DALLOC(s->partials_r, RED_SLOTS * (size_t)MAX_RED_BLOCKS);
  DALLOC(s->sc, 1);
  {
    // block sums of the largest scan (particle slots, binning keys or cells), two levels
    const size_t nmax = std::max({(size_t)s->cap, (size_t)s->K, (size_t)C}) + 1;
    const size_t nb = (nmax + 4095) / 4096;
    s->scan_cap = (int)(2 * nb + (nb + 4095) / 4096 + 2);
    DALLOC(s->scanBlocks, (size_t)s->scan_cap);
  }
  TQ_HIP(hipHostMalloc((void**)&s->sc_host, sizeof(DevScalars), hipHostMallocDefault));
  if (int e = solver_alloc(s)) return e;
  // zero state (the reference allocates zero-filled typed arrays)
  auto z = [&](void* p, size_t bytes) { return hipMemsetAsync(p, 0, bytes, s->st); };
  TQ_HIP(z(s->u, 4 * nu)); TQ_HIP(z(s->v, 4 * nv)); TQ_HIP(z(s->uw, 4 * nu)); TQ_HIP(z(s->vw, 4 * nv));
  TQ_HIP(z(s->uVal, nu)); TQ_HIP(z(s->vVal, nv));
  TQ_HIP(z(s->dens, 4 * C)); TQ_HIP(z(s->dye, 12 * C)); TQ_HIP(z(s->countCell, 4 * C)); TQ_HIP(z(s->tipo, C));
  TQ_HIP(z(s->fase, 4 * C)); TQ_HIP(z(s->nX, 4 * C)); TQ_HIP(z(s->nY, 4 * C)); TQ_HIP(z(s->kappa, 4 * C));
  TQ_HIP(z(s->qAr, 4 * C)); TQ_HIP(z(s->div, 4 * C)); TQ_HIP(z(s->q, 4 * C));
  TQ_HIP(z(s->rot, 4 * C)); TQ_HIP(z(s->cflag, C));
  TQ_HIP(hipMemsetAsync(s->hkeys, 0xFF, 8 * (size_t)s->hsize, s->st));
  TQ_HIP(z(s->hcount, 4 * (size_t)s->hsize));
  TQ_HIP(hipMemsetAsync(s->hfirst, 0xFF, 4 * (size_t)s->hsize, s->st));
  TQ_HIP(z(s->pBolha, 4 * MAX_BOLHAS)); TQ_HIP(z(s->volB, 4 * MAX_BOLHAS)); TQ_HIP(z(s->facesB, 4 * MAX_BOLHAS));
  TQ_HIP(z(s->cellsB, 4 * MAX_BOLHAS)); TQ_HIP(z(s->rhsB, 8 * MAX_BOLHAS)); TQ_HIP(z(s->rhsB32, 4 * MAX_BOLHAS));
  TQ_HIP(z(s->yB32, 4 * MAX_BOLHAS)); TQ_HIP(z(s->volAlvo, 4 * MAX_BOLHAS)); TQ_HIP(z(s->vivo, MAX_BOLHAS));
  TQ_HIP(z(s->bvec, 8 * BV_COUNT * (size_t)MAX_BOLHAS));
  for (int b = 0; b < 2; b++) { TQ_HIP(z(s->zbuf[b], 4 * C)); TQ_HIP(z(s->pbuf[b], 4 * C)); }
  TQ_HIP(z(s->r, 4 * C)); TQ_HIP(z(s->ap, 4 * C)); TQ_HIP(z(s->qlo, 4 * C));
  TQ_HIP(z(s->sc, sizeof(DevScalars)));
  std::memset(s->sc_host, 0, sizeof(DevScalars));
  return 0;
}

// A tank of a table batch enqueues on the batch's stream (shared by every tank of the batch, which
// then orders the tanks' own calls and the batch's table launches): its own stream is released.
int share_stream(tq_sim* s, hipStream_t st) {
  TQ_HIP(hipStreamSynchronize(s->st));
  if (s->own_st) TQ_HIP(hipStreamDestroy(s->st));
  s->st = st;
  s->own_st = false;
  return 0;
}

static void free_all(tq_sim* s) {
  auto f = [](void* p) { if (p) (void)hipFree(p); };
  for (int b = 0; b < 2; b++) { f(s->pos[b]); f(s->vel[b]); f(s->cm[b]); f(s->ext[b]); f(s->seq[b]); }
  f(s->pkey); f(s->prank); f(s->addflag); f(s->stage); f(s->render_img); f(s->sdf); f(s->solid); f(s->countKey); f(s->startKey);
  f(s->countCell); f(s->u); f(s->v); f(s->uw); f(s->vw); f(s->uTmp); f(s->vTmp); f(s->uVal); f(s->vVal);
  f(s->uValTmp); f(s->vValTmp); f(s->dens); f(s->dye); f(s->tipo); f(s->fase); f(s->nX); f(s->nY); f(s->kappa);
  f(s->qAr); f(s->div); f(s->q); f(s->rot); f(s->cflag); f(s->cand); f(s->pai); f(s->comp); f(s->rootFlag);
  f(s->compIdx); f(s->compCells); f(s->compAtm); f(s->compTopo); f(s->compAlvo); f(s->compDono); f(s->compLabel);
  f(s->hkeys); f(s->hcount); f(s->hfirst); f(s->hlist); f(s->ordBuf); f(s->pBolha); f(s->volB); f(s->facesB);
  f(s->cellsB); f(s->rhsB); f(s->rhsB32); f(s->yB32); f(s->volAlvo); f(s->vivo); f(s->vivoNovo); f(s->bvec);
  f(s->activeList); f(s->opMask); f(s->r); f(s->ap); f(s->qlo);
  for (int b = 0; b < 2; b++) { f(s->zbuf[b]); f(s->pbuf[b]); }
  f(s->partials); f(s->partials_r); f(s->sc); f(s->scanBlocks);
  if (s->sc_host) (void)hipHostFree(s->sc_host);
  if (s->status_dev) (void)hipFree(s->status_dev);
  if (s->status_host) (void)hipHostFree(s->status_host);
  if (s->ev_status) (void)hipEventDestroy(s->ev_status);
  if (s->ev_frame0) (void)hipEventDestroy(s->ev_frame0);
  if (s->ev_frame1) (void)hipEventDestroy(s->ev_frame1);
  solver_free(s);
  for (auto& e : s->ev_pool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
  if (s->st && s->own_st) (void)hipStreamDestroy(s->st);
}

// ------------------------------------------------------------------ the substep
// (with a substep record open, the hooks copy what the reverse pass needs)
static int substep(tq_sim* s, double dt) {
  s->substeps_run++;
  if (int e = p2g(s, dt, 1)) return e;           // #paraGrade + #gravidadeEFronteira + #tipos
  s->tipos_prontos = true;
  if (s->rec) if (int e = rec_hook(s, REC_P2G, dt)) return e;
  if (s->d.ar) if (int e = propagate_labels(s)) return e;
  if (int e = surface(s)) return e;
  if (int e = air_pressure_and_div(s, dt, 3)) return e;  // #pressaoDoAr + #classificar
  if (s->d.ar) if (int e = bubble_rhs(s, dt)) return e;
  if (s->rec) if (int e = rec_hook(s, REC_SOLVE, dt)) return e;
  if (int e = pressure_solve(s)) return e;
  if (int e = pressure_post(s, dt)) return e;
  if (int e = gradient(s)) return e;
  if (int e = extrapolate(s)) return e;
  if (s->rec) if (int e = rec_hook(s, REC_EXTRAP, dt)) return e;
  if (int e = g2p(s, dt)) return e;
  if (s->rec) if (int e = rec_hook(s, REC_G2P, dt)) return e;
  return 0;
}

static int update_air(tq_sim* s) {
  if (!s->tipos_prontos) {
    if (int e = p2g(s, 0.0, 0)) return e;
    s->tipos_prontos = true;
  }
  return tracker_update(s);
}

static int reset_air(tq_sim* s) {
  s->tipos_prontos = false;
  return tracker_reset(s);
}

static int sync(tq_sim* s) {
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

static int read_scalars(tq_sim* s) {
  TQ_HIP(hipMemcpyAsync(s->sc_host, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

// ------------------------------------------------------------------ fields
struct FieldRef {
  void* ptr;
  size_t elems;
  int kind;  // 0 f32, 1 u8, 2 u32, 3 u32 on device exposed as f32 (volumeBolha)
};

static bool field_ref(tq_sim* s, int f, FieldRef* r) {
  const size_t C = s->C;
  switch (f) {
    case 0: *r = {s->u, (size_t)s->nu, 0}; return true;
    case 1: *r = {s->v, (size_t)s->nv, 0}; return true;
    case 2: *r = {s->uw, (size_t)s->nu, 0}; return true;
    case 3: *r = {s->vw, (size_t)s->nv, 0}; return true;
    case 4: *r = {s->uVal, (size_t)s->nu, 1}; return true;
    case 5: *r = {s->vVal, (size_t)s->nv, 1}; return true;
    case 6: *r = {s->dens, C, 0}; return true;
    case 7: *r = {s->countCell, C, 2}; return true;
    case 8: *r = {s->dye, 3 * C, 0}; return true;
    case 9: *r = {s->tipo, C, 1}; return true;
    case 10: *r = {s->solid, C, 1}; return true;
    case 11: *r = {s->div, C, 0}; return true;
    case 12: *r = {s->q, C, 0}; return true;
    case 13: *r = {s->fase, C, 0}; return true;
    case 14: *r = {s->nX, C, 0}; return true;
    case 15: *r = {s->nY, C, 0}; return true;
    case 16: *r = {s->kappa, C, 0}; return true;
    case 17: *r = {s->qAr, C, 0}; return true;
    case 18: *r = {s->rot, C, 2}; return true;
    case 19: *r = {s->pBolha, (size_t)MAX_BOLHAS, 0}; return true;
    case 20: *r = {s->volB, (size_t)MAX_BOLHAS, 3}; return true;
    case 21: *r = {s->rhsB32, (size_t)MAX_BOLHAS, 0}; return true;
    case 22: *r = {s->yB32, (size_t)MAX_BOLHAS, 0}; return true;
    case 23: *r = {s->volAlvo, (size_t)MAX_BOLHAS, 0}; return true;
    case 24: *r = {s->vivo, (size_t)MAX_BOLHAS, 1}; return true;
    default: return false;
  }
}

static size_t kind_bytes(int k) { return k == 1 ? 1 : 4; }

}  // namespace tq

using namespace tq;

#define TQ_CHECK_SIM(s) do { if (!(s)) { set_error("null handle"); return TQ_E_ARG; } } while (0)
// calls that change the state other than by the recorded operations (faucet, step) would make the
// recorded frames unreproducible for tq_backward
#define TQ_NOT_TAPING(s, what)                                                                 \
  do {                                                                                         \
    TQ_CHECK_SIM(s);                                                                           \
    if ((s)->tape) { set_error(what ": not allowed while frames are recorded"); return TQ_E_STATE; } \
  } while (0)

namespace tq {
// avancar's work enqueued on the tank's stream without waiting for it (tq_batch_step)
int step_enqueue(tq_sim* s, double dt, int substeps) {
  if (!s) { set_error("null handle"); return TQ_E_ARG; }
  if (!(dt > 0) || substeps < 0) { set_error("tq_step: bad dt or substep count"); return TQ_E_ARG; }
  if (s->d.ar) if (int e = update_air(s)) return e;
  const double d = dt / substeps;
  for (int k = 0; k < substeps; k++)
    if (int e = substep(s, d)) return e;
  return 0;
}
int stream_wait(tq_sim* s) { return sync(s); }

// The launches of the same frame appended to `log` instead of launched: round 0 holds the air
// update, round k (1..substeps) substep k.  The host-side state (buffer parities, iteration
// counters) advances as if the frame had been enqueued, so the log must be launched next.
int step_record(tq_sim* s, double dt, int substeps, LaunchLog* log) {
  if (!s || !log) { set_error("null handle"); return TQ_E_ARG; }
  if (!(dt > 0) || substeps < 0) { set_error("tq_step: bad dt or substep count"); return TQ_E_ARG; }
  if (s->tape || s->rec) { set_error("table batch: a tank is recording frames (tq_tape_begin)"); return TQ_E_STATE; }
  if (s->timing) { set_error("table batch: kernel timing is on for a tank"); return TQ_E_STATE; }
  if (s->d.tol > 0) {
    set_error("table batch: a tank solves to a tolerance (a table frame needs a fixed PCG iteration count, tol = 0)");
    return TQ_E_STATE;
  }
  s->lrec = log;
  log->rounds.push_back((uint32_t)log->ops.size());
  int e = s->d.ar ? update_air(s) : 0;
  const double d = dt / substeps;
  for (int k = 0; k < substeps && !e; k++) {
    log->rounds.push_back((uint32_t)log->ops.size());
    e = substep(s, d);
  }
  s->lrec = nullptr;
  return e;
}
}  // namespace tq

namespace tq {
int emit_faucet_impl(tq_sim* s, const tq_faucet* f, double dt, uint32_t* seed, int* emitted, int* accepted) {
  if (!f || !seed) { set_error("tq_emit_faucet: null argument"); return TQ_E_ARG; }
  if (emitted) *emitted = 0;
  if (accepted) *accepted = 0;
  // emitirTorneira's velocity, flow and count (principal.js:301-306)
  double vx = f->pvx * 0.5, vy = f->pvy * 0.5 - 0.7;
  const double v = js_hypot(vx, vy);
  if (v > 4) { vx *= 4 / v; vy *= 4 / v; }
  const double vazao = 2 * f->raio * std::max(0.7, js_hypot(vx, vy));
  const double nd = std::min(40000.0, js_round((vazao * dt) / (s->h * s->h) * PPC));
  const int n = nd > 0 ? (int)nd : 0;
  if (n <= 0) return 0;
  int keep = n;
  if ((long long)s->n_upper + keep > s->cap) {
    // reclaim dead slots first, then cap the list as MotorCPU does (motor.js:100-104)
    if (int e = bin_and_sort(s)) return e;
    if (int e = read_scalars(s)) return e;
    s->n_upper = (int)s->sc_host->n_slots;
    keep = std::max(0, std::min(keep, s->cap - s->n_upper));
  }
  if ((unsigned long long)s->seq_next + keep >= 0xFFFFFFF0ull) { set_error("tq_emit_faucet: insertion counter exhausted"); return TQ_E_CAPACITY; }
  if (int e = faucet_dev(s, f, vx, vy, n, keep, seed)) return e;
  s->seq_next += (uint32_t)keep;
  s->n_upper += keep;
  if (emitted) *emitted = n;
  if (accepted) {
    if (int e = read_scalars(s)) return e;
    *accepted = keep > 0 ? (int)s->sc_host->add_count : 0;
  }
  return 0;
}
}  // namespace tq

namespace tq {
// downloads the live particles ordered by insertion number (the reference's order)
int download_ordered(tq_sim* s, std::vector<float>& out, std::vector<uint32_t>* seqs) {
  if (int e = read_scalars(s)) return e;
  const size_t n = s->sc_host->n_slots;
  const int o = s->cur;
  std::vector<float2> pos(n), vel(n);
  std::vector<float4> cm(n), ext(n);
  std::vector<uint32_t> sq(n);
  if (n) {
    TQ_HIP(hipMemcpyAsync(pos.data(), s->pos[o], n * sizeof(float2), hipMemcpyDeviceToHost, s->st));
    TQ_HIP(hipMemcpyAsync(vel.data(), s->vel[o], n * sizeof(float2), hipMemcpyDeviceToHost, s->st));
    TQ_HIP(hipMemcpyAsync(cm.data(), s->cm[o], n * sizeof(float4), hipMemcpyDeviceToHost, s->st));
    TQ_HIP(hipMemcpyAsync(ext.data(), s->ext[o], n * sizeof(float4), hipMemcpyDeviceToHost, s->st));
    TQ_HIP(hipMemcpyAsync(sq.data(), s->seq[o], n * sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
    TQ_HIP(hipStreamSynchronize(s->st));
  }
  std::vector<uint32_t> idx;
  idx.reserve(n);
  for (size_t k = 0; k < n; k++) if (pos[k].x == pos[k].x) idx.push_back((uint32_t)k);
  std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return sq[a] < sq[b]; });
  out.resize(idx.size() * 12);
  if (seqs) seqs->resize(idx.size());
  for (size_t m = 0; m < idx.size(); m++) {
    const uint32_t k = idx[m];
    float* p = &out[m * 12];
    p[0] = pos[k].x; p[1] = pos[k].y; p[2] = vel[k].x; p[3] = vel[k].y;
    p[4] = cm[k].x; p[5] = cm[k].y; p[6] = cm[k].z; p[7] = cm[k].w;
    p[8] = ext[k].x; p[9] = ext[k].y; p[10] = ext[k].z; p[11] = ext[k].w;
    if (seqs) (*seqs)[m] = sq[k];
  }
  return 0;
}
}  // namespace tq

extern "C" {

const char* tq_last_error(void) { return g_err.c_str(); }

void tq_desc_default(tq_desc* d, int nx, int ny, double h) {
  std::memset(d, 0, sizeof *d);
  d->nx = nx; d->ny = ny; d->h = h;
  d->gx = 0; d->gy = -GRAVIDADE;
  d->iteracoes = ITERACOES_PCG;
  d->tau_densidade = TAU_DENSIDADE;
  d->tensao = TENSAO_SUPERFICIAL;
  d->ar = 1;
  d->max_particles = 0;
  d->precision = TQ_PREC_FP32;
  d->tol = 0;
  d->mg_min_lds_level = 0;
  d->mg_half_level = -1;
  d->use_graph = 0;
  d->mg_dense = -1;
}

int tq_create(const tq_desc* d, tq_sim** out) {
  if (!d || !out) { set_error("tq_create: null argument"); return TQ_E_ARG; }
  *out = nullptr;
  if (d->nx < 4 || d->ny < 4 || !(d->h > 0)) { set_error("tq_create: grid must be at least 4x4 with h > 0"); return TQ_E_ARG; }
  if ((long long)d->nx * d->ny > (1LL << 26)) { set_error("tq_create: grid too large"); return TQ_E_CAPACITY; }
  if (d->precision < TQ_PREC_FP32 || d->precision > TQ_PREC_MIXED_X) { set_error("tq_create: precision"); return TQ_E_ARG; }
  tq_sim* s = new tq_sim();
  s->d = *d;
  s->nx = d->nx; s->ny = d->ny; s->C = d->nx * d->ny; s->h = d->h;
  s->nu = (d->nx + 1) * d->ny; s->nv = d->nx * (d->ny + 1);
  s->ntx = (d->nx + 15) / 16; s->nty = (d->ny + 15) / 16;
  s->K = s->ntx * s->nty * 256;
  s->cap = d->max_particles > 0 ? d->max_particles
                                : (int)std::ceil((double)d->nx * d->ny * PARTICULAS_MAX_POR_CELULA);
  hipError_t e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
  if (e != hipSuccess) { hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); delete s; return TQ_E_HIP; }
  if (int rc = alloc_all(s)) { free_all(s); delete s; return rc; }
  // the empty tank: frame-only SDF (cena.js:139-144), solid cells, fresh air tracker
  std::vector<float> sdf((size_t)4 * s->C);
  frame_sdf(sdf, s->nx, s->ny);
  if (hipMemcpyAsync(s->sdf, sdf.data(), sdf.size() * 4, hipMemcpyHostToDevice, s->st) != hipSuccess ||
      solid_cells(s) != 0 || tracker_reset(s) != 0 || hipStreamSynchronize(s->st) != hipSuccess) {
    set_error("tq_create: initialisation failed");
    free_all(s);
    delete s;
    return TQ_E_HIP;
  }
  s->tipos_prontos = false;
  tune_apply_cached(s);
  *out = s;
  return 0;
}

int tq_destroy(tq_sim* s) {
  if (!s) return 0;
  (void)hipStreamSynchronize(s->st);
  tape_free(s);
  free_all(s);
  delete s;
  return 0;
}

int tq_set_sdf(tq_sim* s, const float* sdf, int row0, int row1) {
  TQ_NOT_TAPING(s, "tq_set_sdf");
  TQ_CHECK_SIM(s);
  const int mw = 2 * s->nx, mh = 2 * s->ny;
  if (!sdf || row0 < 0 || row1 >= mh || row0 > row1) { set_error("tq_set_sdf: bad rows"); return TQ_E_ARG; }
  TQ_HIP(hipMemcpyAsync(s->sdf + (size_t)row0 * mw, sdf + (size_t)row0 * mw, sizeof(float) * mw * (row1 - row0 + 1),
                        hipMemcpyHostToDevice, s->st));
  if (int e = solid_cells(s)) return e;
  return sync(s);
}

int tq_add_particles(tq_sim* s, const float* list, int count, int check, int* accepted) {
  TQ_NOT_TAPING(s, "tq_add_particles");
  TQ_CHECK_SIM(s);
  if (count < 0 || (count > 0 && !list)) { set_error("tq_add_particles: bad list"); return TQ_E_ARG; }
  if (accepted) *accepted = 0;
  if (count == 0) return 0;
  if ((long long)s->n_upper + count > s->cap) {
    // reclaim dead slots first, then cap the list as MotorCPU does (motor.js:100-104)
    if (int e = bin_and_sort(s)) return e;
    if (int e = read_scalars(s)) return e;
    s->n_upper = (int)s->sc_host->n_slots;
    count = std::min(count, s->cap - s->n_upper);
    if (count <= 0) { if (accepted) *accepted = 0; return 0; }
  }
  if ((unsigned long long)s->seq_next + count >= 0xFFFFFFF0ull) { set_error("tq_add_particles: insertion counter exhausted"); return TQ_E_CAPACITY; }
  TQ_HIP(hipMemcpyAsync(s->stage, list, sizeof(float) * 12 * (size_t)count, hipMemcpyHostToDevice, s->st));
  if (int e = add_particles_dev(s, count, check)) return e;
  s->seq_next += (uint32_t)count;
  s->n_upper += count;
  if (accepted) {
    if (int e = read_scalars(s)) return e;
    *accepted = (int)s->sc_host->add_count;
  } else if (int e = sync(s)) return e;  // the caller may free `list` now (copy is complete)
  return 0;
}

int tq_emit_faucet(tq_sim* s, const tq_faucet* f, double dt, uint32_t* seed, int* emitted, int* accepted) {
  TQ_CHECK_SIM(s);
  if (!f || !seed) { set_error("tq_emit_faucet: null argument"); return TQ_E_ARG; }
  const uint32_t seed0 = *seed, seq0 = s->seq_next;
  if (int e = emit_faucet_impl(s, f, dt, seed, emitted, accepted)) return e;
  if (s->tape) return tape_log_emit(s, f, dt, seed0, seq0, s->seq_next);
  return 0;
}

int tq_apply_tool(tq_sim* s, const tq_tool* t) {
  TQ_NOT_TAPING(s, "tq_apply_tool");
  TQ_CHECK_SIM(s);
  if (!t) return 0;
  // queued behind the tank's work: the tool's parameters travel by value, nothing to wait for
  return apply_tool_dev(s, t);
}

static int status_buffers(tq_sim* s) {
  if (s->status_dev) return 0;
  TQ_HIP(hipMalloc((void**)&s->status_dev, 2 * sizeof(uint64_t)));
  TQ_HIP(hipHostMalloc((void**)&s->status_host, 2 * sizeof(uint64_t), hipHostMallocDefault));
  TQ_HIP(hipEventCreate(&s->ev_status));
  TQ_HIP(hipEventCreate(&s->ev_frame0));
  TQ_HIP(hipEventCreate(&s->ev_frame1));
  return 0;
}

int tq_status_request(tq_sim* s) {
  TQ_CHECK_SIM(s);
  if (int e = status_buffers(s)) return e;
  if (int e = status_enqueue(s)) return e;
  TQ_HIP(hipMemcpyAsync(s->status_host, s->status_dev, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipEventRecord(s->ev_status, s->st));
  s->status_pending = true;
  s->status_has = s->has_vmax;
  return 0;
}

int tq_step_async(tq_sim* s, double dt, int substeps) {
  TQ_NOT_TAPING(s, "tq_step_async");
  TQ_CHECK_SIM(s);
  if (int e = status_buffers(s)) return e;
  TQ_HIP(hipEventRecord(s->ev_frame0, s->st));
  if (int e = tq::step_enqueue(s, dt, substeps)) return e;
  TQ_HIP(hipEventRecord(s->ev_frame1, s->st));
  s->status_timed = true;
  return tq_status_request(s);
}

int tq_status_poll(tq_sim* s, int* ready, int* n, double* vmax, int* has_vmax, double* gpu_ms) {
  TQ_CHECK_SIM(s);
  if (!ready) { set_error("tq_status_poll: null ready"); return TQ_E_ARG; }
  *ready = 0;
  if (!s->status_pending) return 0;
  const hipError_t q = hipEventQuery(s->ev_status);
  if (q == hipErrorNotReady) return 0;
  if (q != hipSuccess) return hip_fail(q, "hipEventQuery", __FILE__, __LINE__);
  *ready = 1;
  if (n) *n = (int)(uint32_t)s->status_host[0];
  if (vmax) {
    const uint64_t b = s->status_host[1];
    std::memcpy(vmax, &b, sizeof(double));
  }
  if (has_vmax) *has_vmax = s->status_has ? 1 : 0;
  if (gpu_ms) {
    float ms = -1.f;
    // the last tq_step_async frame, complete once the request queued after it is
    if (s->status_timed) TQ_HIP(hipEventElapsedTime(&ms, s->ev_frame0, s->ev_frame1));
    *gpu_ms = ms;
  }
  return 0;
}

int tq_substep(tq_sim* s, double dt) {
  TQ_NOT_TAPING(s, "tq_substep");
  TQ_CHECK_SIM(s);
  if (!(dt > 0)) { set_error("tq_substep: dt must be > 0"); return TQ_E_ARG; }
  if (int e = substep(s, dt)) return e;
  return sync(s);
}

int tq_step(tq_sim* s, double dt, int substeps) {
  if (int e = tq::step_enqueue(s, dt, substeps)) return e;
  if (int e = sync(s)) return e;
  if (s->tape) return tape_log_step(s, dt, substeps);
  return 0;
}

int tq_update_air(tq_sim* s) {
  TQ_NOT_TAPING(s, "tq_update_air");
  TQ_CHECK_SIM(s);
  if (int e = update_air(s)) return e;
  return sync(s);
}

int tq_reset_air(tq_sim* s) {
  TQ_NOT_TAPING(s, "tq_reset_air");
  TQ_CHECK_SIM(s);
  if (int e = reset_air(s)) return e;
  return sync(s);
}

int tq_clear_water(tq_sim* s) {
  TQ_NOT_TAPING(s, "tq_clear_water");
  TQ_CHECK_SIM(s);
  const uint32_t zero = 0;
  TQ_HIP(hipMemcpyAsync(&s->sc->n_slots, &zero, sizeof zero, hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemsetAsync(s->countCell, 0, sizeof(uint32_t) * s->C, s->st));
  s->n_upper = 0;
  if (int e = reset_air(s)) return e;
  return sync(s);
}

int tq_count(tq_sim* s, int* n) {
  TQ_CHECK_SIM(s);
  if (!n) { set_error("tq_count: null"); return TQ_E_ARG; }
  return count_live(s, n);
}

int tq_read_particles(tq_sim* s, float* dst, int max, int* n) {
  TQ_CHECK_SIM(s);
  std::vector<float> all;
  if (int e = download_ordered(s, all, nullptr)) return e;
  const int live = (int)(all.size() / 12);
  const int m = std::min(live, std::max(max, 0));
  if (m > 0) {
    if (!dst) { set_error("tq_read_particles: null dst"); return TQ_E_ARG; }
    std::memcpy(dst, all.data(), sizeof(float) * 12 * (size_t)m);
  }
  if (n) *n = live;
  return 0;
}

int tq_write_particles(tq_sim* s, const float* src, int n) {
  TQ_NOT_TAPING(s, "tq_write_particles");
  TQ_CHECK_SIM(s);
  if (n < 0 || n > s->cap || (n > 0 && !src)) { set_error("tq_write_particles: bad count"); return TQ_E_ARG; }
  std::vector<float2> pos(n), vel(n);
  std::vector<float4> cm(n), ext(n);
  std::vector<uint32_t> sq(n);
  for (int k = 0; k < n; k++) {
    const float* p = src + (size_t)k * 12;
    pos[k] = make_float2(p[0], p[1]); vel[k] = make_float2(p[2], p[3]);
    cm[k] = make_float4(p[4], p[5], p[6], p[7]); ext[k] = make_float4(p[8], p[9], p[10], p[11]);
    sq[k] = (uint32_t)k;
  }
  const int o = s->cur;
  if (n) {
    TQ_HIP(hipMemcpyAsync(s->pos[o], pos.data(), n * sizeof(float2), hipMemcpyHostToDevice, s->st));
    TQ_HIP(hipMemcpyAsync(s->vel[o], vel.data(), n * sizeof(float2), hipMemcpyHostToDevice, s->st));
    TQ_HIP(hipMemcpyAsync(s->cm[o], cm.data(), n * sizeof(float4), hipMemcpyHostToDevice, s->st));
    TQ_HIP(hipMemcpyAsync(s->ext[o], ext.data(), n * sizeof(float4), hipMemcpyHostToDevice, s->st));
    TQ_HIP(hipMemcpyAsync(s->seq[o], sq.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s->st));
  }
  const uint32_t nn = (uint32_t)n;
  TQ_HIP(hipMemcpyAsync(&s->sc->n_slots, &nn, sizeof nn, hipMemcpyHostToDevice, s->st));
  s->n_upper = n;
  s->seq_next = (uint32_t)n;
  return sync(s);
}

int tq_truncate(tq_sim* s, int m) {
  TQ_NOT_TAPING(s, "tq_truncate");
  TQ_CHECK_SIM(s);
  if (m < 0) { set_error("tq_truncate: negative"); return TQ_E_ARG; }
  std::vector<float> all;
  if (int e = download_ordered(s, all, nullptr)) return e;
  const int live = (int)(all.size() / 12);
  if (m >= live) return 0;
  return tq_write_particles(s, all.data(), m);
}

int tq_vmax(tq_sim* s, double* v, int* has) {
  TQ_CHECK_SIM(s);
  if (int e = read_scalars(s)) return e;
  unsigned long long b = s->sc_host->vmax_bits;
  double d;
  std::memcpy(&d, &b, sizeof d);
  if (v) *v = d;
  if (has) *has = s->has_vmax ? 1 : 0;
  return 0;
}

int tq_area(tq_sim* s, double* a) {
  TQ_CHECK_SIM(s);
  int n = 0;
  if (int e = count_live(s, &n)) return e;
  if (a) *a = ((double)n / PPC) * s->h * s->h;
  return 0;
}

int tq_set_options(tq_sim* s, const tq_desc* d) {
  TQ_NOT_TAPING(s, "tq_set_options");
  TQ_CHECK_SIM(s);
  if (!d) { set_error("tq_set_options: null"); return TQ_E_ARG; }
  if (d->precision < TQ_PREC_FP32 || d->precision > TQ_PREC_MIXED_X) {
    set_error("tq_set_options: precision");
    return TQ_E_ARG;
  }
  if (d->mg_dense < -1 || d->mg_dense > 1) {
    set_error("tq_set_options: mg_dense is -1, 0 or 1");
    return TQ_E_ARG;
  }
  s->d.gx = d->gx; s->d.gy = d->gy; s->d.iteracoes = d->iteracoes; s->d.tau_densidade = d->tau_densidade;
  s->d.tensao = d->tensao; s->d.ar = d->ar; s->d.precision = d->precision; s->d.tol = d->tol;
  s->d.mg_half_level = d->mg_half_level; s->d.use_graph = d->use_graph; s->d.mg_dense = d->mg_dense;
  return 0;
}

int tq_read_field(tq_sim* s, int f, void* dst, size_t bytes) {
  TQ_CHECK_SIM(s);
  FieldRef r;
  if (!field_ref(s, f, &r)) { set_error("tq_read_field: unknown field"); return TQ_E_ARG; }
  if (!dst || bytes != r.elems * kind_bytes(r.kind)) { set_error("tq_read_field: size mismatch"); return TQ_E_ARG; }
  TQ_HIP(hipMemcpyAsync(dst, r.ptr, bytes, hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  if (r.kind == 3) {
    uint32_t* u = (uint32_t*)dst;
    float* o = (float*)dst;
    for (size_t k = 0; k < r.elems; k++) o[k] = (float)u[k];
  }
  return 0;
}

int tq_write_field(tq_sim* s, int f, const void* src, size_t bytes) {
  TQ_NOT_TAPING(s, "tq_write_field");
  TQ_CHECK_SIM(s);
  FieldRef r;
  if (!field_ref(s, f, &r)) { set_error("tq_write_field: unknown field"); return TQ_E_ARG; }
  if (!src || bytes != r.elems * kind_bytes(r.kind)) { set_error("tq_write_field: size mismatch"); return TQ_E_ARG; }
  if (r.kind == 3) {
    std::vector<uint32_t> u(r.elems);
    for (size_t k = 0; k < r.elems; k++) u[k] = (uint32_t)((const float*)src)[k];
    TQ_HIP(hipMemcpyAsync(r.ptr, u.data(), bytes, hipMemcpyHostToDevice, s->st));
    TQ_HIP(hipStreamSynchronize(s->st));
  } else {
    TQ_HIP(hipMemcpyAsync(r.ptr, src, bytes, hipMemcpyHostToDevice, s->st));
    if (f == 21) {  // rhsBolha: the solver reads the FP64 copy
      std::vector<double> dd(r.elems);
      for (size_t k = 0; k < r.elems; k++) dd[k] = ((const float*)src)[k];
      TQ_HIP(hipMemcpyAsync(s->rhsB, dd.data(), 8 * r.elems, hipMemcpyHostToDevice, s->st));
      TQ_HIP(hipStreamSynchronize(s->st));
    }
    TQ_HIP(hipStreamSynchronize(s->st));
  }
  if (f == 9) s->tipos_prontos = true;
  return 0;
}

int tq_set_tracker_first(tq_sim* s, int primeira) {
  TQ_NOT_TAPING(s, "tq_set_tracker_first");
  TQ_CHECK_SIM(s);
  const int v = primeira ? 1 : 0;
  TQ_HIP(hipMemcpyAsync(&s->sc->primeira, &v, sizeof v, hipMemcpyHostToDevice, s->st));
  return sync(s);
}

// Derived per-bubble counts the solver needs when a test starts from the solve phase: faces
// and cells per label come from the bubble-rhs pass, which a parity test may have skipped.
int tq_run_phase(tq_sim* s, int ph, double dt) {
  TQ_NOT_TAPING(s, "tq_run_phase");
  TQ_CHECK_SIM(s);
  int e = 0;
  switch (ph) {
    case 0: e = p2g(s, dt, 1); s->tipos_prontos = true; break;
    case 1: case 2: break;
    case 3: e = propagate_labels(s); break;
    case 4: e = surface(s); break;
    case 5: e = air_pressure_and_div(s, dt, 1); break;
    case 6: e = air_pressure_and_div(s, dt, 2); break;
    case 7: e = bubble_rhs(s, dt); break;
    case 8:
      if (s->d.ar) e = bubble_counts(s);
      if (!e) e = pressure_solve(s);
      if (!e) e = pressure_post(s, dt);
      break;
    case 9: e = gradient(s); break;
    case 10: e = extrapolate(s); break;
    case 11: e = g2p(s, dt); break;
    case 12: e = update_air(s); break;
    default: set_error("tq_run_phase: unknown phase"); return TQ_E_ARG;
  }
  if (e) return e;
  return sync(s);
}

int tq_debug_vcycle(tq_sim* s, const float* b, float* z) {
  TQ_CHECK_SIM(s);
  if (!b || !z) { set_error("tq_debug_vcycle: null argument"); return TQ_E_ARG; }
  if (s->d.ar)
    if (int e = bubble_counts(s)) return e;
  return debug_vcycle(s, b, z);
}

int tq_solver_stats_get(tq_sim* s, tq_solver_stats* st, int compute_true) {
  TQ_CHECK_SIM(s);
  if (!st) { set_error("tq_solver_stats_get: null"); return TQ_E_ARG; }
  if (int e = read_scalars(s)) return e;
  const DevScalars& h = *s->sc_host;
  st->iterations = h.iters;
  st->norm0 = h.norm0;
  st->norm = h.norm_last;
  st->norm_b = h.normb;
  st->levels = (int)s->lv.size();
  st->active_bubbles = h.active_bubbles;
  st->total_iterations = (long long)h.total_iters;
  st->solves = (long long)h.solves;
  st->operator_builds = (long long)h.op_builds;
  st->true_resid_iterate = h.true_done ? h.true_norm : -1.0;
  st->first_iterations = h.restarts > 0 ? h.iters_first : h.iters;
  st->restarts = h.restarts;
  st->air_components = (int)h.ncomp;
  st->overlap_pairs = (int)h.last_pairs;
  st->bubble_candidates = (int)h.last_cand;
  st->true_resid = -1;
  if (compute_true) {
    double tr = 0;
    if (int e = true_residual(s, &tr)) return e;
    st->true_resid = tr;
  }
  return 0;
}

int tq_timing_enable(tq_sim* s, int enable, const char* only) {
  TQ_CHECK_SIM(s);
  if (s->timing) timing_flush(s);
  s->timing = enable != 0;
  s->timing_only = only ? only : "";
  if (enable) s->ktimes.clear();
  return 0;
}

// launch caps by name: the field and the largest value a grid may take
static int* launch_field(tq_sim* s, const char* name) {
  if (!name) return nullptr;
  if (!std::strcmp(name, "spmv_grid")) return &s->launch_spmv_grid;
  if (!std::strcmp(name, "update_grid")) return &s->launch_update_grid;
  if (!std::strcmp(name, "split_cells")) return &s->launch_split_cells;
  if (!std::strcmp(name, "pass_fold_grid")) return &s->launch_pass_fold_grid;
  return nullptr;
}

int tq_set_launch(tq_sim* s, const char* name, int value) {
  TQ_CHECK_SIM(s);
  int* f = launch_field(s, name);
  if (!f) { set_error("tq_set_launch: unknown launch parameter"); return TQ_E_ARG; }
  const int top = f == &s->launch_split_cells ? (1 << 27) : MAX_RED_BLOCKS;
  if (value < 0 || value > top) { set_error("tq_set_launch: value out of range"); return TQ_E_ARG; }
  if (*f != value) {
    *f = value;
    solver_launch_changed(s);
  }
  return 0;
}

int tq_get_launch(tq_sim* s, const char* name, int* value) {
  TQ_CHECK_SIM(s);
  int* f = launch_field(s, name);
  if (!f || !value) { set_error("tq_get_launch: unknown launch parameter or null"); return TQ_E_ARG; }
  *value = *f;
  return 0;
}

int tq_timing_read(tq_sim* s, tq_kernel_time* out, int max, int* n) {
  TQ_CHECK_SIM(s);
  timing_flush(s);
  const int m = std::min((int)s->ktimes.size(), std::max(max, 0));
  for (int k = 0; k < m; k++) {
    std::memset(&out[k], 0, sizeof out[k]);
    std::snprintf(out[k].name, sizeof out[k].name, "%s", s->ktimes[k].name.c_str());
    out[k].total_ms = s->ktimes[k].total_ms;
    out[k].launches = s->ktimes[k].launches;
    out[k].bytes_per_launch = s->ktimes[k].launches ? s->ktimes[k].total_bytes / s->ktimes[k].launches : 0.0;
  }
  if (n) *n = (int)s->ktimes.size();
  return 0;
}

int tq_sync(tq_sim* s) {
  TQ_CHECK_SIM(s);
  return sync(s);
}

}  // extern "C"
