// tq_batch.hip -- independent tanks stepped together (the batched configuration), in two launch
// models.  The reference steps one SimulacaoCPU per tank (fisica-cpu.js:155-159 for each).
//
// Table mode (tq_batch_create_table), the default: one launch per kernel of the step for all
// tanks.  Every tank enqueues on the batch's one stream.  A step
//   1. records each tank's frame (step_record, on the host threads of the batch): the launches its
//      avancar would make -- body, grid, block, LDS bytes and the packed arguments -- in rounds:
//      round 0 the air update, round k substep k;
//   2. merges, round by round, the launch sequences of the tanks taking part (a tank with fewer
//      substeps sits out the later rounds) into one sequence of groups: launches of the same
//      body and block shape from different tanks are one group, and every tank's launches keep
//      their order (a longest-common-subsequence merge, added one tank at a time; the merge of a
//      round is cached by the tanks' sequences, which repeat from frame to frame);
//   3. uploads every tank's records of the frame in one copy and issues each group as one table
//      launch (k_tab, tq_launch.hpp) of up to TAB_MAX tanks, one after another on the stream (or
//      replays a HIP graph of them captured from an earlier frame with the same launch list,
//      tq_batch_use_graphs), then gathers every tank's vMax into one host buffer (tq_batch_vmax)
//      and waits.
// Each launch covers all tanks, so the device runs every phase across the batch's 2-3 M cells
// instead of tank by tank, and a frame costs one launch per phase rather than one per phase and
// tank.  The table runs each body with the tank's own logical block, grid and arguments, so every
// tank computes exactly what it computes stepped alone.  Table frames need the fixed PCG
// iteration count (tol = 0, the reference's mode): a residual-terminated solve waits on its
// tank's flags between chunks of iterations, which a recorded frame cannot.
//
// Thread mode (tq_batch_create): every tank keeps its own stream and a fixed pool of host threads
// enqueues the tanks' frames (tank k on thread k mod T), then waits for their streams; the
// device overlaps the tanks' kernels on its hardware queues.
//
// After an error in a step the tanks' states are undefined (a table frame may have been recorded
// for some tanks and not launched).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <map>
#include <mutex>
#include <string>
#include <thread>

#include "tq_internal.hpp"

namespace tq {
namespace {

// a group of the merged sequence: one body and block shape, and the launches (participant of the
// round, launch index from that participant's round start) that run as one table launch
struct Member {
  int part, op;
};
struct Group {
  const void* tab;
  dim3 block;
  std::vector<Member> m;
};
using Plan = std::vector<Group>;

// a table launch ready to issue
struct Issue {
  const void* tab;
  dim3 block;
  uint32_t shm, blocks;
  TabHdr hdr;
  size_t rec;  // offset of its records in the frame's upload
};

inline bool same_shape(const void* tab, const dim3& b, const LaunchOp& o) {
  return tab == o.tab && b.x == o.block.x && b.y == o.block.y && b.z == o.block.z;
}

inline uint64_t mix(uint64_t h, uint64_t v) {
  h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
  return h;
}

// One tank's launches of a round added to the merged sequence: a longest common subsequence of
// the groups' shapes and the tank's launches becomes shared groups, the rest stays in order.
Plan merge_in(Plan&& plan, int part, const LaunchOp* ops, int n) {
  const int na = (int)plan.size();
  const int w = n + 1;
  std::vector<uint32_t> L((size_t)(na + 1) * w, 0u);
  for (int i = na - 1; i >= 0; i--)
    for (int j = n - 1; j >= 0; j--)
      L[(size_t)i * w + j] = same_shape(plan[i].tab, plan[i].block, ops[j])
                                 ? L[(size_t)(i + 1) * w + j + 1] + 1
                                 : std::max(L[(size_t)(i + 1) * w + j], L[(size_t)i * w + j + 1]);
  Plan out;
  out.reserve((size_t)na + n);
  int i = 0, j = 0;
  while (i < na || j < n) {
    if (i < na && j < n && same_shape(plan[i].tab, plan[i].block, ops[j])) {
      plan[i].m.push_back(Member{part, j});
      out.push_back(std::move(plan[i]));
      i++;
      j++;
    } else if (j < n && (i == na || L[(size_t)i * w + j + 1] >= L[(size_t)(i + 1) * w + j])) {
      out.push_back(Group{ops[j].tab, ops[j].block, {Member{part, j}}});
      j++;
    } else {
      out.push_back(std::move(plan[i]));
      i++;
    }
  }
  return out;
}

}  // namespace
}  // namespace tq

using tq::Group;
using tq::Issue;
using tq::LaunchLog;
using tq::LaunchOp;
using tq::Plan;

struct tq_batch {
  std::vector<tq_sim*> tanks;
  bool table = false;
  int nthreads = 1;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_go, cv_done;
  unsigned gen = 0;
  int pending = 0;
  bool quit = false;
  double dt = 0;
  std::vector<int> subs;
  int err = 0;
  std::string err_msg;
  // table mode
  hipStream_t st = nullptr;                   // the stream every tank enqueues on
  std::vector<LaunchLog> logs;                // each tank's frame
  std::vector<LaunchLog> tails;               // each tank's vMax gather launch
  std::map<std::vector<uint64_t>, Plan> plans;  // merged rounds, by the tanks' launch sequences
  unsigned char* up_host = nullptr;           // pinned upload of a frame's records
  unsigned char* up_dev = nullptr;
  size_t up_cap = 0;
  unsigned long long* vmax_dev = nullptr;     // every tank's vMax bits after the last table step
  unsigned long long* vmax_host = nullptr;    // pinned
  bool vmax_valid = false;
  std::vector<uint64_t> vmax_epoch;           // each tank's substep count when vmax_host was filled
  tq_batch_stats stats{};                     // host time per part of the steps (tq_batch_stats_get)
  // a frame's table launches captured as a HIP graph and replayed while the launch list (bodies,
  // grids, block offsets, record offsets) repeats; the records themselves are uploaded each frame.
  // Off by default: over the 500-frame maze batch the substep counts vary from tank to tank and
  // frame to frame, so most frames are a new pattern (386 captures in 500 frames), and a replayed
  // frame runs no faster than the same launches issued one by one (the issue overlaps the device).
  bool use_graphs = false;
  struct FrameGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
    uint64_t used;
  };
  std::vector<FrameGraph> graphs;
  uint64_t graph_clock = 0;
};

namespace tq {
// gather of one tank's vMax bits (the word tq_vmax reads)
__device__ __forceinline__ void k_vmax_put(Blk B_, const DevScalars* __restrict__ sc,
                                           unsigned long long* __restrict__ dst) {
  if (threadIdx.x == 0) *dst = sc->vmax_bits;
}
}  // namespace tq

namespace {

// a worker's share of a step: enqueue its tanks' frames (thread mode) or record them (table mode)
int worker_job(tq_batch* b, int w, std::string* msg) {
  const size_t n = b->tanks.size();
  int e = 0;
  for (size_t k = (size_t)w; k < n && !e; k += (size_t)b->nthreads) {
    if (b->table) {
      b->logs[k].clear();
      e = tq::step_record(b->tanks[k], b->dt, b->subs[k], &b->logs[k]);
    } else {
      e = tq::step_enqueue(b->tanks[k], b->dt, b->subs[k]);
    }
    if (e) *msg = "tank " + std::to_string(k) + ": " + tq_last_error();
  }
  if (b->table) return e;
  for (size_t k = (size_t)w; k < n; k += (size_t)b->nthreads) {
    const int e2 = tq::stream_wait(b->tanks[k]);
    if (e2 && !e) {
      e = e2;
      *msg = "tank " + std::to_string(k) + ": " + tq_last_error();
    }
  }
  return e;
}

void worker_main(tq_batch* b, int w) {
  unsigned seen = 0;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(b->mu);
      b->cv_go.wait(lk, [&] { return b->quit || b->gen != seen; });
      if (b->quit) return;
      seen = b->gen;
    }
    std::string msg;
    const int e = worker_job(b, w, &msg);
    std::lock_guard<std::mutex> lk(b->mu);
    if (e && !b->err) {
      b->err = e;
      b->err_msg = msg;
    }
    if (--b->pending == 0) b->cv_done.notify_one();
  }
}

void stop_workers(tq_batch* b) {
  {
    std::lock_guard<std::mutex> lk(b->mu);
    b->quit = true;
  }
  b->cv_go.notify_all();
  for (auto& t : b->workers)
    if (t.joinable()) t.join();
  b->workers.clear();
}

// runs worker_job on every worker and waits
int run_workers(tq_batch* b) {
  std::unique_lock<std::mutex> lk(b->mu);
  b->err = 0;
  b->err_msg.clear();
  b->pending = b->nthreads;
  b->gen++;
  b->cv_go.notify_all();
  b->cv_done.wait(lk, [&] { return b->pending == 0; });
  if (b->err) {
    tq::set_error(b->err_msg);
    return b->err;
  }
  return 0;
}

// the launches [first[p], first[p] + len[p]) of participant p's log, merged (cached by sequence)
const Plan& round_plan(tq_batch* b, const std::vector<int>& tank, const std::vector<const LaunchOp*>& first,
                       const std::vector<int>& len) {
  std::vector<uint64_t> key;
  key.reserve(2 * tank.size());
  for (size_t p = 0; p < tank.size(); p++) {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < len[p]; i++) {
      const LaunchOp& o = first[p][i];
      h = tq::mix(h, (uint64_t)(uintptr_t)o.tab);
      h = tq::mix(h, ((uint64_t)o.block.x << 32) | ((uint64_t)o.block.y << 16) | o.block.z);
    }
    key.push_back((uint64_t)tank[p]);
    key.push_back(h ^ ((uint64_t)len[p] << 48));
  }
  auto it = b->plans.find(key);
  if (it != b->plans.end()) return it->second;
  if (b->plans.size() > 256) b->plans.clear();  // bounded: patterns of substep counts repeat
  Plan plan;
  for (size_t p = 0; p < tank.size(); p++) plan = tq::merge_in(std::move(plan), (int)p, first[p], len[p]);
  return b->plans.emplace(std::move(key), std::move(plan)).first->second;
}

// appends the table launches of one merged round to `out`, copying the records into `up`
int stage_round(tq_batch* b, const Plan& plan, const std::vector<int>& tank, const std::vector<const LaunchOp*>& first,
                const std::vector<LaunchLog>& logs, std::vector<Issue>& out, size_t* up) {
  for (const Group& g : plan) {
    for (size_t c0 = 0; c0 < g.m.size(); c0 += tq::TAB_MAX) {
      const size_t c1 = std::min(g.m.size(), c0 + (size_t)tq::TAB_MAX);
      Issue is{};
      is.tab = g.tab;
      is.block = g.block;
      uint32_t stride = 0;
      for (size_t q = c0; q < c1; q++) {
        const LaunchOp& o = first[g.m[q].part][g.m[q].op];
        stride = std::max(stride, (uint32_t)((tq::TAB_HDR + o.bytes + tq::TAB_ALIGN - 1) & ~(tq::TAB_ALIGN - 1)));
      }
      is.hdr.ntank = (int)(c1 - c0);
      is.hdr.stride = (int)stride;
      is.rec = *up;
      uint64_t blocks = 0;
      for (size_t q = c0; q < c1; q++) {
        const int p = g.m[q].part;
        const LaunchOp& o = first[p][g.m[q].op];
        if (o.grid.z != 1) {
          tq::set_error("table batch: a launch with a 3-D grid");
          return TQ_E_STATE;
        }
        is.hdr.start[q - c0] = (uint32_t)blocks;
        blocks += (uint64_t)o.grid.x * o.grid.y;
        is.shm = std::max(is.shm, o.shm);
        unsigned char* r = b->up_host + *up;
        const int gx = (int)o.grid.x, gy = (int)o.grid.y;
        std::memcpy(r, &gx, 4);
        std::memcpy(r + 4, &gy, 4);
        std::memcpy(r + tq::TAB_HDR, logs[tank[p]].arena.data() + o.off, o.bytes);
        *up += stride;
      }
      if (blocks > 0x7FFFFFFFull) {
        tq::set_error("table batch: a table launch of more than 2^31 blocks");
        return TQ_E_CAPACITY;
      }
      is.hdr.start[c1 - c0] = (uint32_t)blocks;
      is.blocks = (uint32_t)blocks;
      if (blocks > 0) out.push_back(is);
    }
  }
  return 0;
}

int issue(tq_batch* b, Issue& is) {
  const unsigned char* rec = b->up_dev + is.rec;
  void* args[] = {&is.hdr, &rec};
  TQ_HIP(hipLaunchKernel(is.tab, dim3(is.blocks), is.block, args, is.shm, b->st));
  return 0;
}

void drop_graphs(tq_batch* b) {
  for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

constexpr size_t GRAPH_CACHE = 8;  // frame graphs kept (substep-count patterns that recur)

// the frame's launches from a captured graph: replayed when the launch list matches one captured
// before, else captured now (the launches are recorded, not run, then the graph is launched)
int launch_graph(tq_batch* b, std::vector<Issue>& issues) {
  std::vector<uint64_t> key;
  key.reserve(issues.size() * 8 + 1);
  key.push_back((uint64_t)(uintptr_t)b->up_dev);
  for (const Issue& is : issues) {
    key.push_back((uint64_t)(uintptr_t)is.tab);
    key.push_back(((uint64_t)is.block.x << 42) | ((uint64_t)is.block.y << 21) | is.block.z);
    key.push_back(((uint64_t)is.shm << 32) | is.blocks);
    key.push_back(((uint64_t)(uint32_t)is.hdr.ntank << 32) | (uint32_t)is.hdr.stride);
    key.push_back(is.rec);
    for (int t = 0; t <= is.hdr.ntank; t++) key.push_back(is.hdr.start[t]);
  }
  b->graph_clock++;
  for (auto& g : b->graphs)
    if (g.key == key) {
      g.used = b->graph_clock;
      b->stats.graph_replays++;
      TQ_HIP(hipGraphLaunch(g.exec, b->st));
      return 0;
    }
  hipGraph_t graph = nullptr;
  TQ_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  int e = 0;
  for (Issue& is : issues)
    if ((e = issue(b, is)) != 0) break;
  const hipError_t ec = hipStreamEndCapture(b->st, &graph);
  if (e) {
    if (graph) (void)hipGraphDestroy(graph);
    return e;
  }
  if (ec != hipSuccess) return tq::hip_fail(ec, "hipStreamEndCapture", __FILE__, __LINE__);
  hipGraphExec_t exec = nullptr;
  const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return tq::hip_fail(ei, "hipGraphInstantiate", __FILE__, __LINE__);
  if (b->graphs.size() >= GRAPH_CACHE) {
    auto old = std::min_element(b->graphs.begin(), b->graphs.end(),
                                [](const tq_batch::FrameGraph& x, const tq_batch::FrameGraph& y) { return x.used < y.used; });
    (void)hipGraphExecDestroy(old->exec);
    b->graphs.erase(old);
  }
  b->graphs.push_back(tq_batch::FrameGraph{std::move(key), exec, b->graph_clock});
  b->stats.graph_builds++;
  TQ_HIP(hipGraphLaunch(exec, b->st));
  return 0;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int table_step(tq_batch* b) {
  const int n = (int)b->tanks.size();
  b->vmax_valid = false;
  double t0 = now_s();
  if (int e = run_workers(b)) return e;  // every tank's frame recorded
  double t1 = now_s();
  b->stats.record_s += t1 - t0;
  for (int k = 0; k < n; k++) {
    b->tails[k].clear();
    tq::Kernel<64, 0, tq::k_vmax_put>::go(&b->tails[k], nullptr, dim3(1), dim3(64), 0,
                                      (const tq::DevScalars*)b->tanks[k]->sc, b->vmax_dev + k);
  }
  // upload size: every launch's record
  size_t need = 0;
  for (int k = 0; k < n; k++)
    for (const auto* L : {&b->logs[k], &b->tails[k]})
      for (const LaunchOp& o : L->ops) need += (tq::TAB_HDR + o.bytes + tq::TAB_ALIGN - 1) & ~(tq::TAB_ALIGN - 1);
  if (need > b->up_cap) {
    TQ_HIP(hipStreamSynchronize(b->st));
    drop_graphs(b);  // they read the records at the old address
    if (b->up_host) (void)hipHostFree(b->up_host);
    if (b->up_dev) (void)hipFree(b->up_dev);
    b->up_host = nullptr;
    b->up_dev = nullptr;
    b->up_cap = 0;
    const size_t cap = need + need / 4 + 4096;
    TQ_HIP(hipHostMalloc((void**)&b->up_host, cap, hipHostMallocDefault));
    TQ_HIP(hipMalloc((void**)&b->up_dev, cap));
    b->up_cap = cap;
  }
  std::vector<Issue> issues;
  size_t up = 0;
  int rounds = 0;
  for (int k = 0; k < n; k++) rounds = std::max(rounds, (int)b->logs[k].rounds.size());
  std::vector<int> tank, len;
  std::vector<const LaunchOp*> first;
  for (int r = 0; r <= rounds; r++) {  // round `rounds`: the vMax gather
    tank.clear();
    len.clear();
    first.clear();
    for (int k = 0; k < n; k++) {
      const LaunchLog& L = r < rounds ? b->logs[k] : b->tails[k];
      if (r < rounds && r >= (int)L.rounds.size()) continue;
      const uint32_t o0 = r < rounds ? L.rounds[r] : 0;
      const uint32_t o1 = (r < rounds && r + 1 < (int)L.rounds.size()) ? L.rounds[r + 1] : (uint32_t)L.ops.size();
      if (o1 == o0) continue;
      tank.push_back(k);
      first.push_back(L.ops.data() + o0);
      len.push_back((int)(o1 - o0));
    }
    if (tank.empty()) continue;
    const double tp = now_s();
    const Plan& plan = round_plan(b, tank, first, len);
    const double ts = now_s();
    b->stats.plan_s += ts - tp;
    if (int e = stage_round(b, plan, tank, first, r < rounds ? b->logs : b->tails, issues, &up)) return e;
    b->stats.stage_s += now_s() - ts;
  }
  t1 = now_s();
  if (up > 0) TQ_HIP(hipMemcpyAsync(b->up_dev, b->up_host, up, hipMemcpyHostToDevice, b->st));
  if (b->use_graphs) {
    if (int e = launch_graph(b, issues)) return e;
  } else {
    for (Issue& is : issues)
      if (int e = issue(b, is)) return e;
  }
  TQ_HIP(hipMemcpyAsync(b->vmax_host, b->vmax_dev, sizeof(unsigned long long) * n, hipMemcpyDeviceToHost, b->st));
  const double t2 = now_s();
  b->stats.issue_s += t2 - t1;
  TQ_HIP(hipStreamSynchronize(b->st));
  b->stats.wait_s += now_s() - t2;
  b->stats.steps++;
  b->stats.launches = (int)issues.size();
  b->stats.upload_bytes = (double)up;
  b->vmax_valid = true;
  for (int k = 0; k < n; k++) b->vmax_epoch[k] = b->tanks[k]->substeps_run;
  return 0;
}

void release(tq_batch* b) {
  stop_workers(b);
  drop_graphs(b);
  for (tq_sim* s : b->tanks) tq_destroy(s);
  if (b->up_host) (void)hipHostFree(b->up_host);
  if (b->up_dev) (void)hipFree(b->up_dev);
  if (b->vmax_host) (void)hipHostFree(b->vmax_host);
  if (b->vmax_dev) (void)hipFree(b->vmax_dev);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

int create(const tq_desc* descs, int n, int threads, bool table, tq_batch** out) {
  if (!out || !descs || n <= 0) {
    tq::set_error("tq_batch_create: bad arguments");
    return TQ_E_ARG;
  }
  *out = nullptr;
  tq_batch* b = new tq_batch();
  b->table = table;
  auto fail = [&](int e, const std::string& msg) {
    release(b);
    tq::set_error(msg);
    return e;
  };
  if (table) {
    hipError_t he = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipMalloc((void**)&b->vmax_dev, sizeof(unsigned long long) * n);
    if (he == hipSuccess) he = hipHostMalloc((void**)&b->vmax_host, sizeof(unsigned long long) * n, hipHostMallocDefault);
    if (he != hipSuccess) {
      tq::hip_fail(he, "tq_batch_create", __FILE__, __LINE__);
      return fail(TQ_E_HIP, tq_last_error());
    }
  }
  for (int k = 0; k < n; k++) {
    tq_sim* s = nullptr;
    if (int e = tq_create(&descs[k], &s)) return fail(e, "tank " + std::to_string(k) + ": " + tq_last_error());
    b->tanks.push_back(s);
    if (table)
      if (int e = tq::share_stream(s, b->st)) return fail(e, "tank " + std::to_string(k) + ": " + tq_last_error());
  }
  b->nthreads = std::max(1, std::min(threads, std::min(n, 64)));
  b->subs.assign((size_t)n, 1);
  b->logs.resize((size_t)n);
  b->tails.resize((size_t)n);
  b->vmax_epoch.assign((size_t)n, 0);
  for (int w = 0; w < b->nthreads; w++) b->workers.emplace_back(worker_main, b, w);
  *out = b;
  return 0;
}

}  // namespace

extern "C" {

int tq_batch_create(const tq_desc* descs, int n, int threads, tq_batch** out) {
  return create(descs, n, threads, false, out);
}

int tq_batch_create_table(const tq_desc* descs, int n, int threads, tq_batch** out) {
  return create(descs, n, threads, true, out);
}

int tq_batch_size(tq_batch* b, int* n) {
  if (!b || !n) {
    tq::set_error("tq_batch_size: null argument");
    return TQ_E_ARG;
  }
  *n = (int)b->tanks.size();
  return 0;
}

tq_sim* tq_batch_tank(tq_batch* b, int k) {
  if (!b || k < 0 || k >= (int)b->tanks.size()) return nullptr;
  return b->tanks[(size_t)k];
}

int tq_batch_step(tq_batch* b, double dt, const int* substeps) {
  if (!b || !substeps || !(dt > 0)) {
    tq::set_error("tq_batch_step: bad arguments");
    return TQ_E_ARG;
  }
  for (size_t k = 0; k < b->tanks.size(); k++) {
    if (substeps[k] < 0) {
      tq::set_error("tq_batch_step: negative substep count");
      return TQ_E_ARG;
    }
    if (b->tanks[k]->tape) {
      // a recorded tank's frames go through tq_step, which logs them
      tq::set_error("tq_batch_step: a tank is recording frames (tq_tape_begin)");
      return TQ_E_STATE;
    }
  }
  b->dt = dt;
  b->subs.assign(substeps, substeps + b->tanks.size());
  if (b->table) return table_step(b);
  return run_workers(b);
}

int tq_batch_vmax(tq_batch* b, double* v, int* has) {
  if (!b) {
    tq::set_error("tq_batch_vmax: null handle");
    return TQ_E_ARG;
  }
  for (size_t k = 0; k < b->tanks.size(); k++) {
    if (b->vmax_valid && b->vmax_epoch[k] == b->tanks[k]->substeps_run) {
      if (v) std::memcpy(&v[k], &b->vmax_host[k], sizeof(double));
      if (has) has[k] = b->tanks[k]->has_vmax ? 1 : 0;
    } else if (int e = tq_vmax(b->tanks[k], v ? &v[k] : nullptr, has ? &has[k] : nullptr)) {
      return e;
    }
  }
  return 0;
}

int tq_batch_use_graphs(tq_batch* b, int on) {
  if (!b) {
    tq::set_error("tq_batch_use_graphs: null handle");
    return TQ_E_ARG;
  }
  b->use_graphs = on != 0;
  if (!b->use_graphs) drop_graphs(b);
  return 0;
}

int tq_batch_stats_get(tq_batch* b, tq_batch_stats* out, int reset) {
  if (!b || !out) {
    tq::set_error("tq_batch_stats_get: null argument");
    return TQ_E_ARG;
  }
  *out = b->stats;
  if (reset) b->stats = tq_batch_stats{};
  return 0;
}

int tq_batch_destroy(tq_batch* b) {
  if (!b) return 0;
  release(b);
  return 0;
}

}  // extern "C"
