// tq_internal.hpp -- device state of one tank and the launchers shared by the HIP sources.
//
// Layout in HBM (one tank, C = nx*ny cells):
//   particles  SoA, two buffers (sorted by cell each substep): pos float2, vel float2,
//              cm float4 (cux cuy cvx cvy), ext float4 (espuma absR absG absB), seq u32
//              (insertion number: the reference's particle order is ascending seq).
//   MAC grid   u (nx+1)*ny at (i h, (j+1/2) h), v nx*(ny+1) at ((i+1/2) h, j h),
//              cell arrays i + j*nx, y up -- the reference's indexing (fisica-cpu.js:38-70).
//   binning    keys are tile-major (16x16 cell tiles) so one tile's particles are contiguous.
//   solver     per MG level: u8 mask (neighbour bits | type), x/b/tmp in f32 or f16.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tanque.h"
#include "tq_launch.hpp"

namespace tq {

// ------------------------------------------------------------------ constants (js/constantes.js)
constexpr double GRAVIDADE = 9.81;
constexpr double DENSIDADE_AGUA = 1000;
constexpr double TENSAO_SUPERFICIAL = 0.0728;
constexpr int MAX_BOLHAS = TQ_MAX_BOLHAS;
constexpr double VOLTA_VOLUME_BOLHA = 0.05;
constexpr double PRESSAO_BOLHA_MAX = 20000;
constexpr int PPC = 4;
constexpr double SDF_MAX = 10;
constexpr int MOLDURA = 2;
constexpr double RAIO_PARTICULA = 0.3;
constexpr double MICROPASSO = 0.8;
constexpr int MAX_MICROPASSOS = 4;
constexpr double TAU_DENSIDADE = 0.1;
constexpr double EXCESSO_MIN = -0.5, EXCESSO_MAX = 1.0;
constexpr int ITERACOES_PCG = 3;
constexpr double ESPUMA_V0 = 1.1, ESPUMA_V1 = 2.8, ESPUMA_TAXA = 5, ESPUMA_VIDA = 1.2;
constexpr double DIFUSAO_CORANTE = 3;
constexpr double PARTICULAS_MAX_POR_CELULA = 4.6;
constexpr uint8_t SOLIDO = 0, AGUA = 1, AR = 2;
constexpr int MIN_BOLHA_CELULAS = 4;

constexpr int TILE = 16;                 // P2G binning tile and the MG smoother block (pressao-cpu.js:17)
constexpr uint32_t DEAD_KEY = 0xFFFFFFFFu;
// Fixed-point scale of the P2G sums: 2^46.  A particle far from a node can give it a weight of
// ~1e-11 (binary32 positions quantise the fractional offset to ~1e-5 cells), which must stay
// non-zero and keep its relative precision; |sum w v| stays below 2^63/2^46 ~ 1.3e5 m/s for
// the speeds G2P allows (3.2 h/dt per particle, tens of particles per face).
constexpr double FX_SCALE = 70368744177664.0;  // 2^46
constexpr double FX_INV = 1.0 / 70368744177664.0;
constexpr int RED_BLOCK = 256;           // block size of reduction kernels
constexpr int MAX_RED_BLOCKS = 65536;    // partial slots per reduction (one per block)
constexpr int RED_SLOTS = 10;            // reductions with their own partial arrays (solver layout)
constexpr int BV_COUNT = 18;             // solver bubble vectors of MAX_BOLHAS doubles each

// mask byte of a multigrid cell / operator cell
//   bits 0-3: free neighbours (1 left, 2 right, 4 down, 8 up) as prepararVizinhos
//   bits 4-5: cell type (0 solid, 1 water, 2 air)
//   multigrid masks, bit 6: red-black colour (i + j + 1) & 1
//   operator mask, bit 6: tracked bubble cell (air, label >= 2 with water faces);
//                  bit 7: that bubble cell is interior (not on the grid border)
constexpr uint8_t MK_COLOUR = 0x40, MK_BUB = 0x40, MK_BUBIN = 0x80;
__host__ __device__ inline int mask_viz(uint8_t m) { return m & 15; }
__host__ __device__ inline int mask_type(uint8_t m) { return (m >> 4) & 3; }
__host__ __device__ inline int mask_cnt(uint8_t m) { return __builtin_popcount(m & 15); }

// Device-side scalars of one tank (one allocation).
struct DevScalars {
  uint32_t n_slots;     // particle slots in the current buffer (live + dead-marked)
  uint32_t n_live;      // live particles after the last binning
  unsigned long long vmax_bits;  // max |v| of the last G2P (binary64 bits, non-negative)
  uint32_t add_count;   // particles accepted by the last add
  uint32_t cand_count;  // label-propagation candidates
  uint32_t ncomp;       // air components of the last tracker update
  uint32_t npairs;      // (component, old label) pairs of the tracker update in progress
  uint32_t last_pairs;  // pairs of the last tracker update
  uint32_t last_cand;   // bubble candidates (components of >= 4 cells of volume) of the last update
  uint32_t ticket[16];  // last-block counters of the reductions
  int done;             // PCG converged / stopped
  int iters;            // PCG iterations run
  unsigned long long total_iters;  // PCG iterations since creation
  unsigned long long solves;       // pressure solves since creation
  int primeira;         // tracker first snapshot flag
  int active_bubbles;
  double rz, pap, alpha, beta, norm2, norm0, normb, norm_last;
  double tol;
  double proj_mean;     // null-space component removed from b (residual-terminated mode)
  int maxit;
  float mg_scale;       // mixed precision: scale applied to the V-cycle input of this iteration
  float mg_scale_next;  // mg_mult/||r|| of the newest residual, the scale of the next iteration
  float mg_mult;        // numerator of the scale: FP16 values of the input sit this far above 1/||r||
  double rzp[2];        // r.z of the last two iterations, by iteration parity
  int itp[2];           // iterations completed, by the parity of the iteration that completed them
  float mgs[2];         // mixed precision: V-cycle input scale of the iterations of each parity
  int op_dirty;         // this solve's one-workgroup level masks differ from the last operator build
  unsigned long long op_builds;  // operator builds (coarsest and level-lc matrices) since creation
  // residual-terminated mode: the iterate is the pair q + qlo (binary32 each); at the stop its true
  // residual ||b - A[q + qlo; y]|| is recomputed in binary64, and a restart from that residual
  // follows while it is above tol
  double true_norm;     // true residual of the iterate at the last stop (-1: not computed yet)
  int true_done;        // true_norm belongs to the current segment's stop
  int iters_base;       // iterations of the segments before the current one (restarts)
  int iters_first;      // iterations of the first segment
  int restarts;         // restarts of the last solve
};

struct Level {
  int nx = 0, ny = 0, n = 0;
  uint8_t* mask = nullptr;
  void* b = nullptr;
  void* x = nullptr;
  void* tmp = nullptr;
  void* res = nullptr;          // where this level's V-cycle result is (x, or tmp after a two-kernel up path)
  bool hx = false, hb = false;  // FP16 storage of x/tmp and of b on this level (this solve)
};

struct Tape;     // recorded frames of one tank (tq_adjoint.hip)
struct SubRec;   // one substep's record for the reverse pass

struct KTime {
  std::string name;
  double total_ms = 0;
  long launches = 0;
  double total_bytes = 0;  // algorithmic bytes of the timed launches
};

}  // namespace tq

struct tq_sim {
  tq_desc d;
  int nx, ny, C, nu, nv;
  int ntx, nty, K;        // binning tiles and key space (ntx*nty*256)
  double h;
  hipStream_t st = nullptr;

  // particles
  int cap = 0;
  int n_upper = 0;        // host-side upper bound of the slot count
  int cur = 0;
  float2* pos[2] = {};
  float2* vel[2] = {};
  float4* cm[2] = {};
  float4* ext[2] = {};
  uint32_t* seq[2] = {};
  uint32_t* pkey = nullptr;
  uint32_t* prank = nullptr;
  uint32_t* addflag = nullptr;  // add/compaction scratch (cap+1)
  uint32_t seq_next = 0;
  float* stage = nullptr;       // staging for uploads of particle lists (cap*12)
  uint8_t* render_img = nullptr;  // tq_render_water's two RGBA images, allocated on its first call
  size_t render_bytes = 0;

  // scene / grid
  float* sdf = nullptr;
  uint8_t* solid = nullptr;
  uint32_t* countKey = nullptr;  // K
  uint32_t* startKey = nullptr;  // K+1
  uint32_t* countCell = nullptr; // C
  float *u = nullptr, *v = nullptr, *uw = nullptr, *vw = nullptr;
  float *uTmp = nullptr, *vTmp = nullptr;
  uint8_t *uVal = nullptr, *vVal = nullptr, *uValTmp = nullptr, *vValTmp = nullptr;
  float *dens = nullptr, *dye = nullptr;
  uint8_t* tipo = nullptr;
  float *fase = nullptr, *nX = nullptr, *nY = nullptr, *kappa = nullptr;
  float *qAr = nullptr, *div = nullptr, *q = nullptr;

  // air
  uint32_t* rot = nullptr;
  uint8_t* cflag = nullptr;
  uint32_t* cand = nullptr;
  int32_t* pai = nullptr;
  int32_t* comp = nullptr;
  uint32_t* rootFlag = nullptr;  // C+1 (scan input / component numbering)
  uint32_t* compIdx = nullptr;   // C+1
  uint32_t* compCells = nullptr; // per component: cells, atm, topo, (C each)
  uint32_t* compAtm = nullptr;
  uint32_t* compTopo = nullptr;
  float* compAlvo = nullptr;     // novoAlvo per component
  int32_t* compDono = nullptr;
  uint32_t* compLabel = nullptr;
  unsigned long long* hkeys = nullptr;  // (component, old label) hash table
  uint32_t* hcount = nullptr;
  uint32_t* hfirst = nullptr;
  uint32_t* hlist = nullptr;
  int hsize = 0;
  uint32_t* ordBuf = nullptr;   // bookkeeping scratch (sort buffers), 4*C
  // bubbles (MAX_BOLHAS each)
  float* pBolha = nullptr;     // pressaoBolha (Pa)
  uint32_t* volB = nullptr;    // volumeBolha (cell counts)
  uint32_t* facesB = nullptr;  // water faces per bubble
  uint32_t* cellsB = nullptr;  // cells per bubble (celulasDe)
  double* rhsB = nullptr;      // bubble rhs (FP64 sums)
  float* rhsB32 = nullptr;     // rhsBolha as the reference stores it
  float* yB32 = nullptr;       // yBolha (f32 as stored by the reference)
  float* volAlvo = nullptr;    // volumeAlvo
  uint8_t* vivo = nullptr;
  uint8_t* vivoNovo = nullptr;
  double* bvec = nullptr;      // bubble CG vectors (BV_COUNT x MAX_BOLHAS; layout in k_solver.hip)
  int* activeList = nullptr;   // active bubble labels (count in sc->active_bubbles)

  // solver
  std::vector<tq::Level> lv;
  uint8_t* opMask = nullptr;
  float *r = nullptr, *ap = nullptr;
  float* qlo = nullptr;          // low word of the compensated iterate (residual-terminated mode)
  float* zbuf[2] = {};
  float* pbuf[2] = {};
  int zc = 0, pc = 0;
  double* partials = nullptr;   // per-block partial sums: RED_SLOTS x MAX_RED_BLOCKS
  double* bpart = nullptr;      // per-block sums of the slotted bubbles: 2 x BSLOTS x MAX_RED_BLOCKS
  double* partials_r = nullptr; // the same layouts, pre-reduced to a few partials for large grids
  double* bpart_r = nullptr;
  int it_k = -1;                // index of the PCG iteration being launched (-1: the initial V-cycle)
  int nb_spmv = 0, nb_upd = 0, nb_z = 0;  // partial counts of the last spmv, update and z launches
  int launch_spmv_grid = 0;     // most blocks of spmv (0: the built-in cap)
  int launch_update_grid = 0;   // most blocks of the PCG update / level-0 input kernel (0: the built-in cap)
  int launch_split_cells = 0;   // levels above this many cells run each multigrid pass as its own kernel (0: 2^20)
  int launch_pass_fold_grid = 0; // most blocks of a split level 0's folding pass-B kernel (0: the built-in cap)
  int* labelSlot = nullptr;     // bubble label -> reduction slot (-1: 128-bit fixed-point atomics)
  int lc = 0;                   // first multigrid level solved inside the one-workgroup kernel
  int mg_scaled = 0;            // this solve stores some multigrid level in FP16 (input scaled)
  int qcomp = 0;                // this solve keeps the compensated iterate q + qlo (tolerance mode)
  float* dense = nullptr;       // coarsest-level operator (when that level is small)
  float* vop = nullptr;         // level lc's V-cycle as a dense matrix (rows of vop_ld floats)
  int vop_ld = 0;               // 0: level lc has no dense V-cycle operator
  uint8_t* op_masks = nullptr;  // masks of levels lc..L-1 at the last operator build
  int op_cells = 0;             // their total cell count
  bool vop_was_active = false;  // the last solve applied the level-lc matrix
  void* graph_exec = nullptr;   // hipGraphExec_t of two PCG iterations
  int graph_key = -1;
  bool capturing = false;       // a stream capture is recording (no timing events)
  tq::LaunchLog* lrec = nullptr; // launches are appended here instead of launched (tank-table batches)
  bool own_st = true;           // st was created by this tank (a table batch shares its own stream)
  tq::DevScalars* sc = nullptr;
  tq::DevScalars* sc_host = nullptr;  // pinned mirror

  // scan scratch
  uint32_t* scanBlocks = nullptr;
  int scan_cap = 0;             // entries of scanBlocks

  // differentiation (tq_adjoint.hip)
  std::vector<std::pair<void*, size_t>> bufs;  // alloc_all's buffers (pointer, bytes): what a checkpoint copies
  tq::Tape* tape = nullptr;       // frames being recorded for tq_backward
  tq::SubRec* rec = nullptr;      // the substep record the substep's hooks fill (null: none)
  uint16_t* g2p_flags = nullptr;  // where G2P writes each particle's decision word (null: nowhere)

  // non-blocking status readback (tq_step_async / tq_status_request / tq_status_poll)
  uint64_t* status_dev = nullptr;   // [0] live particles (low 32 bits), [1] vMax (binary64 bits)
  uint64_t* status_host = nullptr;  // pinned copy of status_dev
  hipEvent_t ev_status = nullptr;   // recorded after the copy
  hipEvent_t ev_frame0 = nullptr;   // recorded before the last frame tq_step_async queued
  hipEvent_t ev_frame1 = nullptr;   // recorded after it
  bool status_pending = false;      // a request was queued (ev_status belongs to it)
  bool status_timed = false;        // a tq_step_async frame was queued before the last request
  bool status_has = false;          // vMax had been measured when the request was queued

  // host-side state
  bool tipos_prontos = false;
  bool has_vmax = false;
  uint64_t substeps_run = 0;    // substeps enqueued or recorded so far
  bool air_reset_pending = true;

  // timing
  bool timing = false;
  std::string timing_only;      // time only launches of this name (empty: all)
  std::vector<tq::KTime> ktimes;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
  std::vector<std::pair<int, int>> ev_pending;  // (ktime index, event pair index)
  std::vector<double> ev_pending_bytes;         // algorithmic bytes of each pending launch
  int ev_used = 0;
};

namespace tq {

// error helpers (tq_api.cpp)
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
#define TQ_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) return tq::hip_fail(e_, #call, __FILE__, __LINE__); \
  } while (0)
#define TQ_LAUNCH_CHECK()                                                   \
  do {                                                                      \
    hipError_t e_ = hipGetLastError();                                      \
    if (e_ != hipSuccess) return tq::hip_fail(e_, "kernel launch", __FILE__, __LINE__); \
  } while (0)

// whether the solve builds level lc's V-cycle as a dense matrix (mg_dense; -1: when the solve runs
// to a tolerance or for at least DENSE_MIN_ITERS fixed iterations, since the build runs one V-cycle
// per row of the level and pays only over many applications)
constexpr int DENSE_MIN_ITERS = 8;
inline bool dense_wanted(const tq_sim* s) {
  return s->d.mg_dense > 0 || (s->d.mg_dense < 0 && (s->d.tol > 0 || s->d.iteracoes >= DENSE_MIN_ITERS));
}

// a kernel body launched on `st`, or appended to the tank's launch log while one is recorded
template <int LB, int W, auto Body, class... A>
inline void launch(tq_sim* s, hipStream_t st, dim3 g, dim3 b, size_t shm, A&&... a) {
  Kernel<LB, W, Body>::go(s->lrec, st, g, b, shm, std::forward<A>(a)...);
}
// zero `bytes` (a multiple of 4) at p: hipMemsetAsync, or a recorded fill launch
int zero_async(tq_sim* s, hipStream_t st, void* p, size_t bytes);

// timing brackets around launches (tq_api.cpp)
int time_begin(tq_sim* s, const char* name, double bytes);
void time_end(tq_sim* s, int slot);

// ---- launchers (k_particles.hip)
int scan_u32(tq_sim* s, const uint32_t* in, uint32_t* out, int n);  // exclusive, out[n] = total
int bin_and_sort(tq_sim* s);                  // keys, counting sort into the other buffer
int p2g(tq_sim* s, double dt, int gravity);   // fused P2G + gravity/boundary + types
int g2p(tq_sim* s, double dt);
int add_particles_dev(tq_sim* s, int count, int check);
int step_enqueue(tq_sim* s, double dt, int substeps);  // tq_step without the final wait
int step_record(tq_sim* s, double dt, int substeps, LaunchLog* log);  // the same frame's launches, logged
int stream_wait(tq_sim* s);
int faucet_dev(tq_sim* s, const tq_faucet* f, double vx, double vy, int n, int keep, uint32_t* seed);  // from s->stage
int apply_tool_dev(tq_sim* s, const tq_tool* t);
int count_live(tq_sim* s, int* n);
int status_enqueue(tq_sim* s);  // live count and vMax into s->status_dev (queued, no wait)
int solid_cells(tq_sim* s);
// ---- launchers (k_grid.hip)
int surface(tq_sim* s);
int air_pressure_and_div(tq_sim* s, double dt, int which);  // which: 1 qAr, 2 div, 3 both
int propagate_labels(tq_sim* s);
int bubble_rhs(tq_sim* s, double dt);
int bubble_counts(tq_sim* s);
int gradient(tq_sim* s);
int extrapolate(tq_sim* s);
int pressure_post(tq_sim* s, double dt);
// ---- launchers (k_air.hip)
int tracker_update(tq_sim* s);
int tracker_reset(tq_sim* s);
// ---- launchers (k_solver.hip)
int solver_alloc(tq_sim* s);
void solver_free(tq_sim* s);
int pressure_solve(tq_sim* s);
void solver_launch_changed(tq_sim* s);
void tune_apply_cached(tq_sim* s);  // tq_tune.hip: a tq_autotune result cached for this grid
int share_stream(tq_sim* s, hipStream_t st);  // enqueue on a stream the tank does not own
int true_residual(tq_sim* s, double* out);
int debug_vcycle(tq_sim* s, const float* b, float* z);
// ---- tq_api.hip
int download_ordered(tq_sim* s, std::vector<float>& out, std::vector<uint32_t>* seqs);
int emit_faucet_impl(tq_sim* s, const tq_faucet* f, double dt, uint32_t* seed, int* emitted, int* accepted);
// ---- tq_adjoint.hip: substep hooks while a record is open, the operation log, the tape's end
constexpr int REC_P2G = 0, REC_SOLVE = 1, REC_EXTRAP = 2, REC_G2P = 3;
int rec_hook(tq_sim* s, int point, double dt);
int tape_log_emit(tq_sim* s, const tq_faucet* f, double dt, uint32_t seed, uint32_t seq0, uint32_t seq1);
int tape_log_step(tq_sim* s, double dt, int substeps);
void tape_free(tq_sim* s);

}  // namespace tq
