/* tanque.h -- C ABI of the MI355X FLIP/APIC substep library (libtanque.so).
 *
 * This is the drop-in boundary for the reference's CPU engine: the step/state interface that
 * js/motor.js (MotorCPU, motor.js:79-120) drives on SimulacaoCPU (js/fisica-cpu.js:28-671).
 * Every entry point returns 0 on success and a non-zero TQ_E* code on failure; the message of
 * the last failure on the calling thread is available from tq_last_error().  No C++
 * exception crosses this boundary.  Handles are opaque; every host buffer is caller-owned
 * and is copied in or out (the library owns all device memory).  One handle must not be used
 * from two threads at once.  Units follow the reference: metres, seconds, SDF in cells.
 */
#ifndef TANQUE_H
#define TANQUE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TQ_OK 0
#define TQ_E_ARG 1      /* bad argument (null pointer, size out of range, unknown id) */
#define TQ_E_HIP 2      /* a HIP runtime call failed */
#define TQ_E_CAPACITY 3 /* particle capacity or a fixed table size would be exceeded */
#define TQ_E_STATE 4    /* call not valid in the current state */

#define TQ_CAMPOS 12     /* floats per particle: x y vx vy cux cuy cvx cvy espuma absR absG absB (fisica-cpu.js:26) */
#define TQ_MAX_BOLHAS 4096 /* bubble labels (constantes.js:20) */

/* Precision of the multigrid preconditioner inside the FP32 PCG. */
#define TQ_PREC_FP32 0   /* MG vectors stored and computed in FP32 */
#define TQ_PREC_MIXED 1  /* MG vectors stored in FP16 (per-cycle scaled), computed in FP32 */
#define TQ_PREC_MIXED_B 2 /* only the MG right-hand sides stored in FP16; x stays FP32 */
#define TQ_PREC_MIXED_X 3 /* only the MG iterates x stored in FP16; right-hand sides stay FP32 */

typedef struct tq_sim tq_sim;

/* Construction options: SimulacaoCPU(cena, {gravidade, iteracoes, tauDensidade, tensao, ar})
 * (fisica-cpu.js:29-77) plus the grid of the Cena (cena.js:15-24) and device sizing. */
typedef struct tq_desc {
  int nx, ny;             /* grid cells */
  double h;               /* cell size (m) */
  double gx, gy;          /* gravidade, default (0, -9.81) */
  int iteracoes;          /* PCG iterations (fixed count) or the iteration cap when tol > 0 */
  double tau_densidade;   /* TAU_DENSIDADE (s) */
  double tensao;          /* surface tension (N/m); 0 disables */
  int ar;                 /* 1: track enclosed air pockets as pressure unknowns */
  int max_particles;      /* device capacity; 0 = ceil(4.6 * nx * ny) as MotorCPU.maxParticulas */
  int precision;          /* TQ_PREC_FP32 or one of the mixed modes TQ_PREC_MIXED* */
  double tol;             /* 0: fixed `iteracoes` (reference); >0: stop at ||r|| <= tol*||b|| */
  int mg_min_lds_level;   /* multigrid levels >= this run in one workgroup from LDS; 0 = the
                             first level of at most 2048 cells whose suffix fits (set higher
                             to keep levels on the multi-workgroup kernels) */
  int mg_half_level;      /* mixed modes: first level whose multigrid vectors are stored in
                             FP16 (level 0 keeps x in FP32 and stores only b in FP16); levels
                             under 4096 cells and those in LDS stay FP32; -1 = default (1) */
  int use_graph;          /* 1: PCG iterations are replayed from a captured HIP graph */
  int mg_dense;           /* 1: when the first one-workgroup level has at most 2048 cells and is
                             not the coarsest, its V-cycle is built once per solve as a dense
                             matrix (one V-cycle per row) and applied by a many-workgroup
                             product; 0: that V-cycle runs in one workgroup every iteration;
                             -1 (default): the matrix when the solve runs to a tolerance or for
                             at least 8 fixed iterations (the reference's 3 apply the V-cycle
                             too few times for the build to pay) */
} tq_desc;

/* Fills `d` with the reference defaults (constantes.js) for an nx x ny grid of cell size h. */
void tq_desc_default(tq_desc* d, int nx, int ny, double h);

/* new SimulacaoCPU(cena, opcoes) (fisica-cpu.js:29-77).  The SDF starts as the empty tank
 * (the frame only, cena.js:139-144); upload walls with tq_set_sdf. */
int tq_create(const tq_desc* desc, tq_sim** out);
int tq_destroy(tq_sim* s);

/* Upload rows [row0, row1] of the (2nx) x (2ny) SDF (cena.js:20-24, rows as in
 * Cena.pegarLinhasAlteradas, cena.js:95-99) and recompute the solid cells
 * (atualizarParedes, fisica-cpu.js:80-85).  `sdf` points at the FULL host array. */
int tq_set_sdf(tq_sim* s, const float* sdf, int row0, int row1);

/* adicionar(lista, checarCheia) (fisica-cpu.js:92-111): `count` particles of TQ_CAMPOS floats.
 * Particles inside walls (distance <= 0.35 cells) and, when `check`, in cells already holding
 * >= 3 particles are skipped; the rest are appended in order.  `accepted` may be NULL. */
int tq_add_particles(tq_sim* s, const float* list, int count, int check, int* accepted);

/* aplicarFerramenta (fisica-cpu.js:129-152). kind: 0 none, 1 esponja, 2 mao, 3 corante.
 * Queued on the handle's stream (reads of the state afterwards see it). */
typedef struct tq_tool {
  int kind;
  double x, y, raio, vx, vy;
  double absorcao[3];
} tq_tool;
int tq_apply_tool(tq_sim* s, const tq_tool* t);

/* MotorGPU's non-blocking frame (motor.js:61-74, where quadro submits and contagem/vMax are read
 * back later): tq_step's work queued on the handle's stream without waiting, followed by a
 * status readback (tq_status_request).  With tol > 0 the pressure solves still poll the device
 * for their stop test; with the reference's fixed iteration count nothing waits.  Not allowed
 * while frames are recorded. */
int tq_step_async(tq_sim* s, double dt, int substeps);
/* Queues behind the handle's queued work a copy of the live particle count and vMax into pinned
 * host memory (MotorGPU.lerContagem / its vMax readback, js/gpu/fisica.js).  No wait. */
int tq_status_request(tq_sim* s);
/* The last requested status without waiting: *ready = 0 while the device has not reached it;
 * else the live count, vMax (*has_vmax = 0 before the first G2P) and the device time in ms of
 * the last frame tq_step_async queued (-1 if none; MotorGPU's custoPorSubpasso measures the same
 * with timestamp queries).  Any output pointer but `ready` may be NULL. */
int tq_status_poll(tq_sim* s, int* ready, int* n, double* vmax, int* has_vmax, double* gpu_ms);

/* avancar(dt, subpassos) (fisica-cpu.js:155-159): air labelling once, then `substeps`
 * substeps of dt/substeps.  Blocks until the device is done. */
int tq_step(tq_sim* s, double dt, int substeps);
/* subpasso(dt) (fisica-cpu.js:176-189) */
int tq_substep(tq_sim* s, double dt);
/* atualizarAr (fisica-cpu.js:162-167) */
int tq_update_air(tq_sim* s);
/* reiniciarAr (fisica-cpu.js:170-174) */
int tq_reset_air(tq_sim* s);
/* MotorCPU.limparAgua (motor.js:95): n = 0, cell counts cleared, reiniciarAr */
int tq_clear_water(tq_sim* s);
/* sim.n = m for m <= n (keeps the first m particles in reference order) */
int tq_truncate(tq_sim* s, int m);

int tq_count(tq_sim* s, int* n);
/* Copies min(n, max) particles in the reference's order (insertion order, removals keep
 * order; fisica-cpu.js:106, 114-126) into dst (TQ_CAMPOS floats each). */
int tq_read_particles(tq_sim* s, float* dst, int max, int* n);
/* vMax of the last G2P (fisica-cpu.js:637); *has = 0 before the first substep */
int tq_vmax(tq_sim* s, double* v, int* has);
/* area() (fisica-cpu.js:192-194) */
int tq_area(tq_sim* s, double* a);

/* Runtime options (gravity, tension, tau, air, PCG iterations, precision, tolerance). */
int tq_set_options(tq_sim* s, const tq_desc* d);

/* Field access (state the reference exposes as typed arrays; ids as in the oracle):
 *  0 u, 1 v, 2 uPeso, 3 vPeso (f32), 4 uValida, 5 vValida (u8), 6 densidade (f32),
 *  7 contagemCelula (u32), 8 corCelula (f32, 3 per cell), 9 tipo (u8), 10 solidoCelula (u8),
 *  11 div, 12 q, 13 fase, 14 normalX, 15 normalY, 16 curvatura, 17 qAr (f32),
 *  18 rotulo (u32), 19 pressaoBolha, 20 volumeBolha, 21 rhsBolha, 22 yBolha,
 *  23 volumeAlvo (f32 x 4096), 24 vivo (u8 x 4096).
 * `bytes` must equal the field size. Writing is a test hook for phase-by-phase parity. */
int tq_read_field(tq_sim* s, int field, void* dst, size_t bytes);
int tq_write_field(tq_sim* s, int field, const void* src, size_t bytes);
/* Replaces all particles (reference order = array order); test hook. */
int tq_write_particles(tq_sim* s, const float* src, int n);
/* Set / read the tracker's first-snapshot flag (RastreadorDeAr.primeira, ar.js:33). */
int tq_set_tracker_first(tq_sim* s, int primeira);

/* Runs one phase of the substep on the current state (test hook).  Phase ids:
 *  0 P2G, 1 gravity+boundary, 2 types, 3 label propagation, 4 surface, 5 air pressure,
 *  6 divergence, 7 bubble rhs, 8 pressure solve, 9 gradient, 10 extrapolation, 11 G2P,
 *  12 air tracker update.  On this device, phases 0-2 always run fused (ids 1 and 2 are
 *  accepted and do nothing after a phase 0 call). */
int tq_run_phase(tq_sim* s, int phase, double dt);

/* One application of the multigrid preconditioner (cicloV, pressao-cpu.js:241-255) to a
 * level-0 vector b of nx*ny floats on the current cell types, z = M b (test hook; no
 * counterpart in the reference's API).  Runs the pressure solve's setup first; b is meant for
 * water cells (bubble cells take the setup's zero bubble residual). */
int tq_debug_vcycle(tq_sim* s, const float* b, float* z);

/* The faucet tool (emitirTorneira, principal.js:301-321) with the seeded LCG of the reference's
 * tests (testes/apoio.js:4-9) in place of Math.random: velocity v = (pvx/2, pvy/2 - 0.7) from
 * the pointer's, clamped to 4 m/s; n = min(40000, round(2 raio max(0.7, |v|) dt / h^2 * 4))
 * particles uniform in the disc of radius raio around (x, y) (two draws each: angle, then
 * radius), with the dye absorption, added through adicionar's checks (fisica-cpu.js:92-111)
 * under the particle cap.  *seed advances by 2n draws; *emitted = n, *accepted = particles kept
 * (either may be null). */
typedef struct tq_faucet {
  double x, y;            /* pointer position (m) */
  double pvx, pvy;        /* pointer velocity (m/s) */
  double raio;            /* brush radius (m) */
  double absorcao[3];     /* dye absorption of the emitted water */
} tq_faucet;
int tq_emit_faucet(tq_sim* s, const tq_faucet* f, double dt, uint32_t* seed, int* emitted, int* accepted);

/* A batch of independent tanks stepped together (the batched configuration; the reference steps
 * one SimulacaoCPU per tank, fisica-cpu.js:29-77, 155-159, so there is no batch counterpart).
 * Every tank is a tq_sim set up through the tq_* calls on tq_batch_tank(b, k); tq_batch_step does
 * avancar(dt, substeps[k]) on every tank and returns when all are done, each tank ending exactly
 * as if stepped alone.  Two launch models:
 *   tq_batch_create_table: the tanks share one stream; a step records every tank's frame on
 *     `threads` host threads and runs each kernel of the step as ONE table launch over all tanks
 *     (blocks = the concatenation of the tanks' grids).  Needs tol = 0 (the fixed PCG iteration
 *     count) on every tank: a step on a tank with tol > 0 returns TQ_E_STATE.
 *   tq_batch_create: every tank on its own stream, its frame enqueued from `threads` host threads
 *     (tank k on thread k mod threads); the device overlaps the tanks' kernels.
 * Do not call tq_* on a tank while tq_batch_step runs; after a failed step the tanks' states are
 * undefined. */
typedef struct tq_batch tq_batch;
int tq_batch_create_table(const tq_desc* descs, int n, int threads, tq_batch** out);
int tq_batch_create(const tq_desc* descs, int n, int threads, tq_batch** out);
int tq_batch_size(tq_batch* b, int* n);
tq_sim* tq_batch_tank(tq_batch* b, int k);   /* NULL when k is out of range */
int tq_batch_step(tq_batch* b, double dt, const int* substeps);
/* every tank's vMax (tq_vmax; has[k] as its `has`), v and has of n entries (either may be null):
 * a table step leaves them in host memory, so after one they are read without a device wait */
int tq_batch_vmax(tq_batch* b, double* v, int* has);
/* table mode: on = 1 captures a frame's table launches as a HIP graph, replayed by later frames
 * with the same launch list (the same substep counts); 0 (the default) issues them one by one */
int tq_batch_use_graphs(tq_batch* b, int on);
/* where a table batch's steps spend host time (seconds, summed over steps since the last reset):
 * recording the tanks' frames, merging the rounds, copying the records into the upload buffer,
 * issuing the copies and table launches (or the frame graph), waiting for the device; the last
 * step's table launches and uploaded record bytes; frame graphs captured and replayed */
typedef struct tq_batch_stats {
  int steps;
  int launches;
  double upload_bytes;
  double record_s, plan_s, stage_s, issue_s, wait_s;
  int graph_builds, graph_replays;
} tq_batch_stats;
int tq_batch_stats_get(tq_batch* b, tq_batch_stats* out, int reset);
int tq_batch_destroy(tq_batch* b);

/* Differentiation of recorded frames (the reference has no gradient; it is the backward() that
 * a torch.autograd.Function would give, for a JavaScript host).  tq_tape_begin starts recording
 * the frames run from now on: tq_emit_faucet (principal.js:301-321) and tq_step
 * (fisica-cpu.js:155-159) calls; every other call that changes the state returns TQ_E_STATE
 * until tq_tape_end.  checkpoint_every > 0 keeps the tank's state every that many tq_step calls
 * and re-runs the frames in between during tq_backward; 0 keeps every substep's record instead.
 * tq_backward takes dL/d(final state) for the live particles in tq_read_particles order, 8
 * floats each (x y vx vy cux cuy cvx cvy), and returns dL/d(jet velocity) of each recorded
 * tq_emit_faucet call (2 doubles each, the velocity after principal.js:302-304's mapping),
 * dL/d(SDF) (4*nx*ny floats, may be NULL) and dL/d(state at tq_tape_begin) (n0 x 8 floats in
 * that moment's tq_read_particles order, may be NULL).  The derivative is that of the recorded
 * run with its discrete decisions held (cell types, bubble labels, micro-step counts, wall
 * push-outs and velocity projections, removals); the adjoint pressure solves run the same
 * MGPCG to relative residual `tol`.  The tank is left in its end state. */
int tq_tape_begin(tq_sim* s, int checkpoint_every);
int tq_tape_end(tq_sim* s);
int tq_tape_info(tq_sim* s, int* frames, int* faucet_calls, int* checkpoints, int* n_begin);
int tq_backward(tq_sim* s, const float* grad_state, int n, double tol, double* grad_faucet, int n_faucet,
                float* grad_sdf, float* grad_state0, int n0);
/* Water fraction of the cell box [i0,i1) x [j0,j1): the mean of min(1, densidade/PPC) over its
 * cells (fisica-cpu.js:305), densidade the P2G density splat (:241-250) of the current particles;
 * grad (may be NULL) gets d(fraction)/d(state), n x 8 floats in tq_read_particles order. */
int tq_water_fraction(tq_sim* s, int i0, int j0, int i1, int j1, double* frac, float* grad, int n);
/* tq_water_fraction's value (no gradient) without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture, tight optimisation
 * loops).  Its working memory (the box's FP64 density, the block partials, the result) lives in
 * a fixed 8.4 MB device array that each GPU the library runs on holds until the process ends, so
 * boxes are limited to 1048576 cells (TQ_E_CAPACITY above).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (two handles, two threads) race.  Blocks until the result
 * is on the host. */
int tq_water_fraction_noalloc(tq_sim* s, int i0, int j0, int i1, int j1, double* frac);
/* Test hooks: substep records of a keep-all tape (checkpoint_every = 0): their count, and field
 * `field` of record k (0 particle count u32, 1 types C u8, 2 labels C u32, 3 qAr C f32,
 * 4 volumeAlvo and 5 yBolha MAX_BOLHAS f32, 6 decision words n u16 and 7 insertion numbers
 * n u32 in P2G order, 8 dt f64). */
int tq_tape_records(tq_sim* s, int* n);
int tq_tape_record(tq_sim* s, int k, int field, void* dst, size_t bytes);

/* Statistics of the last pressure solve. */
typedef struct tq_solver_stats {
  int iterations;          /* PCG iterations run */
  double norm0;            /* ||[r;rb]|| at the warm start */
  double norm;             /* recursive residual norm after the last iteration */
  double norm_b;           /* ||[-div; rhs]|| */
  double true_resid;       /* ||b - A[q;y]|| recomputed (only when requested) */
  int levels;              /* multigrid levels */
  int active_bubbles;
  long long total_iterations; /* PCG iterations since creation */
  long long solves;           /* pressure solves since creation */
  long long operator_builds;  /* solves that rebuilt the dense coarse operators (their masks changed) */
  /* Residual-terminated mode (tol > 0): the iterate is kept as a compensated pair of binary32
   * words (q + q_lo) and its true residual is recomputed in binary64 when the recursive residual
   * reaches tol; while it is above tol the solve restarts from it (residual replacement, at most
   * 3 times).  true_resid above is that of the binary32 q the solve hands on (what the gradient
   * reads: the storage floor of binary32 pressures). */
  double true_resid_iterate;  /* binary64 true residual of the compensated iterate at the stop; -1 if not computed */
  int first_iterations;       /* iterations of the first segment (before any restart) */
  int restarts;               /* residual replacements of the last solve */
  /* the air tracker's last update (tq_update_air): air components, (component, old label)
   * overlap pairs and bubble candidates (components of at least 4 cells of volume) */
  int air_components, overlap_pairs, bubble_candidates;
} tq_solver_stats;
int tq_solver_stats_get(tq_sim* s, tq_solver_stats* st, int compute_true_residual);

/* Per-kernel timing with HIP events on the library's stream.  enable=1 starts collecting;
 * tq_timing_read copies (name, total ms, launches) records. */
typedef struct tq_kernel_time {
  char name[48];
  double total_ms;
  long launches;
  double bytes_per_launch; /* algorithmic bytes per launch (0 if not modelled) */
} tq_kernel_time;
/* `only`: NULL times every kernel; otherwise only launches whose name equals it. */
int tq_timing_enable(tq_sim* s, int enable, const char* only);
int tq_timing_read(tq_sim* s, tq_kernel_time* out, int max, int* n);

/* Launch shapes of the solver, by name: "spmv_grid" (most blocks of the spmv kernel, default
 * 1024), "update_grid" (most blocks of the PCG update / V-cycle input kernel, default 1280),
 * "split_cells" (multigrid levels above this many cells run each pass as its own kernel instead
 * of the fused down/up kernels, default 2^20) and "pass_fold_grid" (most blocks of a split level
 * 0's folding pass-B kernel, default 2048); 0 restores the default.  Grids larger than a cap
 * walk their cells with fewer blocks.  The cap decides how the dot products' partial sums are grouped, so results
 * are bit-reproducible for a fixed setting and can differ in the last bits between settings.
 * No counterpart in the reference (launch tuning). */
int tq_set_launch(tq_sim* s, const char* name, int value);
int tq_get_launch(tq_sim* s, const char* name, int* value);

/* The solver's autotuner (no counterpart in the reference, whose WebGPU engine fixes its dispatch
 * shapes, js/gpu/multigrid.js:283-287).  Times pressure solves of `iterations` PCG iterations on
 * the tank's current pressure system (call it after a few frames) for each candidate of, in turn,
 * the dense/one-workgroup coarse V-cycle (mg_dense), the levels run one pass per kernel
 * (tq_set_launch "split_cells": levels above that many cells; default 2^20), "spmv_grid",
 * "update_grid" and "pass_fold_grid" (most blocks of a split level 0's folding pass-B kernel,
 * default 2048), keeping a change only when it is faster by more than 1 %; the state a solve
 * changes is restored after every trial.  The choice is applied to `s` and cached for the process
 * by (nx, ny, precision): tq_create applies a cached choice.  `out` (may be NULL) receives it with
 * the mean solve time of the defaults and of the choice. */
typedef struct tq_tune_result {
  int spmv_grid, update_grid, split_cells, pass_fold_grid, mg_dense;  /* 0: the built-in default */
  double default_us;  /* one tuning solve with the defaults (device time, us) */
  double tuned_us;    /* with the chosen shapes */
  int trials;         /* settings timed */
} tq_tune_result;
int tq_autotune(tq_sim* s, int iterations, tq_tune_result* out);

/* The render hand-off: the water layer of the reference's 2D renderer (js/render-2d.js:110-162)
 * computed on the device from the particle buffer, at 3 pixels per cell (w = 3 nx, h = 3 ny,
 * rows from the top).  tint_rgba: the Beer-Lambert water tint multiplied over the panel
 * (255,255,255,255 where the water covers less than the threshold); foam_rgba: (248, 251, 252,
 * alpha) foam over it (all zero where uncovered; the reference leaves those RGB bytes as they
 * were).  Sorts the particles by cell for their current positions (particle order is not part of
 * the state: tq_read_particles returns insertion order).  Not allowed while frames are recorded. */
int tq_render_water(tq_sim* s, int w, int h, uint8_t* tint_rgba, uint8_t* foam_rgba);

/* Waits for all queued device work of this handle. */
int tq_sync(tq_sim* s);

const char* tq_last_error(void);

/* ---------------------------------------------------------------- host helpers
 * Scene construction (js/cena.js) and seeding (fisica-cpu.js:705-727, testes/apoio.js:4-9),
 * on caller-owned host arrays of (2nx)*(2ny) floats. */
/* The tank's walls built on the device from capsule strokes (n x 6 doubles: x0 y0 x1 y1 raio
 * apagar, metres, applied in order): the frame, every stroke as Cena.tracar applies it
 * (cena.js:43-72, binary32 store after each), then the solid cells -- Cena.refazer (cena.js:87-91)
 * followed by atualizarParedes (fisica-cpu.js:80-85), bit-identical to tq_scene_stroke. */
int tq_set_strokes(tq_sim* s, const double* strokes, int n);
/* Chain rule of the same construction: out[5k..5k+4] = sum over lattice points of gsdf times
 * d(sdf)/d(x0, y0, x1, y1, raio) of stroke k, the point's value being that of the last stroke that
 * changed it (points set by the frame or clamped to +-10 cells have none).  For the adjoint's wall
 * parameters: tq_backward's SDF gradient -> stroke gradients. */
int tq_strokes_vjp(tq_sim* s, const double* strokes, int n, const float* gsdf, double* out);
/* The tank's SDF lattice as the device holds it (4*nx*ny floats; Cena.sdf, cena.js:24). */
int tq_read_sdf(tq_sim* s, float* dst, size_t bytes);
void tq_scene_frame(float* sdf, int nx, int ny);                      /* Cena #zerar (cena.js:139-144) */
void tq_scene_stroke(float* sdf, int nx, int ny, double h, double x0, double y0, double x1, double y1,
                     double raio, int apagar);                        /* Cena.tracar (cena.js:43-72) */
/* semearRetangulo with the seeded LCG; returns the particle count; writes when out != NULL
 * and max_out suffices (returns -1 otherwise). `seed` is the LCG state, updated in place. */
int tq_seed_rect(double x0, double y0, double x1, double y1, double h, const double* absorcao,
                 uint32_t* seed, float* out, int max_out);
/* passoMaximo (constantes.js:167-175); has_vmax = 0 means vMax == null */
double tq_passo_maximo(double h, double altura, double vmax, int has_vmax, double dt_quadro);

#ifdef __cplusplus
}
#endif
#endif
