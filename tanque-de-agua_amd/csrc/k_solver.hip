// k_solver.hip -- pressure projection: PCG preconditioned by a multigrid V-cycle, with one
// extra unknown per enclosed air pocket (resolverPCG / cicloV, js/pressao-cpu.js:241-402).
//
// CG vectors r, z, p, Ap are FP32 per cell; every dot product, the residual norm and the
// bubble unknowns (y, r, z, p, Ap per label) are FP64.  Reductions are deterministic and free
// of same-address atomics: each block stores one partial per quantity (and one per slotted
// bubble); the next kernel's blocks each re-sum them in block order -- so every block holds
// bit-identical scalars -- and redo the per-bubble scalar work, block 0 alone storing the
// bubble vectors.  Bubbles beyond BSLOTS go to 128-bit fixed-point atomics.  What one kernel reads and
// rewrites alternates by iteration parity (bubble z, p, r, the atomic accumulators, rz).
//
// Multigrid.  The reference smoother (suavizar, pressao-cpu.js:99-150) runs 16x16 tiles that
// each read a snapshot of x (their halo stays at the snapshot) through 4 red-black sweeps.  One
// wavefront owns one tile: each lane holds four cells in registers (one red and one black in
// each of two rows), so every lane updates one cell per row in every half-sweep; neighbours come
// from the wave's private 18x18 LDS tile, synchronised by an LDS wait (no barrier).  A dependent
// kernel costs ~4-5 us here while one pass over a level costs well under 1 us of arithmetic, so
// consecutive passes are fused and the overlapping tiles recomputed: one kernel per level down
// (pass B + restriction + the coarse level's pass A) and one per level up (prolongation +
// pass C + pass D).  Level 0's pass A also applies the PCG update and builds the V-cycle input;
// its pass D writes z and the r.z dots.  The small levels run in one workgroup from LDS, or as
// a dense matrix built once per solve; the coarsest as a dense operator.  Buffer roles per
// level: pass A writes x, pass B x -> tmp (x_B), the up kernel tmp -> x (the level's result).
//
// Kernel sequence of PCG iteration k (parity par = k mod 2):
//   spmv      beta from the z dots, bubble z; p = z + beta p, Ap, p.Ap
//   in_a0     alpha from p.Ap, bubble y, r; q, r update, |r|^2, V-cycle input, pass A on level 0
//   down_0    norm and stop test from |r|^2; pass B on level 0 ...
//   down_l for 0 < l < lc, the levels >= lc, up_l for l = lc-1..0 (level 0 writes z, r.z)
// Every kernel returns at once when the solve is done.
#include "tq_device.hpp"

namespace tq {

// bubble vectors (MAX_BOLHAS doubles each): y, ap, and by iteration parity r, z, p and the
// accumulators of spmv (A) and of the z dots (Z); I: the accumulator outside the iterations.
// Each accumulator holds the low 64 bits of a 128-bit fixed-point sum; its high 64 bits sit
// BV_FX_HI vectors further on (see fx_add).
enum {
  BV_Y = 0, BV_AP, BV_R0, BV_R1, BV_Z0, BV_Z1, BV_P0, BV_P1,
  BV_ACCA0, BV_ACCA1, BV_ACCZ0, BV_ACCZ1, BV_ACCI,
  BV_ACCA0_HI, BV_ACCA1_HI, BV_ACCZ0_HI, BV_ACCZ1_HI, BV_ACCI_HI, BV_N
};
constexpr int BV_FX_HI = BV_ACCA0_HI - BV_ACCA0;
static_assert(BV_N == BV_COUNT, "bubble vector count");
static_assert(BV_ACCI_HI - BV_ACCI == BV_FX_HI, "accumulator high words");
__device__ __host__ inline double* bv(double* base, int k) { return base + (size_t)k * MAX_BOLHAS; }

// Sums of bubbles without a reduction slot, in 128-bit fixed point with 2^-64 resolution: an
// addend v = hi + lo 2^-64 (hi = floor(v), lo the fraction's top 64 bits) is added with two
// 64-bit integer atomics and the carry out of the low word, so the total does not depend on the
// order of the adds (FP64 atomics would).  |sums| < 2^63; addends below 2^-64 are dropped.
__device__ inline void fx_split(double v, unsigned long long* hi, unsigned long long* lo) {
  double fl = floor(v), fr = v - fl;
  if (fr >= 1.0) {  // v in (-2^-54, 0): 1 + v rounds to 1; the addend is below the resolution
    fl += 1.0;
    fr = 0.0;
  }
  *hi = (unsigned long long)(long long)fl;
  *lo = (unsigned long long)(fr * 18446744073709551616.0);
}
// adds the 128-bit value (hi, lo) to the pair of words (H, L) (global or LDS)
__device__ inline void fx_add_words(unsigned long long* L, unsigned long long* H, unsigned long long hi,
                                    unsigned long long lo) {
  const unsigned long long old = atomicAdd(L, lo);
  atomicAdd(H, hi + (old + lo < old ? 1ull : 0ull));
}
__device__ inline void fx_add(double* acc, uint32_t lab, double v) {
  unsigned long long hi, lo;
  fx_split(v, &hi, &lo);
  fx_add_words(reinterpret_cast<unsigned long long*>(acc) + lab,
               reinterpret_cast<unsigned long long*>(acc + (size_t)BV_FX_HI * MAX_BOLHAS) + lab, hi, lo);
}

// Per-block sums of the bubbles without a reduction slot, on their way to the global fixed-point
// accumulators: an LDS table of BAGG labels, each summing its addends in the same 128-bit integer
// arithmetic, flushed with one global add per label at the end of the block.  A cell of a bubble
// (or a face next to one) then costs LDS atomics instead of two contended global ones.  Integer
// sums are associative, so the global totals are bit-identical to adding every addend there; a
// label that finds no free entry within BAGG_PROBE probes is added there directly.
constexpr int BAGG = 64, BAGG_PROBE = 8;
struct BubLds {
  uint32_t key[BAGG];  // label (bubble labels are >= 2), 0 = free
  unsigned long long lo[BAGG], hi[BAGG];
};
// a kernel's handle: its block's table, or none (every addend goes to `acc` directly: a solve
// with at most BSLOTS bubbles has a slot for each, so nothing reaches the table)
struct BubAgg {
  BubLds* t;
  double* acc;  // the global accumulator of the labels without a slot
};
// every thread of the block, with `on` uniform; when on, a block barrier must follow before the
// first addend
__device__ inline BubAgg bagg_begin(BubLds& tab, double* acc, bool on) {
  if (on)
    for (int e = threadIdx.x; e < BAGG; e += blockDim.x) {
      tab.key[e] = 0u;
      tab.lo[e] = 0ull;
      tab.hi[e] = 0ull;
    }
  return BubAgg{on ? &tab : nullptr, acc};
}
// not inlined: it is the cold branch of the cells' loops (bubbles beyond the slots), and inlining
// its probe loop at every call site of an unrolled stencil slowed spmv by ~2.7 us per launch
// even where it never runs.  The handle's two pointers go by value: a handle passed by reference
// needs a home in private memory, which every wave of the kernel then writes at its start.
__device__ __attribute__((noinline)) void bagg_add_at(BubLds* t, double* acc, uint32_t lab, double v) {
  unsigned long long hi, lo;
  fx_split(v, &hi, &lo);
  if (t) {
    uint32_t e = (lab * 2654435761u) >> 26;  // 6 bits: BAGG entries
    for (int p = 0; p < BAGG_PROBE && lab != 0u; p++, e = (e + 1) & (BAGG - 1)) {
      const uint32_t k = atomicCAS(&t->key[e], 0u, lab);
      if (k == 0u || k == lab) {
        fx_add_words(&t->lo[e], &t->hi[e], hi, lo);
        return;
      }
    }
  }
  fx_add_words(reinterpret_cast<unsigned long long*>(acc) + lab,
               reinterpret_cast<unsigned long long*>(acc + (size_t)BV_FX_HI * MAX_BOLHAS) + lab, hi, lo);
}
__device__ inline void bagg_add(const BubAgg& a, uint32_t lab, double v) { bagg_add_at(a.t, a.acc, lab, v); }
// every thread of the block, after its last addend
__device__ inline void bagg_flush(const BubAgg& a) {
  if (!a.t) return;
  __syncthreads();
  for (int e = threadIdx.x; e < BAGG; e += blockDim.x) {
    const uint32_t lab = a.t->key[e];
    if (lab)
      fx_add_words(reinterpret_cast<unsigned long long*>(a.acc) + lab,
                   reinterpret_cast<unsigned long long*>(a.acc + (size_t)BV_FX_HI * MAX_BOLHAS) + lab, a.t->hi[e],
                   a.t->lo[e]);
  }
}

__device__ inline double fx_read(const double* acc, uint32_t lab) {
  const unsigned long long lo = reinterpret_cast<const unsigned long long*>(acc)[lab];
  const long long hi = reinterpret_cast<const long long*>(acc + (size_t)BV_FX_HI * MAX_BOLHAS)[lab];
  return (double)hi + (double)lo * 5.421010862427522e-20;  // 2^-64
}
__device__ inline void fx_clear(double* acc, uint32_t lab) {
  reinterpret_cast<unsigned long long*>(acc)[lab] = 0;
  reinterpret_cast<unsigned long long*>(acc + (size_t)BV_FX_HI * MAX_BOLHAS)[lab] = 0;
}

// partial-sum arrays (slots of MAX_RED_BLOCKS doubles): 0..4 the solve's set-up and the true
// residual, then the products of one iteration, each read by the following kernel
constexpr int PS_SPMV = 5;   // p.Ap
constexpr int PS_RR = 6;     // |r|^2 after the update
constexpr int PS_Z = 7;      // r.z, r.z_old (two slots)
constexpr int PS_LAST = RED_SLOTS - 1;  // last element: the true residual's result
__host__ __device__ inline size_t ps_off(int slot) { return (size_t)slot * MAX_RED_BLOCKS; }

constexpr int BSLOTS = 8;          // bubbles reduced through per-block slots
constexpr int TS = 18;             // smoother tile with its 1-cell halo
constexpr int TCELLS = TS * TS;
// LDS row pitch of a smoother tile: with 24 floats the lanes' reads of one colour and its
// neighbours (two columns per lane, rows lane / 8 and lane / 8 + 8) fall in distinct banks; the
// natural pitch 18 gives two-way conflicts on every read of a half-sweep
constexpr int TP = 24;
constexpr int TLDS = TS * TP;      // floats of a tile in LDS
constexpr int WG_WAVES = 16;       // one-workgroup kernel: 1024 threads
constexpr int WG_POOL = 136192;    // its LDS pool for the level arrays (x, tmp, b, mask): 160 KiB - 16 tiles
constexpr int WG_AUTO_CELLS = 2048;  // default: levels from the first one this small run in one workgroup
constexpr int DENSE_MAX = 192;       // coarsest levels up to this many cells use the dense operator
constexpr int VOP_MAX = 2048;        // largest level whose V-cycle may be applied as a dense matrix
constexpr int FIN_THREADS = 512;

template <class S> __device__ inline float ld(const S* p, size_t i);
template <> __device__ inline float ld<float>(const float* p, size_t i) { return p[i]; }
template <> __device__ inline float ld<__half>(const __half* p, size_t i) { return __half2float(p[i]); }
template <class S> __device__ inline void st(S* p, size_t i, float v);
template <> __device__ inline void st<float>(float* p, size_t i, float v) { p[i] = v; }
// FP16 stores saturate at the largest finite value: an extreme entry is clipped instead of
// becoming inf (the preconditioner is perturbed, the PCG vectors stay finite)
__device__ inline __half sat_half(float v) { return __float2half(fminf(fmaxf(v, -65504.f), 65504.f)); }
template <> __device__ inline void st<__half>(__half* p, size_t i, float v) { p[i] = sat_half(v); }
// the value a store in S keeps
template <class S> __device__ inline float rnd(float v);
template <> __device__ inline float rnd<float>(float v) { return v; }
template <> __device__ inline float rnd<__half>(float v) { return __half2float(sat_half(v)); }

// Guarded loads: the element at c where `in` holds, else z.  The load itself is unconditional
// (index 0 when `in` fails: every array here has an element 0), so the loads of a tile are
// issued back to back and their latencies overlap.  Written as `in ? p[c] : z`, each load sits
// in a branch of its own and the wave waits for it before issuing the next one.
template <class T>
__device__ __forceinline__ T ld_if(const T* __restrict__ p, size_t c, bool in, T z) {
  const T v = p[in ? c : 0];
  return in ? v : z;
}
template <class S>
__device__ __forceinline__ float ldf_if(const S* __restrict__ p, size_t c, bool in) {
  const float v = ld<S>(p, in ? c : 0);
  return in ? v : 0.f;
}

// Row pairs.  In the smoother tiles a lane's two cells of a row are the adjacent columns 2p,
// 2p + 1 from an even tile origin, so with an even row length the pair starts at an even index
// and moves as one 8-byte (FP16: 4-byte) access: half the memory instructions, and every store
// writes whole 8-byte words instead of every other 4-byte word of a row.
struct F2 {
  float x, y;
};
template <class S> __device__ inline F2 ld2(const S* p, size_t c);
template <> __device__ inline F2 ld2<float>(const float* p, size_t c) {
  const float2 v = *reinterpret_cast<const float2*>(p + c);
  return {v.x, v.y};
}
template <> __device__ inline F2 ld2<__half>(const __half* p, size_t c) {
  const __half2 v = *reinterpret_cast<const __half2*>(p + c);
  return {__low2float(v), __high2float(v)};
}
template <class S> __device__ inline void st2(S* p, size_t c, float a, float b);
template <> __device__ inline void st2<float>(float* p, size_t c, float a, float b) {
  *reinterpret_cast<float2*>(p + c) = make_float2(a, b);
}
template <> __device__ inline void st2<__half>(__half* p, size_t c, float a, float b) {
  *reinterpret_cast<__half2*>(p + c) = __halves2half2(sat_half(a), sat_half(b));
}
__device__ inline uint16_t ld2b(const uint8_t* p, size_t c) { return *reinterpret_cast<const uint16_t*>(p + c); }
// a row pair: the index of its even column (0 outside the grid) and whether it is in the grid
// (both cells are, or neither: even row length, even origin)
struct Pair {
  size_t c;
  bool in;
};


// orders this wavefront's LDS accesses across lanes: a compiler barrier only.  The LDS executes a
// wavefront's DS instructions in issue order, so a read issued after another lane's write sees it
// without waiting for the write to complete; what must not happen is the compiler moving an access
// across (to it, other lanes' cells at other offsets look independent).
__device__ inline void wave_sync() { asm volatile("" ::: "memory"); }

__device__ inline int nslots(int na) { return na < BSLOTS ? na : BSLOTS; }

__device__ inline void slot_add(double (&acc)[BSLOTS], int slot, double val) {
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) acc[k] += (k == slot) ? val : 0.0;
}

// a bubble contribution: through the block's slot accumulators, or the block's LDS table of the
// fixed-point sums of labels without a slot
__device__ inline void bub_add(double (&sl)[BSLOTS], const int* __restrict__ labelSlot, const BubAgg& agg, uint32_t lab,
                               double val) {
  const int s = labelSlot[lab];
  if (s >= 0) slot_add(sl, s, val);
  else bagg_add(agg, lab, val);
}
// the same with the label's slot already read
__device__ inline void bub_add_s(double (&sl)[BSLOTS], int s, const BubAgg& agg, uint32_t lab, double val) {
  if (s >= 0) slot_add(sl, s, val);
  else bagg_add(agg, lab, val);
}


// Block sum of the first nv of v[]; thread 0 stores partial `bid` of each quantity.  nv must
// be uniform over the block; `sh` holds NMAX*16 doubles.
// Block totals of the first nv of v[] in every thread, with one barrier: each wave's sums (DPP)
// go to sh[k * 16 + wave], then every thread adds the waves' sums of each quantity in wave order
// (the order one thread's loop over them would take, so the totals are those of the two-barrier
// form that has thread k sum quantity k and share it).  v[k] = 0 for k >= nv.  The other threads
// may still be reading sh when a thread returns: the caller puts a barrier before writing it
// again, or uses another region.
template <int NV>
__device__ inline void block_total(double (&v)[NV], int nv, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++)
    if (k < nv) {
      const double t = wave_sum(v[k]);
      if (lane == 0) sh[k * 16 + w] = t;
    }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) {
    double t = 0;
    if (k < nv)
      for (int q = 0; q < nw; q++) t += sh[k * 16 + q];
    v[k] = t;
  }
}
// LDS of a kernel's folds: the partials' totals ((2 + BSLOTS) quantities), then two regions for
// the bubbles' sums, so that consecutive reductions need no barrier between them
constexpr int FOLD_SH = (2 + BSLOTS) * 16 + 2 * 16 + 16;
__device__ inline double* fold_sh2(double* sh) { return sh + (2 + BSLOTS) * 16; }
__device__ inline double* fold_sh3(double* sh) { return sh + (2 + BSLOTS) * 16 + 32; }

// the value of slot s (s < BSLOTS) of a register array (no indexed private memory)
__device__ inline double slot_val(const double (&v)[BSLOTS], int s) {
  double r = 0;
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) r = k == s ? v[k] : r;
  return r;
}

// The block partials of a kernel's last reduction: the first na of a[] to part, the first nb of
// b[] to bpart (rows of MAX_RED_BLOCKS), with one barrier; sh ((NA + NB) * 16 doubles) is left
// being read, so nothing may write it afterwards.
template <int NA, int NB>
__device__ inline void store_partials_last(const Blk& B_, double (&a)[NA], int na, double* __restrict__ part,
                                           double (&b)[NB], int nb, double* __restrict__ bpart, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NA; k++)
    if (k < na) {
      const double t = wave_sum(a[k]);
      if (lane == 0) sh[k * 16 + w] = t;
    }
#pragma unroll
  for (int k = 0; k < NB; k++)
    if (k < nb) {
      const double t = wave_sum(b[k]);
      if (lane == 0) sh[(NA + k) * 16 + w] = t;
    }
  __syncthreads();
  const int t = threadIdx.x;
  const int bid = B_.x + B_.y * B_.gx;
  if (t < na || (t >= NA && t < NA + nb)) {
    double s = 0;
    for (int q = 0; q < nw; q++) s += sh[t * 16 + q];
    if (t < NA) part[(size_t)t * MAX_RED_BLOCKS + bid] = s;
    else bpart[(size_t)(t - NA) * MAX_RED_BLOCKS + bid] = s;
  }
}

template <int NA>
__device__ inline void store_partials_last(const Blk& B_, double (&a)[NA], int na, double* __restrict__ part, double* sh) {
  double none[1] = {0};
  store_partials_last<NA, 1>(B_, a, na, part, none, 0, nullptr, sh);
}

template <int NMAX>
__device__ inline void store_partials(const Blk& B_, double (&v)[NMAX], int nv, double* __restrict__ part, double* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int k = 0; k < NMAX; k++)
    if (k < nv) {
      const double t = wave_sum(v[k]);
      if (lane == 0) sh[k * 16 + w] = t;
    }
  __syncthreads();
  if ((int)threadIdx.x < nv) {
    double t = 0;
    for (int q = 0; q < nw; q++) t += sh[threadIdx.x * 16 + q];
    part[(size_t)threadIdx.x * MAX_RED_BLOCKS + B_.x + B_.y * B_.gx] = t;
  }
  __syncthreads();  // `sh` is free again for the caller's next reduction
}

// Totals of nb partials of the first nv quantities, in a fixed order (deterministic), in two
// stages so that a kernel can issue other loads between them: part_load adds this thread's
// partials into acc[], part_reduce sums acc[] over the block (valid in every thread).
template <int NMAX>
__device__ inline void part_load(const double* __restrict__ part, int nb, int nv, double (&acc)[NMAX]) {
#pragma unroll
  for (int k = 0; k < NMAX; k++) acc[k] = 0;
  // partials b, b + blockDim.x, ... in that order; two of them per step, every quantity's load
  // issued before any is added (the rows of quantities k >= nv are read and not added, so the
  // addresses do not wait for nv)
  const int bd = blockDim.x;
  for (int b = threadIdx.x; b - (int)threadIdx.x < nb; b += 2 * bd) {
    const bool in0 = b < nb, in1 = b + bd < nb;
    double x0[NMAX], x1[NMAX];
#pragma unroll
    for (int k = 0; k < NMAX; k++) {
      const size_t row = (size_t)(k < nv ? k : 0) * MAX_RED_BLOCKS;
      x0[k] = part[row + (in0 ? b : 0)];
      x1[k] = part[row + (in1 ? b + bd : 0)];
    }
#pragma unroll
    for (int k = 0; k < NMAX; k++) {
      if (k < nv && in0) acc[k] += x0[k];
      if (k < nv && in1) acc[k] += x1[k];
    }
  }
}
template <int NMAX>
__device__ inline void part_reduce(const double (&acc)[NMAX], int nv, double (&tot)[NMAX], double* sh, double* res) {
  (void)res;
#pragma unroll
  for (int k = 0; k < NMAX; k++) tot[k] = acc[k];
  block_total<NMAX>(tot, nv, sh);
  __syncthreads();  // `sh` is free again
}
template <int NMAX>
__device__ inline void sum_partials(const double* __restrict__ part, int nb, int nv, double (&tot)[NMAX], double* sh,
                                    double* res) {
  double acc[NMAX];
  part_load<NMAX>(part, nb, nv, acc);
  part_reduce<NMAX>(acc, nv, tot, sh, res);
}

// block sum valid in every thread (`res` holds NV doubles of LDS)
template <int NV>
__device__ inline void block_sum_all(double (&v)[NV], double* sh, double* res) {
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) res[k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = res[k];
  __syncthreads();
}

__device__ inline bool first_block(const Blk& B_) { return B_.x == 0 && B_.y == 0; }

// ================================================================== reductions folded into consumers
struct BubArgs {
  double* bvec;
  const int* activeList;
  const int* labelSlot;
  const uint32_t* facesB;
};

// The z dots of iteration k's input (pressao-cpu.js:386-394), at the start of spmv: r.z and
// r.z_old from the partials, the bubbles' new z (slot totals in zs[], else the atomic
// accumulator) and Polak-Ribiere beta = max(0, (r.z - r.z_old) / rz_prev), 0 on the first
// iteration.  Block 0 stores z of this parity, clears the accumulator the next z dots use
// (the other parity) and records rz.
// In three stages, so that the kernel's own loads overlap the fold's chain of dependent loads:
// zf_load issues the partials' loads and this thread's first active bubble (k = threadIdx.x),
// zf_load_bub that bubble's slot and vectors, zf_finish reduces (sums in the order of a
// sequential loop over the partials and the active bubbles).
struct ZBub {  // one active bubble's values read by the z-dot fold
  int b, sl;
  double rc, zo, az, po;
};
// bubble b = activeList[k] (slot k when k < BSLOTS)
__device__ inline ZBub zbub_load(const BubArgs& B, int par, int b, int k) {
  ZBub v;
  v.b = b;
  v.sl = k < BSLOTS ? k : -1;
  v.rc = bv(B.bvec, BV_R0 + 1 - par)[b];
  v.zo = bv(B.bvec, BV_Z0 + 1 - par)[b];
  v.az = fx_read(bv(B.bvec, BV_ACCZ0 + par), b);
  v.po = bv(B.bvec, BV_P0 + 1 - par)[b];
  return v;
}
struct ZFold {
  double a[2], s[BSLOTS];
  double tot[2];
  int na, b0;
  ZBub v0;
};
// (activeList holds MAX_BOLHAS > blockDim.x entries: this thread's entry is read before na is
// known and kept only when threadIdx.x < na)
__device__ inline void zf_load(ZFold& f, const double* __restrict__ part, const double* __restrict__ bpartZ, int nb,
                               const BubArgs& B, int na) {
  f.na = na;
  const int b0 = B.activeList[threadIdx.x];
  part_load<2>(part, nb, 2, f.a);
  part_load<BSLOTS>(bpartZ, nb, nslots(na), f.s);
  f.b0 = (int)threadIdx.x < na ? b0 : 0;
}
__device__ inline void zf_load_bub(ZFold& f, const BubArgs& B, int par) { f.v0 = zbub_load(B, par, f.b0, threadIdx.x); }
// the partials' totals in every thread, in one block reduction (sh: FOLD_SH doubles): r.z,
// r.z_old in f.tot, the slotted bubbles' z sums in zs[]
__device__ inline void zf_reduce(ZFold& f, double* sh, double (&zs)[BSLOTS]) {
  double q[2 + BSLOTS];
  q[0] = f.a[0];
  q[1] = f.a[1];
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) q[2 + k] = f.s[k];
  block_total<2 + BSLOTS>(q, 2 + nslots(f.na), sh);
  f.tot[0] = q[0];
  f.tot[1] = q[1];
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) zs[k] = q[2 + k];
}
// beta; ps[] gets the new p of every slotted bubble (its z plus beta times its old p), valid
// after the block's next barrier
__device__ inline double zf_finish(const ZFold& f, const Blk& B_, int par, int first, const BubArgs& B, double* sh,
                                   const double (&zs)[BSLOTS], float* ps, DevScalars* __restrict__ sc) {
  const int na = f.na;
  const double* tot = f.tot;
  double* Zn = bv(B.bvec, BV_Z0 + par);
  double* accZnext = bv(B.bvec, BV_ACCZ0 + 1 - par);
  const bool b0 = first_block(B_);
  double bacc[2] = {0, 0};
  auto add = [&](const ZBub& v) {
    const double zn = v.sl >= 0 ? slot_val(zs, v.sl) : v.az;
    bacc[0] += v.rc * zn;
    bacc[1] += v.rc * v.zo;
    if (b0) {
      Zn[v.b] = zn;
      fx_clear(accZnext, v.b);
    }
  };
  if ((int)threadIdx.x < na) add(f.v0);
  for (int k = threadIdx.x + blockDim.x; k < na; k += blockDim.x) add(zbub_load(B, par, B.activeList[k], k));
  block_total<2>(bacc, 2, fold_sh2(sh));
  const double rzNovo = tot[0] + bacc[0], cruzado = tot[1] + bacc[1];
  const double rzPrev = sc->rzp[1 - par];
  const double beta = first ? 0.0 : (fabs(rzPrev) < 1e-30 ? 0.0 : fmax(0.0, (rzNovo - cruzado) / rzPrev));
  if (b0 && threadIdx.x == 0) {
    sc->rzp[par] = rzNovo;
    sc->rz = rzNovo;
    sc->beta = beta;
  }
  auto pslot = [&](const ZBub& v) {
    if (v.sl >= 0) ps[v.sl] = (float)(slot_val(zs, v.sl) + beta * v.po);
  };
  if ((int)threadIdx.x < na) pslot(f.v0);
  for (int k = threadIdx.x + blockDim.x; k < na; k += blockDim.x) pslot(zbub_load(B, par, B.activeList[k], k));
  return beta;
}

// alpha = rz / p.Ap (pressao-cpu.js:370-380) at the start of in_a0, with the bubbles' p and Ap
// (slot totals, else the atomic accumulator) and their new residuals (rbs[] for the slotted
// ones).  Block 0 stores p, Ap, y += alpha p, r -= alpha Ap of the bubbles, the bubble part of
// |r|^2 and clears the accumulator of the next spmv (the other parity).  Staged like the z-dot
// fold: af_load, af_load_bub, af_finish.
struct ABub {  // one active bubble's values read by the alpha fold
  int b, sl;
  uint32_t fb, cb;
  double zn, po, az, ro, yb;
};
// bubble b = activeList[k] (slot k when k < BSLOTS); cb its cell count
__device__ inline ABub abub_load(const BubArgs& B, const uint32_t* __restrict__ cellsB, int par, int b, int k) {
  ABub v;
  v.b = b;
  v.sl = k < BSLOTS ? k : -1;
  v.cb = cellsB[b];
  v.fb = B.facesB[b];
  v.zn = bv(B.bvec, BV_Z0 + par)[b];
  v.po = bv(B.bvec, BV_P0 + 1 - par)[b];
  v.az = fx_read(bv(B.bvec, BV_ACCA0 + par), b);
  v.ro = bv(B.bvec, BV_R0 + 1 - par)[b];
  v.yb = bv(B.bvec, BV_Y)[b];
  return v;
}
struct AFold {
  double a[1], s[BSLOTS];
  double tot[1];
  int na, b0;
  double beta, rz;
  ABub v0;
};
__device__ inline void af_load(AFold& f, const double* __restrict__ part, const double* __restrict__ bpartA, int nb,
                               int par, const BubArgs& B, const DevScalars* __restrict__ sc) {
  f.na = sc->active_bubbles;
  f.beta = sc->beta;
  f.rz = sc->rzp[par];
  const int b0 = B.activeList[threadIdx.x];  // kept only when threadIdx.x < na (see zf_load)
  part_load<1>(part, nb, 1, f.a);
  part_load<BSLOTS>(bpartA, nb, nslots(f.na), f.s);
  f.b0 = (int)threadIdx.x < f.na ? b0 : 0;
}
__device__ inline void af_load_bub(AFold& f, const BubArgs& B, const uint32_t* __restrict__ cellsB, int par) {
  f.v0 = abub_load(B, cellsB, par, f.b0, threadIdx.x);
}
// the partials' totals in every thread, in one block reduction (sh: FOLD_SH doubles): p.Ap in
// f.tot, the slotted bubbles' Ap sums in aps[]
__device__ inline void af_reduce(AFold& f, double* sh, double (&aps)[BSLOTS]) {
  double q[1 + BSLOTS];
  q[0] = f.a[0];
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) q[1 + k] = f.s[k];
  block_total<1 + BSLOTS>(q, 1 + nslots(f.na), sh);
  f.tot[0] = q[0];
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) aps[k] = q[1 + k];
}
// alpha; rbs[] gets the slotted bubbles' new residuals and bsl[] their cells' V-cycle input
// (residual over cell count, binary32), valid after the block's next barrier
__device__ inline double af_finish(const AFold& f, const Blk& B_, int par, const BubArgs& B,
                                   const uint32_t* __restrict__ cellsB, double* sh, const double (&aps)[BSLOTS],
                                   double* rbs, float* bsl, DevScalars* __restrict__ sc) {
  const int na = f.na;
  const double* tot = f.tot;
  const double beta = f.beta, rz = f.rz;
  auto pap_of = [&](const ABub& v, double* pn_out) -> double {
    const double pn = v.zn + beta * v.po;
    *pn_out = pn;
    return (double)v.fb * pn + (v.sl >= 0 ? slot_val(aps, v.sl) : v.az);
  };
  double bacc[1] = {0};
  auto add_pap = [&](const ABub& v) {
    double pn;
    const double apb = pap_of(v, &pn);
    bacc[0] += pn * apb;
  };
  const bool mine = (int)threadIdx.x < na;
  if (mine) add_pap(f.v0);
  for (int k = threadIdx.x + blockDim.x; k < na; k += blockDim.x) add_pap(abub_load(B, cellsB, par, B.activeList[k], k));
  block_total<1>(bacc, 1, fold_sh2(sh));
  const double pap = tot[0] + bacc[0];
  const double alpha = fabs(pap) < 1e-30 ? 0.0 : rz / pap;
  const bool b0 = first_block(B_);
  double* Pn = bv(B.bvec, BV_P0 + par);
  double* Rn = bv(B.bvec, BV_R0 + par);
  double* Yb = bv(B.bvec, BV_Y);
  double* APb = bv(B.bvec, BV_AP);
  double* accAnext = bv(B.bvec, BV_ACCA0 + 1 - par);
  double rr[1] = {0};
  auto update = [&](const ABub& v) {
    double pn;
    const double apb = pap_of(v, &pn);
    const double rbn = v.ro - alpha * apb;
    if (v.sl >= 0) {
      rbs[v.sl] = rbn;
      bsl[v.sl] = (float)(rbn / v.cb);
    }
    rr[0] += rbn * rbn;
    if (b0) {
      Pn[v.b] = pn;
      APb[v.b] = apb;
      Yb[v.b] = v.yb + alpha * pn;
      Rn[v.b] = rbn;
      fx_clear(accAnext, v.b);
    }
  };
  if (mine) update(f.v0);
  for (int k = threadIdx.x + blockDim.x; k < na; k += blockDim.x) update(abub_load(B, cellsB, par, B.activeList[k], k));
  block_total<1>(rr, 1, fold_sh3(sh));
  if (b0 && threadIdx.x == 0) {
    sc->alpha = alpha;
    sc->pap = rr[0];  // bubble part of |r|^2
  }
  __syncthreads();  // rbs[], bsl[] written; sh free again
  return alpha;
}

// the new residual of bubble b inside in_a0 (slotted: rbs[]; else recomputed as the alpha fold does)
__device__ inline double bubble_r(const BubArgs& B, int b, int par, double alpha, double beta, const double* rbs) {
  const int sl = B.labelSlot[b];
  if (sl >= 0) return rbs[sl];
  const double pn = bv(B.bvec, BV_Z0 + par)[b] + beta * bv(B.bvec, BV_P0 + 1 - par)[b];
  const double apb = (double)B.facesB[b] * pn + fx_read(bv(B.bvec, BV_ACCA0 + par), b);
  return bv(B.bvec, BV_R0 + 1 - par)[b] - alpha * apb;
}

// the norm and stop test after iteration k's update (pressao-cpu.js:382-386), at the start of
// level 0's down kernel; block 0 records them, the iteration count and the next V-cycle input
// scale.  Returns true when the solve stops.
__device__ inline bool fold_update(const Blk& B_, const double* __restrict__ part, int nb, int par, double* sh, double* res,
                                   DevScalars* __restrict__ sc) {
  (void)res;
  double tot[1];
  part_load<1>(part, nb, 1, tot);
  block_total<1>(tot, 1, sh);  // sh is this fold's own: no barrier after
  const double n2 = tot[0] + sc->pap;
  const double nr = sqrt(n2);
  const int itn = sc->itp[1 - par] + 1;
  bool stop = itn >= sc->maxit;
  if (sc->tol > 0 && (nr == 0 || nr <= sc->tol * sc->normb)) stop = true;
  if (first_block(B_) && threadIdx.x == 0) {
    sc->norm2 = n2;
    sc->norm_last = nr;
    sc->itp[par] = itn;
    sc->mgs[1 - par] = nr > 0 ? (float)(sc->mg_mult / nr) : 1.f;
    if (stop) sc->done = 1;
  }
  return stop;
}

// Large grids: nb per-block partials of nv quantities (and of the slotted bubbles) summed into
// PRE_R partials over contiguous block ranges, so each block of the consumer re-sums only PRE_R
// (the consumers' redundant sums would otherwise grow as the square of the block count).
constexpr int PRE_R = 64;
constexpr int FOLD_DIRECT = 1280;  // consumers re-sum up to this many partials themselves
__device__ __forceinline__ void k_prereduce(Blk B_, const double* __restrict__ src, int nb, int nv,
                                            const double* __restrict__ bsrc, double* __restrict__ dst,
                                            double* __restrict__ bdst, const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[(2 + BSLOTS) * 16];
  const int per = (nb + PRE_R - 1) / PRE_R;
  const int b0 = B_.x * per, b1 = b0 + per < nb ? b0 + per : nb;
  const int ns = bsrc ? nslots(sc->active_bubbles) : 0;
  double v[2 + BSLOTS];
#pragma unroll
  for (int k = 0; k < 2 + BSLOTS; k++) v[k] = 0;
  for (int b = b0 + (int)threadIdx.x; b < b1; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < 2; k++)
      if (k < nv) v[k] += src[ps_off(k) + b];
#pragma unroll
    for (int k = 0; k < BSLOTS; k++)
      if (k < ns) v[2 + k] += bsrc[ps_off(k) + b];
  }
  block_sum<2 + BSLOTS>(v, sh);
  if (threadIdx.x == 0) {
    for (int k = 0; k < nv; k++) dst[ps_off(k) + B_.x] = v[k];
    for (int k = 0; k < ns; k++) bdst[ps_off(k) + B_.x] = v[2 + k];
  }
}

// ================================================================== setup (types and neighbour masks)
// Level 0: operator mask (true types, bubble bits) and MG mask (air with a bubble label as water,
// pressao-cpu.js:305-308; colour bit).  Levels 1..PYR_LEVELS: their type bits (tiposGrossos,
// pressao-cpu.js:63-81: air if any child is air, else water if any is water) as a pyramid in LDS.
// A level-l cell covers level-0 cells [2^l I, 2^l (I+1)) on each axis (clipped to the grid), so
// a 64x64 level-0 tile holds whole cells of levels 1..6.  Cells outside a level are 0xFF.
constexpr int PYR_T = 64;
constexpr int PYR_LEVELS = 6;
struct PyrPlan {
  int L;
  int nx[12], ny[12];
  uint8_t* mask[12];
};

__device__ __forceinline__ void k_setup_pyr(Blk B_, const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                            const uint32_t* __restrict__ facesB, int ar,
                                            uint8_t* __restrict__ opMask, const PyrPlan& P) {
  __shared__ uint8_t ta[PYR_T * PYR_T], tb[PYR_T * PYR_T / 4];
  const int tid = threadIdx.x;
  const int nx = P.nx[0], ny = P.ny[0];
  const int ox = B_.x * PYR_T, oy = B_.y * PYR_T;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = tid + 1024 * k;
    const int i = ox + (e & (PYR_T - 1)), j = oy + e / PYR_T;
    uint8_t tmo = 0xFF;
    if (i < nx && j < ny) {
      const size_t c = (size_t)i + (size_t)j * nx;
      int m = 0;
      if (i > 0 && tipo[c - 1] != SOLIDO) m |= 1;
      if (i < nx - 1 && tipo[c + 1] != SOLIDO) m |= 2;
      if (j > 0 && tipo[c - nx] != SOLIDO) m |= 4;
      if (j < ny - 1 && tipo[c + nx] != SOLIDO) m |= 8;
      const uint8_t t = tipo[c];
      const uint32_t lb = (ar && t == AR) ? rot[c] : 0;
      const bool labelled = lb >= 2 && lb < (uint32_t)MAX_BOLHAS;
      const bool bub = labelled && facesB[lb] > 0;
      const bool inner = i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1;
      const uint8_t tm = labelled ? AGUA : t;
      opMask[c] = (uint8_t)(m | (t << 4) | (bub ? MK_BUB : 0) | (bub && inner ? MK_BUBIN : 0));
      P.mask[0][c] = (uint8_t)(m | (tm << 4) | (((i + j + 1) & 1) ? MK_COLOUR : 0));
      tmo = tm;
    }
    ta[e] = tmo;
  }
  __syncthreads();
  uint8_t* src = ta;
  uint8_t* dst = tb;
  for (int l = 1; l <= PYR_LEVELS && l < P.L; l++) {
    const int edge = PYR_T >> l;
    const int gi0 = ox >> l, gj0 = oy >> l;
    for (int e = tid; e < edge * edge; e += 1024) {
      const int I = e % edge, J = e / edge;
      uint8_t t = 0xFF;
      if (gi0 + I < P.nx[l] && gj0 + J < P.ny[l]) {
        bool temAr = false, temAgua = false;
#pragma unroll
        for (int bb = 0; bb < 2; bb++)
#pragma unroll
          for (int aa = 0; aa < 2; aa++) {
            const uint8_t ct = src[(2 * I + aa) + (2 * J + bb) * (2 * edge)];
            if (ct == AR) temAr = true;
            else if (ct == AGUA) temAgua = true;
          }
        t = temAr ? AR : temAgua ? AGUA : SOLIDO;
        P.mask[l][(size_t)(gi0 + I) + (size_t)(gj0 + J) * P.nx[l]] = (uint8_t)(t << 4);
      }
      dst[e] = t;
    }
    __syncthreads();
    uint8_t* tmp = src;
    src = dst;
    dst = tmp;
  }
}

// tiposGrossos (pressao-cpu.js:63-81): air if any child is air, else water if any is water
__device__ __forceinline__ void k_coarse_type(Blk B_, const uint8_t* __restrict__ fm, int fnx, int fny,
                                              uint8_t* __restrict__ cm, int cnx, int cny) {
  const int C = B_.x * blockDim.x + threadIdx.x;
  if (C >= cnx * cny) return;
  const int J = C / cnx, I = C - J * cnx;
  bool temAr = false, temAgua = false;
  for (int b = 0; b < 2; b++)
    for (int a = 0; a < 2; a++) {
      const int i = 2 * I + a, j = 2 * J + b;
      if (i >= fnx || j >= fny) continue;
      const int t = mask_type(fm[i + j * fnx]);
      if (t == AR) temAr = true; else if (t == AGUA) temAgua = true;
    }
  cm[C] = (uint8_t)((temAr ? AR : temAgua ? AGUA : SOLIDO) << 4);
}

// prepararVizinhos (pressao-cpu.js:44-60) on the type bits of levels v0..L-1 (level 0's are set by
// k_setup_pyr), plus the colour bit; one thread per cell of the levels concatenated.  The dense
// operators (coarsest G, level lc's V-cycle matrix) depend only on the masks of levels lc..L-1:
// those are compared with the copy taken at the last build, which takes the new masks, and the
// solve is flagged when any cell differs or `force`.
struct VizPlan {
  int v0, L, lc;
  int off[13];  // first index of each level in the concatenation (off[L] = total)
  int nx[12], ny[12];
  uint8_t* mask[12];
};
__device__ __forceinline__ void k_viz_all(Blk B_, const VizPlan& P, uint8_t* __restrict__ saved, int force,
                                          DevScalars* __restrict__ sc) {
  const int g = B_.x * blockDim.x + threadIdx.x;
  if (g >= P.off[P.L]) return;
  int l = P.v0;
  while (g >= P.off[l + 1]) l++;
  const int c = g - P.off[l], nx = P.nx[l], ny = P.ny[l];
  uint8_t* m = P.mask[l];
  uint8_t mk = m[c];
  if (l > 0) {
    const int j = c / nx, i = c - j * nx;
    int v = 0;
    if (i > 0 && mask_type(m[c - 1]) != SOLIDO) v |= 1;
    if (i < nx - 1 && mask_type(m[c + 1]) != SOLIDO) v |= 2;
    if (j > 0 && mask_type(m[c - nx]) != SOLIDO) v |= 4;
    if (j < ny - 1 && mask_type(m[c + nx]) != SOLIDO) v |= 8;
    mk = (uint8_t)((mk & 0x30) | v | (((i + j + 1) & 1) ? MK_COLOUR : 0));
    m[c] = mk;
  }
  if (l >= P.lc) {
    const int s = g - P.off[P.lc];
    if (force || saved[s] != mk) {
      saved[s] = mk;
      sc->op_dirty = 1;
    }
  }
}

// bubble vectors, the active list (labels with water faces, in label order) and their slots.
// Block 0 sets y, the list and the slots; block q = 1 .. BV_N - 1 clears bubble vector q.
__device__ __forceinline__ void k_bub_init(Blk B_, const uint32_t* __restrict__ facesB, const float* __restrict__ yB32,
                                           int ar, double* __restrict__ bvec, int* __restrict__ activeList,
                                           int* __restrict__ labelSlot, DevScalars* __restrict__ sc) {
  __shared__ int cnt[16];
  const int tid = threadIdx.x;
  if (B_.x > 0) {
    double* v = bv(bvec, B_.x);
    for (int r = tid; r < MAX_BOLHAS; r += blockDim.x) v[r] = 0.0;
    return;
  }
  int mine = 0;
  for (int k = 0; k < 4; k++) {
    const int r = tid * 4 + k;
    const bool act = ar && r >= 2 && facesB[r] > 0;
    mine += act;
    bv(bvec, BV_Y)[r] = act ? (double)yB32[r] : 0.0;
    labelSlot[r] = -1;
  }
  // exclusive scan of the per-thread counts in label order: within each wavefront by shuffles,
  // then over the 16 wavefront totals
  const int lane = tid & 63, w = tid >> 6;
  int incl = mine;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d, 64);
    if (lane >= d) incl += v;
  }
  if (lane == 63) cnt[w] = incl;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int q = 0; q < 16; q++) { const int c = cnt[q]; cnt[q] = acc; acc += c; }
    sc->active_bubbles = acc;
    sc->op_dirty = 0;
  }
  __syncthreads();
  int o = cnt[w] + incl - mine;
  for (int k = 0; k < 4; k++) {
    const int r = tid * 4 + k;
    if (ar && r >= 2 && facesB[r] > 0) {
      activeList[o] = r;
      labelSlot[r] = o < BSLOTS ? o : -1;
      o++;
    }
  }
}

// a restart of the compensated iterate: the bubble directions and the fixed-point accumulators
// start from zero as at the solve's set-up; y (the bubble unknowns) is kept
__device__ __forceinline__ void k_bub_restart(Blk B_, double* __restrict__ bvec) {
  for (int r = threadIdx.x; r < MAX_BOLHAS; r += blockDim.x)
    for (int q = BV_AP; q < BV_N; q++) bv(bvec, q)[r] = 0.0;
}

// ================================================================== init: warm-start residual
// q := 0 off water, bubble cells := y (difundir); r = -div - A q; bubble rows A_b [q;y].
// qmode 0: q alone (the reference's binary32 iterate); 1: compensated iterate q + qlo starting
// with qlo = 0; 2: a restart of the compensated iterate (its residual from q + qlo in binary64).
__device__ __forceinline__ void k_pcg_init(Blk B_, const uint8_t* __restrict__ opMask, const uint32_t* __restrict__ rot,
                                           const float* __restrict__ div, float* __restrict__ q,
                                           float* __restrict__ qlo, int qmode,
                                           float* __restrict__ r, float* __restrict__ pbuf0, int nx, int C,
                                           double* __restrict__ bvec, const int* __restrict__ labelSlot,
                                           double* __restrict__ partials, double* __restrict__ bpart,
                                           DevScalars* __restrict__ sc) {
  __shared__ double sh[FOLD_SH];
  __shared__ BubLds aggt;
  const int ns = nslots(sc->active_bubbles);
  double v[5] = {0, 0, 0, 0, 0};  // r.r, div.div, sum(b), n_water, anchored
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double* yb = bv(bvec, BV_Y);
  const BubAgg agg = bagg_begin(aggt, bv(bvec, BV_ACCI), sc->active_bubbles > BSLOTS);
  if (agg.t) __syncthreads();
  // the cell's value as the operator sees it: binary32 q (bubble cells: their y), plus qlo on a restart
  auto qe = [&](int d, uint8_t md) -> float {
    if (mask_type(md) == AGUA) return q[d];
    if (md & MK_BUB) return (float)yb[rot[d]];
    return 0.f;
  };
  auto qd = [&](int d, uint8_t md) -> double {
    if (qmode == 2) {  // a restart: water q + qlo, bubble y in binary64
      if (mask_type(md) == AGUA) return (double)q[d] + (double)qlo[d];
      if (md & MK_BUB) return yb[rot[d]];
    }
    return (double)qe(d, md);
  };
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c < C) {
    const uint8_t m = opMask[c];
    const float qc = qe(c, m);
    const double qcd = qd(c, m);
    pbuf0[c] = 0.f;
    if (mask_type(m) == AGUA) {
      double s = 0;
      bool anchored = false;
      const int nb[4] = {c - 1, c + 1, c - nx, c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        const uint8_t md = opMask[nb[k]];
        s += qd(nb[k], md);
        if (mask_type(md) == AR) {
          if (md & MK_BUB) {
            if (md & MK_BUBIN) bub_add(sl, labelSlot, agg, rot[nb[k]], -qcd);
          } else anchored = true;
        }
      }
      const double aq = mask_cnt(m) * qcd - s;
      const float rr = (float)(-(double)div[c] - aq);
      r[c] = rr;
      v[0] = (double)rr * rr;
      v[1] = (double)div[c] * div[c];
      v[2] = -(double)div[c];
      v[3] = 1;
      v[4] = anchored ? 1 : 0;
    } else r[c] = 0.f;
    q[c] = qc;
    if (qmode == 1 || (qmode == 2 && mask_type(m) != AGUA)) qlo[c] = 0.f;
  }
  bagg_flush(agg);
  store_partials_last<5, BSLOTS>(B_, v, 5, partials, sl, ns, bpart, sh);
}

// restart = 0 starts a solve (tolerance, iteration cap and the mixed modes' scale numerator are
// set here, as arguments, so no host copy precedes the solve); restart = 1 restarts the
// compensated iterate from its true residual, adding the finished segment's iterations to the count.
__device__ __forceinline__ void k_fin_init(Blk B_, const uint32_t* __restrict__ facesB, double* __restrict__ bvec,
                                           const double* __restrict__ rhsB,
                                           const int* __restrict__ activeList,
                                           const int* __restrict__ labelSlot,
                                           const double* __restrict__ partials,
                                           const double* __restrict__ bpart, int nb, double tol,
                                           int maxit, float mult, int restart,
                                           DevScalars* __restrict__ sc) {
  if (threadIdx.x == 0) {
    if (restart) {
      const int seg = sc->itp[0] > sc->itp[1] ? sc->itp[0] : sc->itp[1];
      if (sc->restarts == 0) sc->iters_first = seg;
      sc->iters_base += seg;
      sc->restarts += 1;
    } else {
      sc->tol = tol;
      sc->maxit = maxit;
      sc->mg_mult = mult;
      sc->iters_base = 0;
      sc->iters_first = 0;
      sc->restarts = 0;
    }
    sc->true_norm = -1;
    sc->true_done = 0;
  }
  __syncthreads();
  __shared__ double sh[BSLOTS * 16], res[BSLOTS], slotTot[BSLOTS];
  __shared__ double mean_s;
  const int na = sc->active_bubbles;
  double tot[5];
  sum_partials<5>(partials, nb, 5, tot, sh, res);
  double stt[BSLOTS];
  sum_partials<BSLOTS>(bpart, nb, nslots(na), stt, sh, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < BSLOTS; k++) slotTot[k] = stt[k];
  __syncthreads();
  double bacc[3] = {0, 0, 0};  // rb.rb, rhs.rhs, sum(rhs)
  double* rb = bv(bvec, BV_R1);  // iteration 0 (parity 0) reads the residual of parity 1
  double* acc0 = bv(bvec, BV_ACCI);
  const double* yb = bv(bvec, BV_Y);
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const int s = labelSlot[b];
    double accB;
    if (s >= 0) accB = slotTot[s];
    else { accB = fx_read(acc0, b); fx_clear(acc0, b); }
    rb[b] = rhsB[b] - ((double)facesB[b] * yb[b] + accB);
    bacc[0] += rb[b] * rb[b];
    bacc[1] += rhsB[b] * rhsB[b];
    bacc[2] += rhsB[b];
  }
  block_sum<3>(bacc, sh);
  if (threadIdx.x == 0) {
    double rr = tot[0] + bacc[0];
    double bb = tot[1] + bacc[1];
    double mean = 0;
    const double N = tot[3] + na;
    // Residual-terminated mode on a system without any Dirichlet air (every water cell only
    // touches water, walls or tracked bubbles): the operator's null space is the constant over
    // all unknowns; its component is removed from b so the tolerance is reachable.  Then
    // sum(r) = sum(b) (1^T A = 0), so the projected norms follow from the sums.
    if (sc->tol > 0 && tot[4] == 0 && N > 0) {
      mean = (tot[2] + bacc[2]) / N;
      const double sr = tot[2] + bacc[2];
      rr = rr - 2 * mean * sr + N * mean * mean;
      bb = bb - N * mean * mean;
      if (rr < 0) rr = 0;
      if (bb < 0) bb = 0;
    }
    sc->proj_mean = mean;
    mean_s = mean;
    sc->norm2 = rr;
    sc->norm0 = sqrt(rr);
    sc->norm_last = sc->norm0;
    sc->normb = sqrt(bb);
    sc->iters = 0;
    if (!restart) sc->solves += 1;
    sc->rz = 0;
    sc->beta = 0;
    sc->alpha = 0;
    const bool stop = sc->maxit <= 0 || (sc->tol > 0 && (sc->norm0 == 0 || sc->norm0 <= sc->tol * sc->normb));
    sc->done = stop ? 1 : 0;
    sc->mg_scale = sc->norm0 > 0 ? (float)(sc->mg_mult / sc->norm0) : 1.f;
    sc->mg_scale_next = sc->mg_scale;
    // the first cycle (parity 1) and iteration 0 both scale by 1/||r0||
    sc->mgs[0] = sc->mgs[1] = sc->mg_scale;
    sc->rzp[0] = sc->rzp[1] = 0;
    sc->itp[0] = sc->itp[1] = 0;
  }
  __syncthreads();
  if (mean_s != 0)
    for (int k = threadIdx.x; k < na; k += blockDim.x) rb[activeList[k]] -= mean_s;
}

// ================================================================== smoother tiles
// The lane's four cells: in rows ly0 and ly0 + 8 (ly0 = lane >> 3) the pair of columns 2p,
// 2p + 1 (p = lane & 7); one cell of each pair is red and one black.  Cells4 indexes them
// [colour][row], so a half-sweep touches registers at compile-time indices only.
// m: bits 0-3 free neighbours, bit 4 updated by the sweeps (water with a free neighbour),
// bit 5 water (kept in the output; other cells are written as 0).  ik = 1/count and kp = 0 on
// updated cells, ik = 0 and kp = 1 on the others, so one fma gives either the update or the
// unchanged value.
struct Cells4 {
  float v[2][2], b[2][2], ik[2][2], kp[2][2];
  int m[2][2];
};

// column and row in the tile of the lane's cell of colour q in row r (colour (i + j + 1) & 1,
// tile origins even)
__device__ inline int cell_lx(int lane, int q) { return 2 * (lane & 7) + ((q + (lane >> 3) + 1) & 1); }
__device__ inline int cell_ly(int lane, int r) { return (lane >> 3) + 8 * r; }
__device__ inline int cell_e(int lane, int q, int r) { return cell_lx(lane, q) + 1 + (cell_ly(lane, r) + 1) * TP; }
// the lane's row pair r, and the colour of its even column
__device__ inline Pair cell_pair(int lane, int r, int ox, int oy, int nx, int ny) {
  const int i = ox + 2 * (lane & 7), j = oy + cell_ly(lane, r);
  const bool in = i < nx && j < ny && i >= 0 && j >= 0;
  return {in ? (size_t)i + (size_t)j * nx : 0, in};
}
__device__ inline int pair_c0(int lane) { return ((lane >> 3) + 1) & 1; }
// the pair's two values in colour order [c][r] (x is the even column's)
template <class T>
__device__ inline void by_colour(T (&v)[2][2], int r, int c0, T x, T y) {
  v[0][r] = c0 ? y : x;
  v[1][r] = c0 ? x : y;
}

__device__ inline void cell_setup(Cells4& c, int q, int r, uint8_t mk, float b) {
  const int cnt = mask_cnt(mk);
  const bool water = mask_type(mk) == AGUA;
  const bool upd = water && cnt > 0;
  c.m[q][r] = mask_viz(mk) | (upd ? 16 : 0) | (water ? 32 : 0);
  c.b[q][r] = b;
  c.ik[q][r] = upd ? 1.f / (float)cnt : 0.f;
  c.kp[q][r] = upd ? 0.f : 1.f;
}

__device__ inline void cell_clear(Cells4& c, int q, int r) {
  c.m[q][r] = 0;
  c.b[q][r] = 0.f;
  c.ik[q][r] = 0.f;
  c.kp[q][r] = 1.f;
}

// the cells of colour COR: x = (sum of free neighbours + b) / count.  Every cell that is not a
// free neighbour (solid, or outside the grid) holds +0 in the tile: x is 0 on every cell that
// is not water (the passes write 0 there, the tile's border outside the grid is 0), so the sum
// over all four neighbours in the reference's order (left, right, down, up) has the bits of
// the sum over the free ones.  Both rows' neighbours are read before any arithmetic (one LDS
// latency per half-sweep), and the update is branch-free: fma(s + b, 1/count, 0) on updated
// cells, fma(s + b, 0, x) = x on the others.
// One horizontal neighbour of each cell is the lane's own cell of the other colour in that row
// (its value in registers equals the tile's): only the other three come from T.
template <int COR>
__device__ inline void half_sweep(float* T, Cells4& c, int lane) {
  const int off = (COR + (lane >> 3) + 1) & 1;  // the cells of colour COR sit at column 2p + off
  float th[2], td[2], tu[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int e = cell_e(lane, COR, r);
    th[r] = T[off ? e + 1 : e - 1];
    td[r] = T[e - TP];
    tu[r] = T[e + TP];
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const float own = c.v[1 - COR][r];
    const float tl = off ? own : th[r], tr = off ? th[r] : own;
    const float sum = ((tl + tr) + td[r]) + tu[r];
    const float nv = __builtin_fmaf(sum + c.b[COR][r], c.ik[COR][r], c.kp[COR][r] * c.v[COR][r]);
    c.v[COR][r] = nv;
    T[cell_e(lane, COR, r)] = nv;
  }
}

// 4 red-black sweeps (colour order C0, 1 - C0) of the tile in T (18x18, halo fixed); the
// lane's own cells live in registers and are mirrored into T after every update.
template <int C0>
__device__ inline void tile_sweeps(float* T, Cells4& c, int lane) {
#pragma unroll
  for (int it = 0; it < 4; it++) {
    half_sweep<C0>(T, c, lane);
    wave_sync();
    half_sweep<1 - C0>(T, c, lane);
    wave_sync();
  }
}

// T := snapshot of src on the tile + halo (0 outside the grid): the six loads of a lane first,
// then the six LDS stores
// the tile with its halo as row pairs from the even column ox - 2: 10 pairs per row, 180 in all
// (three per lane); a pair's two values go to columns 2k - 1 and 2k of the tile (the first
// pair's first and the last pair's second column lie outside it)
constexpr int TPAIRS = 10 * TS;
__device__ inline Pair tile_pair(int p, int ox, int oy, int nx, int ny, int* lx, int* ly) {
  const int y = p / 10, k = p - 10 * y;
  const int i = ox - 2 + 2 * k, j = oy - 1 + y;
  const bool in = p < TPAIRS && i >= 0 && j >= 0 && i < nx && j < ny;
  *lx = 2 * k - 1;
  *ly = y;
  return {in ? (size_t)i + (size_t)j * nx : 0, in};
}
__device__ inline void tile_put(float* T, int p, int lx, int ly, float a, float b) {
  if (p < TPAIRS) {
    if (lx >= 0) T[lx + ly * TP] = a;
    if (lx < TS - 1) T[lx + 1 + ly * TP] = b;
  }
}

// (PAIRS false: cells only -- the walking pass kernel, whose registers the pair path would take
// past 64, i.e. below 8 waves per SIMD)
template <class SX, bool PAIRS = true>
__device__ inline void fill_tile(float* T, const SX* __restrict__ src, int ox, int oy, int nx, int ny, int lane) {
  if (PAIRS && (nx & 1) == 0) {
    F2 v[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int lx, ly;
      const Pair pr = tile_pair(lane + 64 * q, ox, oy, nx, ny, &lx, &ly);
      const F2 t = ld2<SX>(src, pr.c);
      v[q] = pr.in ? t : F2{0.f, 0.f};
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int lx, ly;
      (void)tile_pair(lane + 64 * q, ox, oy, nx, ny, &lx, &ly);
      tile_put(T, lane + 64 * q, lx, ly, v[q].x, v[q].y);
    }
    return;
  }
  float v[6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int e = lane + 64 * q;
    const int ly = e / TS, lx = e - ly * TS;
    const int i = ox - 1 + lx, j = oy - 1 + ly;
    const bool in = e < TCELLS && i >= 0 && j >= 0 && i < nx && j < ny;
    v[q] = ldf_if<SX>(src, (size_t)i + (size_t)j * nx, in);
  }
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int e = lane + 64 * q;
    const int ly = e / TS, lx = e - ly * TS;
    if (e < TCELLS) T[lx + ly * TP] = v[q];
  }
}

__device__ inline void zero_tile(float* T, int lane) {
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int e = lane + 64 * q;
    const int ly = e / TS, lx = e - ly * TS;
    if (e < TCELLS) T[lx + ly * TP] = 0.f;
  }
}

// the lane's cells: mask and b from global memory (issued before the tile loads so both
// overlap), then the cell state with values from T
struct Own4 {
  uint8_t m[2][2];
  float b[2][2];
};

template <class SB>
__device__ inline void prefetch_cells(Own4& o, const SB* __restrict__ b, const uint8_t* __restrict__ mask, int ox,
                                      int oy, int nx, int ny, int lane) {
  if ((nx & 1) == 0) {
    const int c0 = pair_c0(lane);
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const Pair pr = cell_pair(lane, r, ox, oy, nx, ny);
      const uint16_t m2 = ld2b(mask, pr.c);
      const F2 b2 = ld2<SB>(b, pr.c);
      by_colour(o.m, r, c0, (uint8_t)(pr.in ? m2 & 0xff : 0), (uint8_t)(pr.in ? m2 >> 8 : 0));
      by_colour(o.b, r, c0, pr.in ? b2.x : 0.f, pr.in ? b2.y : 0.f);
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int i = ox + cell_lx(lane, q), j = oy + cell_ly(lane, r);
      const bool in = i < nx && j < ny && i >= 0 && j >= 0;
      const size_t cc = in ? (size_t)i + (size_t)j * nx : 0;
      o.m[q][r] = ld_if<uint8_t>(mask, cc, in, 0);
      o.b[q][r] = ldf_if<SB>(b, cc, in);
    }
}

__device__ inline void init_cells(Cells4& c, const Own4& o, const float* T, int lane) {
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int r = 0; r < 2; r++) {
      cell_setup(c, q, r, o.m[q][r], o.b[q][r]);
      c.v[q][r] = T[cell_e(lane, q, r)];
    }
}

template <class SX>
__device__ inline void store_cells(const Cells4& c, SX* __restrict__ dst, int ox, int oy, int nx, int ny, int lane) {
  if ((nx & 1) == 0) {
    const int c0 = pair_c0(lane);
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const Pair pr = cell_pair(lane, r, ox, oy, nx, ny);
      const float v0 = (c.m[0][r] & 32) ? c.v[0][r] : 0.f, v1 = (c.m[1][r] & 32) ? c.v[1][r] : 0.f;
      if (pr.in) st2<SX>(dst, pr.c, c0 ? v1 : v0, c0 ? v0 : v1);
    }
    return;
  }
#pragma unroll
  for (int q = 0; q < 2; q++)
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int i = ox + cell_lx(lane, q), j = oy + cell_ly(lane, r);
      if (i < nx && j < ny && i >= 0 && j >= 0)
        st<SX>(dst, (size_t)i + (size_t)j * nx, (c.m[q][r] & 32) ? c.v[q][r] : 0.f);
    }
}

// A 256-thread block of the level-0 tile kernels owns a group of GW x GH tiles, one per wave:
// four tiles side by side, so that a block reads 64-cell (256-byte) row segments.
constexpr int GW = 4, GH = 1;
__device__ inline void group_tile(int bx, int by, int w, int* tx, int* ty) {
  *tx = GW * bx + (w % GW);
  *ty = GH * by + (w / GW);
}
// tile of wave w of block (bx, by)
__device__ inline void block_tile(const Blk& B_, int* tx, int* ty) {
  int bx, by;
  xcd_block2(B_, &bx, &by);
  group_tile(bx, by, threadIdx.x >> 6, tx, ty);
}

// ================================================================== level 0 input + first pass
// UPD: q += alpha p, r -= alpha Ap and |r|^2 (PCG update, pressao-cpu.js:373-381); else the
// first cycle (r minus the projected mean).  b0 = scale * r on water, rb/cells on bubble cells
// (pressao-cpu.js:326-332), then pass A (offset 0, colours 0,1) from x = 0 (cicloV,
// pressao-cpu.js:243) when level 0 is not the coarsest.  One-dimensional grid over the tile
// groups (vgx of them per row, nvb in all); large grids launch at most IN_A0_GRID blocks, each
// walking its groups after one fold.
constexpr int IN_A0_GRID = 1024;  // 4 blocks per CU at this kernel's 123 VGPRs
template <class SB, bool UPD, bool PASSA>
__device__ __forceinline__ void k_in_a0(Blk B_, const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ mg0,
                                        const uint32_t* __restrict__ rot, const uint32_t* __restrict__ cellsB,
                                        const float* __restrict__ p, const float* __restrict__ ap,
                                        float* __restrict__ q, float* __restrict__ qlo, int comp,
                                        float* __restrict__ r, int nx, int ny, int ntx,
                                        int nty, int vgx, int nvb, BubArgs B, const double* __restrict__ partA,
                                        const double* __restrict__ bpartA, int nbs, int par,
                                        SB* __restrict__ b0, float* __restrict__ x0,
                                        double* __restrict__ partRR, int scaled, const float* __restrict__ mgs,
                                        DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float Tt[4][TLDS];
  __shared__ double sh[FOLD_SH], rbs[BSLOTS];
  __shared__ float bsl[BSLOTS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int vb = xcd_block(B_.x, B_.gx);
  int ox, oy;
  bool valid;
  const double mean = UPD ? 0.0 : sc->proj_mean;
  const float scale = scaled ? *mgs : 1.f;
  const double* rb = bv(B.bvec, BV_R0 + 1 - par);
  double rr[1] = {0};
  // every load of the four cells first (one memory latency), then the arithmetic
  uint8_t om[2][2], gm[2][2];
  float qv[2][2], pv[2][2], rv[2][2], av[2][2], lv[2][2];
  size_t ci[2][2];
  bool in[2][2];
  const bool pairs = (nx & 1) == 0;  // a lane's two cells of a row move as one access
  const int c0 = pair_c0(lane);
  auto load = [&](int g) {
    const int by = g / vgx, bx = g - by * vgx;
    int tx, ty;
    group_tile(bx, by, w, &tx, &ty);
    valid = tx < ntx && ty < nty;
    ox = tx * 16;
    oy = ty * 16;
    if (pairs) {
      // raw values (pair 0 outside the grid): every use is under in[][]
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const Pair pr = cell_pair(lane, rw, ox, oy, nx, ny);
        const bool inr = valid && pr.in;
        const size_t c = inr ? pr.c : 0;
        in[0][rw] = in[1][rw] = inr;
        by_colour(ci, rw, c0, c, c + 1);
        const uint16_t o2 = ld2b(opMask, c), g2 = ld2b(mg0, c);
        by_colour(om, rw, c0, (uint8_t)(o2 & 0xff), (uint8_t)(o2 >> 8));
        by_colour(gm, rw, c0, (uint8_t)(g2 & 0xff), (uint8_t)(g2 >> 8));
        const F2 r2 = ld2<float>(r, c);
        by_colour(rv, rw, c0, r2.x, r2.y);
        if (UPD) {
          const F2 q2 = ld2<float>(q, c), p2 = ld2<float>(p, c), a2 = ld2<float>(ap, c);
          by_colour(qv, rw, c0, q2.x, q2.y);
          by_colour(pv, rw, c0, p2.x, p2.y);
          by_colour(av, rw, c0, a2.x, a2.y);
          if (comp) {
            const F2 l2 = ld2<float>(qlo, c);
            by_colour(lv, rw, c0, l2.x, l2.y);
          } else {
            lv[0][rw] = lv[1][rw] = 0.f;
          }
        }
      }
      return;
    }
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        in[cq][rw] = valid && i < nx && j < ny;
        const size_t c = in[cq][rw] ? (size_t)i + (size_t)j * nx : 0;
        ci[cq][rw] = c;
        om[cq][rw] = ld_if<uint8_t>(opMask, c, in[cq][rw], 0);
        gm[cq][rw] = ld_if<uint8_t>(mg0, c, in[cq][rw], 0);
        rv[cq][rw] = r[c];
        if (UPD) {
          qv[cq][rw] = q[c];
          pv[cq][rw] = p[c];
          av[cq][rw] = ap[c];
          lv[cq][rw] = comp ? qlo[c] : 0.f;
        }
      }
  };
  // bubble cells: their labels (loaded once the masks have arrived, only in waves that hold
  // one); the first cycle and bubbles without a slot also read their residuals and cell counts
  uint32_t lbv[2][2];
  auto load_bub_labels = [&]() {  // only waves that hold a bubble cell wait for the labels
    bool hb = false;
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) hb = hb || (in[cq][rw] && (om[cq][rw] & MK_BUB));
    if (__ballot(hb)) {
#pragma unroll
      for (int cq = 0; cq < 2; cq++)
#pragma unroll
        for (int rw = 0; rw < 2; rw++)
          lbv[cq][rw] = ld_if<uint32_t>(rot, ci[cq][rw], in[cq][rw] && (om[cq][rw] & MK_BUB), 0u);
    } else {
#pragma unroll
      for (int cq = 0; cq < 2; cq++)
#pragma unroll
        for (int rw = 0; rw < 2; rw++) lbv[cq][rw] = 0u;
    }
  };
  load(vb);
  // UPD: alpha and the bubbles' new residuals from spmv's partials (the alpha fold, its loads
  // beside the tile's); else (the first cycle) the current bubble residuals are those of
  // parity 1 - par
  AFold af;
  if (UPD) af_load(af, partA, bpartA, nbs, par, B, sc);
  if (sc_done) return;
  load_bub_labels();
  double aps[BSLOTS];
  if (UPD) {
    af_load_bub(af, B, cellsB, par);
    af_reduce(af, sh, aps);
  }
  const double alpha = UPD ? af_finish(af, B_, par, B, cellsB, sh, aps, rbs, bsl, sc) : 0.0;
  const double beta = UPD ? af.beta : 0.0;
  const bool wr = UPD || mean != 0;  // r is rewritten
  for (;;) {
    Cells4 cc;
    // the new q, q_lo, r and b0 of each cell (cells other than water keep their q, q_lo and r)
    float qn[2][2], ln[2][2], rn[2][2], bn[2][2];
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        cc.v[cq][rw] = 0.f;
        qn[cq][rw] = UPD ? qv[cq][rw] : 0.f;
        ln[cq][rw] = UPD ? lv[cq][rw] : 0.f;
        rn[cq][rw] = rv[cq][rw];
        bn[cq][rw] = 0.f;
        if (in[cq][rw]) {
          const size_t c = ci[cq][rw];
          const uint8_t m = om[cq][rw];
          float bval = 0.f;
          if (mask_type(m) == AGUA) {
            if (UPD) {
              if (comp) {  // the compensated iterate: q + qlo keeps the update's binary64 sum
                const double sx = (double)qv[cq][rw] + (double)lv[cq][rw] + alpha * (double)pv[cq][rw];
                const float hi = (float)sx;
                qn[cq][rw] = hi;
                ln[cq][rw] = (float)(sx - (double)hi);
              } else {
                qn[cq][rw] = (float)((double)qv[cq][rw] + alpha * (double)pv[cq][rw]);
              }
              rn[cq][rw] = (float)((double)rv[cq][rw] - alpha * (double)av[cq][rw]);
              rr[0] += (double)rn[cq][rw] * rn[cq][rw];
            } else if (mean != 0) {
              rn[cq][rw] = (float)((double)rv[cq][rw] - mean);
            }
            bval = rn[cq][rw];
            if (!pairs) {
              if (UPD) {
                q[c] = qn[cq][rw];
                if (comp) qlo[c] = ln[cq][rw];
              }
              if (wr) r[c] = rn[cq][rw];
            }
          } else if (m & MK_BUB) {
            const uint32_t lb = lbv[cq][rw];
            const int sl = B.labelSlot[lb];
            if (UPD && sl >= 0) bval = bsl[sl];
            else bval = (float)((UPD ? bubble_r(B, lb, par, alpha, beta, rbs) : rb[lb]) / cellsB[lb]);
          }
          bval = rnd<SB>(bval * scale);
          bn[cq][rw] = bval;
          if (!pairs) st<SB>(b0, c, bval);
          cell_setup(cc, cq, rw, gm[cq][rw], bval);
        } else cell_clear(cc, cq, rw);
      }
    if (pairs) {
#pragma unroll
      for (int rw = 0; rw < 2; rw++)
        if (in[0][rw]) {
          const size_t c = c0 ? ci[1][rw] : ci[0][rw];  // the even column
          if (UPD) {
            st2<float>(q, c, c0 ? qn[1][rw] : qn[0][rw], c0 ? qn[0][rw] : qn[1][rw]);
            if (comp) st2<float>(qlo, c, c0 ? ln[1][rw] : ln[0][rw], c0 ? ln[0][rw] : ln[1][rw]);
          }
          if (wr) st2<float>(r, c, c0 ? rn[1][rw] : rn[0][rw], c0 ? rn[0][rw] : rn[1][rw]);
          st2<SB>(b0, c, c0 ? bn[1][rw] : bn[0][rw], c0 ? bn[0][rw] : bn[1][rw]);
        }
    }
    if (PASSA && valid) {
      float* T = Tt[w];
      zero_tile(T, lane);
      wave_sync();
      tile_sweeps<0>(T, cc, lane);
      store_cells<float>(cc, x0, ox, oy, nx, ny, lane);
      wave_sync();  // the wave's tile is free for its next group
    }
    vb += B_.gx;
    if (vb >= nvb) break;
    load(vb);
    load_bub_labels();
  }
  if (UPD) store_partials_last<1>(B_, rr, 1, partRR, sh);
}

struct ZArgs {  // level-0 outputs: z = MG result and the dots r.z, r.z_old
  const uint8_t* opMask;
  const uint32_t* rot;
  const uint32_t* cellsB;
  const int* labelSlot;
  const float* r;
  const float* zold;
  float* znew;
  double* accZ;       // bubble z sums of labels without a slot (128-bit fixed point, fx_add)
  double* partials;   // r.z, r.z_old partials (two slots)
  double* bpart;      // bubble z sums of the slotted labels
  const float* mgs;   // the V-cycle input scale of this cycle
  int scaled;
};

// The bubble cells among a lane's four level-0 cells: their labels (one load in flight each,
// issued once the masks have arrived; other cells read entry 0), then their slots and cell
// counts
struct ZBubCells {
  uint32_t lb[2][2], cb[2][2];
  int sl[2][2];
};
__device__ inline void zbub_labels(ZBubCells& zb, const ZArgs& za, const uint8_t (&om)[2][2], const size_t (&c)[2][2]) {
#pragma unroll
  for (int cq = 0; cq < 2; cq++)
#pragma unroll
    for (int rw = 0; rw < 2; rw++)
      zb.lb[cq][rw] = ld_if<uint32_t>(za.rot, c[cq][rw], (om[cq][rw] & MK_BUB) != 0, 0u);
}
__device__ inline void zbub_vals(ZBubCells& zb, const ZArgs& za) {
#pragma unroll
  for (int cq = 0; cq < 2; cq++)
#pragma unroll
    for (int rw = 0; rw < 2; rw++) {
      zb.sl[cq][rw] = za.labelSlot[zb.lb[cq][rw]];
      zb.cb[cq][rw] = za.cellsB[zb.lb[cq][rw]];
    }
}

// z and its dot/bubble contributions for one cell (x = MG output of level 0; lb, lsl, cb: the
// cell's bubble label, its slot and cell count when it is a bubble cell)
__device__ inline float z_val(uint8_t om, float rc, float zoldc, float xv, float inv, double (&dv)[2],
                             double (&sl)[BSLOTS], const BubAgg& agg, uint32_t lb, int lsl, uint32_t cb) {
  const float x = xv * inv;
  float z = 0.f;
  if (mask_type(om) == AGUA) {
    z = x;
    dv[0] += (double)rc * z;
    dv[1] += (double)rc * zoldc;
  } else if (om & MK_BUB) {
    bub_add_s(sl, lsl, agg, lb, (double)x / cb);
  }
  return z;
}
__device__ inline void z_cell(const ZArgs& za, size_t c, uint8_t om, float rc, float zoldc, float xv, float inv,
                              double (&dv)[2], double (&sl)[BSLOTS], const BubAgg& agg, uint32_t lb, int lsl,
                              uint32_t cb) {
  za.znew[c] = z_val(om, rc, zoldc, xv, inv, dv, sl, agg, lb, lsl, cb);
}

// z and the dots when level 0 itself is solved inside the workgroup (small grids)
__device__ __forceinline__ void k_zdots0(Blk B_, const float* __restrict__ x0, ZArgs za, int C,
                                         const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[FOLD_SH];
  __shared__ BubLds aggt;
  const int ns = nslots(sc->active_bubbles);
  double dv[2] = {0, 0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  const BubAgg agg = bagg_begin(aggt, za.accZ, sc->active_bubbles > BSLOTS);
  if (agg.t) __syncthreads();
  const int c = B_.x * blockDim.x + threadIdx.x;
  const float inv = za.scaled ? 1.f / *za.mgs : 1.f;
  if (c < C) {
    const uint8_t om = za.opMask[c];
    const uint32_t lb = (om & MK_BUB) ? za.rot[c] : 0u;
    z_cell(za, c, om, za.r[c], za.zold[c], x0[c], inv, dv, sl, agg, lb, za.labelSlot[lb], za.cellsB[lb]);
  }
  bagg_flush(agg);
  store_partials_last<2, BSLOTS>(B_, dv, 2, za.partials, sl, ns, za.bpart, sh);
}

// snapshot value of fine cell (i, j) with its prolongation addend: x_B + P e on water cells, from
// the coarse values of the tile's 10x10 patch CP at (PX, PY) (prolongar, pressao-cpu.js:191-214)
__device__ inline float snap_cell(float val, uint8_t mk, int i, int j, const float* CP, int PX, int PY) {
  if (mask_type(mk) == AGUA) {  // in the grid (masks are 0 outside)
    const int I0 = ((i & 1) == 0 ? (i >> 1) - 1 : (i - 1) >> 1) - PX;
    const int J0 = ((j & 1) == 0 ? (j >> 1) - 1 : (j - 1) >> 1) - PY;
    const float wx0 = (i & 1) == 0 ? 0.25f : 0.75f, wy0 = (j & 1) == 0 ? 0.25f : 0.75f;
    const float wx1 = 1.f - wx0, wy1 = 1.f - wy0;
    float e2 = 0.f;
    e2 += CP[I0 + J0 * 10] * wx0 * wy0;
    e2 += CP[I0 + 1 + J0 * 10] * wx1 * wy0;
    e2 += CP[I0 + (J0 + 1) * 10] * wx0 * wy1;
    e2 += CP[I0 + 1 + (J0 + 1) * 10] * wx1 * wy1;
    val += e2;
  }
  return val;
}

// The snapshot of a fine tile with halo: values and masks, as row pairs (even row length) or
// cells; put() then writes each tile cell's snapshot value with its prolongation addend to T.
template <class SX>
struct Snap {
  float v[6];
  uint8_t m[6];
  bool pairs;
  __device__ void load(const SX* __restrict__ src, const uint8_t* __restrict__ mask, int ox, int oy, int nx, int ny,
                       int lane) {
    pairs = (nx & 1) == 0;
    if (pairs) {
#pragma unroll
      for (int q = 0; q < 3; q++) {
        int lx, ly;
        const Pair pr = tile_pair(lane + 64 * q, ox, oy, nx, ny, &lx, &ly);
        const F2 t = ld2<SX>(src, pr.c);
        const uint16_t m2 = ld2b(mask, pr.c);
        v[2 * q] = pr.in ? t.x : 0.f;
        v[2 * q + 1] = pr.in ? t.y : 0.f;
        m[2 * q] = pr.in ? (uint8_t)(m2 & 0xff) : 0;
        m[2 * q + 1] = pr.in ? (uint8_t)(m2 >> 8) : 0;
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int e = lane + 64 * q;
      const int ly = e / TS, lx = e - ly * TS;
      const int i = ox - 1 + lx, j = oy - 1 + ly;
      const bool in = e < TCELLS && i >= 0 && j >= 0 && i < nx && j < ny;
      const size_t c = in ? (size_t)i + (size_t)j * nx : 0;
      v[q] = ldf_if<SX>(src, c, in);
      m[q] = ld_if<uint8_t>(mask, c, in, 0);
    }
  }
  __device__ void put(float* T, const float* CP, int PX, int PY, int ox, int oy, int nx, int ny, int lane) const {
    if (pairs) {
#pragma unroll
      for (int q = 0; q < 3; q++) {
        int lx, ly;
        (void)tile_pair(lane + 64 * q, ox, oy, nx, ny, &lx, &ly);
        const int i = ox - 1 + lx, j = oy - 1 + ly;
        // the columns outside the tile (dropped by tile_put) take no addend: their coarse
        // neighbours lie outside the patch
        const uint8_t m0 = lx >= 0 ? m[2 * q] : 0, m1 = lx + 1 < TS ? m[2 * q + 1] : 0;
        tile_put(T, lane + 64 * q, lx, ly, snap_cell(v[2 * q], m0, i, j, CP, PX, PY),
                 snap_cell(v[2 * q + 1], m1, i + 1, j, CP, PX, PY));
      }
      return;
    }
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int e = lane + 64 * q;
      if (e < TCELLS) {
        const int ly = e / TS, lx = e - ly * TS;
        T[lx + ly * TP] = snap_cell(v[q], m[q], ox - 1 + lx, oy - 1 + ly, CP, PX, PY);
      }
    }
  }
};

// ================================================================== up on level 0 (unfused)
// Level 0 is large enough that recomputing pass C four times costs more than a kernel
// boundary, so its up path stays two kernels: prolongation + pass C, then pass D writing z.
// prolongar (pressao-cpu.js:191-214) folded into the snapshot of the first up pass (offset 8,
// colours 1,0): water cells read x + P e, with the coarse values of the tile's 10x10 patch
template <class SX, class SB, class SC>
__device__ __forceinline__ void k_prolong_pass(Blk B_, const SX* __restrict__ src, SX* __restrict__ dst,
                                               const SB* __restrict__ b, const uint8_t* __restrict__ mask,
                                               int nx, int ny, int ntx, int nty, const SC* __restrict__ cx,
                                               const uint8_t* __restrict__ cmk, int cnx, int cny,
                                               const DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float Tt[4][TLDS];
  __shared__ float CPt[4][100];
  const int lane = threadIdx.x & 63;
  int tx, ty;
  block_tile(B_, &tx, &ty);
  if (tx >= ntx || ty >= nty) return;
  const int ox = tx * 16 - 8, oy = ty * 16 - 8;
  float* T = Tt[threadIdx.x >> 6];
  float* CP = CPt[threadIdx.x >> 6];
  const int PX = (ox >> 1) - 1, PY = (oy >> 1) - 1;
  // all global loads first: own cells, the coarse patch, the snapshot with its masks
  Own4 own;
  prefetch_cells<SB>(own, b, mask, ox, oy, nx, ny, lane);
  float cv[2];
  uint8_t cm[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = lane + 64 * q;
    const int bj = e / 10, ai = e - bj * 10;
    const int I = PX + ai, J = PY + bj;
    const bool in = e < 100 && I >= 0 && J >= 0 && I < cnx && J < cny;
    const size_t C = in ? (size_t)I + (size_t)J * cnx : 0;
    cm[q] = ld_if<uint8_t>(cmk, C, in, 0);
    cv[q] = ldf_if<SC>(cx, C, in);
  }
  Snap<SX> sn;
  sn.load(src, mask, ox, oy, nx, ny, lane);
  if (sc_done) return;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = lane + 64 * q;
    if (e < 100) CP[e] = mask_type(cm[q]) == AGUA ? cv[q] : 0.f;
  }
  wave_sync();
  sn.put(T, CP, PX, PY, ox, oy, nx, ny, lane);
  wave_sync();
  Cells4 cc;
  init_cells(cc, own, T, lane);
  tile_sweeps<1>(T, cc, lane);
  store_cells<SX>(cc, dst, ox, oy, nx, ny, lane);
}

// last up pass of level 0 (offset 0, colours 1,0) writing z and the dot partials
template <class SB>
__device__ __forceinline__ void k_pass_z(Blk B_, const float* __restrict__ src, const SB* __restrict__ b,
                                         const uint8_t* __restrict__ mask, int nx, int ny, int ntx, int nty,
                                         ZArgs za, const DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float Tt[4][TLDS];
  __shared__ double sh[FOLD_SH];
  __shared__ BubLds aggt;
  const int lane = threadIdx.x & 63;
  const int ns = nslots(sc->active_bubbles);
  const BubAgg agg = bagg_begin(aggt, za.accZ, sc->active_bubbles > BSLOTS);
  if (agg.t) __syncthreads();
  int tx, ty;
  block_tile(B_, &tx, &ty);
  double dv[2] = {0, 0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  // every wave loads (a wave past the grid's tiles reads in-bounds cells it never uses), so the
  // stop test below is uniform over the block
  const bool valid = tx < ntx && ty < nty;
  const int ox = tx * 16, oy = ty * 16;
  float* T = Tt[threadIdx.x >> 6];
  Own4 own;
  prefetch_cells<SB>(own, b, mask, ox, oy, nx, ny, lane);
  const float inv = za.scaled ? 1.f / *za.mgs : 1.f;
  uint8_t om[2][2];
  float rv[2][2], zo[2][2];
  size_t cz[2][2];
  const bool pairs = (nx & 1) == 0;  // a lane's two cells of a row move as one access
  const int c0 = pair_c0(lane);
  if (pairs) {
#pragma unroll
    for (int rw = 0; rw < 2; rw++) {
      const Pair pr = cell_pair(lane, rw, ox, oy, nx, ny);
      by_colour(cz, rw, c0, pr.c, pr.in ? pr.c + 1 : (size_t)0);
      const uint16_t o2 = ld2b(za.opMask, pr.c);
      const F2 r2 = ld2<float>(za.r, pr.c), z2 = ld2<float>(za.zold, pr.c);
      by_colour(om, rw, c0, (uint8_t)(pr.in ? o2 & 0xff : 0), (uint8_t)(pr.in ? o2 >> 8 : 0));
      by_colour(rv, rw, c0, r2.x, r2.y);
      by_colour(zo, rw, c0, z2.x, z2.y);
    }
  } else {
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        const bool in = i < nx && j < ny;
        const size_t c = in ? (size_t)i + (size_t)j * nx : 0;
        cz[cq][rw] = c;
        om[cq][rw] = ld_if<uint8_t>(za.opMask, c, in, 0);
        rv[cq][rw] = za.r[c];
        zo[cq][rw] = za.zold[c];
      }
  }
  fill_tile<float>(T, src, ox, oy, nx, ny, lane);
  if (sc_done) return;
  if (valid) {
    // the bubble cells' labels, slots and cell counts load while the tile is smoothed
    ZBubCells zb;
    zbub_labels(zb, za, om, cz);
    zbub_vals(zb, za);
    wave_sync();
    Cells4 cc;
    init_cells(cc, own, T, lane);
    tile_sweeps<1>(T, cc, lane);
    float zz[2][2];
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        zz[cq][rw] = 0.f;
        if (i < nx && j < ny) {
          zz[cq][rw] = z_val(om[cq][rw], rv[cq][rw], zo[cq][rw], (cc.m[cq][rw] & 32) ? cc.v[cq][rw] : 0.f, inv, dv,
                             sl, agg, zb.lb[cq][rw], zb.sl[cq][rw], zb.cb[cq][rw]);
          if (!pairs) za.znew[(size_t)i + (size_t)j * nx] = zz[cq][rw];
        }
      }
    if (pairs) {
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const Pair pr = cell_pair(lane, rw, ox, oy, nx, ny);
        if (pr.in) st2<float>(za.znew, pr.c, c0 ? zz[1][rw] : zz[0][rw], c0 ? zz[0][rw] : zz[1][rw]);
      }
    }
  }
  bagg_flush(agg);
  store_partials_last<2, BSLOTS>(B_, dv, 2, za.partials, sl, ns, za.bpart, sh);
}

// ================================================================== large levels: one kernel per pass
// Above SPLIT_CELLS cells the fused down/up kernels' recomputed halo tiles (9/4 of pass B, 4x
// pass C) cost more than a kernel boundary, so those levels run each pass once: pass B, then
// restriction + the coarse pass A; prolongation + pass C, then pass D.
constexpr int SPLIT_CELLS = 1 << 20;
constexpr int PASS_FOLD_GRID = 2048;  // 8 blocks per CU at k_pass's register count
static int split_cells(const tq_sim* s) { return s->launch_split_cells > 0 ? s->launch_split_cells : SPLIT_CELLS; }
static int pass_fold_grid(const tq_sim* s) {
  return s->launch_pass_fold_grid > 0 ? s->launch_pass_fold_grid : PASS_FOLD_GRID;
}

// one smoothing pass of level l (offset desl, colour order from C0); on level 0 of an iteration
// it first does the norm and stop test (fold_update)
template <class SX, class SB, int C0>
__device__ __forceinline__ void k_pass(Blk B_, const SX* __restrict__ src, SX* __restrict__ dst,
                                       const SB* __restrict__ b, const uint8_t* __restrict__ mask, int nx,
                                       int ny, int desl, int ntx, int nty, int vgx, int nvb,
                                       const double* __restrict__ partRR, int nbu, int par, int fold,
                                       DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float Tt[4][TLDS];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  float* T = Tt[w];
  // one-dimensional grid over the tile groups (vgx per row, nvb in all); a folding launch
  // is capped and its blocks walk several groups after one fold
  int vb = xcd_block(B_.x, B_.gx);
  bool valid;
  int ox, oy;
  Own4 own;
  auto load = [&](int g) {
    const int by = g / vgx, bx = g - by * vgx;
    int tx, ty;
    group_tile(bx, by, w, &tx, &ty);
    valid = tx < ntx && ty < nty;
    ox = tx * 16 - desl;
    oy = ty * 16 - desl;
    if (valid) {
      prefetch_cells<SB>(own, b, mask, ox, oy, nx, ny, lane);
      fill_tile<SX, false>(T, src, ox, oy, nx, ny, lane);
    }
  };
  load(vb);
  if (sc_done) return;
  if (fold) {
    __shared__ double fsh[16], fres[1];
    if (fold_update(B_, partRR, nbu, par, fsh, fres, sc)) return;
  }
  for (;;) {
    if (valid) {
      wave_sync();
      Cells4 cc;
      init_cells(cc, own, T, lane);
      tile_sweeps<C0>(T, cc, lane);
      store_cells<SX>(cc, dst, ox, oy, nx, ny, lane);
      wave_sync();  // the wave's tile is free for its next group
    }
    vb += B_.gx;
    if (vb >= nvb) break;
    load(vb);
  }
}

// restringir (pressao-cpu.js:159-188): the block owns a 16x16 coarse tile; fine residuals
// r = b - A x_B on the 34x34 fine cells it gathers from (2I-1 .. 2I+2 per axis, weights
// 1/4 3/4 3/4 1/4), from x_B on 36x36, in LDS.  Then wave 0 runs the coarse level's pass A
// (offset 0 = this tile) from x = 0 unless the coarse level is solved otherwise.
template <class SXF, class SBF, class SCB, class SCX, bool PASSA>
__device__ __forceinline__ void k_restrict_a(Blk B_, const SXF* __restrict__ fx, const SBF* __restrict__ fb,
                                             const uint8_t* __restrict__ fm, int fnx, int fny,
                                             SCB* __restrict__ cb, SCX* __restrict__ cxo,
                                             const uint8_t* __restrict__ cmk, int cnx, int cny,
                                             const DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float X[36 * 36], R[34 * 34], Bt[256], T[TLDS];
  const int tid = threadIdx.x;
  int bxl, byl;
  xcd_block2(B_, &bxl, &byl);
  const int CI0 = bxl * 16, CJ0 = byl * 16;
  const int fx0 = 2 * CI0 - 2, fy0 = 2 * CJ0 - 2;  // X origin
  // every global load first: the residual cells' masks and b, the coarse masks, then x_B
  uint8_t mr[5];
  float br[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int e = tid + 256 * q;
    const int ly = e / 34, lx = e - ly * 34;
    const int i = fx0 + 1 + lx, j = fy0 + 1 + ly;
    const bool in = e < 34 * 34 && i >= 0 && j >= 0 && i < fnx && j < fny;
    const size_t c = in ? (size_t)i + (size_t)j * fnx : 0;
    mr[q] = ld_if<uint8_t>(fm, c, in, 0);
    br[q] = ldf_if<SBF>(fb, c, in);
  }
  const int Im = CI0 + (tid & 15), Jm = CJ0 + (tid >> 4);
  const bool cin = Im < cnx && Jm < cny;
  const size_t Cm = cin ? (size_t)Im + (size_t)Jm * cnx : 0;
  const uint8_t cmine = ld_if<uint8_t>(cmk, Cm, cin, 0);
  uint8_t cma[2][2];
#pragma unroll
  for (int cq = 0; cq < 2; cq++)
#pragma unroll
    for (int rw = 0; rw < 2; rw++) {
      const int I = CI0 + cell_lx(tid & 63, cq), J = CJ0 + cell_ly(tid & 63, rw);
      cma[cq][rw] = ld_if<uint8_t>(cmk, (size_t)I + (size_t)J * cnx, PASSA && tid < 64 && I < cnx && J < cny, 0);
    }
  float xv[6];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int e = tid + 256 * q;
    const int ly = e / 36, lx = e - ly * 36;
    const int i = fx0 + lx, j = fy0 + ly;
    xv[q] = ldf_if<SXF>(fx, (size_t)i + (size_t)j * fnx, e < 36 * 36 && i >= 0 && j >= 0 && i < fnx && j < fny);
  }
  if (sc_done) return;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int e = tid + 256 * q;
    if (e < 36 * 36) X[e] = xv[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const int e = tid + 256 * q;
    if (e < 34 * 34) {
      const int ly = e / 34, lx = e - ly * 34;
      const uint8_t m = mr[q];
      float rr = 0.f;
      if (mask_type(m) == AGUA) {
        const int ex = (lx + 1) + (ly + 1) * 36;
        float sn = (m & 1) ? X[ex - 1] : 0.f;
        sn += (m & 2) ? X[ex + 1] : 0.f;
        sn += (m & 4) ? X[ex - 36] : 0.f;
        sn += (m & 8) ? X[ex + 36] : 0.f;
        rr = br[q] - ((float)mask_cnt(m) * X[ex] - sn);
      }
      R[e] = rr;
    }
  }
  __syncthreads();
  {
    float bc = 0.f;
    if (cin) {
      if (mask_type(cmine) == AGUA) {
        const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        // fine cell 2I-1+bi sits at R index (2I-1+bi) - (fx0+1) = 2(I-CI0) + bi
        const int rx = 2 * (tid & 15), ry = 2 * (tid >> 4);
        float acc = 0.f;
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
          float racc = 0.f;
#pragma unroll
          for (int bi = 0; bi < 4; bi++) racc += R[(rx + bi) + (ry + bj) * 34] * w[bi];
          acc += racc * w[bj];
        }
        bc = rnd<SCB>(acc);
      }
      st<SCB>(cb, Cm, bc);
    }
    Bt[tid] = bc;
  }
  if (!PASSA) return;
  __syncthreads();
  if (tid >= 64) return;
  const int lane = tid;
  zero_tile(T, lane);
  Cells4 cc;
#pragma unroll
  for (int cq = 0; cq < 2; cq++)
#pragma unroll
    for (int rw = 0; rw < 2; rw++) {
      const int li = cell_lx(lane, cq), lj = cell_ly(lane, rw);
      const int I = CI0 + li, J = CJ0 + lj;
      cc.v[cq][rw] = 0.f;
      if (I < cnx && J < cny) cell_setup(cc, cq, rw, cma[cq][rw], Bt[li + 16 * lj]);
      else cell_clear(cc, cq, rw);
    }
  wave_sync();
  tile_sweeps<0>(T, cc, lane);
  store_cells<SCX>(cc, cxo, CI0, CJ0, cnx, cny, lane);
}

// ================================================================== down: pass B + restriction + pass A
// One block owns a 16x16 tile of the coarse level l+1.  Its restriction (restringir,
// pressao-cpu.js:159-188) reads the fine residual r = b - A x_B on the 34x34 fine cells
// 2I-1 .. 2I+2 around it, i.e. x_B on a 36x36 window, which nine of level l's pass-B tiles
// (offset 8, suavizar pressao-cpu.js:99-150) cover: nine wavefronts compute them (each B tile is
// recomputed by the blocks whose windows overlap it, cheap next to a kernel boundary) into a
// 48x48 LDS window.  The block writes x_B of its own 32x32 fine cells (the up path needs it),
// then the coarse b, then wave 0 runs the coarse level's pass A (offset 0 = this tile) from x = 0.
constexpr int BRA_WAVES = 9;
constexpr int XBW = 48;  // x_B window: fine cells 2*CI0 - 8 .. 2*CI0 + 39

template <class SXF, class SBF, class SCB, class SCX, bool PASSA>
__device__ __forceinline__ void k_bra(Blk B_, const SXF* __restrict__ fx, SXF* __restrict__ fxb,
                                      const SBF* __restrict__ fb, const uint8_t* __restrict__ fm,
                                      int fnx, int fny, SCB* __restrict__ cb, SCX* __restrict__ cxo,
                                      const uint8_t* __restrict__ cmk, int cnx, int cny,
                                      const double* __restrict__ partRR, int nbu, int par, int fold,
                                      DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float Tt[BRA_WAVES][TLDS];
  __shared__ float XB[XBW * XBW];
  __shared__ float R[34 * 34];
  __shared__ float Bt[256];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int bxl, byl;
  xcd_block2(B_, &bxl, &byl);
  const int CI0 = bxl * 16, CJ0 = byl * 16;
  const int wx0 = 2 * CI0 - 8, wy0 = 2 * CJ0 - 8;   // window origin (fine)
  const int ox = wx0 + 16 * (wave % 3), oy = wy0 + 16 * (wave / 3);  // this wave's B tile
  float* T = Tt[wave];
  // ---- every global load first: the B tile's cells, the residual cells' masks and b, the
  // coarse masks (this thread's coarse cell, wave 0's pass-A cells), the B tile's snapshot
  Own4 own;
  prefetch_cells<SBF>(own, fb, fm, ox, oy, fnx, fny, lane);
  uint8_t mr[3];
  float br[3];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int e = tid + 64 * BRA_WAVES * q;
    const int ly = e / 34, lx = e - ly * 34;
    const int i = 2 * CI0 - 1 + lx, j = 2 * CJ0 - 1 + ly;
    const bool in = e < 34 * 34 && i >= 0 && j >= 0 && i < fnx && j < fny;
    const size_t c = in ? (size_t)i + (size_t)j * fnx : 0;
    mr[q] = ld_if<uint8_t>(fm, c, in, 0);
    br[q] = ldf_if<SBF>(fb, c, in);
  }
  const int Im = CI0 + (tid & 15), Jm = CJ0 + ((tid >> 4) & 15);
  const bool cin = tid < 256 && Im < cnx && Jm < cny;
  const size_t Cm = cin ? (size_t)Im + (size_t)Jm * cnx : 0;
  const uint8_t cmine = ld_if<uint8_t>(cmk, Cm, cin, 0);
  uint8_t cma[2][2];
#pragma unroll
  for (int cq = 0; cq < 2; cq++)
#pragma unroll
    for (int rw = 0; rw < 2; rw++) {
      const int I = CI0 + cell_lx(lane, cq), J = CJ0 + cell_ly(lane, rw);
      cma[cq][rw] = ld_if<uint8_t>(cmk, (size_t)I + (size_t)J * cnx, PASSA && wave == 0 && I < cnx && J < cny, 0);
    }
  fill_tile<SXF>(T, fx, ox, oy, fnx, fny, lane);
  if (sc_done) return;
  if (fold) {  // level 0 of an iteration: the norm and stop test (fold_update) while the loads land
    __shared__ double fsh[16], fres[1];
    if (fold_update(B_, partRR, nbu, par, fsh, fres, sc)) return;
  }
  wave_sync();
  // ---- pass B on this wave's tile (colours 0,1), into the window and (own cells) to memory
  {
    Cells4 cc;
    init_cells(cc, own, T, lane);
    tile_sweeps<0>(T, cc, lane);
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        const bool in = i >= 0 && j >= 0 && i < fnx && j < fny;
        const float v = (in && (cc.m[cq][rw] & 32)) ? rnd<SXF>(cc.v[cq][rw]) : 0.f;
        XB[(i - wx0) + (j - wy0) * XBW] = v;
        if (in && i >= 2 * CI0 && i < 2 * CI0 + 32 && j >= 2 * CJ0 && j < 2 * CJ0 + 32)
          st<SXF>(fxb, (size_t)i + (size_t)j * fnx, v);
      }
  }
  __syncthreads();
  // ---- fine residual on the 34x34 cells (R index (lx, ly) = fine (2CI0-1+lx, 2CJ0-1+ly))
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int e = tid + 64 * BRA_WAVES * q;
    if (e < 34 * 34) {
      const int ly = e / 34, lx = e - ly * 34;
      const uint8_t m = mr[q];
      float rr = 0.f;
      if (mask_type(m) == AGUA) {
        const int ex = (lx + 7) + (ly + 7) * XBW;
        float sn = (m & 1) ? XB[ex - 1] : 0.f;
        sn += (m & 2) ? XB[ex + 1] : 0.f;
        sn += (m & 4) ? XB[ex - XBW] : 0.f;
        sn += (m & 8) ? XB[ex + XBW] : 0.f;
        rr = br[q] - ((float)mask_cnt(m) * XB[ex] - sn);
      }
      R[e] = rr;
    }
  }
  __syncthreads();
  // ---- coarse b = P^T r (weights 1/4 3/4 3/4 1/4 per axis)
  if (tid < 256) {
    float bc = 0.f;
    if (cin) {
      if (mask_type(cmine) == AGUA) {
        const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        const int rx = 2 * (tid & 15), ry = 2 * (tid >> 4);
        float acc = 0.f;
#pragma unroll
        for (int bj = 0; bj < 4; bj++) {
          float racc = 0.f;
#pragma unroll
          for (int bi = 0; bi < 4; bi++) racc += R[(rx + bi) + (ry + bj) * 34] * w[bi];
          acc += racc * w[bj];
        }
        bc = rnd<SCB>(acc);
      }
      st<SCB>(cb, Cm, bc);
    }
    Bt[tid] = bc;
  }
  if (!PASSA) return;
  __syncthreads();
  if (wave != 0) return;
  // ---- pass A of the coarse level on this tile, from x = 0 (cicloV, pressao-cpu.js:243)
  zero_tile(T, lane);
  Cells4 cc;
#pragma unroll
  for (int cq = 0; cq < 2; cq++)
#pragma unroll
    for (int rw = 0; rw < 2; rw++) {
      const int li = cell_lx(lane, cq), lj = cell_ly(lane, rw);
      const int I = CI0 + li, J = CJ0 + lj;
      cc.v[cq][rw] = 0.f;
      if (I < cnx && J < cny) cell_setup(cc, cq, rw, cma[cq][rw], Bt[li + 16 * lj]);
      else cell_clear(cc, cq, rw);
    }
  wave_sync();
  tile_sweeps<0>(T, cc, lane);
  store_cells<SCX>(cc, cxo, CI0, CJ0, cnx, cny, lane);
}

// ================================================================== up: prolongation + pass C + pass D
// One block owns a 16x16 tile T of level l at offset 0.  Pass D (offset 0, colours 1,0) on T
// reads pass C's output on T's 18x18 window, which the four offset-8 tiles overlapping T cover:
// four wavefronts run pass C (colours 1,0) on them from a snapshot with the prolongation folded
// in (x_B + P e on water cells, prolongar pressao-cpu.js:191-214; coarse values from a 10x10 patch
// per tile) into a 32x32 LDS window; then wave 0 runs pass D on T.  The result goes to `xout`
// (a different array from the x_B the neighbouring blocks read), or with ZD (level 0) becomes z
// and the r.z, r.z_old partials.
constexpr int CW = 32;  // pass-C window: cells 16a - 8 .. 16a + 23

template <class SX, class SB, class SC, bool ZD>
__device__ __forceinline__ void k_pcd(Blk B_, const SX* __restrict__ xb, SX* __restrict__ xout,
                                      const SB* __restrict__ b, const uint8_t* __restrict__ mask, int nx,
                                      int ny, const SC* __restrict__ cx, const uint8_t* __restrict__ cmk,
                                      int cnx, int cny, ZArgs za, const DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ float Tt[4][TLDS];
  __shared__ float CPt[4][100];
  __shared__ float W[CW * CW];
  __shared__ double sh[FOLD_SH];
  __shared__ BubLds aggt;
  // (the barrier before pass D orders the table's clearing before the adds)
  const BubAgg agg = bagg_begin(aggt, za.accZ, ZD && sc->active_bubbles > BSLOTS);
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  int bxl, byl;
  xcd_block2(B_, &bxl, &byl);
  const int ax = bxl * 16, ay = byl * 16;  // tile T
  const int wx0 = ax - 8, wy0 = ay - 8;                      // window origin
  const int ox = wx0 + 16 * (wave & 1), oy = wy0 + 16 * (wave >> 1);  // this wave's C tile
  float* T = Tt[wave];
  float* CP = CPt[wave];
  const int PX = (ox >> 1) - 1, PY = (oy >> 1) - 1;
  // ---- every global load first
  Own4 own, ownD;
  prefetch_cells<SB>(own, b, mask, ox, oy, nx, ny, lane);
  if (wave == 0) prefetch_cells<SB>(ownD, b, mask, ax, ay, nx, ny, lane);
  uint8_t om[2][2] = {{0, 0}, {0, 0}};
  float rv[2][2] = {{0, 0}, {0, 0}}, zo[2][2] = {{0, 0}, {0, 0}};
  size_t cz[2][2] = {{0, 0}, {0, 0}};
  const float inv = (ZD && za.scaled) ? 1.f / *za.mgs : 1.f;
  if (ZD && wave == 0) {
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ax + cell_lx(lane, cq), j = ay + cell_ly(lane, rw);
        const bool in = i < nx && j < ny;
        const size_t c = in ? (size_t)i + (size_t)j * nx : 0;
        cz[cq][rw] = c;
        om[cq][rw] = ld_if<uint8_t>(za.opMask, c, in, 0);
        rv[cq][rw] = za.r[c];
        zo[cq][rw] = za.zold[c];
      }
  }
  float cv[2];
  uint8_t cm[2];
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = lane + 64 * q;
    const int bj = e / 10, ai = e - bj * 10;
    const int I = PX + ai, J = PY + bj;
    const bool in = e < 100 && I >= 0 && J >= 0 && I < cnx && J < cny;
    const size_t C = in ? (size_t)I + (size_t)J * cnx : 0;
    cm[q] = ld_if<uint8_t>(cmk, C, in, 0);
    cv[q] = ldf_if<SC>(cx, C, in);
  }
  Snap<SX> sn;
  sn.load(xb, mask, ox, oy, nx, ny, lane);
  if (sc_done) return;
#pragma unroll
  for (int q = 0; q < 2; q++) {
    const int e = lane + 64 * q;
    if (e < 100) CP[e] = mask_type(cm[q]) == AGUA ? cv[q] : 0.f;
  }
  // z's bubble cells: labels now (the masks have arrived), slots and cell counts after pass C
  ZBubCells zb;
  if (ZD && wave == 0) zbub_labels(zb, za, om, cz);
  wave_sync();
  // ---- snapshot x_B + P e, pass C on this wave's tile, into the window
  sn.put(T, CP, PX, PY, ox, oy, nx, ny, lane);
  wave_sync();
  {
    Cells4 cc;
    init_cells(cc, own, T, lane);
    tile_sweeps<1>(T, cc, lane);
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        const bool in = i >= 0 && j >= 0 && i < nx && j < ny;
        W[(i - wx0) + (j - wy0) * CW] = (in && (cc.m[cq][rw] & 32)) ? rnd<SX>(cc.v[cq][rw]) : 0.f;
      }
  }
  if (ZD && wave == 0) zbub_vals(zb, za);
  __syncthreads();
  double dv[2] = {0, 0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (wave == 0) {
    // ---- pass D on T from the window (T's 18x18 = window cells 7 .. 24)
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int e = lane + 64 * q;
      if (e < TCELLS) {
        const int ly = e / TS, lx = e - ly * TS;
        T[lx + ly * TP] = W[(lx + 7) + (ly + 7) * CW];
      }
    }
    wave_sync();
    Cells4 cc;
    init_cells(cc, ownD, T, lane);
    tile_sweeps<1>(T, cc, lane);
    if (!ZD) {
      store_cells<SX>(cc, xout, ax, ay, nx, ny, lane);
    } else {
#pragma unroll
      for (int cq = 0; cq < 2; cq++)
#pragma unroll
        for (int rw = 0; rw < 2; rw++) {
          const int i = ax + cell_lx(lane, cq), j = ay + cell_ly(lane, rw);
          if (i < nx && j < ny)
            z_cell(za, (size_t)i + (size_t)j * nx, om[cq][rw], rv[cq][rw], zo[cq][rw],
                   (cc.m[cq][rw] & 32) ? cc.v[cq][rw] : 0.f, inv, dv, sl, agg, zb.lb[cq][rw], zb.sl[cq][rw],
                   zb.cb[cq][rw]);
        }
    }
  }
  if (ZD) {
    const int ns = nslots(sc->active_bubbles);
    bagg_flush(agg);
    store_partials_last<2, BSLOTS>(B_, dv, 2, za.partials, sl, ns, za.bpart, sh);
  }
}

// ================================================================== coarsest-level operator
// resolverGrosso (pressao-cpu.js:217-238) runs 64 red-black sweeps from x = 0, so its result is
// x = G b for a matrix G fixed by the level's masks.  Column j is the result for b = e_j: one
// wavefront per column runs the sweeps on its own copy of the level (LDS), cells in registers.
// Columns of cells that are not water are zero (their b is never read).
__device__ __forceinline__ void k_coarsest_op(Blk B_, const uint8_t* __restrict__ mask, int nx, int n,
                                              float* __restrict__ G, const DevScalars* __restrict__ sc) {
  if (!sc->op_dirty) return;  // masks as at the last build: G is current
  __shared__ float x[DENSE_MAX];
  const int j = B_.x, lane = threadIdx.x;
  for (int c = lane; c < n; c += 64) x[c] = 0.f;
  if (mask_type(mask[j]) != AGUA) {
    for (int c = lane; c < n; c += 64) G[(size_t)c * n + j] = 0.f;
    return;
  }
  int mk[3];
  float ikk[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int c = lane + 64 * k;
    const uint8_t m = c < n ? mask[c] : 0;
    const int q = mask_cnt(m);
    const bool act = mask_type(m) == AGUA && q > 0;
    mk[k] = act ? (mask_viz(m) | 16 | ((m & MK_COLOUR) ? 32 : 0)) : 0;
    ikk[k] = act ? 1.f / (float)q : 0.f;
  }
  wave_sync();
  for (int k2 = 0; k2 < 64; k2++)
#pragma unroll
    for (int o = 0; o < 2; o++) {
      const int cor = (k2 & 1) ? 1 - o : o;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int c = lane + 64 * k;
        const int m = mk[k];
        if ((m & 16) && ((m >> 5) & 1) == cor) {
          float sum = (m & 1) ? x[c - 1] : 0.f;
          sum += (m & 2) ? x[c + 1] : 0.f;
          sum += (m & 4) ? x[c - nx] : 0.f;
          sum += (m & 8) ? x[c + nx] : 0.f;
          x[c] = (sum + (c == j ? 1.f : 0.f)) * ikk[k];
        }
      }
      wave_sync();
    }
  for (int c = lane; c < n; c += 64) G[(size_t)c * n + j] = x[c];
}

// ================================================================== coarse levels in one workgroup
struct WGPlan {
  int lc, L;
  const float* dense;  // coarsest-level operator G (n x n, row-major) or null
  int nx[12], ny[12];
  int offx[12], offt[12], offb[12], offm[12];  // byte offsets into the LDS pool
  int pool;                                     // bytes of the pool in use
  const uint8_t* mask[12];
};

// one smoothing pass of an LDS-resident level: S -> W, every wave owning whole tiles
template <int C0, int NW>
__device__ __forceinline__ void wg_pass(const float* S, float* W, const float* B, const uint8_t* M, int nx, int ny, int desl,
                               float* T) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ntx = (nx + desl + 15) >> 4, nty = (ny + desl + 15) >> 4;
  for (int t = wave; t < ntx * nty; t += NW) {
    const int ty = t / ntx, tx = t - ty * ntx;
    const int ox = tx * 16 - desl, oy = ty * 16 - desl;
#pragma unroll
    for (int q = 0; q < 6; q++) {
      const int e = lane + 64 * q;
      if (e < TCELLS) {
        const int ly = e / TS, lx = e - ly * TS;
        const int i = ox - 1 + lx, j = oy - 1 + ly;
        T[lx + ly * TP] = (i >= 0 && j >= 0 && i < nx && j < ny) ? S[i + j * nx] : 0.f;
      }
    }
    wave_sync();
    Cells4 cc;
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        if (i >= 0 && j >= 0 && i < nx && j < ny) cell_setup(cc, cq, rw, M[i + j * nx], B[i + j * nx]);
        else cell_clear(cc, cq, rw);
        cc.v[cq][rw] = T[cell_e(lane, cq, rw)];
      }
    tile_sweeps<C0>(T, cc, lane);
#pragma unroll
    for (int cq = 0; cq < 2; cq++)
#pragma unroll
      for (int rw = 0; rw < 2; rw++) {
        const int i = ox + cell_lx(lane, cq), j = oy + cell_ly(lane, rw);
        if (i >= 0 && j >= 0 && i < nx && j < ny) W[i + j * nx] = (cc.m[cq][rw] & 32) ? cc.v[cq][rw] : 0.f;
      }
    wave_sync();
  }
  __syncthreads();
}

// The V-cycle of levels lc..L-1 on LDS-resident arrays, NW waves per workgroup.  On entry every
// level's mask is in LDS, level lc holds b and x (x after level lc's pass A when `passA_done`,
// else zero) and the deeper levels hold b = x = 0.  Returns where level lc's result is.
template <int NW>
__device__ __forceinline__ const float* wg_vcycle(const WGPlan& P, char* pool, float* T, bool passA_done) {
  // level arrays live at plan offsets; bit l of `sw` says x and tmp of level l are swapped
  unsigned sw = 0;
  auto Xp = [&](int l) { return (float*)(pool + (((sw >> l) & 1) ? P.offt[l] : P.offx[l])); };
  auto Tp = [&](int l) { return (float*)(pool + (((sw >> l) & 1) ? P.offx[l] : P.offt[l])); };
  auto Bp = [&](int l) { return (float*)(pool + P.offb[l]); };
  auto Mp = [&](int l) { return (uint8_t*)(pool + P.offm[l]); };
  // ---- down
  for (int l = P.lc; l < P.L - 1; l++) {
    const int nx = P.nx[l], ny = P.ny[l];
    if (l > P.lc || !passA_done) {  // pass A
      wg_pass<0, NW>(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 0, T);
      sw ^= 1u << l;
    }
    wg_pass<0, NW>(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 8, T);
    sw ^= 1u << l;
    // restriction to l+1: fine residual into the free array, then b = P^T r, x = 0
    const float* X = Xp(l);
    const float* B = Bp(l);
    const uint8_t* M = Mp(l);
    float* R = Tp(l);
    for (int c = threadIdx.x; c < nx * ny; c += blockDim.x) {
      const uint8_t m = M[c];
      float rr = 0.f;
      if (mask_type(m) == AGUA) {
        float s = (m & 1) ? X[c - 1] : 0.f;
        s += (m & 2) ? X[c + 1] : 0.f;
        s += (m & 4) ? X[c - nx] : 0.f;
        s += (m & 8) ? X[c + nx] : 0.f;
        rr = B[c] - ((float)mask_cnt(m) * X[c] - s);
      }
      R[c] = rr;
    }
    __syncthreads();
    float* Bc = Bp(l + 1);
    float* Xc = Xp(l + 1);
    const uint8_t* Mc = Mp(l + 1);
    const int cnx = P.nx[l + 1], cny = P.ny[l + 1];
    for (int C = threadIdx.x; C < cnx * cny; C += blockDim.x) {
      const int J = C / cnx, I = C - J * cnx;
      float acc = 0.f;
      if (mask_type(Mc[C]) == AGUA) {
        const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
        for (int bj = 0; bj < 4; bj++) {
          const int j = 2 * J - 1 + bj;
          if (j < 0 || j >= ny) continue;
          float racc = 0.f;
          for (int bi = 0; bi < 4; bi++) {
            const int i = 2 * I - 1 + bi;
            if (i < 0 || i >= nx) continue;
            racc += R[i + j * nx] * w[bi];
          }
          acc += racc * w[bj];
        }
      }
      Bc[C] = acc;
      Xc[C] = 0.f;
    }
    __syncthreads();
  }
  // ---- coarsest: 64 red-black sweeps in place, colour order alternating (resolverGrosso,
  // pressao-cpu.js:217-238)
  {
    const int l = P.L - 1, nx = P.nx[l], n = nx * P.ny[l];
    float* x = Xp(l);
    const float* b = Bp(l);
    const uint8_t* M = Mp(l);
    if (P.dense) {
      // x = G b with G the 64-sweep operator of this solve (k_coarsest_op)
      const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
      for (int c = wave; c < n; c += NW) {
        float acc = 0.f;
        for (int j = lane; j < n; j += 64) acc += P.dense[(size_t)c * n + j] * b[j];
        acc = wave_sumf(acc);
        if (lane == 0) x[c] = acc;
      }
    } else if (n <= 4 * 64) {
      // one wavefront, each lane's (up to) four cells in registers; neighbours from LDS
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int mk[4];
        float bk[4], ikk[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int c = lane + 64 * k;
          const uint8_t m = c < n ? M[c] : 0;
          const int q = mask_cnt(m);
          const bool act = mask_type(m) == AGUA && q > 0;
          mk[k] = act ? (mask_viz(m) | 16 | ((m & MK_COLOUR) ? 32 : 0)) : 0;
          bk[k] = act ? b[c] : 0.f;
          ikk[k] = act ? 1.f / (float)q : 0.f;
        }
        for (int k2 = 0; k2 < 64; k2++)
#pragma unroll
          for (int o = 0; o < 2; o++) {
            const int cor = (k2 & 1) ? 1 - o : o;
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const int c = lane + 64 * k;
              const int m = mk[k];
              if ((m & 16) && ((m >> 5) & 1) == cor) {
                float sum = (m & 1) ? x[c - 1] : 0.f;
                sum += (m & 2) ? x[c + 1] : 0.f;
                sum += (m & 4) ? x[c - nx] : 0.f;
                sum += (m & 8) ? x[c + nx] : 0.f;
                x[c] = (sum + bk[k]) * ikk[k];
              }
            }
            wave_sync();
          }
      }
    } else {
      for (int k2 = 0; k2 < 64; k2++)
        for (int o = 0; o < 2; o++) {
          const int cor = (k2 & 1) ? 1 - o : o;
          for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const uint8_t m = M[c];
            const int q = mask_cnt(m);
            if (mask_type(m) != AGUA || ((m & MK_COLOUR) ? 1 : 0) != cor || !q) continue;
            float sum = (m & 1) ? x[c - 1] : 0.f;
            sum += (m & 2) ? x[c + 1] : 0.f;
            sum += (m & 4) ? x[c - nx] : 0.f;
            sum += (m & 8) ? x[c + nx] : 0.f;
            x[c] = (sum + b[c]) / (float)q;
          }
          __syncthreads();
        }
    }
    __syncthreads();
  }
  // ---- up
  for (int l = P.L - 2; l >= P.lc; l--) {
    const int nx = P.nx[l], ny = P.ny[l], cnx = P.nx[l + 1], cny = P.ny[l + 1];
    float* X = Xp(l);
    const uint8_t* M = Mp(l);
    const float* Xc = Xp(l + 1);
    const uint8_t* Mc = Mp(l + 1);
    for (int c = threadIdx.x; c < nx * ny; c += blockDim.x) {
      if (mask_type(M[c]) != AGUA) continue;
      const int j = c / nx, i = c - j * nx;
      const int I0 = (i & 1) == 0 ? (i >> 1) - 1 : (i - 1) >> 1;
      const int J0 = (j & 1) == 0 ? (j >> 1) - 1 : (j - 1) >> 1;
      const float wx0 = (i & 1) == 0 ? 0.25f : 0.75f, wy0 = (j & 1) == 0 ? 0.25f : 0.75f;
      float e = 0.f;
      for (int b = 0; b < 2; b++) {
        const int J = J0 + b;
        if (J < 0 || J >= cny) continue;
        const float wy = b ? 1.f - wy0 : wy0;
        for (int a = 0; a < 2; a++) {
          const int I = I0 + a;
          if (I < 0 || I >= cnx) continue;
          const int C = I + J * cnx;
          if (mask_type(Mc[C]) == AGUA) e += Xc[C] * (a ? 1.f - wx0 : wx0) * wy;
        }
      }
      X[c] += e;
    }
    __syncthreads();
    wg_pass<1, NW>(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 8, T);
    sw ^= 1u << l;
    wg_pass<1, NW>(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 0, T);
    sw ^= 1u << l;
  }
  return Xp(P.lc);
}

// copies each LDS level's mask in and zeroes its vectors
__device__ __forceinline__ void wg_load_masks(const WGPlan& P, char* pool) {
  for (int l = P.lc; l < P.L; l++) {
    const int n = P.nx[l] * P.ny[l];
    float* X = (float*)(pool + P.offx[l]);
    float* B = (float*)(pool + P.offb[l]);
    uint8_t* M = (uint8_t*)(pool + P.offm[l]);
    const uint8_t* gm = P.mask[l];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      M[c] = gm[c];
      X[c] = 0.f;
      B[c] = 0.f;
    }
  }
}

// Levels lc..L-1 from LDS in one workgroup.  Level lc arrives with b and (unless it is the
// coarsest) x after its pass A; the result of level lc's V-cycle is written to out_lc, which may
// be x_lc (every input is in LDS before the first output is stored).
__device__ __forceinline__ void k_coarse_wg(Blk B_, const WGPlan& P, const float* __restrict__ b_lc, const float* x_lc,
                                            float* out_lc, const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ __attribute__((aligned(16))) char pool[WG_POOL];
  __shared__ float tiles[WG_WAVES][TLDS];
  wg_load_masks(P, pool);
  const bool xin = P.lc < P.L - 1;
  const int n = P.nx[P.lc] * P.ny[P.lc];
  float* X0 = (float*)(pool + P.offx[P.lc]);
  float* B0 = (float*)(pool + P.offb[P.lc]);
  for (int c = threadIdx.x; c < n; c += blockDim.x) {
    X0[c] = xin ? x_lc[c] : 0.f;
    B0[c] = b_lc[c];
  }
  __syncthreads();
  const float* X = wg_vcycle<WG_WAVES>(P, pool, tiles[threadIdx.x >> 6], true);
  for (int c = threadIdx.x; c < n; c += blockDim.x) out_lc[c] = X[c];
}

// ---------------------------------------------------------- level lc's V-cycle as a dense matrix
// The V-cycle from x = 0 is a linear map of b, x = V b, with V symmetric (the up passes repeat
// the down passes in reverse colour order and prolongation is the transpose of restriction).
// Row j of C is V e_j, one workgroup per row; the product below then reads C by rows,
// (C b)_i = sum_j (V e_i)_j b_j = (V b)_i.  Rows of cells that are not water are zero: V
// neither reads b nor writes x there.
constexpr int VOP_WAVES = 4;
__device__ __forceinline__ void k_vcycle_op(Blk B_, const WGPlan& P, int ld, float* __restrict__ C,
                                            DevScalars* __restrict__ sc) {
  if (!sc->op_dirty) return;  // masks as at the last build: C is current
  extern __shared__ __attribute__((aligned(16))) char vpool[];
  const int lc = P.lc, n = P.nx[lc] * P.ny[lc], j = B_.x;
  float* row = C + (size_t)j * ld;
  if (mask_type(P.mask[lc][j]) != AGUA) {
    for (int c = threadIdx.x; c < ld; c += blockDim.x) row[c] = 0.f;
    return;
  }
  wg_load_masks(P, vpool);
  __syncthreads();
  if (threadIdx.x == 0) ((float*)(vpool + P.offb[lc]))[j] = 1.f;
  __syncthreads();
  float* T = (float*)(vpool + P.pool) + (threadIdx.x >> 6) * TLDS;
  const float* X = wg_vcycle<VOP_WAVES>(P, vpool, T, false);
  for (int c = threadIdx.x; c < ld; c += blockDim.x) row[c] = c < n ? X[c] : 0.f;
}

__device__ __forceinline__ void k_op_count(Blk B_, DevScalars* __restrict__ sc) {
  if (threadIdx.x == 0 && sc->op_dirty) sc->op_builds += 1;
}

// x = C b on level lc: VOP_ROWS rows per wavefront, each lane reading float4 columns
constexpr int VOP_ROWS = 2;
__device__ __forceinline__ void k_vcycle_apply(Blk B_, const float* __restrict__ C, int ld, int n,
                                               const float* __restrict__ b, float* __restrict__ x,
                                               const DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  const int lane = threadIdx.x & 63;
  const int r0 = (B_.x * 4 + (threadIdx.x >> 6)) * VOP_ROWS;
  if (r0 >= n) return;
  const int nk = ld >> 8;  // float4 columns per lane (ld is a multiple of 256, at most 2048)
  float4 cv[VOP_ROWS][8];
  float4 bv[8];
  // every load unconditional (columns beyond ld read column block 0 and are not summed), so all
  // of them are in flight together
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int c = 4 * lane + 256 * (k < nk ? k : 0);
#pragma unroll
    for (int r = 0; r < VOP_ROWS; r++) {
      const float4 v = *(const float4*)(C + (size_t)(r0 + r < n ? r0 + r : r0) * ld + c);
      cv[r][k] = r0 + r < n ? v : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    bv[k].x = ld_if<float>(b, c, c < n, 0.f);
    bv[k].y = ld_if<float>(b, c + 1, c + 1 < n, 0.f);
    bv[k].z = ld_if<float>(b, c + 2, c + 2 < n, 0.f);
    bv[k].w = ld_if<float>(b, c + 3, c + 3 < n, 0.f);
  }
  if (sc_done) return;
#pragma unroll
  for (int r = 0; r < VOP_ROWS; r++) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (k < nk) acc += cv[r][k].x * bv[k].x + cv[r][k].y * bv[k].y + cv[r][k].z * bv[k].z + cv[r][k].w * bv[k].w;
    acc = wave_sumf(acc);
    if (lane == 0 && r0 + r < n) x[r0 + r] = acc;
  }
}


// First kernel of iteration k: the z dots and beta (fold_zdots), then per cell p = z + beta p_old
// (bubble cells: their new z and old p), Ap and the p.Ap partials (produto / aplicar,
// pressao-cpu.js:84-93, 257-261).  A block owns a 64x16 tile of cells (4 per thread, one row
// segment of 64 per wavefront) and stages the tile with its one-cell halo in LDS: mask, z and
// p_old are read once per cell (1.16x the tile, not 5x), p is formed once per cell and the
// stencil reads it from LDS.  Large grids launch at most SPMV_GRID blocks (the ones that fit on
// the chip at once) and each block walks its tiles after one fold: the fold's chain of dependent
// loads and barriers is paid once per block, not once per tile.
constexpr int SPMV_TW = 64, SPMV_TH = 16;            // tile (cells)
constexpr int SPMV_HW = SPMV_TW + 2, SPMV_HN = SPMV_HW * (SPMV_TH + 2);  // with its halo: 66 x 18
constexpr int SPMV_LOADS = (SPMV_HN + 255) / 256;    // halo-tile cells per thread (5)
constexpr int SPMV_GRID = 1024;  // 4 blocks per CU: the kernel is held to 128 VGPRs (4 waves per SIMD)
static int spmv_tiles(int nx, int ny) { return ((nx + SPMV_TW - 1) / SPMV_TW) * ((ny + SPMV_TH - 1) / SPMV_TH); }
__device__ __forceinline__ void k_spmv(Blk B_, const uint8_t* __restrict__ opMask, const uint32_t* __restrict__ rot,
                                       const float* __restrict__ z, const float* __restrict__ pold,
                                       float* __restrict__ pnew, float* __restrict__ ap, int nx, int ny,
                                       BubArgs B, const double* __restrict__ partZ,
                                       const double* __restrict__ bpartZ, int nbz, int par, int first,
                                       double* __restrict__ partA, double* __restrict__ bpartA,
                                       DevScalars* __restrict__ sc) {
  // the stop flag is read beside the first loads (one memory latency for both); a stopped solve
  // returns once they are issued
  const int sc_done = sc->done;
  __shared__ double sh[FOLD_SH];
  __shared__ float P[SPMV_HN];
  __shared__ uint8_t M[SPMV_HN];
  __shared__ float pslt[BSLOTS];    // new p of the slotted bubbles
  __shared__ BubLds aggt;
  const int tid = threadIdx.x;
  const int ntx = (nx + SPMV_TW - 1) / SPMV_TW, nch = ntx * ((ny + SPMV_TH - 1) / SPMV_TH);
  int ch = xcd_block(B_.x, B_.gx);
  // every load of the halo tile first: it does not depend on beta, so its latency overlaps the
  // fold's reductions (cells outside the grid: mask 0, never a free neighbour)
  int i0 = 0, j0 = 0;
  int gidx[SPMV_LOADS];
  uint8_t mk[SPMV_LOADS];
  float zv[SPMV_LOADS], pv[SPMV_LOADS];
  auto load = [&](int chunk) {
    const int ty = chunk / ntx, tx = chunk - ty * ntx;
    i0 = tx * SPMV_TW;
    j0 = ty * SPMV_TH;
#pragma unroll
    for (int q = 0; q < SPMV_LOADS; q++) {
      const int e = tid + 256 * q;
      const int hy = e / SPMV_HW, hx = e - hy * SPMV_HW;
      const int i = i0 - 1 + hx, j = j0 - 1 + hy;
      const bool in = e < SPMV_HN && i >= 0 && j >= 0 && i < nx && j < ny;
      const int c = in ? i + j * nx : -1;
      gidx[q] = c;
      mk[q] = ld_if<uint8_t>(opMask, (size_t)c, in, 0);
      zv[q] = ld_if<float>(z, (size_t)c, in, 0.f);
      pv[q] = ld_if<float>(pold, (size_t)c, in, 0.f);
    }
  };
  load(ch);
  ZFold zf;
  zf_load(zf, partZ, bpartZ, nbz, B, sc->active_bubbles);
  if (sc_done) return;
  zf_load_bub(zf, B, par);
  double zs[BSLOTS];
  zf_reduce(zf, sh, zs);
  const double beta = zf_finish(zf, B_, par, first, B, sh, zs, pslt, sc);
  const int ns = nslots(zf.na);
  // (the loop's first barrier orders the table's clearing before the adds)
  const BubAgg agg = bagg_begin(aggt, bv(B.bvec, BV_ACCA0 + par), zf.na > BSLOTS);
  // p of a halo cell from its loaded z, p_old and mask (bubble cells: the bubble's new z and
  // old p, the slotted ones from pslt[])
  auto peff = [&](int q) -> float {
    const uint8_t md = mk[q];
    if (mask_type(md) == AGUA) return (float)((double)zv[q] + beta * (double)pv[q]);
    if (md & MK_BUB) {
      const uint32_t lb = rot[gidx[q]];
      const int sl = B.labelSlot[lb];
      if (sl >= 0) return pslt[sl];
      return (float)(fx_read(bv(B.bvec, BV_ACCZ0 + par), lb) + beta * bv(B.bvec, BV_P0 + 1 - par)[lb]);
    }
    return 0.f;
  };
  double v[1] = {0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int lx = tid & (SPMV_TW - 1), w = tid >> 6;
  for (;;) {
    __syncthreads();  // the previous tile's P and M are no longer read
#pragma unroll
    for (int q = 0; q < SPMV_LOADS; q++) {
      const int e = tid + 256 * q;
      if (e < SPMV_HN) {
        P[e] = gidx[q] >= 0 ? peff(q) : 0.f;
        M[e] = mk[q];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SPMV_TH / 4; r++) {
      const int ly = w + 4 * r;
      const int i = i0 + lx, j = j0 + ly;
      if (i < nx && j < ny) {
        const int c = i + j * nx;
        const int e = (lx + 1) + (ly + 1) * SPMV_HW;
        const uint8_t m = M[e];
        const float pc = P[e];
        pnew[c] = pc;
        if (mask_type(m) == AGUA) {
          const int ne[4] = {e - 1, e + 1, e - SPMV_HW, e + SPMV_HW};
          const int nc[4] = {c - 1, c + 1, c - nx, c + nx};
          double s = 0;
#pragma unroll
          for (int k = 0; k < 4; k++) {
            if (!(m & (1 << k))) continue;
            s += (double)P[ne[k]];
            if (M[ne[k]] & MK_BUBIN) {
              const uint32_t lb = rot[nc[k]];
              bub_add(sl, B.labelSlot, agg, lb, -(double)pc);
            }
          }
          const float a = (float)(mask_cnt(m) * (double)pc - s);
          ap[c] = a;
          v[0] += (double)pc * a;
        } else ap[c] = 0.f;
      }
    }
    ch += B_.gx;
    if (ch >= nch) break;
    load(ch);
  }
  bagg_flush(agg);
  store_partials_last<1, BSLOTS>(B_, v, 1, partA, sl, ns, bpartA, sh);
}

// the update fold as a kernel of its own, when level 0 has no down kernel (one-workgroup grids)
__device__ __forceinline__ void k_fold_update(Blk B_, const double* __restrict__ part, int nb, int par,
                                              DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16], res[1];
  (void)fold_update(B_, part, nb, par, sh, res, sc);
}

// ================================================================== end of the solve
__device__ __forceinline__ void k_pcg_final(Blk B_, const uint8_t* __restrict__ tipo, float* __restrict__ q, int C) {
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c < C && tipo[c] != AGUA) q[c] = 0.f;
}
__device__ __forceinline__ void k_pcg_final_b(Blk B_, const double* __restrict__ bvec, const int* __restrict__ activeList,
                                              float* __restrict__ yB32, DevScalars* __restrict__ sc) {
  if (threadIdx.x == 0) {
    const int seg = sc->itp[0] > sc->itp[1] ? sc->itp[0] : sc->itp[1];
    if (sc->restarts == 0) sc->iters_first = seg;
    const int it = sc->iters_base + seg;
    sc->iters = it;
    sc->total_iters += (unsigned long long)it;
  }
  for (int k = threadIdx.x; k < sc->active_bubbles; k += blockDim.x) {
    const int b = activeList[k];
    yB32[b] = (float)bv((double*)bvec, BV_Y)[b];
  }
}

// true residual ||b' - A[q;y]|| in binary64 (b' is the projected rhs that was solved), of the
// binary32 q or, with qlo, of the compensated iterate q + qlo.  gated: the stop check of the
// residual-terminated mode, appended to every chunk of iterations -- it runs once, after the
// segment's stop, and returns at once otherwise.
__device__ inline bool true_res_skip(const DevScalars* sc, int gated) { return gated && (!sc->done || sc->true_done); }

__device__ __forceinline__ void k_true_res(Blk B_, const uint8_t* __restrict__ opMask, const uint32_t* __restrict__ rot,
                                           const float* __restrict__ div, const float* __restrict__ q,
                                           const float* __restrict__ qlo, int nx, int C,
                                           double* __restrict__ bvec, const int* __restrict__ labelSlot,
                                           double* __restrict__ partials, double* __restrict__ bpart, int gated,
                                           const DevScalars* __restrict__ sc) {
  if (true_res_skip(sc, gated)) return;
  __shared__ double sh[FOLD_SH];
  __shared__ BubLds aggt;
  const int ns = nslots(sc->active_bubbles);
  const double* yb = bv(bvec, BV_Y);
  const BubAgg agg = bagg_begin(aggt, bv(bvec, BV_ACCI), sc->active_bubbles > BSLOTS);
  if (agg.t) __syncthreads();
  const double mean = sc->proj_mean;
  double v[1] = {0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto qe = [&](int d, uint8_t md) -> double {
    if (mask_type(md) == AGUA) return qlo ? (double)q[d] + (double)qlo[d] : (double)q[d];
    if (md & MK_BUB) return qlo ? yb[rot[d]] : (double)(float)yb[rot[d]];  // compensated: y in binary64
    return 0.0;
  };
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c < C) {
    const uint8_t m = opMask[c];
    if (mask_type(m) == AGUA) {
      const double qc = qe(c, m);
      double s = 0;
      const int nb[4] = {c - 1, c + 1, c - nx, c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        const uint8_t md = opMask[nb[k]];
        s += qe(nb[k], md);
        if (md & MK_BUBIN) bub_add(sl, labelSlot, agg, rot[nb[k]], -qc);
      }
      const double e = (-(double)div[c] - mean) - (mask_cnt(m) * qc - s);
      v[0] = e * e;
    }
  }
  bagg_flush(agg);
  store_partials_last<1, BSLOTS>(B_, v, 1, partials, sl, ns, bpart, sh);
}

__device__ __forceinline__ void k_fin_true_res(Blk B_, const uint32_t* __restrict__ facesB,
                                               double* __restrict__ bvec, const double* __restrict__ rhsB,
                                               const int* __restrict__ activeList,
                                               const int* __restrict__ labelSlot,
                                               const double* __restrict__ partials,
                                               const double* __restrict__ bpart, int nb, int gated,
                                               DevScalars* __restrict__ sc,
                                               double* __restrict__ out) {
  if (true_res_skip(sc, gated)) return;
  __shared__ double sh[BSLOTS * 16], res[BSLOTS], slotTot[BSLOTS];
  const int na = sc->active_bubbles;
  double tot[1];
  sum_partials<1>(partials, nb, 1, tot, sh, res);
  double stt[BSLOTS];
  sum_partials<BSLOTS>(bpart, nb, nslots(na), stt, sh, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < BSLOTS; k++) slotTot[k] = stt[k];
  __syncthreads();
  const double* yb = bv(bvec, BV_Y);
  double* acc0 = bv(bvec, BV_ACCI);
  double bacc[1] = {0};
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const int s = labelSlot[b];
    double accB;
    if (s >= 0) accB = slotTot[s];
    else { accB = fx_read(acc0, b); fx_clear(acc0, b); }
    const double e = (rhsB[b] - sc->proj_mean) - ((double)facesB[b] * yb[b] + accB);
    bacc[0] += e * e;
  }
  block_sum<1>(bacc, sh);
  if (threadIdx.x == 0) {
    const double t = sqrt(tot[0] + bacc[0]);
    if (out) *out = t;
    if (gated) {
      sc->true_norm = t;
      sc->true_done = 1;
    }
  }
}

// ================================================================== host side
static long wg_bytes(int n) { return 12L * n + ((n + 15) & ~15); }

// level lc's V-cycle is applied as the dense matrix of this solve
static bool vop_active(const tq_sim* s) { return s->vop_ld > 0 && dense_wanted(s); }

int solver_alloc(tq_sim* s) {
  s->lv.clear();
  {
    Level a;
    a.nx = s->nx; a.ny = s->ny;
    s->lv.push_back(a);
  }
  while (std::min(s->lv.back().nx, s->lv.back().ny) > 12) {
    Level a;
    a.nx = (s->lv.back().nx + 1) / 2;
    a.ny = (s->lv.back().ny + 1) / 2;
    s->lv.push_back(a);
  }
  if (s->lv.size() > 12) { set_error("grid too large: more than 12 multigrid levels"); return TQ_E_CAPACITY; }
  for (auto& L : s->lv) {
    L.n = L.nx * L.ny;
    TQ_HIP(hipMalloc(&L.mask, L.n));
    // sized for FP32 so the precision can switch at run time
    TQ_HIP(hipMalloc(&L.b, sizeof(float) * L.n));
    TQ_HIP(hipMalloc(&L.x, sizeof(float) * L.n));
    TQ_HIP(hipMalloc(&L.tmp, sizeof(float) * L.n));
  }
  // the one-workgroup suffix: every level from lc on must fit the LDS pool
  const int L = (int)s->lv.size();
  long bytes = 0;
  int fit = L;
  for (int l = L - 1; l >= 0; l--) {
    if (bytes + wg_bytes(s->lv[l].n) > WG_POOL) break;
    bytes += wg_bytes(s->lv[l].n);
    fit = l;
  }
  if (fit == L) {
    set_error("grid too elongated: the coarsest multigrid level does not fit the one-workgroup solver");
    return TQ_E_CAPACITY;
  }
  int lc = fit;
  if (s->d.mg_min_lds_level <= 0) {
    while (lc < L - 1 && s->lv[lc].n > WG_AUTO_CELLS) lc++;
  } else {
    lc = std::max(fit, std::min(s->d.mg_min_lds_level, L - 1));
  }
  s->lc = lc;
  s->dense = nullptr;
  if (s->lv.back().n <= DENSE_MAX) TQ_HIP(hipMalloc(&s->dense, sizeof(float) * DENSE_MAX * DENSE_MAX));
  s->op_cells = 0;
  for (int l = lc; l < L; l++) s->op_cells += s->lv[l].n;
  TQ_HIP(hipMalloc(&s->op_masks, s->op_cells));
  TQ_HIP(hipMemset(s->op_masks, 0xFF, s->op_cells));  // no mask byte is 0xFF: the first solve builds
  s->vop_was_active = false;
  // level lc's V-cycle as a matrix: lc above the coarsest, at most VOP_MAX cells, the
  // builder's LDS (levels lc.. plus one tile per wave) under 64 KiB
  s->vop = nullptr;
  s->vop_ld = 0;
  if (lc > 0 && lc < L - 1 && s->lv[lc].n <= VOP_MAX) {
    long pool = 0;
    for (int l = lc; l < L; l++) pool += wg_bytes(s->lv[l].n);
    if (pool + VOP_WAVES * TLDS * 4 <= 65536) {
      s->vop_ld = (s->lv[lc].n + 255) & ~255;
      TQ_HIP(hipMalloc(&s->vop, sizeof(float) * (size_t)s->vop_ld * s->lv[lc].n));
    }
  }
  TQ_HIP(hipMalloc(&s->bpart, sizeof(double) * 2 * BSLOTS * MAX_RED_BLOCKS));  // spmv's, then the z dots'
  TQ_HIP(hipMalloc(&s->bpart_r, sizeof(double) * 2 * BSLOTS * MAX_RED_BLOCKS));
  TQ_HIP(hipMalloc(&s->labelSlot, sizeof(int) * MAX_BOLHAS));
  return 0;
}

static void graph_drop(tq_sim* s) {
  if (s->graph_exec) (void)hipGraphExecDestroy((hipGraphExec_t)s->graph_exec);
  s->graph_exec = nullptr;
  s->graph_key = -1;
}

// a launch cap changed: the captured PCG iterations hold the old grids
void solver_launch_changed(tq_sim* s) { graph_drop(s); }

void solver_free(tq_sim* s) {
  graph_drop(s);
  for (auto& L : s->lv) {
    (void)hipFree(L.mask); (void)hipFree(L.b); (void)hipFree(L.x); (void)hipFree(L.tmp);
  }
  s->lv.clear();
  if (s->bpart) (void)hipFree(s->bpart);
  if (s->bpart_r) (void)hipFree(s->bpart_r);
  s->bpart_r = nullptr;
  if (s->labelSlot) (void)hipFree(s->labelSlot);
  if (s->dense) (void)hipFree(s->dense);
  s->dense = nullptr;
  if (s->vop) (void)hipFree(s->vop);
  s->vop = nullptr;
  s->vop_ld = 0;
  if (s->op_masks) (void)hipFree(s->op_masks);
  s->op_masks = nullptr;
  s->bpart = nullptr;
  s->labelSlot = nullptr;
}

// per-level storage for this solve: the mixed modes store b (TQ_PREC_MIXED: and x) of the
// multi-workgroup levels from mg_half_level on in FP16 (level 0: b only); small levels and the
// LDS levels stay FP32
static bool plan_precision(tq_sim* s) {
  const int pr = s->d.precision;
  const bool mixed = pr == TQ_PREC_MIXED || pr == TQ_PREC_MIXED_B || pr == TQ_PREC_MIXED_X;
  const bool half_x = pr == TQ_PREC_MIXED || pr == TQ_PREC_MIXED_X;
  const bool half_b = pr == TQ_PREC_MIXED || pr == TQ_PREC_MIXED_B;
  const int h0 = s->d.mg_half_level < 0 ? 1 : s->d.mg_half_level;
  bool any = false;
  for (int l = 0; l < (int)s->lv.size(); l++) {
    Level& L = s->lv[l];
    const bool half = mixed && l < s->lc && l >= h0 && L.n >= 4096;
    L.hb = half && half_b;
    L.hx = half && half_x && l > 0;
    any = any || L.hb || L.hx;
  }
  return any;
}

static int tiles_of(int n, int desl) { return (n + desl + 15) / 16; }
static dim3 tile_grid(int nx, int ny, int desl) {
  return dim3((tiles_of(nx, desl) + GW - 1) / GW, (tiles_of(ny, desl) + GH - 1) / GH);
}
static int blocks_of(int C) { return (C + 255) / 256; }

static double vbytes(bool h) { return h ? 2.0 : 4.0; }

// bubble-slot partials: spmv (and the solve's set-up) use the first half, the z dots the second
static double* bpart_a(tq_sim* s) { return s->bpart; }
static double* bpart_z(tq_sim* s) { return s->bpart + (size_t)BSLOTS * MAX_RED_BLOCKS; }
static BubArgs bubargs(tq_sim* s) {
  BubArgs B;
  B.bvec = s->bvec;
  B.activeList = s->activeList;
  B.labelSlot = s->labelSlot;
  B.facesB = s->facesB;
  return B;
}
// Parities of the launch in progress (iteration k = s->it_k; k = -1: the first cycle).  The
// cycle of iteration k scales its input by mgs[k & 1] and leaves its z sums for iteration k+1
// (accumulator and partials of parity (k+1) & 1).
static int it_par(const tq_sim* s) { return s->it_k & 1; }
static int z_par(const tq_sim* s) { return (s->it_k + 1) & 1; }

// The partials a folding consumer reads: the producer's own, or (many blocks) PRE_R sums of them
struct RedSrc {
  const double* part;
  const double* bpart;
  int nb;
};
static RedSrc red_src(tq_sim* s, int slot, int nv, int bhalf, int nb) {
  const double* src = s->partials + ps_off(slot);
  const double* bsrc = bhalf < 0 ? nullptr : s->bpart + (size_t)bhalf * BSLOTS * MAX_RED_BLOCKS;
  if (nb <= FOLD_DIRECT) return {src, bsrc, nb};
  double* dst = s->partials_r + ps_off(slot);
  double* bdst = bhalf < 0 ? nullptr : s->bpart_r + (size_t)bhalf * BSLOTS * MAX_RED_BLOCKS;
  int t = time_begin(s, "pcg_prereduce", 8.0 * nv * nb);
  launch<256, 0, k_prereduce>(s, s->st, dim3(PRE_R), dim3(256), 0, src, nb, nv, bsrc, dst, bdst,
                              (const DevScalars*)s->sc);
  time_end(s, t);
  return {dst, bdst, PRE_R};
}

static void launch_in_a0(tq_sim* s, bool upd) {
  Level& F = s->lv[0];
  const int ntx = tiles_of(F.nx, 0), nty = tiles_of(F.ny, 0);
  const dim3 tg = tile_grid(F.nx, F.ny, 0);
  const int vgx = (int)tg.x, nvb = (int)(tg.x * tg.y);
  const dim3 g(std::min(nvb, s->launch_update_grid > 0 ? s->launch_update_grid : IN_A0_GRID));
  const bool pa = s->lv.size() > 1;
  const float* pp = s->pbuf[s->pc];
  // bytes: opMask, mg0, b0 write, x0 write; update: q, p, r, Ap read + q, r write (and q_lo read
  // and written for the compensated iterate)
  const double by = (2 + vbytes(F.hb) + (pa ? 4 : 0) + (upd ? (s->qcomp ? 32 : 24) : 4)) * (double)F.n;
  const RedSrc rs = upd ? red_src(s, PS_SPMV, 1, 0, s->nb_spmv) : RedSrc{s->partials + ps_off(PS_SPMV), bpart_a(s), 0};
  int t = time_begin(s, upd ? "pcg_update_mg_in_l0" : "pcg_mg_in_l0", by);
  // the first cycle reads the bubble residuals of parity 1 (par 0 selects them)
  const int par = upd ? it_par(s) : 0;
  const BubArgs B = bubargs(s);
  const float* mgs = &s->sc->mgs[it_par(s)];
#define TQ_IN_A0(SB, U, P)                                                                                 \
  launch<256, 0, k_in_a0<SB, U, P>>(s, s->st, g, dim3(256), 0, s->opMask, F.mask, s->rot, s->cellsB, pp, s->ap, \
                                    s->q, s->qlo, s->qcomp, s->r, F.nx, F.ny, ntx, nty, vgx, nvb, B, rs.part,   \
                                    rs.bpart, rs.nb, par, (SB*)F.b, (float*)F.x, s->partials + ps_off(PS_RR),  \
                                    s->mg_scaled, mgs, s->sc)
  if (F.hb) {
    if (upd) { if (pa) TQ_IN_A0(__half, true, true); else TQ_IN_A0(__half, true, false); }
    else { if (pa) TQ_IN_A0(__half, false, true); else TQ_IN_A0(__half, false, false); }
  } else {
    if (upd) { if (pa) TQ_IN_A0(float, true, true); else TQ_IN_A0(float, true, false); }
    else { if (pa) TQ_IN_A0(float, false, true); else TQ_IN_A0(float, false, false); }
  }
#undef TQ_IN_A0
  time_end(s, t);
  s->nb_upd = (int)(g.x * g.y);
}

// storage-type dispatch: fn(x tag, b tag) with the level's element types
template <class T> struct Tag { using type = T; };
template <class Fn>
static void with_types(const Level& L, Fn fn) {
  if (L.hx) {
    if (L.hb) fn(Tag<__half>{}, Tag<__half>{}); else fn(Tag<__half>{}, Tag<float>{});
  } else {
    if (L.hb) fn(Tag<float>{}, Tag<__half>{}); else fn(Tag<float>{}, Tag<float>{});
  }
}
#define TQ_TYPES(tx, tb) using SX = typename decltype(tx)::type; using SB = typename decltype(tb)::type

// pass B of level l, restriction to l+1 and pass A of l+1 (l+1 not the coarsest):
// x(l) -> tmp(l) (x_B), b(l+1), x(l+1)
static void launch_bra(tq_sim* s, int l) {
  Level& F = s->lv[l];
  Level& G = s->lv[l + 1];
  const bool pa = l + 1 < (int)s->lv.size() - 1 && !(l + 1 == s->lc && vop_active(s));
  const dim3 g((G.nx + 15) / 16, (G.ny + 15) / 16);
  // level 0 of an iteration (not the first cycle) starts with the norm and stop test
  const int fold = (l == 0 && s->it_k >= 0) ? 1 : 0;
  const RedSrc rsu = fold ? red_src(s, PS_RR, 1, -1, s->nb_upd) : RedSrc{s->partials + ps_off(PS_RR), nullptr, 0};
  const double* rr = rsu.part;
  // x(l) and b(l) and mask(l) in, x_B(l) out; coarse mask in, b (and x) out
  if (F.n > split_cells(s)) {  // pass B, then restriction + the coarse pass A
    static const char* const pnames[12] = {"mg_pass_b_l0", "mg_pass_b_l1", "mg_pass_b_l2", "mg_pass_b_l3",
                                           "mg_pass_b_l4", "mg_pass_b_l5", "mg_pass_b_l6", "mg_pass_b_l7",
                                           "mg_pass_b_l8", "mg_pass_b_l9", "mg_pass_b_l10", "mg_pass_b_l11"};
    static const char* const rnames[12] = {"mg_restrict_l0", "mg_restrict_l1", "mg_restrict_l2", "mg_restrict_l3",
                                           "mg_restrict_l4", "mg_restrict_l5", "mg_restrict_l6", "mg_restrict_l7",
                                           "mg_restrict_l8", "mg_restrict_l9", "mg_restrict_l10", "mg_restrict_l11"};
    const int ntx8 = tiles_of(F.nx, 8), nty8 = tiles_of(F.ny, 8);
    int t = time_begin(s, pnames[l], (2 * vbytes(F.hx) + vbytes(F.hb) + 1) * (double)F.n);
    with_types(F, [&](auto tx, auto tb) {
      TQ_TYPES(tx, tb);
      const dim3 tg = tile_grid(F.nx, F.ny, 8);
      const int nvb = (int)(tg.x * tg.y);
      // the folding launch (level 0 of an iteration) walks several groups per block after one fold
      const int nbk = fold ? std::min(nvb, pass_fold_grid(s)) : nvb;
      launch<256, 0, k_pass<SX, SB, 0>>(s, s->st, dim3(nbk), dim3(256), 0, (const SX*)F.x, (SX*)F.tmp,
          (const SB*)F.b, F.mask, F.nx, F.ny, 8, ntx8, nty8, (int)tg.x, nvb, rr, rsu.nb, it_par(s),
          fold, s->sc);
    });
    time_end(s, t);
    t = time_begin(s, rnames[l], (vbytes(F.hx) + vbytes(F.hb) + 1) * (double)F.n +
                                     (1 + vbytes(G.hb) + (pa ? vbytes(G.hx) : 0)) * (double)G.n);
    with_types(F, [&](auto fx, auto fb) {
      using SXF = typename decltype(fx)::type;
      using SBF = typename decltype(fb)::type;
      with_types(G, [&](auto cx, auto cb) {
        using SCX = typename decltype(cx)::type;
        using SCB = typename decltype(cb)::type;
        if (pa)
          launch<256, 0, k_restrict_a<SXF, SBF, SCB, SCX, true>>(s, s->st, g, dim3(256), 0, (const SXF*)F.tmp,
              (const SBF*)F.b, F.mask, F.nx, F.ny, (SCB*)G.b, (SCX*)G.x, G.mask, G.nx, G.ny, s->sc);
        else
          launch<256, 0, k_restrict_a<SXF, SBF, SCB, SCX, false>>(s, s->st, g, dim3(256), 0, (const SXF*)F.tmp,
              (const SBF*)F.b, F.mask, F.nx, F.ny, (SCB*)G.b, (SCX*)G.x, G.mask, G.nx, G.ny, s->sc);
      });
    });
    time_end(s, t);
    return;
  }
  static const char* const names[12] = {"mg_down_l0", "mg_down_l1", "mg_down_l2", "mg_down_l3",
                                        "mg_down_l4", "mg_down_l5", "mg_down_l6", "mg_down_l7",
                                        "mg_down_l8", "mg_down_l9", "mg_down_l10", "mg_down_l11"};
  int t = time_begin(s, names[l],
                     (2 * vbytes(F.hx) + vbytes(F.hb) + 1) * (double)F.n +
                         (1 + vbytes(G.hb) + (pa ? vbytes(G.hx) : 0)) * (double)G.n);
  with_types(F, [&](auto fx, auto fb) {
    using SXF = typename decltype(fx)::type;
    using SBF = typename decltype(fb)::type;
    with_types(G, [&](auto cx, auto cb) {
      using SCX = typename decltype(cx)::type;
      using SCB = typename decltype(cb)::type;
      if (pa)
        launch<64 * BRA_WAVES, 0, k_bra<SXF, SBF, SCB, SCX, true>>(s, s->st, g, dim3(64 * BRA_WAVES), 0, (const SXF*)F.x,
            (SXF*)F.tmp, (const SBF*)F.b, F.mask, F.nx, F.ny, (SCB*)G.b, (SCX*)G.x, G.mask, G.nx, G.ny,
            rr, rsu.nb, it_par(s), fold, s->sc);
      else
        launch<64 * BRA_WAVES, 0, k_bra<SXF, SBF, SCB, SCX, false>>(s, s->st, g, dim3(64 * BRA_WAVES), 0, (const SXF*)F.x,
            (SXF*)F.tmp, (const SBF*)F.b, F.mask, F.nx, F.ny, (SCB*)G.b, (SCX*)G.x, G.mask, G.nx, G.ny,
            rr, s->nb_upd, it_par(s), fold, s->sc);
    });
  });
  time_end(s, t);
}

static ZArgs zargs(tq_sim* s);

// prolongation from level l+1 (its final x in x(l+1)), passes C and D of level l:
// tmp(l) (x_B) -> x(l); on level 0 the result becomes z and the dot partials (returns their count)
static int launch_pcd(tq_sim* s, int l) {
  Level& F = s->lv[l];
  Level& G = s->lv[l + 1];
  const bool zd = l == 0;
  const dim3 g((F.nx + 15) / 16, (F.ny + 15) / 16);
  const ZArgs za = zargs(s);
  // x_B, b, mask in; x out (or z with r, z_old, opMask in); coarse x and mask in
  const double fine = zd ? (vbytes(F.hx) + vbytes(F.hb) + 1 + 1 + 4 + 4 + 4) : (2 * vbytes(F.hx) + vbytes(F.hb) + 1);
  static const char* const names[12] = {"mg_up_z_l0", "mg_up_l1", "mg_up_l2", "mg_up_l3", "mg_up_l4", "mg_up_l5",
                                        "mg_up_l6", "mg_up_l7", "mg_up_l8", "mg_up_l9", "mg_up_l10", "mg_up_l11"};
  if (!zd && F.n > split_cells(s)) {  // prolongation + pass C (tmp -> x), then pass D (x -> tmp)
    static const char* const cnames[12] = {"mg_prolong_c_l0", "mg_prolong_c_l1", "mg_prolong_c_l2",
                                           "mg_prolong_c_l3", "mg_prolong_c_l4", "mg_prolong_c_l5",
                                           "mg_prolong_c_l6", "mg_prolong_c_l7", "mg_prolong_c_l8",
                                           "mg_prolong_c_l9", "mg_prolong_c_l10", "mg_prolong_c_l11"};
    static const char* const dnames[12] = {"mg_pass_d_l0", "mg_pass_d_l1", "mg_pass_d_l2", "mg_pass_d_l3",
                                           "mg_pass_d_l4", "mg_pass_d_l5", "mg_pass_d_l6", "mg_pass_d_l7",
                                           "mg_pass_d_l8", "mg_pass_d_l9", "mg_pass_d_l10", "mg_pass_d_l11"};
    const int ntx8 = tiles_of(F.nx, 8), nty8 = tiles_of(F.ny, 8);
    const int ntx = tiles_of(F.nx, 0), nty = tiles_of(F.ny, 0);
    int t = time_begin(s, cnames[l], (2 * vbytes(F.hx) + vbytes(F.hb) + 1) * (double)F.n +
                                         (vbytes(G.hx) + 1) * (double)G.n);
    with_types(F, [&](auto tx, auto tb) {
      TQ_TYPES(tx, tb);
      if (G.hx)
        launch<256, 0, k_prolong_pass<SX, SB, __half>>(s, s->st, tile_grid(F.nx, F.ny, 8), dim3(256), 0,
            (const SX*)F.tmp, (SX*)F.x, (const SB*)F.b, F.mask, F.nx, F.ny, ntx8, nty8,
            (const __half*)G.res, G.mask, G.nx, G.ny, s->sc);
      else
        launch<256, 0, k_prolong_pass<SX, SB, float>>(s, s->st, tile_grid(F.nx, F.ny, 8), dim3(256), 0,
            (const SX*)F.tmp, (SX*)F.x, (const SB*)F.b, F.mask, F.nx, F.ny, ntx8, nty8,
            (const float*)G.res, G.mask, G.nx, G.ny, s->sc);
    });
    time_end(s, t);
    t = time_begin(s, dnames[l], (2 * vbytes(F.hx) + vbytes(F.hb) + 1) * (double)F.n);
    with_types(F, [&](auto tx, auto tb) {
      TQ_TYPES(tx, tb);
      const dim3 tg = tile_grid(F.nx, F.ny, 0);
      launch<256, 0, k_pass<SX, SB, 1>>(s, s->st, dim3(tg.x * tg.y), dim3(256), 0, (const SX*)F.x,
          (SX*)F.tmp, (const SB*)F.b, F.mask, F.nx, F.ny, 0, ntx, nty, (int)tg.x, (int)(tg.x * tg.y),
          (const double*)nullptr, 0, 0, 0, s->sc);
    });
    time_end(s, t);
    F.res = F.tmp;
    return 0;
  }
  int t = time_begin(s, names[l], fine * F.n + (vbytes(G.hx) + 1) * (double)G.n);
  with_types(F, [&](auto tx, auto tb) {
    TQ_TYPES(tx, tb);
    auto go = [&](auto tc) {
      using SC = typename decltype(tc)::type;
      if (zd)
        launch<256, 0, k_pcd<SX, SB, SC, true>>(s, s->st, g, dim3(256), 0, (const SX*)F.tmp, (SX*)F.x,
            (const SB*)F.b, F.mask, F.nx, F.ny, (const SC*)G.res, G.mask, G.nx, G.ny, za, s->sc);
      else
        launch<256, 0, k_pcd<SX, SB, SC, false>>(s, s->st, g, dim3(256), 0, (const SX*)F.tmp, (SX*)F.x,
            (const SB*)F.b, F.mask, F.nx, F.ny, (const SC*)G.res, G.mask, G.nx, G.ny, za, s->sc);
    };
    if (G.hx) go(Tag<__half>{}); else go(Tag<float>{});
  });
  time_end(s, t);
  F.res = F.x;
  return (int)(g.x * g.y);
}

static ZArgs zargs(tq_sim* s) {
  ZArgs za;
  za.opMask = s->opMask; za.rot = s->rot; za.cellsB = s->cellsB;
  za.labelSlot = s->labelSlot; za.r = s->r; za.zold = s->zbuf[s->zc]; za.znew = s->zbuf[1 - s->zc];
  za.accZ = bv(s->bvec, BV_ACCZ0 + z_par(s));
  za.partials = s->partials + ps_off(PS_Z);
  za.bpart = bpart_z(s);
  za.mgs = &s->sc->mgs[it_par(s)];
  za.scaled = s->mg_scaled;
  return za;
}

// LDS layout of levels lc..L-1: x, tmp, b (FP32) per level, then the masks
static WGPlan wg_plan(tq_sim* s) {
  WGPlan P{};
  const int L = (int)s->lv.size();
  P.lc = s->lc;
  P.L = L;
  int off = 0;
  for (int l = s->lc; l < L; l++) {
    const int n = s->lv[l].n;
    P.nx[l] = s->lv[l].nx; P.ny[l] = s->lv[l].ny;
    P.offx[l] = off; off += 4 * n;
    P.offt[l] = off; off += 4 * n;
    P.offb[l] = off; off += 4 * n;
    P.mask[l] = s->lv[l].mask;
  }
  for (int l = s->lc; l < L; l++) { P.offm[l] = off; off += (s->lv[l].n + 15) & ~15; }
  P.pool = off;
  P.dense = s->dense;
  return P;
}

static void launch_coarse_wg(tq_sim* s) {
  const int L = (int)s->lv.size();
  Level& G = s->lv[s->lc];
  if (vop_active(s)) {
    // x(lc) = C b(lc): the matrix and b in, x out
    const int waves = (G.n + VOP_ROWS - 1) / VOP_ROWS;
    int t = time_begin(s, "mg_coarse_mv", 4.0 * s->vop_ld * G.n + 8.0 * G.n);
    launch<256, 0, k_vcycle_apply>(s, s->st, dim3((waves + 3) / 4), dim3(256), 0, (const float*)s->vop, s->vop_ld,
                                   G.n, (const float*)G.b, (float*)G.x, s->sc);
    time_end(s, t);
    return;
  }
  const WGPlan P = wg_plan(s);
  // b and x of level lc in, its result out, every LDS level's mask, the dense operator
  double bytes = 12.0 * G.n;
  for (int l = s->lc; l < L; l++) bytes += s->lv[l].n;
  if (s->dense) bytes += 4.0 * s->lv[L - 1].n * s->lv[L - 1].n;
  int t = time_begin(s, "mg_coarse_wg", bytes);
  launch<1024, 0, k_coarse_wg>(s, s->st, dim3(1), dim3(1024), 0, P, (const float*)G.b, (const float*)G.x,
                               (float*)G.x, s->sc);
  time_end(s, t);
}

// builds this solve's matrix of level lc's V-cycle (after the level masks and the coarsest operator),
// on the tank's own stream: a second stream that ran the build beside the solve's set-up and first
// kernels slowed those kernels by more than the ~90 us of overlap it won
static int launch_vcycle_op(tq_sim* s) {
  const WGPlan P = wg_plan(s);
  const Level& G = s->lv[s->lc];
  const size_t lds = P.pool + VOP_WAVES * TLDS * sizeof(float);
  int t = time_begin(s, "mg_vcycle_op", 4.0 * s->vop_ld * G.n);
  launch<64 * VOP_WAVES, 0, k_vcycle_op>(s, s->st, dim3(G.n), dim3(64 * VOP_WAVES), lds, P, s->vop_ld, s->vop, s->sc);
  time_end(s, t);
  return 0;
}

// level 0 up: prolongation from level 1 (x(1)) + pass C, tmp(0) -> x(0); then pass D on x(0)
// writing z and the dot partials (returns their count)
static int launch_up0(tq_sim* s) {
  Level& F = s->lv[0];
  Level& G = s->lv[1];
  const int ntx8 = tiles_of(F.nx, 8), nty8 = tiles_of(F.ny, 8);
  const dim3 g8 = tile_grid(F.nx, F.ny, 8);
  int t = time_begin(s, "mg_prolong_c_l0", (8 + vbytes(F.hb) + 1) * (double)F.n + (vbytes(G.hx) + 1) * (double)G.n);
  with_types(F, [&](auto tx, auto tb) {
    TQ_TYPES(tx, tb);
    if (G.hx)
      launch<256, 0, k_prolong_pass<SX, SB, __half>>(s, s->st, g8, dim3(256), 0, (const SX*)F.tmp, (SX*)F.x,
          (const SB*)F.b, F.mask, F.nx, F.ny, ntx8, nty8, (const __half*)G.res, G.mask, G.nx, G.ny,
          s->sc);
    else
      launch<256, 0, k_prolong_pass<SX, SB, float>>(s, s->st, g8, dim3(256), 0, (const SX*)F.tmp, (SX*)F.x,
          (const SB*)F.b, F.mask, F.nx, F.ny, ntx8, nty8, (const float*)G.res, G.mask, G.nx, G.ny,
          s->sc);
  });
  time_end(s, t);
  const int ntx = tiles_of(F.nx, 0), nty = tiles_of(F.ny, 0);
  const dim3 g = tile_grid(F.nx, F.ny, 0);
  const ZArgs za = zargs(s);
  t = time_begin(s, "mg_pass_z_l0", (4 + vbytes(F.hb) + 1 + 1 + 4 + 4 + 4) * (double)F.n);
  if (F.hb)
    launch<256, 0, k_pass_z<__half>>(s, s->st, g, dim3(256), 0, (const float*)F.x, (const __half*)F.b, F.mask, F.nx,
                                     F.ny, ntx, nty, za, s->sc);
  else
    launch<256, 0, k_pass_z<float>>(s, s->st, g, dim3(256), 0, (const float*)F.x, (const float*)F.b, F.mask, F.nx,
                                    F.ny, ntx, nty, za, s->sc);
  time_end(s, t);
  return (int)(g.x * g.y);
}

// V-cycle on r (level 0 input already built by in_a0, whose pass A is done) -> z and the z dots
static void precondition_rest(tq_sim* s) {
  const int lc = s->lc;
  if (lc == 0 && s->it_k >= 0) {  // no level-0 down kernel to carry the norm and stop test
    int t = time_begin(s, "pcg_update_fold", 0);
    launch<FIN_THREADS, 0, k_fold_update>(s, s->st, dim3(1), dim3(FIN_THREADS), 0,
                                          (const double*)(s->partials + ps_off(PS_RR)), s->nb_upd, it_par(s), s->sc);
    time_end(s, t);
  }
  for (int l = 0; l < lc; l++) launch_bra(s, l);
  launch_coarse_wg(s);
  s->lv[lc].res = s->lv[lc].x;
  if (lc > 0) {
    for (int l = lc - 1; l >= 1; l--) launch_pcd(s, l);
    s->nb_z = launch_up0(s);
  } else {
    s->nb_z = blocks_of(s->C);
    int t = time_begin(s, "pcg_zdots", 16.0 * s->C);
    launch<256, 0, k_zdots0>(s, s->st, dim3(s->nb_z), dim3(256), 0, (const float*)s->lv[0].x, zargs(s), s->C, s->sc);
    time_end(s, t);
  }
  s->zc = 1 - s->zc;
}

// PCG iteration k = s->it_k: spmv (z dots, beta, p, Ap), in_a0 (alpha, update, V-cycle input),
// the V-cycle (level 0's down kernel starts with the norm and stop test)
static void pcg_iteration(tq_sim* s) {
  const int C = s->C;
  const int nbs = std::min(spmv_tiles(s->nx, s->ny), s->launch_spmv_grid > 0 ? s->launch_spmv_grid : SPMV_GRID);
  const int pn = 1 - s->pc;
  const RedSrc rz = red_src(s, PS_Z, 2, 1, s->nb_z);
  int t = time_begin(s, "pcg_spmv", (1 + 4 + 4 + 4 + 4) * (double)C);
  launch<256, 4, k_spmv>(s, s->st, dim3(nbs), dim3(256), 0, s->opMask, s->rot, s->zbuf[s->zc], s->pbuf[s->pc],
                         s->pbuf[pn], s->ap, s->nx, s->ny, bubargs(s), rz.part, rz.bpart, rz.nb, it_par(s),
                         s->it_k == 0 ? 1 : 0, s->partials + ps_off(PS_SPMV), bpart_a(s), s->sc);
  time_end(s, t);
  s->nb_spmv = nbs;
  s->pc = pn;
  launch_in_a0(s, true);
  precondition_rest(s);
  s->it_k++;
}

// Two PCG iterations (the buffer ping-pong period) captured once into a HIP graph and
// replayed; the capture depends on the level plan, so a plan change recaptures.
static int graph_pair(tq_sim* s) {
  const int key = (s->qcomp << 16) | (s->d.precision << 8) | ((s->d.mg_half_level + 1) << 2) |
                  (vop_active(s) ? 2 : 0) | (s->mg_scaled ? 1 : 0);
  if (!s->graph_exec || s->graph_key != key) {
    graph_drop(s);
    hipGraph_t g = nullptr;
    const int k0 = s->it_k;
    s->capturing = true;
    TQ_HIP(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    pcg_iteration(s);
    pcg_iteration(s);
    const hipError_t e = hipStreamEndCapture(s->st, &g);
    s->it_k = k0;
    s->capturing = false;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    hipGraphExec_t ex = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e2 != hipSuccess) return hip_fail(e2, "hipGraphInstantiate", __FILE__, __LINE__);
    s->graph_exec = ex;
    s->graph_key = key;
    // the capture flipped the ping-pong state twice: back where it started
  }
  TQ_HIP(hipGraphLaunch((hipGraphExec_t)s->graph_exec, s->st));
  s->it_k += 2;
  return 0;
}

// The captured pair's parity-dependent buffers are fixed, so pairs start at odd iterations;
// iteration 0 (beta = 0) and any unpaired iteration run as eager launches.
static int run_iterations(tq_sim* s, int n) {
  const bool graph = s->d.use_graph && !s->timing && !s->lrec;
  while (n > 0) {
    if (graph && n >= 2 && (s->it_k & 1)) {
      if (const int rc = graph_pair(s)) return rc;
      n -= 2;
    } else {
      pcg_iteration(s);
      n--;
    }
  }
  TQ_LAUNCH_CHECK();
  return 0;
}

// masks of every level, bubble vectors and slots, the dense operators when their masks changed
static int solve_setup(tq_sim* s) {
  const int C = s->C;
  const int L = (int)s->lv.size();
  int t = time_begin(s, "pcg_setup", (1 + 4 + 1 + 1) * (double)C);
  {
    PyrPlan Pp{};
    Pp.L = L;
    for (int l = 0; l < L; l++) {
      Pp.nx[l] = s->lv[l].nx;
      Pp.ny[l] = s->lv[l].ny;
      Pp.mask[l] = s->lv[l].mask;
    }
    launch<1024, 0, k_setup_pyr>(s, s->st, dim3((s->nx + PYR_T - 1) / PYR_T, (s->ny + PYR_T - 1) / PYR_T), dim3(1024), 0, s->tipo, s->rot, s->facesB, s->d.ar, s->opMask, Pp);
  }
  for (int l = PYR_LEVELS + 1; l < L; l++) {  // levels past the pyramid (grids over 768 cells a side)
    const Level& F = s->lv[l - 1];
    const Level& G = s->lv[l];
    launch<256, 0, k_coarse_type>(s, s->st, dim3(blocks_of(G.n)), dim3(256), 0, F.mask, F.nx, F.ny, G.mask, G.nx, G.ny);
  }
  launch<1024, 0, k_bub_init>(s, s->st, dim3(BV_N), dim3(1024), 0, s->facesB, s->yB32, s->d.ar, s->bvec,
                              s->activeList, s->labelSlot, s->sc);
  {
    // neighbour bits of levels 1.., and whether the dense operators must be rebuilt (their masks
    // changed, or the level-lc matrix was not kept current while it was switched off)
    VizPlan V{};
    V.v0 = std::min(1, s->lc);
    V.L = L;
    V.lc = s->lc;
    int off = 0;
    for (int l = V.v0; l < L; l++) {
      V.off[l] = off;
      V.nx[l] = s->lv[l].nx;
      V.ny[l] = s->lv[l].ny;
      V.mask[l] = s->lv[l].mask;
      off += s->lv[l].n;
    }
    V.off[L] = off;
    const bool force = vop_active(s) && !s->vop_was_active;
    if (off > 0)
      launch<256, 0, k_viz_all>(s, s->st, dim3(blocks_of(off)), dim3(256), 0, V, s->op_masks, force ? 1 : 0, s->sc);
    s->vop_was_active = vop_active(s);
  }
  if (s->dense) {
    const Level& G = s->lv[L - 1];
    launch<64, 0, k_coarsest_op>(s, s->st, dim3(G.n), dim3(64), 0, G.mask, G.nx, G.n, s->dense,
                                 (const DevScalars*)s->sc);
  }
  time_end(s, t);
  if (vop_active(s))
    if (int e = launch_vcycle_op(s)) return e;
  launch<64, 0, k_op_count>(s, s->st, dim3(1), dim3(64), 0, s->sc);
  TQ_LAUNCH_CHECK();
  return 0;
}

// the stop check of the residual-terminated mode, queued after every chunk of iterations: the
// binary64 true residual of the compensated iterate, computed once after the segment stops
static void launch_true_check(tq_sim* s) {
  const int C = s->C;
  const int nb = blocks_of(C);
  launch<256, 0, k_true_res>(s, s->st, dim3(nb), dim3(256), 0, s->opMask, s->rot, s->div, s->q, (const float*)s->qlo,
                             s->nx, C, s->bvec, s->labelSlot, s->partials, s->bpart, 1, (const DevScalars*)s->sc);
  launch<FIN_THREADS, 0, k_fin_true_res>(s, s->st, dim3(1), dim3(FIN_THREADS), 0, s->facesB, s->bvec, s->rhsB, s->activeList,
                                         s->labelSlot, s->partials, s->bpart, nb, 1, s->sc, (double*)nullptr);
}

constexpr int MAX_RESTARTS = 3;  // residual replacements per solve (each one re-runs the set-up residual)

int pressure_solve(tq_sim* s) {
  const int C = s->C;
  s->mg_scaled = plan_precision(s) ? 1 : 0;
  const double tol = s->d.tol;
  const int maxit = s->d.iteracoes;
  // residual-terminated mode: the iterate is kept compensated (q + qlo) so its true residual can
  // reach tol; the fixed-iteration mode keeps the reference's binary32 iterate exactly
  s->qcomp = tol > 0 ? 1 : 0;
  if (int e = solve_setup(s)) return e;
  // FP16 right-hand sides with FP32 x can sit 2^12 above unit norm (more of the residual stays
  // in the normal FP16 range); FP16 x keeps unit-norm input so the V-cycle output stays in range
  const float mult = s->d.precision == TQ_PREC_MIXED_B ? 4096.f : 1.f;
  const int nb = blocks_of(C);
  for (int restart = 0;; restart++) {
    s->pc = 0;
    s->zc = 0;
    int t = time_begin(s, "pcg_init", (1 + 1 + 4 + 4 + 4 + 4 + 4 + (s->qcomp ? 4 : 0)) * (double)C);
    if (restart) launch<1024, 0, k_bub_restart>(s, s->st, dim3(1), dim3(1024), 0, s->bvec);
    launch<256, 0, k_pcg_init>(s, s->st, dim3(nb), dim3(256), 0, s->opMask, s->rot, s->div, s->q, s->qlo,
                               s->qcomp ? (restart ? 2 : 1) : 0, s->r, s->pbuf[s->pc], s->nx, C, s->bvec, s->labelSlot,
                               s->partials, s->bpart, s->sc);
    launch<FIN_THREADS, 0, k_fin_init>(s, s->st, dim3(1), dim3(FIN_THREADS), 0, s->facesB, s->bvec, s->rhsB, s->activeList,
        s->labelSlot, s->partials, s->bpart, nb, tol, maxit, mult, restart ? 1 : 0, s->sc);
    time_end(s, t);
    TQ_LAUNCH_CHECK();
    s->it_k = -1;
    launch_in_a0(s, false);
    precondition_rest(s);
    s->it_k = 0;
    TQ_LAUNCH_CHECK();
    if (tol <= 0) {
      const int rc = run_iterations(s, maxit);
      if (rc) return rc;
      break;
    }
    // launch in chunks, each followed by the gated stop check, and poll the device flags
    // (kernels after the stop exit at once); the first chunk is the last solve's first segment
    int launched = 0;
    int chunk = restart ? 4 : std::max(2, s->sc_host->iters_first > 0 ? (s->sc_host->iters_first + 1) & ~1 : 8);
    while (launched < maxit) {
      const int n = std::min(chunk, maxit - launched);
      const int rc = run_iterations(s, n);
      if (rc) return rc;
      launched += n;
      launch_true_check(s);
      TQ_LAUNCH_CHECK();
      TQ_HIP(hipMemcpyAsync(s->sc_host, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->st));
      TQ_HIP(hipStreamSynchronize(s->st));
      if (s->sc_host->done) break;
      chunk = 4;
    }
    const DevScalars& h = *s->sc_host;
    const bool converged = h.norm_last <= tol * h.normb;  // the segment stopped on its recursive residual
    if (!converged || !(h.true_norm > tol * h.normb) || restart >= MAX_RESTARTS) break;
  }
  launch<256, 0, k_pcg_final>(s, s->st, dim3(blocks_of(C)), dim3(256), 0, s->tipo, s->q, C);
  launch<1024, 0, k_pcg_final_b>(s, s->st, dim3(1), dim3(1024), 0, s->bvec, s->activeList, s->yB32, s->sc);
  TQ_LAUNCH_CHECK();
  return 0;
}

int true_residual(tq_sim* s, double* out) {
  const int C = s->C;
  const int nb = blocks_of(C);
  double* dres = s->partials + ps_off(PS_LAST) + MAX_RED_BLOCKS - 1;
  launch<256, 0, k_true_res>(s, s->st, dim3(nb), dim3(256), 0, s->opMask, s->rot, s->div, s->q, (const float*)nullptr,
                             s->nx, C, s->bvec, s->labelSlot, s->partials, s->bpart, 0, (const DevScalars*)s->sc);
  launch<FIN_THREADS, 0, k_fin_true_res>(s, s->st, dim3(1), dim3(FIN_THREADS), 0, s->facesB, s->bvec, s->rhsB, s->activeList,
                                         s->labelSlot, s->partials, s->bpart, nb, 0, s->sc, dres);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(out, dres, sizeof(double), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

// One application of the preconditioner to a level-0 vector (test hook): this solve's setup on
// the current types, then the V-cycle of r = b from x = 0, z out.  Bubble cells take the
// bubble residual of the setup (zero), so inputs are meant for water cells.
int debug_vcycle(tq_sim* s, const float* hb, float* hz) {
  const int C = s->C;
  s->mg_scaled = plan_precision(s) ? 1 : 0;
  s->pc = 0;
  s->zc = 0;
  if (int e = solve_setup(s)) return e;
  double n2 = 0;
  for (int c = 0; c < C; c++) n2 += (double)hb[c] * hb[c];
  const float mult = s->d.precision == TQ_PREC_MIXED_B ? 4096.f : 1.f;
  const float scale = n2 > 0 ? (float)(mult / std::sqrt(n2)) : 1.f;
  const int zero = 0;
  const double dzero = 0;
  TQ_HIP(hipMemcpyAsync(s->r, hb, sizeof(float) * C, hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(&s->sc->done, &zero, sizeof(int), hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(&s->sc->proj_mean, &dzero, sizeof(double), hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(&s->sc->mgs[1], &scale, sizeof(float), hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(&s->sc->mg_mult, &mult, sizeof(float), hipMemcpyHostToDevice, s->st));
  s->it_k = -1;
  launch_in_a0(s, false);
  precondition_rest(s);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(hz, s->zbuf[s->zc], sizeof(float) * C, hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

}  // namespace tq
