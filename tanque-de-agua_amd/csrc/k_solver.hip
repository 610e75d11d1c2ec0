// k_solver.hip -- pressure projection: PCG preconditioned by a multigrid V-cycle, with one
// extra unknown per enclosed air pocket (resolverPCG / cicloV, js/pressao-cpu.js:241-402).
//
// CG vectors r, z, p, Ap are FP32 per cell; every dot product, the residual norm and the
// bubble unknowns (y, r, z, p, Ap per label) are FP64.  Reductions are deterministic and free
// of same-address atomics: grid-stride kernels store one partial per block (and one per block
// for each of up to BSLOTS active bubbles), and a one-block "fin" kernel sums them in block
// order and does the scalar/bubble work.  Bubbles beyond BSLOTS fall back to FP64 atomics.
//
// The V-cycle is launch-lean: per large level one kernel for both down-smoothing passes (the
// half-block-shifted second pass recomputes the four first-pass tiles it overlaps), one
// LDS-tiled restriction, and one up kernel (prolongation folded into the snapshot load, both
// passes); the coarse levels that fit in LDS are solved by ONE workgroup; level 0's up kernel
// also produces z and the r.z dots.  Mixed mode stores b in FP16 on level 0 and x, b in FP16
// on the other large levels (per-cycle scaled by 1/||r||), with FP32 arithmetic.
//
// Kernel sequence of one PCG iteration:
//   spmv (p = z + beta p fused, Ap, p.Ap) -> fin_spmv (alpha, bubble y, r)
//   update (q, r, |r|^2) -> fin_update (norm, stop test)
//   mg_input, [down_l, restrict_l] for l < lc, coarse_wg (levels >= lc), up_l for l = lc-1..0
//   (up_0 writes z and the dots) -> fin_zdots (beta)
// Every kernel returns at once when the solve is done.
#include "tq_device.hpp"

namespace tq {

enum { BV_Y = 0, BV_R, BV_Z, BV_ZV, BV_P, BV_AP, BV_ACC0, BV_ACC1, BV_N };
__device__ __host__ inline double* bv(double* base, int k) { return base + (size_t)k * MAX_BOLHAS; }

constexpr int NBR = 512;        // blocks of the grid-stride CG kernels
constexpr int BSLOTS = 8;       // bubbles reduced through per-block slots
constexpr int WG_LDS = 147456;  // LDS pool of the one-workgroup coarse kernel (144 KiB)
constexpr int WG_BYTES_PER_CELL = 13;

template <class S> __device__ inline float ld(const S* p, size_t i);
template <> __device__ inline float ld<float>(const float* p, size_t i) { return p[i]; }
template <> __device__ inline float ld<__half>(const __half* p, size_t i) { return __half2float(p[i]); }
template <class S> __device__ inline void st(S* p, size_t i, float v);
template <> __device__ inline void st<float>(float* p, size_t i, float v) { p[i] = v; }
template <> __device__ inline void st<__half>(__half* p, size_t i, float v) { p[i] = __float2half(v); }

// all LDS operations of this wavefront are complete (and not reordered across by the compiler)
__device__ inline void wave_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

#define GRID_STRIDE(c, n) for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < (n); c += gridDim.x * blockDim.x)

__device__ inline void slot_add(double (&acc)[BSLOTS], int slot, double val) {
#pragma unroll
  for (int k = 0; k < BSLOTS; k++) acc[k] += (k == slot) ? val : 0.0;
}

template <int NV>
__device__ inline void write_partials(double (&v)[NV], double* part, double* sh) {
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0) {
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
    for (int k = 0; k < NV; k++) part[(size_t)k * MAX_RED_BLOCKS + bid] = v[k];
  }
}

// Sum of nb partials per quantity in a fixed assignment (deterministic); result in all threads.
template <int NV>
__device__ inline void read_partials(const double* part, int nb, double (&tot)[NV], double* sh, double* res) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] += part[(size_t)k * MAX_RED_BLOCKS + b];
  block_sum<NV>(acc, sh);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) res[k] = acc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) tot[k] = res[k];
  __syncthreads();
}

__device__ inline bool is_bub_cell(const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                   const uint32_t* __restrict__ facesB, int ar, size_t c, uint32_t* lab) {
  if (!ar || tipo[c] != AR) return false;
  const uint32_t r = rot[c];
  if (r < 2 || facesB[r] == 0) return false;
  *lab = r;
  return true;
}

__device__ inline bool interior(int i, int j, int nx, int ny) { return i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1; }

// a bubble contribution: through the block's slot accumulators, or an FP64 atomic when the
// label has no slot
__device__ inline void bub_add(double (&sl)[BSLOTS], const int* __restrict__ labelSlot, double* accG, uint32_t lab,
                               double val) {
  const int s = labelSlot[lab];
  if (s >= 0) slot_add(sl, s, val);
  else atomicAdd(&accG[lab], val);
}

// ================================================================== setup (types and neighbour masks)
// level 0: operator mask (true types) and MG mask (bubble air as water, pressao-cpu.js:305-308)
__global__ __launch_bounds__(256) void k_setup0(const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot, int nx,
                                                int ny, int ar, uint8_t* __restrict__ opMask, uint8_t* __restrict__ mg0) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  int m = 0;
  if (i > 0 && tipo[c - 1] != SOLIDO) m |= 1;
  if (i < nx - 1 && tipo[c + 1] != SOLIDO) m |= 2;
  if (j > 0 && tipo[c - nx] != SOLIDO) m |= 4;
  if (j < ny - 1 && tipo[c + nx] != SOLIDO) m |= 8;
  const uint8_t t = tipo[c];
  const uint8_t tm = (ar && t == AR && rot[c] >= 2) ? AGUA : t;
  opMask[c] = (uint8_t)(m | (t << 4));
  mg0[c] = (uint8_t)(m | (tm << 4));
}

// tiposGrossos (pressao-cpu.js:63-81): air if any child is air, else water if any is water
__global__ __launch_bounds__(256) void k_coarse_type(const uint8_t* __restrict__ fm, int fnx, int fny,
                                                     uint8_t* __restrict__ cm, int cnx, int cny) {
  const int C = blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= cnx * cny) return;
  const int I = C % cnx, J = C / cnx;
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

// prepararVizinhos (pressao-cpu.js:44-60) on the type bits of a level
__global__ __launch_bounds__(256) void k_viz(uint8_t* __restrict__ m, int nx, int ny) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  int v = 0;
  if (i > 0 && mask_type(m[c - 1]) != SOLIDO) v |= 1;
  if (i < nx - 1 && mask_type(m[c + 1]) != SOLIDO) v |= 2;
  if (j > 0 && mask_type(m[c - nx]) != SOLIDO) v |= 4;
  if (j < ny - 1 && mask_type(m[c + nx]) != SOLIDO) v |= 8;
  m[c] = (uint8_t)((m[c] & 0xF0) | v);
}

// bubble vectors, the active list (labels with water faces, in label order) and their slots
__global__ __launch_bounds__(1024) void k_bub_init(const uint32_t* __restrict__ facesB, const float* __restrict__ yB32,
                                                   int ar, double* __restrict__ bvec, int* __restrict__ activeList,
                                                   int* __restrict__ labelSlot, DevScalars* __restrict__ sc) {
  __shared__ int cnt[1024];
  const int tid = threadIdx.x;
  int mine = 0;
  for (int k = 0; k < 4; k++) {
    const int r = tid * 4 + k;
    const bool act = ar && r >= 2 && facesB[r] > 0;
    mine += act;
    for (int q = 0; q < BV_N; q++) bv(bvec, q)[r] = 0.0;
    if (act) bv(bvec, BV_Y)[r] = (double)yB32[r];
    labelSlot[r] = -1;
  }
  cnt[tid] = mine;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int t = 0; t < 1024; t++) { int c = cnt[t]; cnt[t] = acc; acc += c; }
    sc->active_bubbles = acc;
  }
  __syncthreads();
  int o = cnt[tid];
  for (int k = 0; k < 4; k++) {
    const int r = tid * 4 + k;
    if (ar && r >= 2 && facesB[r] > 0) {
      activeList[o] = r;
      labelSlot[r] = o < BSLOTS ? o : -1;
      o++;
    }
  }
}

// ================================================================== init: warm-start residual
// q := 0 off water, bubble cells := y (difundir); r = -div - A q; bubble rows A_b [q;y].
__global__ __launch_bounds__(256) void k_pcg_init(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                                  const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                                  const float* __restrict__ div, float* __restrict__ q,
                                                  float* __restrict__ r, float* __restrict__ pbuf0, int nx, int ny,
                                                  int ar, double* __restrict__ bvec, const int* __restrict__ labelSlot,
                                                  double* __restrict__ partials, double* __restrict__ bpart,
                                                  DevScalars* __restrict__ sc) {
  __shared__ double sh[BSLOTS * 16];
  const int C = nx * ny;
  const int na = sc->active_bubbles;
  double v[5] = {0, 0, 0, 0, 0};  // r.r, div.div, sum(b), n_water, anchored
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double* yb = bv(bvec, BV_Y);
  double* acc0 = bv(bvec, BV_ACC0);
  auto qe = [&](size_t d) -> float {
    const uint8_t t = tipo[d];
    if (t == AGUA) return q[d];
    uint32_t lb;
    if (is_bub_cell(tipo, rot, facesB, ar, d, &lb)) return (float)yb[lb];
    return 0.f;
  };
  GRID_STRIDE(c, C) {
    pbuf0[c] = 0.f;
    const float qc = qe(c);
    const uint8_t m = opMask[c];
    if (mask_type(m) == AGUA) {
      double s = 0;
      bool anchored = false;
      const size_t nb[4] = {(size_t)c - 1, (size_t)c + 1, (size_t)c - nx, (size_t)c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        s += (double)qe(nb[k]);
        if (tipo[nb[k]] == AR) {
          uint32_t lb;
          if (is_bub_cell(tipo, rot, facesB, ar, nb[k], &lb)) {
            if (interior((int)(nb[k] % nx), (int)(nb[k] / nx), nx, ny)) bub_add(sl, labelSlot, acc0, lb, -(double)qc);
          } else anchored = true;
        }
      }
      const double aq = mask_cnt(m) * (double)qc - s;
      const float rr = (float)(-(double)div[c] - aq);
      r[c] = rr;
      v[0] += (double)rr * rr;
      v[1] += (double)div[c] * div[c];
      v[2] += -(double)div[c];
      v[3] += 1;
      v[4] += anchored ? 1 : 0;
    } else r[c] = 0.f;
    q[c] = qc;
  }
  write_partials<5>(v, partials, sh);
  if (na > 0) write_partials<BSLOTS>(sl, bpart, sh);
}

__global__ __launch_bounds__(256) void k_fin_init(const uint32_t* __restrict__ facesB, double* __restrict__ bvec,
                                                  const double* __restrict__ rhsB, const int* __restrict__ activeList,
                                                  const int* __restrict__ labelSlot, const double* __restrict__ partials,
                                                  const double* __restrict__ bpart, int nb, DevScalars* __restrict__ sc) {
  __shared__ double sh[BSLOTS * 16], res[BSLOTS], slotTot[BSLOTS];
  __shared__ double mean_s;
  const int na = sc->active_bubbles;
  double tot[5];
  read_partials<5>(partials, nb, tot, sh, res);
  double stt[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (na > 0) read_partials<BSLOTS>(bpart, nb, stt, sh, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < BSLOTS; k++) slotTot[k] = stt[k];
  __syncthreads();
  double bacc[3] = {0, 0, 0};  // rb.rb, rhs.rhs, sum(rhs)
  double* rb = bv(bvec, BV_R);
  double* acc0 = bv(bvec, BV_ACC0);
  const double* yb = bv(bvec, BV_Y);
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const int s = labelSlot[b];
    double accB;
    if (s >= 0) accB = slotTot[s];
    else { accB = acc0[b]; acc0[b] = 0; }
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
    sc->solves += 1;
    sc->rz = 0;
    sc->beta = 0;
    sc->alpha = 0;
    const bool stop = sc->maxit <= 0 || (sc->tol > 0 && (sc->norm0 == 0 || sc->norm0 <= sc->tol * sc->normb));
    sc->done = stop ? 1 : 0;
    sc->mg_scale = sc->norm0 > 0 ? (float)(1.0 / sc->norm0) : 1.f;
  }
  __syncthreads();
  if (mean_s != 0)
    for (int k = threadIdx.x; k < na; k += blockDim.x) rb[activeList[k]] -= mean_s;
}

// ================================================================== MG right-hand side
template <class S>
__global__ __launch_bounds__(256) void k_mg_input(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                                  const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                                  const uint32_t* __restrict__ cellsB, float* __restrict__ r, int C,
                                                  int ar, const double* __restrict__ bvec, int first, int mixed,
                                                  S* __restrict__ b0, const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const float scale = mixed ? sc->mg_scale : 1.f;
  float val = 0.f;
  if (mask_type(opMask[c]) == AGUA) {
    val = r[c];
    if (first && sc->proj_mean != 0) { val = (float)((double)val - sc->proj_mean); r[c] = val; }
  } else {
    uint32_t lb;
    if (is_bub_cell(tipo, rot, facesB, ar, c, &lb)) val = (float)(bv((double*)bvec, BV_R)[lb] / cellsB[lb]);
  }
  st<S>(b0, c, val * scale);
}

// ================================================================== fused block smoothers (suavizar, pressao-cpu.js:99-150)
// A tile of the reference's smoother reads a fixed snapshot: its interior starts from the
// snapshot and its 1-cell halo stays at the snapshot during its 4 red-black sweeps; the output
// keeps water cells and zeroes the rest.  Two passes (tile offsets 0 and 8, or 8 and 0) are
// fused: the block owns one second-pass tile T2, recomputes the four first-pass tiles that
// cover T2's halo (a 32x32 "centre" region), then runs the second pass on T2.
//
// Centre coordinates: X = cx0 + lx, lx in [0, 32); first-pass tile a = lx / 16 (and b = ly / 16).
// LDS: S snapshot of pass 1 on 34x34 (centre + 1 border), W pass-1 work values, L pass-2 tile.
__device__ inline bool pass1_same_tile(int lx, int ly, int nlx, int nly) {
  return nlx >= 0 && nly >= 0 && nlx < 32 && nly < 32 && (nlx >> 4) == (lx >> 4) && (nly >> 4) == (ly >> 4);
}

// The four pass-1 tiles on the centre: W starts as S in every tile interior; sweeps use W inside
// the cell's own tile and S elsewhere.  Thread t owns cells (t%16 + 16a, t/16 + 16b).
__device__ inline void pass1_sweeps(const float* S, float* W, const float* B, const uint8_t* M, int cx0, int cy0, int nx,
                                    int ny, int c0, int c1) {
  const int tid = threadIdx.x;
  for (int it = 0; it < 4; it++)
    for (int o = 0; o < 2; o++) {
      const int cor = o ? c1 : c0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int lx = (tid & 15) + 16 * (k & 1), ly = (tid >> 4) + 16 * (k >> 1);
        const int i = cx0 + lx, j = cy0 + ly;
        if (i < 0 || j < 0 || i >= nx || j >= ny || ((i + j + 1) & 1) != cor) continue;
        const int e = lx + ly * 32;
        const uint8_t m = M[e];
        const int kk = mask_cnt(m);
        if (mask_type(m) != AGUA || kk == 0) continue;
        const int es = (lx + 1) + (ly + 1) * 34;
        float s = 0.f;
        if (m & 1) s += pass1_same_tile(lx, ly, lx - 1, ly) ? W[es - 1] : S[es - 1];
        if (m & 2) s += pass1_same_tile(lx, ly, lx + 1, ly) ? W[es + 1] : S[es + 1];
        if (m & 4) s += pass1_same_tile(lx, ly, lx, ly - 1) ? W[es - 34] : S[es - 34];
        if (m & 8) s += pass1_same_tile(lx, ly, lx, ly + 1) ? W[es + 34] : S[es + 34];
        const float qv = W[es];
        W[es] = qv + ((s + B[e]) / (float)kk - qv);
      }
      __syncthreads();
    }
}

// pass 2 on T2 (centre offsets 8..23) from the masked pass-1 values; returns the cell's value
__device__ inline float pass2_sweeps(const float* W, float* L, const float* B, const uint8_t* M, int cx0, int cy0,
                                     int nx, int ny, int c0, int c1, bool* in_out) {
  const int tid = threadIdx.x;
  for (int e = tid; e < 18 * 18; e += 256) {
    const int lx = 7 + e % 18, ly = 7 + e / 18;  // centre coords
    const int i = cx0 + lx, j = cy0 + ly;
    float val = 0.f;
    if (i >= 0 && j >= 0 && i < nx && j < ny && mask_type(M[lx + ly * 32]) == AGUA) val = W[(lx + 1) + (ly + 1) * 34];
    L[e] = val;
  }
  const int tx = tid & 15, ty = tid >> 4;
  const int lx = 8 + tx, ly = 8 + ty;
  const int i = cx0 + lx, j = cy0 + ly;
  const bool in = i >= 0 && j >= 0 && i < nx && j < ny;
  const uint8_t m = in ? M[lx + ly * 32] : 0;
  const int kk = mask_cnt(m);
  const bool water = in && mask_type(m) == AGUA;
  const bool act = water && kk > 0;
  const float bb = act ? B[lx + ly * 32] : 0.f;
  const int e = (tx + 1) + (ty + 1) * 18;
  const int myc = (i + j + 1) & 1;
  __syncthreads();
  for (int it = 0; it < 4; it++)
    for (int o = 0; o < 2; o++) {
      const int cor = o ? c1 : c0;
      if (act && myc == cor) {
        float s = 0.f;
        if (m & 1) s += L[e - 1];
        if (m & 2) s += L[e + 1];
        if (m & 4) s += L[e - 18];
        if (m & 8) s += L[e + 18];
        const float qv = L[e];
        L[e] = qv + ((s + bb) / (float)kk - qv);
      }
      __syncthreads();
    }
  *in_out = in;
  return water ? L[e] : 0.f;
}

// load b and the mask of the 32x32 centre
template <class SB>
__device__ inline void load_centre(const SB* __restrict__ b, const uint8_t* __restrict__ mask, int cx0, int cy0, int nx,
                                   int ny, float* B, uint8_t* M) {
  for (int e = threadIdx.x; e < 32 * 32; e += 256) {
    const int i = cx0 + (e & 31), j = cy0 + (e >> 5);
    const bool in = i >= 0 && j >= 0 && i < nx && j < ny;
    const size_t c = in ? (size_t)i + (size_t)j * nx : 0;
    M[e] = in ? mask[c] : 0;
    B[e] = in ? ld<SB>(b, c) : 0.f;
  }
}

// Down: passes (0, [red, black]) then (8, [red, black]) from x = 0 (cicloV starts every level
// at zero, pressao-cpu.js:243 and 168).  Grid = offset-8 tiles.
template <class SX, class SB>
__global__ __launch_bounds__(256) void k_down(SX* __restrict__ dst, const SB* __restrict__ b,
                                              const uint8_t* __restrict__ mask, int nx, int ny,
                                              const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ float S[34 * 34], W[34 * 34], B[32 * 32], L[18 * 18];
  __shared__ uint8_t M[32 * 32];
  const int ox2 = blockIdx.x * 16 - 8, oy2 = blockIdx.y * 16 - 8;
  const int cx0 = ox2 - 8, cy0 = oy2 - 8;
  load_centre<SB>(b, mask, cx0, cy0, nx, ny, B, M);
  for (int e = threadIdx.x; e < 34 * 34; e += 256) { S[e] = 0.f; W[e] = 0.f; }
  __syncthreads();
  pass1_sweeps(S, W, B, M, cx0, cy0, nx, ny, 0, 1);
  bool in;
  const float v = pass2_sweeps(W, L, B, M, cx0, cy0, nx, ny, 0, 1, &in);
  if (in) {
    const int i = ox2 + (threadIdx.x & 15), j = oy2 + (threadIdx.x >> 4);
    st<SX>(dst, (size_t)i + (size_t)j * nx, v);
  }
}

// prolongation term of a fine cell (prolongar, pressao-cpu.js:191-214)
template <class SC>
__device__ inline float prolong_at(const SC* __restrict__ cx, const uint8_t* __restrict__ cmk, int cnx, int cny, int i,
                                   int j) {
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
      const size_t C = (size_t)I + (size_t)J * cnx;
      if (mask_type(cmk[C]) == AGUA) e += ld<SC>(cx, C) * (a ? 1.f - wx0 : wx0) * wy;
    }
  }
  return e;
}

struct ZArgs {  // level-0 up kernel: z = MG result and the dots r.z, r.z_old
  const uint8_t* opMask;
  const uint8_t* tipo;
  const uint32_t* rot;
  const uint32_t* facesB;
  const uint32_t* cellsB;
  const int* labelSlot;
  const float* r;
  const float* zold;
  float* znew;
  double* bvec;
  double* partials;
  double* bpart;
  int ar;
  int mixed;
};

// Up: prolongation folded into the snapshot, then passes (8, [black, red]) and (0, [black, red]).
// Grid = offset-0 tiles.  With ZD the block writes z and its dot/bubble partials instead of x.
template <class SX, class SB, class SC, bool ZD>
__global__ __launch_bounds__(256) void k_up(const SX* __restrict__ src, SX* __restrict__ dst, const SB* __restrict__ b,
                                            const uint8_t* __restrict__ mask, int nx, int ny, const SC* __restrict__ cx,
                                            const uint8_t* __restrict__ cmk, int cnx, int cny, ZArgs za,
                                            const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ float S[34 * 34], W[34 * 34], B[32 * 32], L[18 * 18];
  __shared__ uint8_t M[32 * 32];
  __shared__ double sh[BSLOTS * 16];
  const int ox2 = blockIdx.x * 16, oy2 = blockIdx.y * 16;
  const int cx0 = ox2 - 8, cy0 = oy2 - 8;
  load_centre<SB>(b, mask, cx0, cy0, nx, ny, B, M);
  for (int e = threadIdx.x; e < 34 * 34; e += 256) {
    const int i = cx0 - 1 + e % 34, j = cy0 - 1 + e / 34;
    float val = 0.f;
    if (i >= 0 && j >= 0 && i < nx && j < ny) {
      const size_t c = (size_t)i + (size_t)j * nx;
      val = ld<SX>(src, c);
      if (mask_type(mask[c]) == AGUA) val += prolong_at<SC>(cx, cmk, cnx, cny, i, j);
    }
    S[e] = val;
    W[e] = val;
  }
  __syncthreads();
  pass1_sweeps(S, W, B, M, cx0, cy0, nx, ny, 1, 0);
  bool in;
  const float v = pass2_sweeps(W, L, B, M, cx0, cy0, nx, ny, 1, 0, &in);
  const int i = ox2 + (threadIdx.x & 15), j = oy2 + (threadIdx.x >> 4);
  if (!ZD) {
    if (in) st<SX>(dst, (size_t)i + (size_t)j * nx, v);
    return;
  }
  // ---- z and the dots (level 0)
  double dv[2] = {0, 0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  const int na = sc->active_bubbles;
  if (in) {
    const size_t c = (size_t)i + (size_t)j * nx;
    const float xv = v * (za.mixed ? 1.f / sc->mg_scale : 1.f);
    float z = 0.f;
    if (mask_type(za.opMask[c]) == AGUA) {
      z = xv;
      dv[0] = (double)za.r[c] * z;
      dv[1] = (double)za.r[c] * za.zold[c];
    } else {
      uint32_t lb;
      if (is_bub_cell(za.tipo, za.rot, za.facesB, za.ar, c, &lb))
        bub_add(sl, za.labelSlot, bv(za.bvec, BV_ACC1), lb, (double)xv / za.cellsB[lb]);
    }
    za.znew[c] = z;
  }
  write_partials<2>(dv, za.partials, sh);
  if (na > 0) write_partials<BSLOTS>(sl, za.bpart, sh);
}

// ================================================================== restriction (restringir, pressao-cpu.js:159-188)
// Coarse tile 16x16: the fine residual r = b - A x on the 34x34 fine cells it gathers from
// (2I-1 .. 2I+2 per axis, weights 1/4 3/4 3/4 1/4), from x on 36x36, all in LDS.
template <class SXF, class SBF, class SBC>
__global__ __launch_bounds__(256) void k_restrict(const SXF* __restrict__ fx, const SBF* __restrict__ fb,
                                                  const uint8_t* __restrict__ fm, int fnx, int fny, SBC* __restrict__ cb,
                                                  const uint8_t* __restrict__ cmk, int cnx, int cny,
                                                  const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ float X[36 * 36], R[34 * 34];
  const int CI0 = blockIdx.x * 16, CJ0 = blockIdx.y * 16;
  const int fx0 = 2 * CI0 - 2, fy0 = 2 * CJ0 - 2;  // X origin
  for (int e = threadIdx.x; e < 36 * 36; e += 256) {
    const int i = fx0 + e % 36, j = fy0 + e / 36;
    X[e] = (i >= 0 && j >= 0 && i < fnx && j < fny) ? ld<SXF>(fx, (size_t)i + (size_t)j * fnx) : 0.f;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 34 * 34; e += 256) {
    const int lx = e % 34, ly = e / 34;
    const int i = fx0 + 1 + lx, j = fy0 + 1 + ly;
    float rr = 0.f;
    if (i >= 0 && j >= 0 && i < fnx && j < fny) {
      const size_t c = (size_t)i + (size_t)j * fnx;
      const uint8_t m = fm[c];
      if (mask_type(m) == AGUA) {
        const int ex = (lx + 1) + (ly + 1) * 36;
        float s = 0.f;
        if (m & 1) s += X[ex - 1];
        if (m & 2) s += X[ex + 1];
        if (m & 4) s += X[ex - 36];
        if (m & 8) s += X[ex + 36];
        rr = ld<SBF>(fb, c) - ((float)mask_cnt(m) * X[ex] - s);
      }
    }
    R[e] = rr;
  }
  __syncthreads();
  const int I = CI0 + (threadIdx.x & 15), J = CJ0 + (threadIdx.x >> 4);
  if (I >= cnx || J >= cny) return;
  const size_t Cc = (size_t)I + (size_t)J * cnx;
  if (mask_type(cmk[Cc]) != AGUA) { st<SBC>(cb, Cc, 0.f); return; }
  const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
  // fine cell 2I-1+bi sits at R index (2I-1+bi) - (fx0+1) = 2(I-CI0) + bi
  const int rx = 2 * (I - CI0), ry = 2 * (J - CJ0);
  float acc = 0.f;
  for (int bj = 0; bj < 4; bj++) {
    float racc = 0.f;
    for (int bi = 0; bi < 4; bi++) racc += R[(rx + bi) + (ry + bj) * 34] * w[bi];
    acc += racc * w[bj];
  }
  st<SBC>(cb, Cc, acc);
}

// ================================================================== coarse levels in one workgroup
struct WGPlan {
  int lc, L;
  int nx[12], ny[12];
  int offx[12], offt[12], offb[12], offm[12];  // byte offsets into the LDS pool
  const uint8_t* mask[12];
};

// one smoothing pass of an LDS-resident level: S -> W (each wave owns whole tiles and syncs
// only itself inside a tile); the caller swaps S and W afterwards
__device__ inline void wg_smooth(const float* S, float* W, const float* B, const uint8_t* M, int nx, int ny, int desl,
                                 int c0, int c1) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
  const int tnx = (nx + desl + 15) / 16, tny = (ny + desl + 15) / 16;
  for (int t = wave; t < tnx * tny; t += nw) {
    const int ox = (t % tnx) * 16 - desl, oy = (t / tnx) * 16 - desl;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = lane + 64 * k, i = ox + (e & 15), j = oy + (e >> 4);
      if (i >= 0 && j >= 0 && i < nx && j < ny) W[i + j * nx] = S[i + j * nx];
    }
    wave_sync();
    for (int it = 0; it < 4; it++)
      for (int o = 0; o < 2; o++) {
        const int cor = o ? c1 : c0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const int e = lane + 64 * k, i = ox + (e & 15), j = oy + (e >> 4);
          if (i < 0 || j < 0 || i >= nx || j >= ny || ((i + j + 1) & 1) != cor) continue;
          const int c = i + j * nx;
          const uint8_t m = M[c];
          const int kk = mask_cnt(m);
          if (mask_type(m) != AGUA || kk == 0) continue;
          const bool li = i - 1 >= ox, ri = i + 1 < ox + 16, di = j - 1 >= oy, ui = j + 1 < oy + 16;
          float s = 0.f;
          if (m & 1) s += li ? W[c - 1] : S[c - 1];
          if (m & 2) s += ri ? W[c + 1] : S[c + 1];
          if (m & 4) s += di ? W[c - nx] : S[c - nx];
          if (m & 8) s += ui ? W[c + nx] : S[c + nx];
          const float qv = W[c];
          W[c] = qv + ((s + B[c]) / (float)kk - qv);
        }
        wave_sync();
      }
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int e = lane + 64 * k, i = ox + (e & 15), j = oy + (e >> 4);
      if (i >= 0 && j >= 0 && i < nx && j < ny && mask_type(M[i + j * nx]) != AGUA) W[i + j * nx] = 0.f;
    }
    wave_sync();
  }
  __syncthreads();
}

__global__ __launch_bounds__(1024) void k_coarse_wg(WGPlan P, const float* __restrict__ b_lc, float* __restrict__ x_lc,
                                                    const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ __attribute__((aligned(16))) char pool[WG_LDS];
  // level arrays live at plan offsets; bit l of `sw` says x and tmp of level l are swapped
  unsigned sw = 0;
  auto Xp = [&](int l) { return (float*)(pool + (((sw >> l) & 1) ? P.offt[l] : P.offx[l])); };
  auto Tp = [&](int l) { return (float*)(pool + (((sw >> l) & 1) ? P.offx[l] : P.offt[l])); };
  auto Bp = [&](int l) { return (float*)(pool + P.offb[l]); };
  auto Mp = [&](int l) { return (uint8_t*)(pool + P.offm[l]); };
  for (int l = P.lc; l < P.L; l++) {
    const int n = P.nx[l] * P.ny[l];
    float* X = Xp(l);
    float* B = Bp(l);
    uint8_t* M = Mp(l);
    const uint8_t* gm = P.mask[l];
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
      M[c] = gm[c];
      X[c] = 0.f;
      if (l == P.lc) B[c] = b_lc[c];
    }
  }
  __syncthreads();
  // ---- down
  for (int l = P.lc; l < P.L - 1; l++) {
    const int nx = P.nx[l], ny = P.ny[l];
    wg_smooth(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 0, 0, 1);
    sw ^= 1u << l;
    wg_smooth(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 8, 0, 1);
    sw ^= 1u << l;
    // restriction to l+1 (b = P^T r, x = 0)
    const float* X = Xp(l);
    const float* B = Bp(l);
    const uint8_t* M = Mp(l);
    float* Bc = Bp(l + 1);
    float* Xc = Xp(l + 1);
    const uint8_t* Mc = Mp(l + 1);
    const int cnx = P.nx[l + 1], cny = P.ny[l + 1];
    for (int C = threadIdx.x; C < cnx * cny; C += blockDim.x) {
      const int I = C % cnx, J = C / cnx;
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
            const int c = i + j * nx;
            const uint8_t m = M[c];
            if (mask_type(m) != AGUA) continue;
            float s = 0.f;
            if (m & 1) s += X[c - 1];
            if (m & 2) s += X[c + 1];
            if (m & 4) s += X[c - nx];
            if (m & 8) s += X[c + nx];
            racc += (B[c] - ((float)mask_cnt(m) * X[c] - s)) * w[bi];
          }
          acc += racc * w[bj];
        }
      }
      Bc[C] = acc;
      Xc[C] = 0.f;
    }
    __syncthreads();
  }
  // ---- coarsest: 64 red-black sweeps in place, colour order alternating (resolverGrosso)
  {
    const int l = P.L - 1, nx = P.nx[l], ny = P.ny[l], n = nx * ny;
    float* x = Xp(l);
    const float* b = Bp(l);
    const uint8_t* M = Mp(l);
    const bool single_wave = n <= 64 * 32;
    if (!single_wave || threadIdx.x < 64) {
      const int stride = single_wave ? 64 : (int)blockDim.x;
      for (int k = 0; k < 64; k++)
        for (int o = 0; o < 2; o++) {
          const int cor = (k & 1) ? 1 - o : o;
          for (int c = threadIdx.x; c < n; c += stride) {
            const int i = c % nx, j = c / nx;
            const uint8_t m = M[c];
            if (mask_type(m) != AGUA || ((i + j + 1) & 1) != cor) continue;
            const int q = mask_cnt(m);
            if (!q) continue;
            float s = 0.f;
            if (m & 1) s += x[c - 1];
            if (m & 2) s += x[c + 1];
            if (m & 4) s += x[c - nx];
            if (m & 8) s += x[c + nx];
            x[c] = (s + b[c]) / (float)q;
          }
          if (single_wave) wave_sync(); else __syncthreads();
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
      const int i = c % nx, j = c / nx;
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
    wg_smooth(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 8, 1, 0);
    sw ^= 1u << l;
    wg_smooth(Xp(l), Tp(l), Bp(l), Mp(l), nx, ny, 0, 1, 0);
    sw ^= 1u << l;
  }
  const int n = P.nx[P.lc] * P.ny[P.lc];
  const float* X = Xp(P.lc);
  for (int c = threadIdx.x; c < n; c += blockDim.x) x_lc[c] = X[c];
}

// z and the dots when level 0 itself is solved inside the workgroup (small grids)
__global__ __launch_bounds__(256) void k_zdots0(const float* __restrict__ x0, ZArgs za, int C,
                                                const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[BSLOTS * 16];
  const int na = sc->active_bubbles;
  double dv[2] = {0, 0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  GRID_STRIDE(c, C) {
    const float xv = x0[c];
    float z = 0.f;
    if (mask_type(za.opMask[c]) == AGUA) {
      z = xv;
      dv[0] += (double)za.r[c] * z;
      dv[1] += (double)za.r[c] * za.zold[c];
    } else {
      uint32_t lb;
      if (is_bub_cell(za.tipo, za.rot, za.facesB, za.ar, c, &lb))
        bub_add(sl, za.labelSlot, bv(za.bvec, BV_ACC1), lb, (double)xv / za.cellsB[lb]);
    }
    za.znew[c] = z;
  }
  write_partials<2>(dv, za.partials, sh);
  if (na > 0) write_partials<BSLOTS>(sl, za.bpart, sh);
}

__global__ __launch_bounds__(256) void k_fin_zdots(double* __restrict__ bvec, const int* __restrict__ activeList,
                                                   const int* __restrict__ labelSlot, const double* __restrict__ partials,
                                                   const double* __restrict__ bpart, int nb, int first,
                                                   DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[BSLOTS * 16], res[BSLOTS], slotTot[BSLOTS];
  const int na = sc->active_bubbles;
  double tot[2];
  read_partials<2>(partials, nb, tot, sh, res);
  double stt[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (na > 0) read_partials<BSLOTS>(bpart, nb, stt, sh, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < BSLOTS; k++) slotTot[k] = stt[k];
  __syncthreads();
  double* rb = bv(bvec, BV_R);
  double* zb = bv(bvec, BV_Z);
  double* zvb = bv(bvec, BV_ZV);
  double* acc1 = bv(bvec, BV_ACC1);
  double bacc[2] = {0, 0};
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const int s = labelSlot[b];
    double znew;
    if (s >= 0) znew = slotTot[s];
    else { znew = acc1[b]; acc1[b] = 0; }
    zvb[b] = zb[b];
    zb[b] = znew;
    bacc[0] += rb[b] * zb[b];
    bacc[1] += rb[b] * zvb[b];
  }
  block_sum<2>(bacc, sh);
  if (threadIdx.x == 0) {
    const double rzNovo = tot[0] + bacc[0];
    const double cruzado = tot[1] + bacc[1];
    if (first) sc->beta = 0;
    else {
      const double rz = sc->rz;
      sc->beta = fabs(rz) < 1e-30 ? 0 : fmax(0.0, (rzNovo - cruzado) / rz);
    }
    sc->rz = rzNovo;
  }
}

// ================================================================== p = z + beta p, Ap, p.Ap
__global__ __launch_bounds__(256) void k_spmv(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                              const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                              const float* __restrict__ z, const float* __restrict__ pold,
                                              float* __restrict__ pnew, float* __restrict__ ap, int nx, int ny, int ar,
                                              double* __restrict__ bvec, const int* __restrict__ labelSlot,
                                              double* __restrict__ partials, double* __restrict__ bpart,
                                              const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[BSLOTS * 16];
  const int C = nx * ny;
  const double beta = sc->beta;
  const int na = sc->active_bubbles;
  const double* zb = bv(bvec, BV_Z);
  const double* pb = bv(bvec, BV_P);
  double* acc0 = bv(bvec, BV_ACC0);
  double v[1] = {0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto peff = [&](size_t d) -> float {
    if (mask_type(opMask[d]) == AGUA) return (float)((double)z[d] + beta * (double)pold[d]);
    uint32_t lb;
    if (is_bub_cell(tipo, rot, facesB, ar, d, &lb)) return (float)(zb[lb] + beta * pb[lb]);
    return 0.f;
  };
  GRID_STRIDE(c, C) {
    const float pc = peff(c);
    pnew[c] = pc;
    const uint8_t m = opMask[c];
    if (mask_type(m) == AGUA) {
      double s = 0;
      const size_t nb[4] = {(size_t)c - 1, (size_t)c + 1, (size_t)c - nx, (size_t)c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        s += (double)peff(nb[k]);
        uint32_t lb;
        if (na > 0 && is_bub_cell(tipo, rot, facesB, ar, nb[k], &lb) &&
            interior((int)(nb[k] % nx), (int)(nb[k] / nx), nx, ny))
          bub_add(sl, labelSlot, acc0, lb, -(double)pc);
      }
      const float a = (float)(mask_cnt(m) * (double)pc - s);
      ap[c] = a;
      v[0] += (double)pc * a;
    } else ap[c] = 0.f;
  }
  write_partials<1>(v, partials, sh);
  if (na > 0) write_partials<BSLOTS>(sl, bpart, sh);
}

__global__ __launch_bounds__(256) void k_fin_spmv(const uint32_t* __restrict__ facesB, double* __restrict__ bvec,
                                                  const int* __restrict__ activeList, const int* __restrict__ labelSlot,
                                                  const double* __restrict__ partials, const double* __restrict__ bpart,
                                                  int nb, DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[BSLOTS * 16], res[BSLOTS], slotTot[BSLOTS];
  __shared__ double alpha_s;
  const int na = sc->active_bubbles;
  double tot[1];
  read_partials<1>(partials, nb, tot, sh, res);
  double stt[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (na > 0) read_partials<BSLOTS>(bpart, nb, stt, sh, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < BSLOTS; k++) slotTot[k] = stt[k];
  __syncthreads();
  const double beta = sc->beta;
  double* zb = bv(bvec, BV_Z);
  double* pb = bv(bvec, BV_P);
  double* apb = bv(bvec, BV_AP);
  double* acc0 = bv(bvec, BV_ACC0);
  double bacc[1] = {0};
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const int s = labelSlot[b];
    double accB;
    if (s >= 0) accB = slotTot[s];
    else { accB = acc0[b]; acc0[b] = 0; }
    const double pn = zb[b] + beta * pb[b];
    pb[b] = pn;
    apb[b] = (double)facesB[b] * pn + accB;
    bacc[0] += pn * apb[b];
  }
  block_sum<1>(bacc, sh);
  if (threadIdx.x == 0) {
    const double pap = tot[0] + bacc[0];
    alpha_s = fabs(pap) < 1e-30 ? 0 : sc->rz / pap;
    sc->alpha = alpha_s;
  }
  __syncthreads();
  const double alpha = alpha_s;
  double* yb = bv(bvec, BV_Y);
  double* rb = bv(bvec, BV_R);
  double rbs[1] = {0};
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    yb[b] += alpha * pb[b];
    rb[b] -= alpha * apb[b];
    rbs[0] += rb[b] * rb[b];
  }
  block_sum<1>(rbs, sh);
  if (threadIdx.x == 0) sc->pap = rbs[0];  // bubble part of |r|^2 for fin_update
}

// ================================================================== q += alpha p, r -= alpha Ap
__global__ __launch_bounds__(256) void k_update(const uint8_t* __restrict__ opMask, const float* __restrict__ p,
                                                const float* __restrict__ ap, float* __restrict__ q,
                                                float* __restrict__ r, int C, double* __restrict__ partials,
                                                const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16];
  const double alpha = sc->alpha;
  double v[1] = {0};
  GRID_STRIDE(c, C) {
    if (mask_type(opMask[c]) != AGUA) continue;
    q[c] = (float)((double)q[c] + alpha * (double)p[c]);
    const float rn = (float)((double)r[c] - alpha * (double)ap[c]);
    r[c] = rn;
    v[0] += (double)rn * rn;
  }
  write_partials<1>(v, partials, sh);
}

__global__ __launch_bounds__(256) void k_fin_update(const double* __restrict__ partials, int nb,
                                                    DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16], res[1];
  double tot[1];
  read_partials<1>(partials, nb, tot, sh, res);
  if (threadIdx.x == 0) {
    const double n2 = tot[0] + sc->pap;
    const double nr = sqrt(n2);
    sc->norm2 = n2;
    sc->norm_last = nr;
    sc->iters += 1;
    sc->total_iters += 1;
    bool stop = sc->iters >= sc->maxit;
    if (sc->tol > 0 && (nr == 0 || nr <= sc->tol * sc->normb)) stop = true;
    sc->done = stop ? 1 : 0;
    sc->mg_scale = nr > 0 ? (float)(1.0 / nr) : 1.f;
  }
}

// ================================================================== end of the solve
__global__ __launch_bounds__(256) void k_pcg_final(const uint8_t* __restrict__ tipo, float* __restrict__ q, int C) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C && tipo[c] != AGUA) q[c] = 0.f;
}
__global__ __launch_bounds__(1024) void k_pcg_final_b(const double* __restrict__ bvec, const int* __restrict__ activeList,
                                                      float* __restrict__ yB32, const DevScalars* __restrict__ sc) {
  for (int k = threadIdx.x; k < sc->active_bubbles; k += blockDim.x) {
    const int b = activeList[k];
    yB32[b] = (float)bv((double*)bvec, BV_Y)[b];
  }
}

// true residual ||b' - A[q;y]|| (diagnostics; b' is the projected rhs that was solved)
__global__ __launch_bounds__(256) void k_true_res(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                                  const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                                  const float* __restrict__ div, const float* __restrict__ q, int nx,
                                                  int ny, int ar, double* __restrict__ bvec,
                                                  const int* __restrict__ labelSlot, double* __restrict__ partials,
                                                  double* __restrict__ bpart, const DevScalars* __restrict__ sc) {
  __shared__ double sh[BSLOTS * 16];
  const int C = nx * ny;
  const int na = sc->active_bubbles;
  const double* yb = bv(bvec, BV_Y);
  double* acc0 = bv(bvec, BV_ACC0);
  const double mean = sc->proj_mean;
  double v[1] = {0};
  double sl[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  auto qe = [&](size_t d) -> float {
    if (tipo[d] == AGUA) return q[d];
    uint32_t lb;
    if (is_bub_cell(tipo, rot, facesB, ar, d, &lb)) return (float)yb[lb];
    return 0.f;
  };
  GRID_STRIDE(c, C) {
    const uint8_t m = opMask[c];
    if (mask_type(m) != AGUA) continue;
    const float qc = qe(c);
    double s = 0;
    const size_t nb[4] = {(size_t)c - 1, (size_t)c + 1, (size_t)c - nx, (size_t)c + nx};
    for (int k = 0; k < 4; k++) {
      if (!(m & (1 << k))) continue;
      s += (double)qe(nb[k]);
      uint32_t lb;
      if (is_bub_cell(tipo, rot, facesB, ar, nb[k], &lb) && interior((int)(nb[k] % nx), (int)(nb[k] / nx), nx, ny))
        bub_add(sl, labelSlot, acc0, lb, -(double)qc);
    }
    const double e = (-(double)div[c] - mean) - (mask_cnt(m) * (double)qc - s);
    v[0] += e * e;
  }
  write_partials<1>(v, partials, sh);
  if (na > 0) write_partials<BSLOTS>(sl, bpart, sh);
}

__global__ __launch_bounds__(256) void k_fin_true_res(const uint32_t* __restrict__ facesB, double* __restrict__ bvec,
                                                      const double* __restrict__ rhsB, const int* __restrict__ activeList,
                                                      const int* __restrict__ labelSlot,
                                                      const double* __restrict__ partials,
                                                      const double* __restrict__ bpart, int nb,
                                                      const DevScalars* __restrict__ sc, double* __restrict__ out) {
  __shared__ double sh[BSLOTS * 16], res[BSLOTS], slotTot[BSLOTS];
  const int na = sc->active_bubbles;
  double tot[1];
  read_partials<1>(partials, nb, tot, sh, res);
  double stt[BSLOTS] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (na > 0) read_partials<BSLOTS>(bpart, nb, stt, sh, res);
  if (threadIdx.x == 0)
    for (int k = 0; k < BSLOTS; k++) slotTot[k] = stt[k];
  __syncthreads();
  const double* yb = bv(bvec, BV_Y);
  double* acc0 = bv(bvec, BV_ACC0);
  double bacc[1] = {0};
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const int s = labelSlot[b];
    double accB;
    if (s >= 0) accB = slotTot[s];
    else { accB = acc0[b]; acc0[b] = 0; }
    const double e = (rhsB[b] - sc->proj_mean) - ((double)facesB[b] * yb[b] + accB);
    bacc[0] += e * e;
  }
  block_sum<1>(bacc, sh);
  if (threadIdx.x == 0) *out = sqrt(tot[0] + bacc[0]);
}

// ================================================================== host side
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
  // first level kept in LDS: the suffix of levels whose arrays fit the workgroup pool
  const int L = (int)s->lv.size();
  long cells = 0, bytes = 0;
  int lc = L;
  for (int l = L - 1; l >= 0; l--) {
    const long add = 12L * s->lv[l].n + ((s->lv[l].n + 15) & ~15);
    if (bytes + add > WG_LDS) break;
    bytes += add;
    cells += s->lv[l].n;
    lc = l;
  }
  if (lc == L) {
    set_error("grid too elongated: the coarsest multigrid level does not fit the one-workgroup solver");
    return TQ_E_CAPACITY;
  }
  s->lc = std::max(lc, std::min(s->d.mg_min_lds_level, L - 1));
  TQ_HIP(hipMalloc(&s->bpart, sizeof(double) * BSLOTS * MAX_RED_BLOCKS));
  TQ_HIP(hipMalloc(&s->labelSlot, sizeof(int) * MAX_BOLHAS));
  return 0;
}

void solver_free(tq_sim* s) {
  for (auto& L : s->lv) {
    (void)hipFree(L.mask); (void)hipFree(L.b); (void)hipFree(L.x); (void)hipFree(L.tmp);
  }
  s->lv.clear();
  if (s->bpart) (void)hipFree(s->bpart);
  if (s->labelSlot) (void)hipFree(s->labelSlot);
  s->bpart = nullptr;
  s->labelSlot = nullptr;
}

// per-level storage for this solve (mixed: level 0 keeps x in FP32, large levels use FP16)
static void plan_precision(tq_sim* s) {
  const bool mixed = s->d.precision == TQ_PREC_MIXED;
  for (int l = 0; l < (int)s->lv.size(); l++) {
    Level& L = s->lv[l];
    const bool big = mixed && l < s->lc && L.n >= 4096;
    L.hb = big;
    L.hx = big && l > 0;
  }
}

static int nbr_blocks(int C) { return std::max(1, std::min(NBR, (C + 255) / 256)); }

static void launch_down(tq_sim* s, int l) {
  Level& F = s->lv[l];
  dim3 g((F.nx + 8 + 15) / 16, (F.ny + 8 + 15) / 16);
  int t = time_begin(s, l == 0 ? "mg_down_l0" : "mg_down", ((F.hx ? 2 : 4) + (F.hb ? 2 : 4) + 1) * (double)F.n);
  if (!F.hx && !F.hb)
    hipLaunchKernelGGL((k_down<float, float>), g, dim3(256), 0, s->st, (float*)F.x, (const float*)F.b, F.mask, F.nx,
                       F.ny, s->sc);
  else if (!F.hx)
    hipLaunchKernelGGL((k_down<float, __half>), g, dim3(256), 0, s->st, (float*)F.x, (const __half*)F.b, F.mask, F.nx,
                       F.ny, s->sc);
  else
    hipLaunchKernelGGL((k_down<__half, __half>), g, dim3(256), 0, s->st, (__half*)F.x, (const __half*)F.b, F.mask, F.nx,
                       F.ny, s->sc);
  time_end(s, t);
}

template <class SXF, class SBF>
static void restrict_c(tq_sim* s, Level& F, Level& G, dim3 g) {
  if (G.hb)
    hipLaunchKernelGGL((k_restrict<SXF, SBF, __half>), g, dim3(256), 0, s->st, (const SXF*)F.x, (const SBF*)F.b, F.mask,
                       F.nx, F.ny, (__half*)G.b, G.mask, G.nx, G.ny, s->sc);
  else
    hipLaunchKernelGGL((k_restrict<SXF, SBF, float>), g, dim3(256), 0, s->st, (const SXF*)F.x, (const SBF*)F.b, F.mask,
                       F.nx, F.ny, (float*)G.b, G.mask, G.nx, G.ny, s->sc);
}

static void launch_restrict(tq_sim* s, int l) {
  Level& F = s->lv[l];
  Level& G = s->lv[l + 1];
  dim3 g((G.nx + 15) / 16, (G.ny + 15) / 16);
  int t = time_begin(s, l == 0 ? "mg_restrict_l0" : "mg_restrict",
                     ((F.hx ? 2 : 4) + (F.hb ? 2 : 4) + 1) * (double)F.n + (G.hb ? 2 : 4) * (double)G.n);
  if (!F.hx && !F.hb) restrict_c<float, float>(s, F, G, g);
  else if (!F.hx) restrict_c<float, __half>(s, F, G, g);
  else restrict_c<__half, __half>(s, F, G, g);
  time_end(s, t);
}

template <class SX, class SB, class SC>
static void up_c(tq_sim* s, Level& F, Level& G, dim3 g, bool zd, const ZArgs& za) {
  if (zd)
    hipLaunchKernelGGL((k_up<SX, SB, SC, true>), g, dim3(256), 0, s->st, (const SX*)F.x, (SX*)F.tmp, (const SB*)F.b,
                       F.mask, F.nx, F.ny, (const SC*)G.x, G.mask, G.nx, G.ny, za, s->sc);
  else
    hipLaunchKernelGGL((k_up<SX, SB, SC, false>), g, dim3(256), 0, s->st, (const SX*)F.x, (SX*)F.tmp, (const SB*)F.b,
                       F.mask, F.nx, F.ny, (const SC*)G.x, G.mask, G.nx, G.ny, za, s->sc);
}

static ZArgs zargs(tq_sim* s) {
  ZArgs za;
  za.opMask = s->opMask; za.tipo = s->tipo; za.rot = s->rot; za.facesB = s->facesB; za.cellsB = s->cellsB;
  za.labelSlot = s->labelSlot; za.r = s->r; za.zold = s->zbuf[s->zc]; za.znew = s->zbuf[1 - s->zc];
  za.bvec = s->bvec; za.partials = s->partials; za.bpart = s->bpart; za.ar = s->d.ar;
  za.mixed = s->lv[0].hb ? 1 : 0;
  return za;
}

static void launch_up(tq_sim* s, int l) {
  Level& F = s->lv[l];
  Level& G = s->lv[l + 1];
  dim3 g((F.nx + 15) / 16, (F.ny + 15) / 16);
  const bool zd = l == 0;
  const ZArgs za = zargs(s);
  int t = time_begin(s, l == 0 ? "mg_up_l0" : "mg_up",
                     (2 * (F.hx ? 2 : 4) + (F.hb ? 2 : 4) + 1) * (double)F.n + (G.hx ? 2 : 4) * (double)G.n);
  const bool gx = G.hx;  // FP16 coarse x (the LDS-resident level always hands over FP32)
  if (!F.hx && !F.hb) {
    if (gx) up_c<float, float, __half>(s, F, G, g, zd, za); else up_c<float, float, float>(s, F, G, g, zd, za);
  } else if (!F.hx) {
    if (gx) up_c<float, __half, __half>(s, F, G, g, zd, za); else up_c<float, __half, float>(s, F, G, g, zd, za);
  } else {
    if (gx) up_c<__half, __half, __half>(s, F, G, g, zd, za); else up_c<__half, __half, float>(s, F, G, g, zd, za);
  }
  time_end(s, t);
  if (!zd) std::swap(F.x, F.tmp);
}

static void launch_coarse_wg(tq_sim* s) {
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
  Level& G = s->lv[s->lc];
  int t = time_begin(s, "mg_coarse_wg", 9.0 * G.n);
  hipLaunchKernelGGL(k_coarse_wg, dim3(1), dim3(1024), 0, s->st, P, (const float*)G.b, (float*)G.x, s->sc);
  time_end(s, t);
}

static void precondition(tq_sim* s, int first) {
  const int C = s->C;
  Level& L0 = s->lv[0];
  int t = time_begin(s, "pcg_mg_input", (1 + 4 + (L0.hb ? 2 : 4)) * (double)C);
  if (L0.hb)
    hipLaunchKernelGGL(k_mg_input<__half>, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot,
                       s->facesB, s->cellsB, s->r, C, s->d.ar, s->bvec, first, 1, (__half*)L0.b, s->sc);
  else
    hipLaunchKernelGGL(k_mg_input<float>, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot,
                       s->facesB, s->cellsB, s->r, C, s->d.ar, s->bvec, first, 0, (float*)L0.b, s->sc);
  time_end(s, t);
  for (int l = 0; l < s->lc; l++) {
    launch_down(s, l);
    launch_restrict(s, l);
  }
  launch_coarse_wg(s);
  int nb;
  if (s->lc > 0) {
    for (int l = s->lc - 1; l >= 0; l--) launch_up(s, l);
    nb = ((L0.nx + 15) / 16) * ((L0.ny + 15) / 16);
  } else {
    nb = nbr_blocks(C);
    t = time_begin(s, "pcg_zdots", 16.0 * C);
    hipLaunchKernelGGL(k_zdots0, dim3(nb), dim3(256), 0, s->st, (const float*)L0.x, zargs(s), C, s->sc);
    time_end(s, t);
  }
  s->zc = 1 - s->zc;
  t = time_begin(s, "pcg_fin_zdots", 0);
  hipLaunchKernelGGL(k_fin_zdots, dim3(1), dim3(256), 0, s->st, s->bvec, s->activeList, s->labelSlot, s->partials,
                     s->bpart, nb, first, s->sc);
  time_end(s, t);
}

static void pcg_iteration(tq_sim* s) {
  const int C = s->C;
  const int nb = nbr_blocks(C);
  const int pn = 1 - s->pc;
  int t = time_begin(s, "pcg_spmv", (1 + 4 + 4 + 4 + 4) * (double)C);
  hipLaunchKernelGGL(k_spmv, dim3(nb), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot, s->facesB, s->zbuf[s->zc],
                     s->pbuf[s->pc], s->pbuf[pn], s->ap, s->nx, s->ny, s->d.ar, s->bvec, s->labelSlot, s->partials,
                     s->bpart, s->sc);
  time_end(s, t);
  t = time_begin(s, "pcg_fin_spmv", 0);
  hipLaunchKernelGGL(k_fin_spmv, dim3(1), dim3(256), 0, s->st, s->facesB, s->bvec, s->activeList, s->labelSlot,
                     s->partials, s->bpart, nb, s->sc);
  time_end(s, t);
  s->pc = pn;
  t = time_begin(s, "pcg_update", (1 + 8 + 8 + 4 + 4) * (double)C);
  hipLaunchKernelGGL(k_update, dim3(nb), dim3(256), 0, s->st, s->opMask, s->pbuf[s->pc], s->ap, s->q, s->r, C,
                     s->partials, s->sc);
  time_end(s, t);
  t = time_begin(s, "pcg_fin_update", 0);
  hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(256), 0, s->st, s->partials, nb, s->sc);
  time_end(s, t);
  precondition(s, 0);
}

int pressure_solve(tq_sim* s) {
  const int C = s->C;
  const int L = (int)s->lv.size();
  plan_precision(s);
  // ---- setup: masks of every level, bubble vectors and slots
  int t = time_begin(s, "pcg_setup", (1 + 4 + 1 + 1) * (double)C);
  hipLaunchKernelGGL(k_setup0, dim3((C + 255) / 256), dim3(256), 0, s->st, s->tipo, s->rot, s->nx, s->ny, s->d.ar,
                     s->opMask, s->lv[0].mask);
  for (int l = 1; l < L; l++) {
    const Level& F = s->lv[l - 1];
    const Level& G = s->lv[l];
    hipLaunchKernelGGL(k_coarse_type, dim3((G.n + 255) / 256), dim3(256), 0, s->st, F.mask, F.nx, F.ny, G.mask, G.nx, G.ny);
    hipLaunchKernelGGL(k_viz, dim3((G.n + 255) / 256), dim3(256), 0, s->st, G.mask, G.nx, G.ny);
  }
  hipLaunchKernelGGL(k_bub_init, dim3(1), dim3(1024), 0, s->st, s->facesB, s->yB32, s->d.ar, s->bvec, s->activeList,
                     s->labelSlot, s->sc);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  const double tol = s->d.tol;
  const int maxit = s->d.iteracoes;
  TQ_HIP(hipMemcpyAsync(&s->sc->tol, &tol, sizeof(double), hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(&s->sc->maxit, &maxit, sizeof(int), hipMemcpyHostToDevice, s->st));
  const int nb = nbr_blocks(C);
  t = time_begin(s, "pcg_init", (1 + 1 + 4 + 4 + 4 + 4 + 4) * (double)C);
  hipLaunchKernelGGL(k_pcg_init, dim3(nb), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot, s->facesB, s->div, s->q, s->r,
                     s->pbuf[s->pc], s->nx, s->ny, s->d.ar, s->bvec, s->labelSlot, s->partials, s->bpart, s->sc);
  hipLaunchKernelGGL(k_fin_init, dim3(1), dim3(256), 0, s->st, s->facesB, s->bvec, s->rhsB, s->activeList, s->labelSlot,
                     s->partials, s->bpart, nb, s->sc);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  precondition(s, 1);
  TQ_LAUNCH_CHECK();
  if (tol <= 0) {
    for (int it = 0; it < maxit; it++) pcg_iteration(s);
    TQ_LAUNCH_CHECK();
  } else {
    // launch in chunks and poll the device stop flag (kernels after the stop exit at once)
    int launched = 0;
    int chunk = std::max(2, s->sc_host->iters > 0 ? s->sc_host->iters : 8);
    while (launched < maxit) {
      const int n = std::min(chunk, maxit - launched);
      for (int k = 0; k < n; k++) pcg_iteration(s);
      launched += n;
      TQ_LAUNCH_CHECK();
      TQ_HIP(hipMemcpyAsync(s->sc_host, s->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s->st));
      TQ_HIP(hipStreamSynchronize(s->st));
      if (s->sc_host->done) break;
      chunk = 4;
    }
  }
  hipLaunchKernelGGL(k_pcg_final, dim3((C + 255) / 256), dim3(256), 0, s->st, s->tipo, s->q, C);
  hipLaunchKernelGGL(k_pcg_final_b, dim3(1), dim3(1024), 0, s->st, s->bvec, s->activeList, s->yB32, s->sc);
  TQ_LAUNCH_CHECK();
  return 0;
}

int true_residual(tq_sim* s, double* out) {
  const int C = s->C;
  const int nb = nbr_blocks(C);
  double* dres = s->partials + 8 * (size_t)MAX_RED_BLOCKS - 1;
  hipLaunchKernelGGL(k_true_res, dim3(nb), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot, s->facesB, s->div, s->q,
                     s->nx, s->ny, s->d.ar, s->bvec, s->labelSlot, s->partials, s->bpart, s->sc);
  hipLaunchKernelGGL(k_fin_true_res, dim3(1), dim3(256), 0, s->st, s->facesB, s->bvec, s->rhsB, s->activeList,
                     s->labelSlot, s->partials, s->bpart, nb, s->sc, dres);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(out, dres, sizeof(double), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

}  // namespace tq
