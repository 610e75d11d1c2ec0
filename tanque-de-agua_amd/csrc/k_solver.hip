// k_solver.hip -- pressure projection: PCG preconditioned by a multigrid V-cycle, with one
// extra unknown per enclosed air pocket (resolverPCG / cicloV, js/pressao-cpu.js:241-402).
//
// CG vectors r, z, p, Ap are FP32 per cell; every dot product, the residual norm and the
// bubble unknowns (y, r, z, p, Ap per label) are FP64.  Reductions are deterministic: each
// block stores its partial and the last block to finish sums them in block order and also
// performs the per-bubble scalar work.  The multigrid vectors are FP32 or, in the mixed mode,
// FP16 storage (per-cycle scaled by 1/||r||) on the large levels with FP32 arithmetic.
//
// Kernel sequence of one PCG iteration (every kernel exits at once when the solve is done):
//   spmv   p = z + beta p (fused), Ap, p.Ap, bubble rows      -> alpha (last block)
//   update q += alpha p, r -= alpha Ap, |r|^2                 -> norm, stop test
//   input  MG right-hand side (r, bubble residual spread over its cells)
//   V-cycle (2 smoothing passes, restriction per level down; coarse solve; prolongation and
//          2 smoothing passes per level up)
//   zdots  z = MG result (bubble cells averaged), r.z, r.z_old   -> beta (last block)
#include "tq_device.hpp"

namespace tq {

// bubble CG vectors (MAX_BOLHAS doubles each) inside s->bvec
enum { BV_Y = 0, BV_R, BV_Z, BV_ZV, BV_P, BV_AP, BV_ACC0, BV_ACC1, BV_N };
__device__ __host__ inline double* bv(double* base, int k) { return base + (size_t)k * MAX_BOLHAS; }

template <class S> __device__ inline float ld(const S* p, size_t i);
template <> __device__ inline float ld<float>(const float* p, size_t i) { return p[i]; }
template <> __device__ inline float ld<__half>(const __half* p, size_t i) { return __half2float(p[i]); }
template <class S> __device__ inline void st(S* p, size_t i, float v);
template <> __device__ inline void st<float>(float* p, size_t i, float v) { p[i] = v; }
template <> __device__ inline void st<__half>(__half* p, size_t i, float v) { p[i] = __float2half(v); }

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

// bubble vectors, the active list (labels with water faces, in label order)
__global__ __launch_bounds__(1024) void k_bub_init(const uint32_t* __restrict__ facesB, const float* __restrict__ yB32,
                                                   int ar, double* __restrict__ bvec, int* __restrict__ activeList,
                                                   DevScalars* __restrict__ sc) {
  __shared__ int cnt[1024];
  const int tid = threadIdx.x;
  int mine = 0;
  for (int k = 0; k < 4; k++) {
    const int r = tid * 4 + k;
    const bool act = ar && r >= 2 && facesB[r] > 0;
    mine += act;
    for (int q = 0; q < BV_N; q++) bv(bvec, q)[r] = 0.0;
    if (act) bv(bvec, BV_Y)[r] = (double)yB32[r];
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
    if (ar && r >= 2 && facesB[r] > 0) activeList[o++] = r;
  }
}

// wave-aggregated FP64 atomic add into a per-label accumulator (label 0 = nothing)
__device__ inline void label_atomic_add(double* acc, uint32_t label, double val) {
  const int lane = threadIdx.x & 63;
  bool pending = label != 0;
  while (true) {
    unsigned long long b = __ballot(pending);
    if (!b) break;
    const int leader = __ffsll((long long)b) - 1;
    const uint32_t ll = __shfl(label, leader, 64);
    const bool m = pending && label == ll;
    const double s = wave_sum(m ? val : 0.0);
    if (lane == leader) atomicAdd(&acc[ll], s);
    if (m) pending = false;
  }
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

// ================================================================== init: warm-start residual
// q := 0 off water, bubble cells := y (difundir); r = -div - A q; bubble r = rhs - A_b [q;y].
__global__ __launch_bounds__(256) void k_pcg_init(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                                  const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                                  const float* __restrict__ div, float* __restrict__ q,
                                                  float* __restrict__ r, float* __restrict__ pbuf0, int nx, int ny,
                                                  int ar, double* __restrict__ bvec, const double* __restrict__ rhsB,
                                                  const int* __restrict__ activeList, double* __restrict__ partials,
                                                  DevScalars* __restrict__ sc) {
  __shared__ double sh[5 * 16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double v[5] = {0, 0, 0, 0, 0};  // r.r, div.div, sum(b), n_water, anchored
  const double* yb = bv(bvec, BV_Y);
  uint32_t pairLab[4] = {0, 0, 0, 0};
  float qc = 0.f;
  if (c < nx * ny) {
    const int i = c % nx, j = c / nx;
    pbuf0[c] = 0.f;
    auto qe = [&](size_t d) -> float {
      const uint8_t t = tipo[d];
      if (t == AGUA) return q[d];
      uint32_t lb;
      if (is_bub_cell(tipo, rot, facesB, ar, d, &lb)) return (float)yb[lb];
      return 0.f;
    };
    qc = qe(c);
    const uint8_t m = opMask[c];
    if (mask_type(m) == AGUA) {
      double s = 0;
      bool anchored = false;
      const size_t nb[4] = {(size_t)c - 1, (size_t)c + 1, (size_t)c - nx, (size_t)c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        s += (double)qe(nb[k]);
        const uint8_t tn = tipo[nb[k]];
        uint32_t lb;
        if (tn == AR) {
          if (is_bub_cell(tipo, rot, facesB, ar, nb[k], &lb)) {
            const int ni = (int)(nb[k] % nx), nj = (int)(nb[k] / nx);
            if (interior(ni, nj, nx, ny)) pairLab[k] = lb;
          } else anchored = true;
        }
      }
      const double aq = mask_cnt(m) * (double)qc - s;
      const float rr = (float)(-(double)div[c] - aq);
      r[c] = rr;
      v[0] = (double)rr * rr;
      v[1] = (double)div[c] * div[c];
      v[2] = -(double)div[c];
      v[3] = 1;
      v[4] = anchored ? 1 : 0;
    } else r[c] = 0.f;
    q[c] = qc;
  }
  // bubble rows: aq_b = faces*y_b - sum of q over the water cells at its faces
  for (int k = 0; k < 4; k++) label_atomic_add(bv(bvec, BV_ACC0), pairLab[k], pairLab[k] ? -(double)qc : 0.0);
  double tot[5];
  if (!grid_sum_last<5>(v, partials, &sc->ticket[0], tot, sh)) return;
  // ---- last block: bubble residuals, norms, null-space handling, stop test
  __shared__ double bs[3 * 16];
  double bacc[3] = {0, 0, 0};  // rb.rb, rhs.rhs, sum(rhs)
  const int na = sc->active_bubbles;
  double* rb = bv(bvec, BV_R);
  double* acc0 = bv(bvec, BV_ACC0);
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const double aq = (double)facesB[b] * yb[b] + acc0[b];
    acc0[b] = 0;
    rb[b] = rhsB[b] - aq;
    bacc[0] += rb[b] * rb[b];
    bacc[1] += rhsB[b] * rhsB[b];
    bacc[2] += rhsB[b];
  }
  block_sum<3>(bacc, bs);
  if (threadIdx.x == 0) {
    double rr = tot[0] + bacc[0];
    double bb = tot[1] + bacc[1];
    double mean = 0;
    const double N = tot[3] + na;
    // Residual-terminated mode on a system without any Dirichlet air (every water cell only
    // touches water, walls or tracked bubbles): the operator's null space is the constant
    // over all unknowns; remove its component from b so the 1e-6 target is reachable.
    if (sc->tol > 0 && tot[4] == 0 && N > 0) {
      mean = (tot[2] + bacc[2]) / N;
      // r' = r - mean, b' = b - mean over all N unknowns (sum(r) = sum(b) - sum(A q) = sum(b) here)
      const double sr = tot[2] + bacc[2];
      rr = rr - 2 * mean * sr + N * mean * mean;
      bb = bb - N * mean * mean;
      if (rr < 0) rr = 0;
      if (bb < 0) bb = 0;
    }
    sc->proj_mean = mean;  // applied to r by the first MG input (r -= mean)
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
  const double mean = sc->proj_mean;
  if (mean != 0)
    for (int k = threadIdx.x; k < na; k += blockDim.x) rb[activeList[k]] -= mean;
}

// ================================================================== MG right-hand side
template <class S>
__global__ __launch_bounds__(256) void k_mg_input(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                                  const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                                  const uint32_t* __restrict__ cellsB, float* __restrict__ r, int C,
                                                  int ar, const double* __restrict__ bvec, int first, int mixed,
                                                  S* __restrict__ b0, DevScalars* __restrict__ sc) {
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

// ================================================================== block smoother (suavizar, pressao-cpu.js:99-150)
// 16x16 tile + 1-cell halo read from the snapshot `src`, 4 red-black sweeps in LDS, interior
// written to `dst`.  desl = 0 or 8 shifts the tiling by half a block; (c0, c1) is the colour
// order, mirrored on the way up so the preconditioner stays symmetric.
template <class S>
__global__ __launch_bounds__(256) void k_smooth(const S* __restrict__ src, S* __restrict__ dst, const S* __restrict__ b,
                                                const uint8_t* __restrict__ mask, int nx, int ny, int desl, int c0,
                                                int c1, int srcZero, const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ float loc[18 * 18];
  const int tid = threadIdx.x;
  const int ox = blockIdx.x * 16 - desl, oy = blockIdx.y * 16 - desl;
  for (int e = tid; e < 18 * 18; e += 256) {
    const int i = ox + e % 18 - 1, j = oy + e / 18 - 1;
    loc[e] = (!srcZero && i >= 0 && j >= 0 && i < nx && j < ny) ? ld<S>(src, (size_t)i + (size_t)j * nx) : 0.f;
  }
  const int tx = tid & 15, ty = tid >> 4;
  const int i = ox + tx, j = oy + ty;
  const bool in = i >= 0 && j >= 0 && i < nx && j < ny;
  const size_t c = in ? (size_t)i + (size_t)j * nx : 0;
  const uint8_t m = in ? mask[c] : 0;
  const int k = mask_cnt(m);
  const bool act = in && mask_type(m) == AGUA && k > 0;
  const float bb = act ? ld<S>(b, c) : 0.f;
  const float invk = act ? 1.f / (float)k : 0.f;
  const int e = (tx + 1) + (ty + 1) * 18;
  const int myc = (i + j + 1) & 1;
  __syncthreads();
  for (int it = 0; it < 4; it++) {
    for (int o = 0; o < 2; o++) {
      const int cor = o ? c1 : c0;
      if (act && myc == cor) {
        float s = 0.f;
        if (m & 1) s += loc[e - 1];
        if (m & 2) s += loc[e + 1];
        if (m & 4) s += loc[e - 18];
        if (m & 8) s += loc[e + 18];
        const float qv = loc[e];
        loc[e] = qv + ((s + bb) * invk - qv);
      }
      __syncthreads();
    }
  }
  if (in) st<S>(dst, c, mask_type(m) == AGUA ? loc[e] : 0.f);
}

// ================================================================== restriction (restringir, pressao-cpu.js:159-188)
// Coarse b = P^T r with the fine residual r = b - A x recomputed on the fly; the 4x4 fine
// cells 2I-1..2I+2 carry weights (1/4, 3/4, 3/4, 1/4) per axis.
template <class SF, class SC>
__global__ __launch_bounds__(256) void k_restrict(const SF* __restrict__ fx, const SF* __restrict__ fb,
                                                  const uint8_t* __restrict__ fm, int fnx, int fny,
                                                  SC* __restrict__ cb, const uint8_t* __restrict__ cmk, int cnx, int cny,
                                                  const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  const int C = blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= cnx * cny) return;
  if (mask_type(cmk[C]) != AGUA) { st<SC>(cb, C, 0.f); return; }
  const int I = C % cnx, J = C / cnx;
  const float w[4] = {0.25f, 0.75f, 0.75f, 0.25f};
  float acc = 0.f;
  for (int bj = 0; bj < 4; bj++) {
    const int j = 2 * J - 1 + bj;
    if (j < 0 || j >= fny) continue;
    float racc = 0.f;
    for (int bi = 0; bi < 4; bi++) {
      const int i = 2 * I - 1 + bi;
      if (i < 0 || i >= fnx) continue;
      const size_t c = (size_t)i + (size_t)j * fnx;
      const uint8_t m = fm[c];
      if (mask_type(m) != AGUA) continue;
      float s = 0.f;
      if (m & 1) s += ld<SF>(fx, c - 1);
      if (m & 2) s += ld<SF>(fx, c + 1);
      if (m & 4) s += ld<SF>(fx, c - fnx);
      if (m & 8) s += ld<SF>(fx, c + fnx);
      const float rr = ld<SF>(fb, c) - ((float)mask_cnt(m) * ld<SF>(fx, c) - s);
      racc += rr * w[bi];
    }
    acc += racc * w[bj];
  }
  st<SC>(cb, C, acc);
}

// ================================================================== prolongation (prolongar, pressao-cpu.js:191-214)
template <class SC, class SF>
__global__ __launch_bounds__(256) void k_prolong(const SC* __restrict__ cx, const uint8_t* __restrict__ cmk, int cnx,
                                                 int cny, SF* __restrict__ fx, const uint8_t* __restrict__ fm, int fnx,
                                                 int fny, const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= fnx * fny) return;
  if (mask_type(fm[c]) != AGUA) return;
  const int i = c % fnx, j = c / fnx;
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
  st<SF>(fx, c, ld<SF>(fx, c) + e);
}

// ================================================================== coarsest level (resolverGrosso, pressao-cpu.js:217-238)
// 64 red-black Gauss-Seidel sweeps (colour order alternating) from x = 0, one workgroup, LDS.
constexpr int COARSE_MAX = 12288;
__global__ __launch_bounds__(1024) void k_coarse(float* __restrict__ x, const float* __restrict__ b,
                                                 const uint8_t* __restrict__ mask, int nx, int ny,
                                                 const DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ float lx[COARSE_MAX], lb[COARSE_MAX];
  __shared__ uint8_t lm[COARSE_MAX];
  const int n = nx * ny;
  for (int c = threadIdx.x; c < n; c += blockDim.x) { lx[c] = 0.f; lb[c] = b[c]; lm[c] = mask[c]; }
  __syncthreads();
  for (int k = 0; k < 64; k++)
    for (int o = 0; o < 2; o++) {
      const int cor = (k & 1) ? 1 - o : o;
      for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int i = c % nx, j = c / nx;
        const uint8_t m = lm[c];
        if (mask_type(m) != AGUA || ((i + j + 1) & 1) != cor) continue;
        const int q = mask_cnt(m);
        if (!q) continue;
        float s = 0.f;
        if (m & 1) s += lx[c - 1];
        if (m & 2) s += lx[c + 1];
        if (m & 4) s += lx[c - nx];
        if (m & 8) s += lx[c + nx];
        lx[c] = (s + lb[c]) / (float)q;
      }
      __syncthreads();
    }
  for (int c = threadIdx.x; c < n; c += blockDim.x) x[c] = lx[c];
}

// ================================================================== z, r.z, r.z_old -> beta
template <class S>
__global__ __launch_bounds__(256) void k_zdots(const S* __restrict__ x0, const uint8_t* __restrict__ opMask,
                                               const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                               const uint32_t* __restrict__ facesB, const uint32_t* __restrict__ cellsB,
                                               const float* __restrict__ r, const float* __restrict__ zold,
                                               float* __restrict__ znew, int C, int ar, int mixed, int first,
                                               double* __restrict__ bvec, const int* __restrict__ activeList,
                                               double* __restrict__ partials, DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[2 * 16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const float inv = mixed ? 1.f / sc->mg_scale : 1.f;
  double v[2] = {0, 0};
  uint32_t lab = 0;
  double bval = 0;
  if (c < C) {
    const float xv = ld<S>(x0, c) * inv;
    float z = 0.f;
    if (mask_type(opMask[c]) == AGUA) {
      z = xv;
      v[0] = (double)r[c] * z;
      v[1] = (double)r[c] * zold[c];
    } else {
      uint32_t lb;
      if (is_bub_cell(tipo, rot, facesB, ar, c, &lb)) { lab = lb; bval = (double)xv / cellsB[lb]; }
    }
    znew[c] = z;
  }
  label_atomic_add(bv(bvec, BV_ACC1), lab, bval);
  double tot[2];
  if (!grid_sum_last<2>(v, partials, &sc->ticket[1], tot, sh)) return;
  __shared__ double bs[2 * 16];
  double bacc[2] = {0, 0};
  const int na = sc->active_bubbles;
  double* rb = bv(bvec, BV_R);
  double* zb = bv(bvec, BV_Z);
  double* zvb = bv(bvec, BV_ZV);
  double* acc1 = bv(bvec, BV_ACC1);
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    zvb[b] = zb[b];
    zb[b] = acc1[b];
    acc1[b] = 0;
    bacc[0] += rb[b] * zb[b];
    bacc[1] += rb[b] * zvb[b];
  }
  block_sum<2>(bacc, bs);
  if (threadIdx.x == 0) {
    const double rzNovo = tot[0] + bacc[0];
    const double cruzado = tot[1] + bacc[1];
    if (first) {
      sc->beta = 0;
    } else {
      const double rz = sc->rz;
      sc->beta = fabs(rz) < 1e-30 ? 0 : fmax(0.0, (rzNovo - cruzado) / rz);
    }
    sc->rz = rzNovo;
  }
}

// ================================================================== p = z + beta p, Ap, p.Ap -> alpha
__global__ __launch_bounds__(256) void k_spmv(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                              const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                              const float* __restrict__ z, const float* __restrict__ pold,
                                              float* __restrict__ pnew, float* __restrict__ ap, int nx, int ny, int ar,
                                              double* __restrict__ bvec, const int* __restrict__ activeList,
                                              double* __restrict__ partials, DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const double beta = sc->beta;
  const double* zb = bv(bvec, BV_Z);
  const double* pb = bv(bvec, BV_P);
  double v[1] = {0};
  uint32_t pairLab[4] = {0, 0, 0, 0};
  float pc = 0.f;
  if (c < nx * ny) {
    const int i = c % nx, j = c / nx;
    auto peff = [&](size_t d) -> float {
      if (mask_type(opMask[d]) == AGUA) return (float)((double)z[d] + beta * (double)pold[d]);
      uint32_t lb;
      if (is_bub_cell(tipo, rot, facesB, ar, d, &lb)) return (float)(zb[lb] + beta * pb[lb]);
      return 0.f;
    };
    pc = peff(c);
    pnew[c] = pc;
    const uint8_t m = opMask[c];
    if (mask_type(m) == AGUA) {
      double s = 0;
      const size_t nb[4] = {(size_t)c - 1, (size_t)c + 1, (size_t)c - nx, (size_t)c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        s += (double)peff(nb[k]);
        uint32_t lb;
        if (is_bub_cell(tipo, rot, facesB, ar, nb[k], &lb)) {
          const int ni = (int)(nb[k] % nx), nj = (int)(nb[k] / nx);
          if (interior(ni, nj, nx, ny)) pairLab[k] = lb;
        }
      }
      const float a = (float)(mask_cnt(m) * (double)pc - s);
      ap[c] = a;
      v[0] = (double)pc * a;
    } else ap[c] = 0.f;
  }
  for (int k = 0; k < 4; k++) label_atomic_add(bv(bvec, BV_ACC0), pairLab[k], pairLab[k] ? -(double)pc : 0.0);
  double tot[1];
  if (!grid_sum_last<1>(v, partials, &sc->ticket[2], tot, sh)) return;
  __shared__ double bs[2 * 16];
  double bacc[2] = {0, 0};
  const int na = sc->active_bubbles;
  double* pbw = bv(bvec, BV_P);
  double* apb = bv(bvec, BV_AP);
  double* acc0 = bv(bvec, BV_ACC0);
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    const double pn = zb[b] + beta * pbw[b];
    pbw[b] = pn;
    apb[b] = (double)facesB[b] * pn + acc0[b];
    acc0[b] = 0;
    bacc[0] += pn * apb[b];
  }
  block_sum<2>(bacc, bs);
  __shared__ double alpha_s;
  if (threadIdx.x == 0) {
    const double pap = tot[0] + bacc[0];
    alpha_s = fabs(pap) < 1e-30 ? 0 : sc->rz / pap;
    sc->alpha = alpha_s;
  }
  __syncthreads();
  // bubble unknowns: y += alpha p, r -= alpha Ap
  const double alpha = alpha_s;
  double* yb = bv(bvec, BV_Y);
  double* rb = bv(bvec, BV_R);
  double rbs[1] = {0};
  for (int k = threadIdx.x; k < na; k += blockDim.x) {
    const int b = activeList[k];
    yb[b] += alpha * pbw[b];
    rb[b] -= alpha * apb[b];
    rbs[0] += rb[b] * rb[b];
  }
  block_sum<1>(rbs, bs);
  if (threadIdx.x == 0) sc->pap = rbs[0];  // bubble part of |r|^2 for the update kernel
}

// ================================================================== q += alpha p, r -= alpha Ap
__global__ __launch_bounds__(256) void k_update(const uint8_t* __restrict__ opMask, const float* __restrict__ p,
                                                const float* __restrict__ ap, float* __restrict__ q,
                                                float* __restrict__ r, int C, double* __restrict__ partials,
                                                DevScalars* __restrict__ sc) {
  if (sc->done) return;
  __shared__ double sh[16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = sc->alpha;
  double v[1] = {0};
  if (c < C && mask_type(opMask[c]) == AGUA) {
    q[c] = (float)((double)q[c] + alpha * (double)p[c]);
    const float rn = (float)((double)r[c] - alpha * (double)ap[c]);
    r[c] = rn;
    v[0] = (double)rn * rn;
  }
  double tot[1];
  if (!grid_sum_last<1>(v, partials, &sc->ticket[3], tot, sh)) return;
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

// true residual ||b - A[q;y]|| (diagnostics): q holds the solved pressures on water cells
__global__ __launch_bounds__(256) void k_true_res(const uint8_t* __restrict__ opMask, const uint8_t* __restrict__ tipo,
                                                  const uint32_t* __restrict__ rot, const uint32_t* __restrict__ facesB,
                                                  const float* __restrict__ div, const float* __restrict__ q, int nx,
                                                  int ny, int ar, double* __restrict__ bvec,
                                                  const double* __restrict__ rhsB, const int* __restrict__ activeList,
                                                  double* __restrict__ partials, DevScalars* __restrict__ sc,
                                                  double* __restrict__ out) {
  __shared__ double sh[16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const double* yb = bv(bvec, BV_Y);
  const double mean = sc->proj_mean;  // residual of the (projected) system that was solved
  double v[1] = {0};
  uint32_t pairLab[4] = {0, 0, 0, 0};
  float qc = 0.f;
  if (c < nx * ny) {
    auto qe = [&](size_t d) -> float {
      if (tipo[d] == AGUA) return q[d];
      uint32_t lb;
      if (is_bub_cell(tipo, rot, facesB, ar, d, &lb)) return (float)yb[lb];
      return 0.f;
    };
    qc = qe(c);
    const uint8_t m = opMask[c];
    if (mask_type(m) == AGUA) {
      double s = 0;
      const size_t nb[4] = {(size_t)c - 1, (size_t)c + 1, (size_t)c - nx, (size_t)c + nx};
      for (int k = 0; k < 4; k++) {
        if (!(m & (1 << k))) continue;
        s += (double)qe(nb[k]);
        uint32_t lb;
        if (is_bub_cell(tipo, rot, facesB, ar, nb[k], &lb)) {
          const int ni = (int)(nb[k] % nx), nj = (int)(nb[k] / nx);
          if (interior(ni, nj, nx, ny)) pairLab[k] = lb;
        }
      }
      const double e = (-(double)div[c] - mean) - (mask_cnt(m) * (double)qc - s);
      v[0] = e * e;
    }
  }
  for (int k = 0; k < 4; k++) label_atomic_add(bv(bvec, BV_ACC0), pairLab[k], pairLab[k] ? -(double)qc : 0.0);
  double tot[1];
  if (!grid_sum_last<1>(v, partials, &sc->ticket[4], tot, sh)) return;
  __shared__ double bs[16];
  double bacc[1] = {0};
  double* acc0 = bv(bvec, BV_ACC0);
  for (int k = threadIdx.x; k < sc->active_bubbles; k += blockDim.x) {
    const int b = activeList[k];
    const double e = (rhsB[b] - mean) - ((double)facesB[b] * yb[b] + acc0[b]);
    acc0[b] = 0;
    bacc[0] += e * e;
  }
  block_sum<1>(bacc, bs);
  if (threadIdx.x == 0) *out = sqrt(tot[0] + bacc[0]);
}

// ================================================================== host side
// FP16 storage on the large levels only (never the coarsest, which the coarse solver reads as FP32)
static bool level_half(const tq_sim* s, int l) {
  return s->d.precision == TQ_PREC_MIXED && s->lv[l].n >= 4096 && l + 1 < (int)s->lv.size();
}

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
  for (auto& L : s->lv) {
    L.n = L.nx * L.ny;
    TQ_HIP(hipMalloc(&L.mask, L.n));
    // always sized for FP32 so the precision can switch at run time
    TQ_HIP(hipMalloc(&L.b, sizeof(float) * L.n));
    TQ_HIP(hipMalloc(&L.x, sizeof(float) * L.n));
    TQ_HIP(hipMalloc(&L.tmp, sizeof(float) * L.n));
  }
  if (s->lv.back().n > COARSE_MAX) {
    set_error("grid too elongated: the coarsest multigrid level exceeds the one-workgroup coarse solver");
    return TQ_E_CAPACITY;
  }
  return 0;
}

void solver_free(tq_sim* s) {
  for (auto& L : s->lv) {
    (void)hipFree(L.mask); (void)hipFree(L.b); (void)hipFree(L.x); (void)hipFree(L.tmp);
  }
  s->lv.clear();
}

template <class S>
static void launch_smooth(tq_sim* s, int l, int desl, int c0, int c1, int srcZero) {
  Level& L = s->lv[l];
  dim3 g((L.nx + desl + 15) / 16, (L.ny + desl + 15) / 16);
  hipLaunchKernelGGL(k_smooth<S>, g, dim3(256), 0, s->st, (const S*)L.x, (S*)L.tmp, (const S*)L.b, L.mask, L.nx, L.ny,
                     desl, c0, c1, srcZero, s->sc);
  std::swap(L.x, L.tmp);
}

static void smooth(tq_sim* s, int l, int desl, int c0, int c1, int srcZero) {
  const double bpc = level_half(s, l) ? 7.0 : 13.0;
  int t = time_begin(s, l == 0 ? "mg_smooth_l0" : "mg_smooth", bpc * s->lv[l].n);
  if (level_half(s, l)) launch_smooth<__half>(s, l, desl, c0, c1, srcZero);
  else launch_smooth<float>(s, l, desl, c0, c1, srcZero);
  time_end(s, t);
}

static void restrict_to(tq_sim* s, int l) {
  Level& F = s->lv[l];
  Level& G = s->lv[l + 1];
  const int n = G.n;
  const bool hf = level_half(s, l), hc = level_half(s, l + 1);
  int t = time_begin(s, "mg_restrict", (hf ? 5.0 : 9.0) * F.n + (hc ? 3.0 : 5.0) * G.n);
  dim3 g((n + 255) / 256), b(256);
  if (hf && hc)
    hipLaunchKernelGGL((k_restrict<__half, __half>), g, b, 0, s->st, (const __half*)F.x, (const __half*)F.b, F.mask, F.nx,
                       F.ny, (__half*)G.b, G.mask, G.nx, G.ny, s->sc);
  else if (hf)
    hipLaunchKernelGGL((k_restrict<__half, float>), g, b, 0, s->st, (const __half*)F.x, (const __half*)F.b, F.mask, F.nx,
                       F.ny, (float*)G.b, G.mask, G.nx, G.ny, s->sc);
  else
    hipLaunchKernelGGL((k_restrict<float, float>), g, b, 0, s->st, (const float*)F.x, (const float*)F.b, F.mask, F.nx,
                       F.ny, (float*)G.b, G.mask, G.nx, G.ny, s->sc);
  time_end(s, t);
}

static void prolong_from(tq_sim* s, int l) {
  Level& F = s->lv[l];
  Level& G = s->lv[l + 1];
  const bool hf = level_half(s, l), hc = level_half(s, l + 1);
  int t = time_begin(s, "mg_prolong", (hf ? 5.0 : 9.0) * F.n + (hc ? 3.0 : 5.0) * G.n);
  dim3 g((F.n + 255) / 256), b(256);
  if (hf && hc)
    hipLaunchKernelGGL((k_prolong<__half, __half>), g, b, 0, s->st, (const __half*)G.x, G.mask, G.nx, G.ny, (__half*)F.x,
                       F.mask, F.nx, F.ny, s->sc);
  else if (hf)
    hipLaunchKernelGGL((k_prolong<float, __half>), g, b, 0, s->st, (const float*)G.x, G.mask, G.nx, G.ny, (__half*)F.x,
                       F.mask, F.nx, F.ny, s->sc);
  else
    hipLaunchKernelGGL((k_prolong<float, float>), g, b, 0, s->st, (const float*)G.x, G.mask, G.nx, G.ny, (float*)F.x,
                       F.mask, F.nx, F.ny, s->sc);
  time_end(s, t);
}

// cicloV (pressao-cpu.js:241-255)
static void vcycle(tq_sim* s) {
  const int L = (int)s->lv.size();
  for (int l = 0; l < L - 1; l++) {
    smooth(s, l, 0, 0, 1, 1);
    smooth(s, l, 8, 0, 1, 0);
    restrict_to(s, l);
  }
  {
    Level& G = s->lv[L - 1];
    int t = time_begin(s, "mg_coarse", 9.0 * G.n);
    hipLaunchKernelGGL(k_coarse, dim3(1), dim3(1024), 0, s->st, (float*)G.x, (const float*)G.b, G.mask, G.nx, G.ny, s->sc);
    time_end(s, t);
  }
  for (int l = L - 2; l >= 0; l--) {
    prolong_from(s, l);
    smooth(s, l, 8, 1, 0, 0);
    smooth(s, l, 0, 1, 0, 0);
  }
}

static void precondition(tq_sim* s, int first) {
  const int C = s->C;
  const bool h0 = level_half(s, 0);
  const int mixed = h0 ? 1 : 0;
  int t = time_begin(s, "pcg_mg_input", (1 + 4 + (h0 ? 2 : 4)) * (double)C);
  if (h0)
    hipLaunchKernelGGL(k_mg_input<__half>, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot,
                       s->facesB, s->cellsB, s->r, C, s->d.ar, s->bvec, first, mixed, (__half*)s->lv[0].b, s->sc);
  else
    hipLaunchKernelGGL(k_mg_input<float>, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot,
                       s->facesB, s->cellsB, s->r, C, s->d.ar, s->bvec, first, 0, (float*)s->lv[0].b, s->sc);
  time_end(s, t);
  vcycle(s);
  const int zn = 1 - s->zc;
  t = time_begin(s, "pcg_zdots", (1 + (h0 ? 2 : 4) + 4 + 4 + 4) * (double)C);
  if (h0)
    hipLaunchKernelGGL(k_zdots<__half>, dim3((C + 255) / 256), dim3(256), 0, s->st, (const __half*)s->lv[0].x, s->opMask,
                       s->tipo, s->rot, s->facesB, s->cellsB, s->r, s->zbuf[s->zc], s->zbuf[zn], C, s->d.ar, 1, first,
                       s->bvec, s->activeList, s->partials, s->sc);
  else
    hipLaunchKernelGGL(k_zdots<float>, dim3((C + 255) / 256), dim3(256), 0, s->st, (const float*)s->lv[0].x, s->opMask,
                       s->tipo, s->rot, s->facesB, s->cellsB, s->r, s->zbuf[s->zc], s->zbuf[zn], C, s->d.ar, 0, first,
                       s->bvec, s->activeList, s->partials, s->sc);
  time_end(s, t);
  s->zc = zn;
}

static void pcg_iteration(tq_sim* s) {
  const int C = s->C;
  const int pn = 1 - s->pc;
  int t = time_begin(s, "pcg_spmv", (1 + 4 + 4 + 4 + 4 + 1) * (double)C);
  hipLaunchKernelGGL(k_spmv, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot, s->facesB,
                     s->zbuf[s->zc], s->pbuf[s->pc], s->pbuf[pn], s->ap, s->nx, s->ny, s->d.ar, s->bvec, s->activeList,
                     s->partials, s->sc);
  time_end(s, t);
  s->pc = pn;
  t = time_begin(s, "pcg_update", (1 + 8 + 8 + 4 + 4) * (double)C);
  hipLaunchKernelGGL(k_update, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->pbuf[s->pc], s->ap, s->q, s->r,
                     C, s->partials, s->sc);
  time_end(s, t);
  precondition(s, 0);
}

int pressure_solve(tq_sim* s) {
  const int C = s->C;
  const int L = (int)s->lv.size();
  if ((C + 255) / 256 > MAX_RED_BLOCKS) { set_error("grid too large for the reduction buffers"); return TQ_E_CAPACITY; }
  // ---- setup: masks of every level, bubble vectors
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
                     s->sc);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  const double tol = s->d.tol;
  const int maxit = s->d.iteracoes;
  TQ_HIP(hipMemcpyAsync(&s->sc->tol, &tol, sizeof(double), hipMemcpyHostToDevice, s->st));
  TQ_HIP(hipMemcpyAsync(&s->sc->maxit, &maxit, sizeof(int), hipMemcpyHostToDevice, s->st));
  t = time_begin(s, "pcg_init", (1 + 1 + 4 + 4 + 4 + 4 + 4) * (double)C);
  hipLaunchKernelGGL(k_pcg_init, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot, s->facesB, s->div,
                     s->q, s->r, s->pbuf[s->pc], s->nx, s->ny, s->d.ar, s->bvec, s->rhsB, s->activeList, s->partials,
                     s->sc);
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
  double* dres = s->partials + 8 * (size_t)MAX_RED_BLOCKS - 1;
  hipLaunchKernelGGL(k_true_res, dim3((C + 255) / 256), dim3(256), 0, s->st, s->opMask, s->tipo, s->rot, s->facesB, s->div,
                     s->q, s->nx, s->ny, s->d.ar, s->bvec, s->rhsB, s->activeList, s->partials, s->sc, dres);
  TQ_LAUNCH_CHECK();
  TQ_HIP(hipMemcpyAsync(out, dres, sizeof(double), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  return 0;
}

}  // namespace tq
