// k_grid.hip -- cell/face kernels of the substep: surface curvature (one LDS-tiled pass),
// air-cell pressure + divergence, label propagation, per-bubble sums, the pressure gradient
// and the velocity extrapolation.
#include "tq_device.hpp"

namespace tq {

// ================================================================== #superficie (fisica-cpu.js:306-374)
// One workgroup computes a 32x32 tile from a 58x58 LDS region (halo 13 = the dependency radius
// of ext3, blur, ext3, blur, ext3, normal, curvature), so the 15 full-grid passes of the
// reference become one launch.  Each pass computes only the box of cells the rest of the chain
// reads -- its radius around the owned tile shrinks from 12 (the first extension layer) to 1
// (the normals) -- so the halo's cells are computed no more often than the owned tile needs.
constexpr int SF_T = 32, SF_H = 13, SF_R = SF_T + 2 * SF_H;  // 58
constexpr int SF_N = SF_R * SF_R;                            // 3364
constexpr int SF_THREADS = 1024;
constexpr int SF_K = (SF_N + SF_THREADS - 1) / SF_THREADS;   // region cells per thread (4)
static_assert(SF_T * SF_T == SF_THREADS, "one owned cell per thread");

// fn(e, lx, ly) for the region cells within RX columns and RY rows of the owned tile; the loop
// over a thread's cells is unrolled (its trip count is known), so their LDS reads are issued
// together instead of one cell's chain after the other's
template <int RX, int RY, class Fn>
__device__ __forceinline__ void sf_box(Fn fn) {
  constexpr int W = SF_T + 2 * RX, H = SF_T + 2 * RY, N = W * H;
#pragma unroll
  for (int k = 0; k < (N + SF_THREADS - 1) / SF_THREADS; k++) {
    const int i = (int)threadIdx.x + k * SF_THREADS;
    if (i < N) {
      const int by = i / W, bx = i - by * W;
      const int lx = SF_H - RX + bx, ly = SF_H - RY + by;
      fn(lx + ly * SF_R, lx, ly);
    }
  }
}

// one extension layer (#estenderNaParede): an unfilled cell takes the mean of its filled
// 4-neighbours; computed on the box of radius R (it reads radius R + 1)
template <int R>
__device__ __forceinline__ void sf_layer(float* f, float* t, uint8_t* fl, const uint8_t* ty, int I0, int J0, int nx,
                                         int ny) {
#pragma clang fp contract(off)
  sf_box<R, R>([&](int e, int lx, int ly) {
    if (fl[e] || ty[e] == 255) return;
    const int gi = I0 + lx, gj = J0 + ly;
    double sum = 0;
    int n = 0;
    if (gi > 0 && fl[e - 1] == 1) { sum += (double)f[e - 1]; n++; }
    if (gi < nx - 1 && fl[e + 1] == 1) { sum += (double)f[e + 1]; n++; }
    if (gj > 0 && fl[e - SF_R] == 1) { sum += (double)f[e - SF_R]; n++; }
    if (gj < ny - 1 && fl[e + SF_R] == 1) { sum += (double)f[e + SF_R]; n++; }
    if (n) { t[e] = (float)(sum / n); fl[e] = 2; }
  });
  __syncthreads();
  sf_box<R, R>([&](int e, int, int) {
    if (fl[e] == 2) { f[e] = t[e]; fl[e] = 1; }
  });
  __syncthreads();
}

// the three layers of a passada whose last layer writes radius R (the filled flags first, on
// the radius R + 3 the first layer reads)
template <int R>
__device__ __forceinline__ void sf_extend(float* f, float* t, uint8_t* fl, const uint8_t* ty, int I0, int J0, int nx,
                                          int ny) {
  sf_box<R + 3, R + 3>([&](int e, int, int) { fl[e] = (ty[e] != SOLIDO && ty[e] != 255) ? 1 : 0; });
  __syncthreads();
  sf_layer<R + 2>(f, t, fl, ty, I0, J0, nx, ny);
  sf_layer<R + 1>(f, t, fl, ty, I0, J0, nx, ny);
  sf_layer<R>(f, t, fl, ty, I0, J0, nx, ny);
}

// the separable (1 2 1) blur, writing radius R (the row pass on radius R columns and R + 1 rows)
template <int R>
__device__ __forceinline__ void sf_blur(float* f, float* t, int I0, int J0, int nx, int ny) {
#pragma clang fp contract(off)
  sf_box<R, R + 1>([&](int e, int lx, int) {
    const int gi = I0 + lx;
    const int em = gi > 0 ? e - 1 : e, ep = gi < nx - 1 ? e + 1 : e;
    t[e] = (float)(((double)f[em] + 2 * (double)f[e] + (double)f[ep]) * 0.25);
  });
  __syncthreads();
  sf_box<R, R>([&](int e, int, int ly) {
    const int gj = J0 + ly;
    const int em = gj > 0 ? e - SF_R : e, ep = gj < ny - 1 ? e + SF_R : e;
    f[e] = (float)(((double)t[em] + 2 * (double)t[e] + (double)t[ep]) * 0.25);
  });
  __syncthreads();
}

__device__ __forceinline__ void k_surface(Blk B_, const uint8_t* __restrict__ tipo, const float* __restrict__ dens,
                                          int nx, int ny, double h, float* __restrict__ fase,
                                          float* __restrict__ nXo, float* __restrict__ nYo,
                                          float* __restrict__ kap) {
#pragma clang fp contract(off)
  __shared__ float f[SF_N], t[SF_N], ga[SF_N], gb[SF_N];
  __shared__ uint8_t ty[SF_N], fl[SF_N];
  const int tid = threadIdx.x;
  int bxl, byl;
  xcd_block2(B_, &bxl, &byl);
  const int I0 = bxl * SF_T - SF_H, J0 = byl * SF_T - SF_H;
  int solid_here = 0, fluid_here = 0;  // solid / filled cells in this thread's part of the region
#pragma unroll
  for (int k = 0; k < SF_K; k++) {
    const int e = tid + SF_THREADS * k;
    if (e >= SF_N) continue;
    const int ly = e / SF_R, lx = e - ly * SF_R;
    const int gi = I0 + lx, gj = J0 + ly;
    if (gi >= 0 && gj >= 0 && gi < nx && gj < ny) {
      const size_t c = (size_t)gi + (size_t)gj * nx;
      const uint8_t tt = tipo[c];
      ty[e] = tt;
      f[e] = tt == SOLIDO ? 0.f : (float)fmin(1.0, (double)dens[c] / PPC);
      solid_here |= tt == SOLIDO;
      fluid_here |= tt != SOLIDO;
    } else {
      ty[e] = 255;
      f[e] = 0.f;
    }
  }
  // The wall extension fills solid cells (cells outside the grid are never filled) from filled
  // neighbours: a region without solid cells, or without a filled one, is left as it is by all its
  // layers, which are skipped (the tank's interior, the inside of thick walls).  Both flags of
  // the region in one barrier: each wave's two ballots, then the 16 waves' words.
  __shared__ int wflag[SF_THREADS / 64];
  {
    const bool ws = __ballot(solid_here) != 0, wf = __ballot(fluid_here) != 0;
    if ((tid & 63) == 0) wflag[tid >> 6] = (ws ? 1 : 0) | (wf ? 2 : 0);
  }
  __syncthreads();
  int any = 0;
#pragma unroll
  for (int k = 0; k < SF_THREADS / 64; k++) any |= wflag[k];
  const bool walls = any == 3;
  // ext3 (radius 10), blur (9), ext3 (6), blur (5), ext3 (2): what the normals (1) read
  if (walls) sf_extend<10>(f, t, fl, ty, I0, J0, nx, ny);
  sf_blur<9>(f, t, I0, J0, nx, ny);
  if (walls) sf_extend<6>(f, t, fl, ty, I0, J0, nx, ny);
  sf_blur<5>(f, t, I0, J0, nx, ny);
  if (walls) sf_extend<2>(f, t, fl, ty, I0, J0, nx, ny);
  // normals (interior cells of the domain only; the reference never writes the border), on the
  // owned tile and its one-cell ring (what the curvature reads)
  sf_box<1, 1>([&](int e, int lx, int ly) {
    const int gi = I0 + lx, gj = J0 + ly;
    float a = 0.f, b = 0.f;
    if (gi >= 1 && gj >= 1 && gi < nx - 1 && gj < ny - 1) {
      const double gx = ((double)f[e + 1] - (double)f[e - 1]) * 0.5, gy = ((double)f[e + SF_R] - (double)f[e - SF_R]) * 0.5;
      const double g = js_hypot(gx, gy);
      a = g > 1e-3 ? (float)(gx / g) : 0.f;
      b = g > 1e-3 ? (float)(gy / g) : 0.f;
    }
    ga[e] = a;
    gb[e] = b;
  });
  __syncthreads();
  // the owned 32x32 cells: one per thread
  {
    const int lx = SF_H + (tid & (SF_T - 1)), ly = SF_H + tid / SF_T;
    const int gi = I0 + lx, gj = J0 + ly;
    if (gi < nx && gj < ny) {
      const int q = lx + ly * SF_R;
      const size_t c = (size_t)gi + (size_t)gj * nx;
      fase[c] = f[q];
      nXo[c] = ga[q];
      nYo[c] = gb[q];
      float kk = 0.f;
      if (gi >= 1 && gj >= 1 && gi < nx - 1 && gj < ny - 1)
        kk = (float)(-(((double)ga[q + 1] - (double)ga[q - 1]) + ((double)gb[q + SF_R] - (double)gb[q - SF_R])) /
                     (2 * h));
      kap[c] = kk;
    }
  }
}

int surface(tq_sim* s) {
  dim3 g((s->nx + SF_T - 1) / SF_T, (s->ny + SF_T - 1) / SF_T);
  int t = time_begin(s, "surface", (1 + 4 + 4 * 4) * (double)s->C);
  launch<SF_THREADS, 0, k_surface>(s, s->st, g, dim3(SF_THREADS), 0, s->tipo, s->dens, s->nx, s->ny, s->h, s->fase, s->nX,
                                   s->nY, s->kappa);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== #pressaoDoAr + #classificar
// qAr (fisica-cpu.js:381-403) of an air cell, in units of q
__device__ inline float qar_at(const uint8_t* __restrict__ tipo, const float* __restrict__ kap, int nx, int ny, int i, int j,
                               double tensao, double kMax, double escala) {
#pragma clang fp contract(off)
  if (i < 1 || j < 1 || i >= nx - 1 || j >= ny - 1) return 0.f;
  const size_t c = (size_t)i + (size_t)j * nx;
  if (tipo[c] != AR) return 0.f;
  double p = 0;
  if (tensao > 0) {
    double s = kap[c];
    int k = 1;
    if (tipo[c - 1] == AGUA) { s += (double)kap[c - 1]; k++; }
    if (tipo[c + 1] == AGUA) { s += (double)kap[c + 1]; k++; }
    if (tipo[c - nx] == AGUA) { s += (double)kap[c - nx]; k++; }
    if (tipo[c + nx] == AGUA) { s += (double)kap[c + nx]; k++; }
    if (k > 1) p += tensao * fmax(-kMax, fmin(kMax, s / k));
  }
  return (float)(p * escala);
}

__device__ __forceinline__ void k_air_div(Blk B_, const uint8_t* __restrict__ tipo, const float* __restrict__ kap,
                                          const float* __restrict__ u, const float* __restrict__ v,
                                          const float* __restrict__ dens, const uint32_t* __restrict__ count,
                                          const uint8_t* __restrict__ solid, int nx, int ny, double tensao,
                                          double kMax, double escala, double alvo, int which,
                                          float* __restrict__ qAr, float* __restrict__ div,
                                          float* __restrict__ q) {
#pragma clang fp contract(off)
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx;
  if (which & 1) qAr[c] = qar_at(tipo, kap, nx, ny, i, j, tensao, kMax, escala);
  if (!(which & 2)) return;
  if (tipo[c] != AGUA) { div[c] = 0.f; q[c] = 0.f; return; }
  const double d = (double)u[(size_t)i + 1 + (size_t)j * (nx + 1)] - (double)u[(size_t)i + (size_t)j * (nx + 1)] +
                   (double)v[(size_t)i + (size_t)(j + 1) * nx] - (double)v[(size_t)i + (size_t)j * nx];
  double excesso = (double)dens[c] / PPC - 1;
  if (excesso < 0) {
    // #miolo (fisica-cpu.js:484-494)
    bool miolo = i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1;
    for (int b = -1; b <= 1 && miolo; b++)
      for (int a = -1; a <= 1; a++) {
        size_t cc = (size_t)(i + a) + (size_t)(j + b) * nx;
        if (solid[cc] || count[cc] == 0) { miolo = false; break; }
      }
    if (!miolo) excesso = 0;
  }
  excesso = fmin(EXCESSO_MAX, fmax(EXCESSO_MIN, excesso));
  double ar = 0;
  if (i > 0 && tipo[c - 1] == AR) ar += (double)qar_at(tipo, kap, nx, ny, i - 1, j, tensao, kMax, escala);
  if (i < nx - 1 && tipo[c + 1] == AR) ar += (double)qar_at(tipo, kap, nx, ny, i + 1, j, tensao, kMax, escala);
  if (j > 0 && tipo[c - nx] == AR) ar += (double)qar_at(tipo, kap, nx, ny, i, j - 1, tensao, kMax, escala);
  if (j < ny - 1 && tipo[c + nx] == AR) ar += (double)qar_at(tipo, kap, nx, ny, i, j + 1, tensao, kMax, escala);
  div[c] = (float)(d - alvo * excesso - ar);
}

int air_pressure_and_div(tq_sim* s, double dt, int which) {
  const double escala = dt / (DENSIDADE_AGUA * s->h), kMax = 1 / s->h;
  const double alvo = s->h / std::max(s->d.tau_densidade, dt);
  int t = time_begin(s, "air_div", (1 + 4 + 4 + 4 + 4 + 1 + 4 + 4 + 4) * (double)s->C);
  launch<256, 0, k_air_div>(s, s->st, dim3((s->C + 255) / 256), dim3(256), 0, s->tipo, s->kappa, s->u, s->v, s->dens,
      s->countCell, s->solid, s->nx, s->ny, s->d.tensao, kMax, escala, alvo, which, s->qAr, s->div, s->q);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== #propagarRotulos (fisica-cpu.js:409-422)
// The reference runs two raster-order in-place passes, so a new air cell can take its label
// through a chain of other new cells.  Candidates (unlabelled interior air) are compacted; one
// workgroup then iterates each pass to its fixpoint: rot[c] = L || R || D || U with L, D the
// live values (already processed in raster order) and R, U the values at the start of the
// pass (0 for candidates).  The dependency graph is acyclic (raster order), so the fixpoint is
// unique and equals the sequential result.
__device__ __forceinline__ void k_label_prep(Blk B_, const uint8_t* __restrict__ tipo, uint32_t* __restrict__ rot,
                                             uint8_t* __restrict__ cflag, uint32_t* __restrict__ cand,
                                             uint32_t* __restrict__ candCount, int nx, int ny) {
  const int c = B_.x * blockDim.x + threadIdx.x;
  bool is_c = false;
  if (c < nx * ny) {
    const int i = c % nx, j = c / nx;
    if (tipo[c] != AR) rot[c] = 0;
    else if (i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1 && rot[c] == 0) is_c = true;
  }
  unsigned long long b = __ballot(is_c);
  if (!b) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)b) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(candCount, (uint32_t)__popcll(b));
  base = __shfl(base, leader, 64);
  if (is_c) {
    uint32_t off = (uint32_t)__popcll(b & ((1ull << lane) - 1));
    cand[base + off] = (uint32_t)c;
    cflag[c] = 1;
  }
}

__device__ __forceinline__ void k_label_prop(Blk B_, uint32_t* __restrict__ rot, uint8_t* __restrict__ cflag,
                                             const uint32_t* __restrict__ cand,
                                             const uint32_t* __restrict__ candCount, int nx) {
  __shared__ int changed;
  const int tid = threadIdx.x;
  const uint32_t m = *candCount;
  for (int pass = 0; pass < 2; pass++) {
    const uint8_t bit = (uint8_t)(1 << pass);
    while (true) {
      if (tid == 0) changed = 0;
      __syncthreads();
      for (uint32_t k = tid; k < m; k += 1024) {
        const uint32_t c = cand[k];
        if (!(cflag[c] & bit)) continue;
        uint32_t val = rot[c - 1];
        if (!val) val = (cflag[c + 1] & bit) ? 0u : rot[c + 1];
        if (!val) val = rot[c - nx];
        if (!val) val = (cflag[c + nx] & bit) ? 0u : rot[c + nx];
        if (val != rot[c]) { rot[c] = val; changed = 1; }
      }
      __syncthreads();
      const int ch = changed;
      __syncthreads();
      if (!ch) break;
    }
    if (pass == 0)
      for (uint32_t k = tid; k < m; k += 1024) {
        const uint32_t c = cand[k];
        if (rot[c] == 0) cflag[c] |= 2;
      }
    __syncthreads();
  }
  for (uint32_t k = tid; k < m; k += 1024) cflag[cand[k]] = 0;
}

int propagate_labels(tq_sim* s) {
  if (int e = zero_async(s, s->st, &s->sc->cand_count, sizeof(uint32_t))) return e;
  int t = time_begin(s, "label_prep", 6.0 * s->C);
  launch<256, 0, k_label_prep>(s, s->st, dim3((s->C + 255) / 256), dim3(256), 0, s->tipo, s->rot, s->cflag, s->cand,
                               &s->sc->cand_count, s->nx, s->ny);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  t = time_begin(s, "label_prop", 0);
  launch<1024, 0, k_label_prop>(s, s->st, dim3(1), dim3(1024), 0, s->rot, s->cflag, s->cand, &s->sc->cand_count, s->nx);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== #ladoDireitoDasBolhas (fisica-cpu.js:429-454)
// Per-bubble sums: interior air cells of a bubble add their count, their water faces and the
// inflow (u - qAr) of those faces.  Lanes of a wavefront that share a label are reduced first,
// then one atomic per label.  The inflow is summed in 64-bit fixed point (2^-40 m/s steps), so
// the total does not depend on the order of the atomics.
constexpr double RHS_FX = 1099511627776.0;  // 2^40

__device__ inline long long wave_sum_i64(long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
// Per-bubble sums over the air cells: cell count, interior cells, faces to water and the face
// terms of the rhs (#ladoDireitoDasBolhas, fisica-cpu.js:429-454) as 2^-40 fixed point, all
// integers, so the order of the adds does not matter.  Each wave sums per label (ballot
// segments); the wave leaders add into a block table of BRHS_SLOTS labels in LDS; the block
// flushes the table once.  One bubble of millions of cells is then a few thousand global
// atomics instead of one per wave.  Labels the full table cannot take go to global atomics.
constexpr int BRHS_SLOTS = 32;
constexpr int BRHS_GRID = 2048;
__device__ __forceinline__ void k_bubble_rhs(Blk B_, const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                             const float* __restrict__ u, const float* __restrict__ v,
                                             const float* __restrict__ qAr, int nx, int ny,
                                             uint32_t* __restrict__ volB, uint32_t* __restrict__ facesB,
                                             uint32_t* __restrict__ cellsB, double* __restrict__ rhsB,
                                             int counts_only) {
#pragma clang fp contract(off)
  __shared__ uint32_t key[BRHS_SLOTS], kc[BRHS_SLOTS], kv[BRHS_SLOTS], kf[BRHS_SLOTS];
  __shared__ unsigned long long ks[BRHS_SLOTS];
  for (int k = threadIdx.x; k < BRHS_SLOTS; k += blockDim.x) {
    key[k] = 0;  // labels of bubbles are >= 2: 0 marks a free slot
    kc[k] = kv[k] = kf[k] = 0;
    ks[k] = 0;
  }
  __syncthreads();
  const int C = nx * ny;
  const int lane = threadIdx.x & 63;
  for (int base = B_.x * blockDim.x; base < C; base += B_.gx * blockDim.x) {
    const int c = base + threadIdx.x;
    uint32_t r = 0;
    int interior = 0, faces = 0;
    double rhs = 0;
    if (c < C) {
      const int i = c % nx, j = c / nx;
      if (tipo[c] == AR && rot[c] >= 2) {
        r = rot[c];
        if (i >= 1 && j >= 1 && i < nx - 1 && j < ny - 1) {
          interior = 1;
          const double tq = qAr[c];
          if (tipo[c - 1] == AGUA) { rhs += (double)u[(size_t)i + (size_t)j * (nx + 1)] - tq; faces++; }
          if (tipo[c + 1] == AGUA) { rhs += -(double)u[(size_t)i + 1 + (size_t)j * (nx + 1)] - tq; faces++; }
          if (tipo[c - nx] == AGUA) { rhs += (double)v[(size_t)i + (size_t)j * nx] - tq; faces++; }
          if (tipo[c + nx] == AGUA) { rhs += -(double)v[(size_t)i + (size_t)(j + 1) * nx] - tq; faces++; }
        }
      }
    }
    bool pending = r >= 2;
    while (true) {
      unsigned long long b = __ballot(pending);
      if (!b) break;
      const int leader = __ffsll((long long)b) - 1;
      const uint32_t lr = __shfl(r, leader, 64);
      const bool m = pending && r == lr;
      const long long sr = wave_sum_i64(m ? (long long)__builtin_rint(rhs * RHS_FX) : 0LL);
      unsigned long long bm = __ballot(m);
      unsigned long long bi = __ballot(m && interior);
      int sf = (int)wave_sum(m ? (double)faces : 0.0);
      if (lane == leader) {
        int sl = -1;
        for (int k = 0; k < BRHS_SLOTS; k++) {  // open addressing from the label's home slot
          const int hs = (int)((lr + (uint32_t)k) % BRHS_SLOTS);
          const uint32_t old = atomicCAS(&key[hs], 0u, lr);
          if (old == 0u || old == lr) { sl = hs; break; }
        }
        if (sl >= 0) {
          atomicAdd(&kc[sl], (uint32_t)__popcll(bm));
          if (bi) {
            atomicAdd(&kv[sl], (uint32_t)__popcll(bi));
            atomicAdd(&kf[sl], (uint32_t)sf);
            atomicAdd(&ks[sl], (unsigned long long)sr);
          }
        } else {
          atomicAdd(&cellsB[lr], (uint32_t)__popcll(bm));
          if (bi) {
            atomicAdd(&volB[lr], (uint32_t)__popcll(bi));
            atomicAdd(&facesB[lr], (uint32_t)sf);
            if (!counts_only) atomicAdd((unsigned long long*)&rhsB[lr], (unsigned long long)sr);
          }
        }
      }
      if (m) pending = false;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < BRHS_SLOTS; k += blockDim.x) {
    const uint32_t lr = key[k];
    if (!lr) continue;
    atomicAdd(&cellsB[lr], kc[k]);
    if (kv[k]) {
      atomicAdd(&volB[lr], kv[k]);
      atomicAdd(&facesB[lr], kf[k]);
      if (!counts_only) atomicAdd((unsigned long long*)&rhsB[lr], ks[k]);
    }
  }
}

__device__ __forceinline__ void k_bubble_finish(Blk B_, const uint32_t* __restrict__ volB, double* __restrict__ rhsB,
                                                float* __restrict__ rhsB32, float* __restrict__ yB32,
                                                const float* __restrict__ pBolha,
                                                const float* __restrict__ volAlvo, double h, double dt) {
#pragma clang fp contract(off)
  const double paraQ = dt / (DENSIDADE_AGUA * h);
  for (int r = threadIdx.x; r < MAX_BOLHAS; r += blockDim.x) {
    // the fixed-point sum of k_bubble_rhs becomes the binary64 rhs
    const double sum = (double)__double_as_longlong(rhsB[r]) / RHS_FX;
    rhsB[r] = sum;
    if (r < 2 || !volB[r]) { rhsB32[r] = (float)sum; continue; }
    double x = sum + VOLTA_VOLUME_BOLHA * ((double)volAlvo[r] - (double)volB[r]) * h / dt;
    rhsB[r] = x;
    rhsB32[r] = (float)x;
    yB32[r] = (float)((double)pBolha[r] * paraQ);
  }
}

// the per-bubble sums of k_bubble_rhs start at zero (one launch instead of four fills)
__device__ __forceinline__ void k_bubble_clear(Blk B_, uint32_t* __restrict__ volB, uint32_t* __restrict__ facesB,
                                               uint32_t* __restrict__ cellsB, double* __restrict__ rhsB) {
  const int r = B_.x * blockDim.x + threadIdx.x;
  if (r >= MAX_BOLHAS) return;
  volB[r] = 0;
  facesB[r] = 0;
  cellsB[r] = 0;
  if (rhsB) rhsB[r] = 0.0;
}

int bubble_rhs(tq_sim* s, double dt) {
  launch<256, 0, k_bubble_clear>(s, s->st, dim3((MAX_BOLHAS + 255) / 256), dim3(256), 0, s->volB, s->facesB, s->cellsB,
                                 s->rhsB);
  int t = time_begin(s, "bubble_rhs", (1 + 4 + 4 + 4 + 4) * (double)s->C);
  launch<256, 0, k_bubble_rhs>(s, s->st, dim3(std::min((s->C + 255) / 256, BRHS_GRID)), dim3(256), 0, s->tipo, s->rot,
                               s->u, s->v, s->qAr, s->nx, s->ny, s->volB, s->facesB, s->cellsB, s->rhsB, 0);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  launch<1024, 0, k_bubble_finish>(s, s->st, dim3(1), dim3(1024), 0, s->volB, s->rhsB, s->rhsB32, s->yB32, s->pBolha,
                                   s->volAlvo, s->h, dt);
  TQ_LAUNCH_CHECK();
  return 0;
}

// Per-bubble cell and face counts only (the rhs is kept): lets a solve start from a state whose
// bubble sums were computed elsewhere.
int bubble_counts(tq_sim* s) {
  launch<256, 0, k_bubble_clear>(s, s->st, dim3((MAX_BOLHAS + 255) / 256), dim3(256), 0, s->volB, s->facesB, s->cellsB,
                                 (double*)nullptr);
  launch<256, 0, k_bubble_rhs>(s, s->st, dim3(std::min((s->C + 255) / 256, BRHS_GRID)), dim3(256), 0, s->tipo, s->rot,
                               s->u, s->v, s->qAr, s->nx, s->ny, s->volB, s->facesB, s->cellsB, s->rhsB, 1);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== #pressao post (fisica-cpu.js:503-514)
__device__ __forceinline__ void k_pressure_post(Blk B_, const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                                const float* __restrict__ qAr, const uint32_t* __restrict__ volB,
                                                const float* __restrict__ yB32, int C, int ar,
                                                float* __restrict__ q) {
#pragma clang fp contract(off)
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c >= C || tipo[c] != AR) return;
  const uint32_t r = rot[c];
  q[c] = (float)((double)qAr[c] + ((ar && r >= 2 && volB[r]) ? (double)yB32[r] : 0.0));
}

__device__ __forceinline__ void k_bubble_pressure(Blk B_, const uint32_t* __restrict__ volB, const float* __restrict__ yB32,
                                                  float* __restrict__ pBolha, double paraPa) {
#pragma clang fp contract(off)
  for (int r = threadIdx.x; r < MAX_BOLHAS; r += blockDim.x)
    if (r >= 2 && volB[r])
      pBolha[r] = (float)fmax(-PRESSAO_BOLHA_MAX, fmin(PRESSAO_BOLHA_MAX, (double)yB32[r] * paraPa));
}

int pressure_post(tq_sim* s, double dt) {
  launch<256, 0, k_pressure_post>(s, s->st, dim3((s->C + 255) / 256), dim3(256), 0, s->tipo, s->rot, s->qAr, s->volB,
                                  s->yB32, s->C, s->d.ar, s->q);
  TQ_LAUNCH_CHECK();
  if (s->d.ar) {
    launch<1024, 0, k_bubble_pressure>(s, s->st, dim3(1), dim3(1024), 0, s->volB, s->yB32, s->pBolha,
                                       (DENSIDADE_AGUA * s->h) / dt);
    TQ_LAUNCH_CHECK();
  }
  return 0;
}

// ================================================================== #gradiente (fisica-cpu.js:519-545)
__device__ __forceinline__ void k_gradient(Blk B_, const uint8_t* __restrict__ tipo, const float* __restrict__ q, int nx,
                                           int ny, float* __restrict__ u, uint8_t* __restrict__ uVal,
                                           float* __restrict__ v, uint8_t* __restrict__ vVal) {
#pragma clang fp contract(off)
  const int k = B_.x * blockDim.x + threadIdx.x;
  const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
  if (k < nu) {
    const int i = k % (nx + 1), j = k / (nx + 1);
    const uint8_t tl = i > 0 ? tipo[(size_t)(i - 1) + (size_t)j * nx] : SOLIDO;
    const uint8_t tr = i < nx ? tipo[(size_t)i + (size_t)j * nx] : SOLIDO;
    if (tl == SOLIDO || tr == SOLIDO) { u[k] = 0.f; uVal[k] = 1; }
    else if (tl == AGUA || tr == AGUA) {
      u[k] = (float)((double)u[k] - ((double)q[(size_t)i + (size_t)j * nx] - (double)q[(size_t)(i - 1) + (size_t)j * nx]));
      uVal[k] = 1;
    } else uVal[k] = 0;
  }
  if (k < nv) {
    const int i = k % nx, j = k / nx;
    const uint8_t tb = j > 0 ? tipo[(size_t)i + (size_t)(j - 1) * nx] : SOLIDO;
    const uint8_t tt = j < ny ? tipo[(size_t)i + (size_t)j * nx] : SOLIDO;
    if (tb == SOLIDO || tt == SOLIDO) { v[k] = 0.f; vVal[k] = 1; }
    else if (tb == AGUA || tt == AGUA) {
      v[k] = (float)((double)v[k] - ((double)q[(size_t)i + (size_t)j * nx] - (double)q[(size_t)i + (size_t)(j - 1) * nx]));
      vVal[k] = 1;
    } else vVal[k] = 0;
  }
}

int gradient(tq_sim* s) {
  const int n = std::max(s->nu, s->nv);
  int t = time_begin(s, "gradient", 9.0 * (s->nu + s->nv));
  launch<256, 0, k_gradient>(s, s->st, dim3((n + 255) / 256), dim3(256), 0, s->tipo, s->q, s->nx, s->ny, s->u, s->uVal,
                             s->v, s->vVal);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== #extrapolar (fisica-cpu.js:547-552, 673-690)
// One layer for both components: an invalid face with valid 4-neighbours takes their mean.
__device__ __forceinline__ void k_extrap(Blk B_, const float* __restrict__ ua, const uint8_t* __restrict__ uva,
                                         float* __restrict__ ub, uint8_t* __restrict__ uvb,
                                         const float* __restrict__ va, const uint8_t* __restrict__ vva,
                                         float* __restrict__ vb, uint8_t* __restrict__ vvb, int nx, int ny) {
#pragma clang fp contract(off)
  const int k = B_.x * blockDim.x + threadIdx.x;
  const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
  for (int comp = 0; comp < 2; comp++) {
    const float* a = comp ? va : ua;
    const uint8_t* val = comp ? vva : uva;
    float* o = comp ? vb : ub;
    uint8_t* ov = comp ? vvb : uvb;
    const int w = comp ? nx : nx + 1, hh = comp ? ny + 1 : ny, n = comp ? nv : nu;
    if (k >= n) continue;
    float res = a[k];
    uint8_t vres = val[k];
    if (!vres) {
      const int i = k % w, j = k / w;
      double soma = 0;
      int m = 0;
      if (i > 0 && val[k - 1]) { soma += (double)a[k - 1]; m++; }
      if (i < w - 1 && val[k + 1]) { soma += (double)a[k + 1]; m++; }
      if (j > 0 && val[k - w]) { soma += (double)a[k - w]; m++; }
      if (j < hh - 1 && val[k + w]) { soma += (double)a[k + w]; m++; }
      if (m > 0) { res = (float)(soma / m); vres = 1; }
    }
    o[k] = res;
    ov[k] = vres;
  }
}

int extrapolate(tq_sim* s) {
  const int n = std::max(s->nu, s->nv);
  int t = time_begin(s, "extrapolate", 10.0 * (s->nu + s->nv));
  launch<256, 0, k_extrap>(s, s->st, dim3((n + 255) / 256), dim3(256), 0, s->u, s->uVal, s->uTmp, s->uValTmp, s->v,
                           s->vVal, s->vTmp, s->vValTmp, s->nx, s->ny);
  launch<256, 0, k_extrap>(s, s->st, dim3((n + 255) / 256), dim3(256), 0, s->uTmp, s->uValTmp, s->u, s->uVal, s->vTmp,
                           s->vValTmp, s->v, s->vVal, s->nx, s->ny);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

}  // namespace tq
