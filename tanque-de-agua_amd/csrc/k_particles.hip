// k_particles.hip -- particle-side kernels of the substep:
//   counting sort by (tile-major) cell, the fused P2G (APIC scatter + gravity/boundary + types),
//   G2P with micro-steps and wall collision, particle adds, tools, and the scan they share.
#include "tq_device.hpp"

namespace tq {

// ================================================================== exclusive scan (u32)
// Three kernels: per-block sums (4096 items per 1024-thread block), one-block scan of the block
// sums, per-block local scan plus offset.  out[n] receives the total.
constexpr int SCAN_T = 1024, SCAN_V = 4, SCAN_B = SCAN_T * SCAN_V;

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) { uint32_t t = sh[w]; sh[w] = acc; acc += t; }
    sh[16] = acc;
  }
  __syncthreads();
  uint32_t r = sh[wid] + x - v;
  *total = sh[16];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void k_scan_reduce(Blk B_, const uint32_t* __restrict__ in, int n, uint32_t* __restrict__ blocks) {
  __shared__ uint32_t sh[17];
  const size_t base = (size_t)B_.x * SCAN_B + (size_t)threadIdx.x * SCAN_V;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_V; k++) if (base + k < (size_t)n) s += in[base + k];
  uint32_t tot;
  block_excl_scan(s, sh, &tot);
  if (threadIdx.x == 0) blocks[B_.x] = tot;
}

__device__ __forceinline__ void k_scan_blocks(Blk B_, uint32_t* __restrict__ blocks, int nb, uint32_t* __restrict__ out_total) {
  __shared__ uint32_t sh[17];
  uint32_t v[SCAN_V], s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_V; k++) {
    int i = threadIdx.x * SCAN_V + k;
    v[k] = i < nb ? blocks[i] : 0;
    s += v[k];
  }
  uint32_t tot;
  uint32_t off = block_excl_scan(s, sh, &tot);
#pragma unroll
  for (int k = 0; k < SCAN_V; k++) {
    int i = threadIdx.x * SCAN_V + k;
    if (i < nb) blocks[i] = off;
    off += v[k];
  }
  if (threadIdx.x == 0) *out_total = tot;
}

__device__ __forceinline__ void k_scan_apply(Blk B_, const uint32_t* __restrict__ in, int n, const uint32_t* __restrict__ blocks,
                                             uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[17];
  const size_t base = (size_t)B_.x * SCAN_B + (size_t)threadIdx.x * SCAN_V;
  uint32_t v[SCAN_V], s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_V; k++) {
    v[k] = base + k < (size_t)n ? in[base + k] : 0;
    s += v[k];
  }
  uint32_t tot;
  uint32_t off = block_excl_scan(s, sh, &tot) + blocks[B_.x];
#pragma unroll
  for (int k = 0; k < SCAN_V; k++) {
    if (base + k < (size_t)n) out[base + k] = off;
    off += v[k];
  }
}

// Exclusive scan of n values (out[n] = total): per-block sums, a scan of the block sums
// (itself two-level past 4096 blocks), then the per-block scans with their offsets.
int scan_u32(tq_sim* s, const uint32_t* in, uint32_t* out, int n) {
  if (n <= 0) {
    if (int e = zero_async(s, s->st, out, sizeof(uint32_t))) return e;
    return 0;
  }
  const int nb = (n + SCAN_B - 1) / SCAN_B;
  const int nb2 = (nb + SCAN_B - 1) / SCAN_B;
  if (nb2 > SCAN_B || 2 * nb + nb2 + 2 > s->scan_cap) { set_error("scan: input too large"); return TQ_E_CAPACITY; }
  uint32_t* blocks = s->scanBlocks;            // nb block sums
  uint32_t* blocks2 = blocks + nb;             // nb2 sums of those (two-level case)
  uint32_t* offs = blocks2 + nb2 + 1;          // nb block offsets (two-level case)
  int t = time_begin(s, "scan", 8.0 * n);
  launch<SCAN_T, 0, k_scan_reduce>(s, s->st, dim3(nb), dim3(SCAN_T), 0, in, n, blocks);
  if (nb <= SCAN_B) {
    launch<SCAN_T, 0, k_scan_blocks>(s, s->st, dim3(1), dim3(SCAN_T), 0, blocks, nb, out + n);
  } else {
    launch<SCAN_T, 0, k_scan_reduce>(s, s->st, dim3(nb2), dim3(SCAN_T), 0, blocks, nb, blocks2);
    launch<SCAN_T, 0, k_scan_blocks>(s, s->st, dim3(1), dim3(SCAN_T), 0, blocks2, nb2, out + n);
    launch<SCAN_T, 0, k_scan_apply>(s, s->st, dim3(nb2), dim3(SCAN_T), 0, blocks, nb, blocks2, offs);
    blocks = offs;
  }
  launch<SCAN_T, 0, k_scan_apply>(s, s->st, dim3(nb), dim3(SCAN_T), 0, in, n, blocks, out);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== binning + counting sort
// Cell key and rank within the cell of every particle.  Particles arrive in cell order (the
// previous substep's sort), so runs of lanes share a key: the run's first lane takes the run's
// ranks with one atomic and the others count from it, instead of one same-address atomic per
// lane.  Ranks within a cell depend on the order the atomics land; every consumer of the sorted
// order sums in fixed point or works per particle, so results do not.
__device__ __forceinline__ void k_bin(Blk B_, const float2* __restrict__ pos, const DevScalars* __restrict__ sc, int nx, int ny,
                                      double inv, int ntx, uint32_t* __restrict__ countKey,
                                      uint32_t* __restrict__ pkey, uint32_t* __restrict__ prank) {
  const uint32_t k = B_.x * blockDim.x + threadIdx.x;
  const bool slot = k < sc->n_slots;
  uint32_t key = DEAD_KEY;
  if (slot) {
    const float2 p = pos[k];
    if (p.x == p.x) {
      int ci, cj;
      count_cell(p.x, p.y, inv, nx, ny, &ci, &cj);
      key = cell_key(ci, cj, ntx);
    }
    pkey[k] = key;
  }
  const int lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || key != prev;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1);
  const int hl = 63 - __clzll(heads & upto);  // this lane's run head
  uint32_t base = 0;
  if (head && key != DEAD_KEY) {
    const unsigned long long after = heads & ~upto;
    const int next = after ? __ffsll((long long)after) - 1 : 64;
    base = atomicAdd(&countKey[key], (uint32_t)(next - lane));
  }
  base = __shfl(base, hl, 64);
  if (slot && key != DEAD_KEY) prank[k] = base + (uint32_t)(lane - hl);
}

__device__ __forceinline__ void k_scatter(Blk B_, const float2* __restrict__ pos, const float2* __restrict__ vel,
                                          const float4* __restrict__ cm, const float4* __restrict__ ext,
                                          const uint32_t* __restrict__ seq, const DevScalars* __restrict__ sc,
                                          const uint32_t* __restrict__ pkey, const uint32_t* __restrict__ prank,
                                          const uint32_t* __restrict__ startKey, float2* __restrict__ pos2,
                                          float2* __restrict__ vel2, float4* __restrict__ cm2,
                                          float4* __restrict__ ext2, uint32_t* __restrict__ seq2) {
  const uint32_t k = B_.x * blockDim.x + threadIdx.x;
  if (k >= sc->n_slots) return;
  uint32_t key = pkey[k];
  if (key == DEAD_KEY) return;
  uint32_t d = startKey[key] + prank[k];
  pos2[d] = pos[k];
  vel2[d] = vel[k];
  cm2[d] = cm[k];
  ext2[d] = ext[k];
  seq2[d] = seq[k];
}

__device__ __forceinline__ void k_set_slots(Blk B_, DevScalars* sc, const uint32_t* __restrict__ total) {
  if (threadIdx.x == 0 && B_.x == 0) { sc->n_slots = *total; sc->n_live = *total; }
}

int bin_and_sort(tq_sim* s) {
  const int n = s->n_upper;
  if (int e = zero_async(s, s->st, s->countKey, sizeof(uint32_t) * s->K)) return e;
  if (n > 0) {
    int t = time_begin(s, "bin", 16.0 * n);
    launch<256, 0, k_bin>(s, s->st, dim3((n + 255) / 256), dim3(256), 0, s->pos[s->cur], s->sc, s->nx, s->ny,
                          1.0 / s->h, s->ntx, s->countKey, s->pkey, s->prank);
    time_end(s, t);
    TQ_LAUNCH_CHECK();
  }
  if (int e = scan_u32(s, s->countKey, s->startKey, s->K)) return e;
  if (n > 0) {
    int o = s->cur, d = 1 - s->cur;
    int t = time_begin(s, "scatter", 2.0 * 60 * n);
    launch<256, 0, k_scatter>(s, s->st, dim3((n + 255) / 256), dim3(256), 0, s->pos[o], s->vel[o], s->cm[o], s->ext[o],
        s->seq[o], s->sc, s->pkey, s->prank, s->startKey, s->pos[d], s->vel[d], s->cm[d], s->ext[d],
        s->seq[d]);
    time_end(s, t);
    TQ_LAUNCH_CHECK();
    s->cur = d;
  }
  launch<64, 0, k_set_slots>(s, s->st, dim3(1), dim3(1), 0, s->sc, s->startKey + s->K);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== P2G (fused)
// One workgroup per 16x16 tile.  The tile OWNS the u faces (i, j) of its cells (and face nx
// for the last tile column), the v faces likewise (and face ny for the last tile row), and its
// cells' density/dye.  It scatters its own particles plus those of the 1-cell ring around it
// (the only other particles whose stencils reach owned nodes) into LDS with 64-bit
// fixed-point integer atomics, so the sums are exact and order-independent, then finalises
// the owned nodes: u = sum(w v)/sum(w), + g dt where weighted, zero at solid/domain faces
// (#paraGrade + #gravidadeEFronteira + #tipos, fisica-cpu.js:198-299).
__device__ __forceinline__ void k_p2g(Blk B_, const float2* __restrict__ pos, const float2* __restrict__ vel,
                                      const float4* __restrict__ cm, const float4* __restrict__ ext,
                                      const uint32_t* __restrict__ startKey, int nx, int ny, int ntx, int K,
                                      double h, const uint8_t* __restrict__ solid, double gdx, double gdy,
                                      int gravity, float* __restrict__ u, float* __restrict__ uw,
                                      float* __restrict__ v, float* __restrict__ vw, float* __restrict__ dens,
                                      float* __restrict__ dye, uint32_t* __restrict__ countCell,
                                      uint8_t* __restrict__ tipo, DevScalars* __restrict__ sc) {
#pragma clang fp contract(off)
  __shared__ unsigned long long aU[17 * 16], wU[17 * 16], aV[16 * 17], wV[16 * 17], aD[256], aC[3 * 256];
  __shared__ uint32_t rStart[68], rOff[69];
  __shared__ uint32_t ownStart, ownCount;
  const int tid = threadIdx.x;
  const int I0 = B_.x * TILE, J0 = B_.y * TILE;
  const int I1 = min(I0 + TILE, nx), J1 = min(J0 + TILE, ny);
  const double inv = 1 / h;

  for (int e = tid; e < 17 * 16; e += 256) { aU[e] = 0; wU[e] = 0; aV[e] = 0; wV[e] = 0; }
  aD[tid] = 0;
  aC[tid] = 0; aC[tid + 256] = 0; aC[tid + 512] = 0;
  if (tid < 68) {
    int gi, gj;
    if (tid < 18) { gj = J0 - 1; gi = I0 - 1 + tid; }
    else if (tid < 36) { gj = J0 + TILE; gi = I0 - 1 + (tid - 18); }
    else if (tid < 52) { gi = I0 - 1; gj = J0 + (tid - 36); }
    else { gi = I0 + TILE; gj = J0 + (tid - 52); }
    uint32_t st = 0, cnt = 0;
    if (gi >= 0 && gj >= 0 && gi < nx && gj < ny) {
      uint32_t key = cell_key(gi, gj, ntx);
      st = startKey[key];
      cnt = startKey[key + 1] - st;
    }
    rStart[tid] = st;
    rOff[tid] = cnt;
  }
  if (tid == 0) {
    uint32_t base = (uint32_t)(B_.y * ntx + B_.x) * 256u;
    ownStart = startKey[base];
    ownCount = startKey[base + 256] - ownStart;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (int r = 0; r < 68; r++) { uint32_t c = rOff[r]; rOff[r] = acc; acc += c; }
    rOff[68] = acc;
  }
  __syncthreads();
  const uint32_t total = ownCount + rOff[68];
  for (uint32_t w = tid; w < total; w += 256) {
    uint32_t k;
    if (w < ownCount) k = ownStart + w;
    else {
      uint32_t t = w - ownCount;
      int lo = 0, hi = 67;  // last r with rOff[r] <= t
      while (lo < hi) { int mid = (lo + hi + 1) >> 1; if (rOff[mid] <= t) lo = mid; else hi = mid - 1; }
      k = rStart[lo] + (t - rOff[lo]);
    }
    const float2 P = pos[k];
    const float2 V = vel[k];
    const float4 Cm = cm[k];
    const double x = P.x, y = P.y, vx = V.x, vy = V.y;
    const double cux = Cm.x, cuy = Cm.y, cvx = Cm.z, cvy = Cm.w;
    // u faces: nodes at (i h, (j + 1/2) h)
    {
      double fx = x * inv, fy = y * inv - 0.5;
      double i0 = clampd(floor(fx), 0, nx - 1), j0 = clampd(floor(fy), 0, ny - 2);
      double tx = fx - i0, ty = fy - j0;
      int ii = (int)i0, jj = (int)j0;
      for (int b = 0; b < 2; b++) {
        double wy = b ? ty : 1 - ty;
        double dy = (j0 + b + 0.5) * h - y;
        int j = jj + b;
        if (j < J0 || j >= J1) continue;
        for (int a = 0; a < 2; a++) {
          int i = ii + a;
          if (!((i >= I0 && i < I1) || (i == nx && I1 == nx))) continue;
          double wgt = (a ? tx : 1 - tx) * wy;
          double dx = (i0 + a) * h - x;
          double val = wgt * (vx + cux * dx + cuy * dy);
          int e = (i - I0) + (j - J0) * 17;
          atomicAdd(&aU[e], (unsigned long long)to_fx(val));
          atomicAdd(&wU[e], (unsigned long long)to_fx(wgt));
        }
      }
    }
    // v faces: nodes at ((i + 1/2) h, j h)
    {
      double fx = x * inv - 0.5, fy = y * inv;
      double i0 = clampd(floor(fx), 0, nx - 2), j0 = clampd(floor(fy), 0, ny - 1);
      double tx = fx - i0, ty = fy - j0;
      int ii = (int)i0, jj = (int)j0;
      for (int b = 0; b < 2; b++) {
        double wy = b ? ty : 1 - ty;
        double dy = (j0 + b) * h - y;
        int j = jj + b;
        if (!((j >= J0 && j < J1) || (j == ny && J1 == ny))) continue;
        for (int a = 0; a < 2; a++) {
          int i = ii + a;
          if (i < I0 || i >= I1) continue;
          double wgt = (a ? tx : 1 - tx) * wy;
          double dx = (i0 + a + 0.5) * h - x;
          double val = wgt * (vy + cvx * dx + cvy * dy);
          int e = (i - I0) + (j - J0) * 16;
          atomicAdd(&aV[e], (unsigned long long)to_fx(val));
          atomicAdd(&wV[e], (unsigned long long)to_fx(wgt));
        }
      }
    }
    // density at cell centres
    {
      double fx = x * inv - 0.5, fy = y * inv - 0.5;
      double i0 = clampd(floor(fx), 0, nx - 2), j0 = clampd(floor(fy), 0, ny - 2);
      double tx = fx - i0, ty = fy - j0;
      int ii = (int)i0, jj = (int)j0;
      double wv[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
      for (int q = 0; q < 4; q++) {
        int i = ii + (q & 1), j = jj + (q >> 1);
        if (i < I0 || i >= I1 || j < J0 || j >= J1) continue;
        atomicAdd(&aD[(i - I0) + (j - J0) * 16], (unsigned long long)to_fx(wv[q]));
      }
    }
    // dye in the count cell (own particles only: their count cell is in this tile)
    if (w < ownCount) {
      int ci, cj;
      count_cell(P.x, P.y, inv, nx, ny, &ci, &cj);
      const float4 E = ext[k];
      int e = (ci - I0) + (cj - J0) * 16;
      atomicAdd(&aC[e], (unsigned long long)to_fx(E.y));
      atomicAdd(&aC[256 + e], (unsigned long long)to_fx(E.z));
      atomicAdd(&aC[512 + e], (unsigned long long)to_fx(E.w));
    }
  }
  __syncthreads();
  // ---- finalise owned u faces
  for (int e = tid; e < 17 * 16; e += 256) {
    int li = e % 17, lj = e / 17;
    int i = I0 + li, j = J0 + lj;
    if (j >= J1 || !((i < I1) || (i == nx && I1 == nx))) continue;
    long long wsum = (long long)wU[e], asum = (long long)aU[e];
    float uwf = (float)((double)wsum * FX_INV);
    float uf = uwf > 0 ? (float)((double)asum / (double)wsum) : 0.f;
    if (gravity) {
      if (uwf > 0) uf = (float)((double)uf + gdx);
      bool sl = i == 0 || i == nx || solid[i - 1 + j * nx] || solid[i + j * nx];
      if (sl) uf = 0;
    }
    size_t idx = (size_t)i + (size_t)j * (nx + 1);
    u[idx] = uf;
    uw[idx] = uwf;
  }
  // ---- finalise owned v faces
  for (int e = tid; e < 16 * 17; e += 256) {
    int li = e % 16, lj = e / 16;
    int i = I0 + li, j = J0 + lj;
    if (i >= I1 || !((j < J1) || (j == ny && J1 == ny))) continue;
    long long wsum = (long long)wV[e], asum = (long long)aV[e];
    float vwf = (float)((double)wsum * FX_INV);
    float vf = vwf > 0 ? (float)((double)asum / (double)wsum) : 0.f;
    if (gravity) {
      if (vwf > 0) vf = (float)((double)vf + gdy);
      bool sl = j == 0 || j == ny || solid[i + (j - 1) * nx] || solid[i + j * nx];
      if (sl) vf = 0;
    }
    size_t idx = (size_t)i + (size_t)j * nx;
    v[idx] = vf;
    vw[idx] = vwf;
  }
  // ---- cells: density, dye, count, type
  {
    int li = tid & 15, lj = tid >> 4;
    int i = I0 + li, j = J0 + lj;
    if (i < I1 && j < J1) {
      size_t c = (size_t)i + (size_t)j * nx;
      dens[c] = (float)((double)(long long)aD[tid] * FX_INV);
      dye[3 * c] = (float)((double)(long long)aC[tid] * FX_INV);
      dye[3 * c + 1] = (float)((double)(long long)aC[256 + tid] * FX_INV);
      dye[3 * c + 2] = (float)((double)(long long)aC[512 + tid] * FX_INV);
      uint32_t key = cell_key(i, j, ntx);
      uint32_t cnt = startKey[key + 1] - startKey[key];
      countCell[c] = cnt;
      tipo[c] = solid[c] ? SOLIDO : (cnt > 0 ? AGUA : AR);
    }
  }
}

int p2g(tq_sim* s, double dt, int gravity) {
  if (int e = bin_and_sort(s)) return e;
  const int o = s->cur;
  // algorithmic bytes: particle read 32 B (+12 B dye) per particle; u,v,uw,vw out; dens, dye,
  // count, type out; solid in
  double bytes = 44.0 * s->n_upper + 8.0 * (s->nu + s->nv) + (4 + 12 + 4 + 1 + 1) * (double)s->C;
  int t = time_begin(s, "p2g", bytes);
  launch<256, 0, k_p2g>(s, s->st, dim3(s->ntx, s->nty), dim3(256), 0, s->pos[o], s->vel[o], s->cm[o], s->ext[o],
                        s->startKey, s->nx, s->ny, s->ntx, s->K, s->h, s->solid, s->d.gx * dt, s->d.gy * dt, gravity,
                        s->u, s->uw, s->v, s->vw, s->dens, s->dye, s->countCell, s->tipo, s->sc);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== G2P (#paraParticulas, fisica-cpu.js:556-661)
__device__ inline double suave(double a, double b, double x) {
#pragma clang fp contract(off)
  double t = fmin(1.0, fmax(0.0, (x - a) / (b - a)));
  return t * t * (3 - 2 * t);
}

__device__ inline int safe_idx(double v) { return v == v ? (int)v : 0; }

// #velocidadeEm (fisica-cpu.js:641-661)
__device__ inline void velocidade_em(const float* __restrict__ u, const float* __restrict__ v, int nx, int ny, double h,
                                     double x, double y, double vmax, double* wx, double* wy) {
#pragma clang fp contract(off)
  const double inv = 1 / h;
  double vx, vy;
  {
    double fx = x * inv, fy = y * inv - 0.5;
    double i0 = clampd(floor(fx), 0, nx - 1), j0 = clampd(floor(fy), 0, ny - 2);
    double tx = fx - i0, ty = fy - j0;
    size_t k = (size_t)safe_idx(i0) + (size_t)safe_idx(j0) * (nx + 1);
    vx = ((double)u[k] * (1 - tx) + (double)u[k + 1] * tx) * (1 - ty) +
         ((double)u[k + nx + 1] * (1 - tx) + (double)u[k + nx + 2] * tx) * ty;
  }
  {
    double fx = x * inv - 0.5, fy = y * inv;
    double i0 = clampd(floor(fx), 0, nx - 2), j0 = clampd(floor(fy), 0, ny - 1);
    double tx = fx - i0, ty = fy - j0;
    size_t k = (size_t)safe_idx(i0) + (size_t)safe_idx(j0) * nx;
    vy = ((double)v[k] * (1 - tx) + (double)v[k + 1] * tx) * (1 - ty) +
         ((double)v[k + nx] * (1 - tx) + (double)v[k + nx + 1] * tx) * ty;
  }
  double vv = js_hypot(vx, vy);
  if (vv > vmax) { *wx = vx * vmax / vv; *wy = vy * vmax / vv; } else { *wx = vx; *wy = vy; }
}

__device__ __forceinline__ void k_g2p(Blk B_, float2* __restrict__ pos, float2* __restrict__ vel, float4* __restrict__ cm,
                                      float4* __restrict__ ext, const DevScalars* __restrict__ scr,
                                      const float* __restrict__ u, const float* __restrict__ v,
                                      const float* __restrict__ dens, const uint32_t* __restrict__ countCell,
                                      const float* __restrict__ dye, const float* __restrict__ sdf, int nx,
                                      int ny, double h, double dt, double decai, double mistura,
                                      unsigned long long* __restrict__ vmax_bits,
                                      uint16_t* __restrict__ flags) {
#pragma clang fp contract(off)
  // particles are in cell order: consecutive blocks on one XCD gather from one band of rows
  const uint32_t k = xcd_block(B_.x, B_.gx) * blockDim.x + threadIdx.x;
  double maior = 0;
  if (k < scr->n_slots) {
    const float2 P = pos[k];
    if (P.x == P.x) {
      const double inv = 1 / h;
      const int w = 2 * nx, hh = 2 * ny;
      const double vmax = (MICROPASSO * MAX_MICROPASSOS * h) / dt;
      double x = P.x, y = P.y;
      double vx = 0, vy = 0, cux = 0, cuy = 0, cvx = 0, cvy = 0;
      {
        double fx = x * inv, fy = y * inv - 0.5;
        double i0 = clampd(floor(fx), 0, nx - 1), j0 = clampd(floor(fy), 0, ny - 2);
        double tx = fx - i0, ty = fy - j0;
        int ii = safe_idx(i0), jj = safe_idx(j0);
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            double val = u[(size_t)(ii + a) + (size_t)(jj + b) * (nx + 1)];
            double wx = a ? tx : 1 - tx, wy = b ? ty : 1 - ty;
            vx += wx * wy * val;
            cux += (a ? 1 : -1) * wy * inv * val;
            cuy += wx * (b ? 1 : -1) * inv * val;
          }
      }
      {
        double fx = x * inv - 0.5, fy = y * inv;
        double i0 = clampd(floor(fx), 0, nx - 2), j0 = clampd(floor(fy), 0, ny - 1);
        double tx = fx - i0, ty = fy - j0;
        int ii = safe_idx(i0), jj = safe_idx(j0);
        for (int b = 0; b < 2; b++)
          for (int a = 0; a < 2; a++) {
            double val = v[(size_t)(ii + a) + (size_t)(jj + b) * nx];
            double wx = a ? tx : 1 - tx, wy = b ? ty : 1 - ty;
            vy += wx * wy * val;
            cvx += (a ? 1 : -1) * wy * inv * val;
            cvy += wx * (b ? 1 : -1) * inv * val;
          }
      }
      const double velm = js_hypot(vx, vy);
      int ci, cj;
      count_cell(P.x, P.y, inv, nx, ny, &ci, &cj);
      const size_t cidx = (size_t)ci + (size_t)cj * nx;
      const double rarefeita = suave(0.95, 0.45, (double)dens[cidx] / PPC);
      const double fonte = suave(ESPUMA_V0, ESPUMA_V1, velm) * rarefeita;
      float4 E = ext[k];
      const double espuma = fmin(1.0, (double)E.x * decai + fonte * ESPUMA_TAXA * dt);
      const uint32_t nCel = countCell[cidx];
      if (nCel > 1) {
        const size_t cc = cidx * 3;
        E.y = (float)((double)E.y + ((double)dye[cc] / nCel - (double)E.y) * mistura);
        E.z = (float)((double)E.z + ((double)dye[cc + 1] / nCel - (double)E.z) * mistura);
        E.w = (float)((double)E.w + ((double)dye[cc + 2] / nCel - (double)E.w) * mistura);
      }
      if (velm > vmax) { vx *= vmax / velm; vy *= vmax / velm; }
      if (velm > maior) maior = velm;
      const int micro = (int)fmin((double)MAX_MICROPASSOS, fmax(1.0, ceil((fmin(velm, vmax) * dt) / (MICROPASSO * h))));
      const double dm = dt / micro;
      bool morta = false;
      // decision word for the reverse pass: micro - 1, push-out steps (bit 2 + m), inward
      // velocity removed (bit 6 + m), removal (bit 10)
      uint32_t fw = (uint32_t)(micro - 1);
      for (int m = 0; m < micro; m++) {
        double wx = vx, wy = vy;
        if (micro > 1) velocidade_em(u, v, nx, ny, h, x, y, vmax, &wx, &wy);
        x += wx * dm;
        y += wy * dm;
        double d = sdf_distancia(sdf, w, hh, h, x, y);
        if (d < RAIO_PARTICULA) {
          fw |= 1u << (2 + m);
          if (d < -SDF_MAX + 0.5) { morta = true; fw |= 1u << 10; break; }
          double nX, nY;
          sdf_normal(sdf, w, hh, h, x, y, &nX, &nY);
          double empurra = (RAIO_PARTICULA - d) * h;
          x += nX * empurra;
          y += nY * empurra;
          double vn = vx * nX + vy * nY;
          if (vn < 0) { vx -= vn * nX; vy -= vn * nY; fw |= 1u << (6 + m); }
        }
      }
      if (flags) flags[k] = (uint16_t)fw;
      if (morta) {
        pos[k] = make_float2(__builtin_nanf(""), P.y);
      } else {
        pos[k] = make_float2((float)x, (float)y);
        vel[k] = make_float2((float)vx, (float)vy);
        cm[k] = make_float4((float)cux, (float)cuy, (float)cvx, (float)cvy);
        E.x = (float)espuma;
        ext[k] = E;
      }
    }
  }
  // top speed: block maximum, then one atomic per block and only when it beats the value already
  // there (every block's atomic on the one word would serialise in the cache)
  __shared__ double wmax[4];
  maior = wave_maxd(maior);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = maior;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) m = fmax(m, wmax[q]);
    const unsigned long long cur = __hip_atomic_load(vmax_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (m > 0 && (unsigned long long)__double_as_longlong(m) > cur)
      atomicMax(vmax_bits, (unsigned long long)__double_as_longlong(m));
  }
}

int g2p(tq_sim* s, double dt) {
  const int n = s->n_upper;
  if (int e = zero_async(s, s->st, &s->sc->vmax_bits, sizeof(unsigned long long))) return e;
  if (n > 0) {
    const int o = s->cur;
    double decai = std::exp(-dt / ESPUMA_VIDA);
    double mistura = std::min(1.0, dt / DIFUSAO_CORANTE);
    // algorithmic bytes per particle: read pos 8 + ext 16, write pos/vel 16 + cm 16 + ext 16;
    // grid reads (u, v 4x4 B each, cell 24 B) are cache-resident gathers
    double bytes = (8 + 16 + 16 + 16 + 16) * (double)n + 24.0 * s->C;
    int t = time_begin(s, "g2p", bytes);
    launch<256, 0, k_g2p>(s, s->st, dim3((n + 255) / 256), dim3(256), 0, s->pos[o], s->vel[o], s->cm[o], s->ext[o],
                          s->sc, s->u, s->v, s->dens, s->countCell, s->dye, s->sdf, s->nx, s->ny, s->h, dt, decai,
                          mistura, &s->sc->vmax_bits, s->g2p_flags);
    time_end(s, t);
    TQ_LAUNCH_CHECK();
  }
  s->has_vmax = true;
  return 0;
}

// ================================================================== adicionar (fisica-cpu.js:92-111)
__device__ __forceinline__ void k_add_flags(Blk B_, const float* __restrict__ list, int count, int check,
                                            const float* __restrict__ sdf, int nx, int ny, double h,
                                            const uint32_t* __restrict__ countCell, uint32_t* __restrict__ flag) {
#pragma clang fp contract(off)
  const int k = B_.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const double x = list[(size_t)k * 12], y = list[(size_t)k * 12 + 1];
  uint32_t ok = 1;
  double d = sdf_distancia(sdf, 2 * nx, 2 * ny, h, x, y);
  if (!(d > RAIO_PARTICULA + 0.05)) ok = 0;
  if (ok && check) {
    double i = floor(x / h), j = floor(y / h);
    if (i < 0 || j < 0 || i >= nx || j >= ny) ok = 0;
    else if (countCell[(size_t)i + (size_t)j * nx] >= (uint32_t)(PPC - 1)) ok = 0;
  }
  flag[k] = ok;
}

__device__ __forceinline__ void k_add_append(Blk B_, const float* __restrict__ list, int count,
                                             const uint32_t* __restrict__ flag, const uint32_t* __restrict__ off,
                                             DevScalars* __restrict__ sc, uint32_t seq0, int cap,
                                             float2* __restrict__ pos, float2* __restrict__ vel,
                                             float4* __restrict__ cm, float4* __restrict__ ext,
                                             uint32_t* __restrict__ seq) {
  const int k = B_.x * blockDim.x + threadIdx.x;
  if (k >= count || !flag[k]) return;
  const uint32_t base = sc->n_slots;
  const uint32_t d = base + off[k];
  if (d >= (uint32_t)cap) return;
  const float* o = list + (size_t)k * 12;
  pos[d] = make_float2(o[0], o[1]);
  vel[d] = make_float2(o[2], o[3]);
  cm[d] = make_float4(o[4], o[5], o[6], o[7]);
  ext[d] = make_float4(o[8], o[9], o[10], o[11]);
  seq[d] = seq0 + off[k];
}

__device__ __forceinline__ void k_add_finish(Blk B_, DevScalars* sc, const uint32_t* __restrict__ total, int cap) {
  uint32_t t = *total;
  uint32_t n = sc->n_slots + t;
  sc->n_slots = n < (uint32_t)cap ? n : (uint32_t)cap;
  sc->add_count = t;
}

int add_particles_dev(tq_sim* s, int count, int check) {
  if (count <= 0) return 0;
  launch<256, 0, k_add_flags>(s, s->st, dim3((count + 255) / 256), dim3(256), 0, s->stage, count, check, s->sdf,
                              s->nx, s->ny, s->h, s->countCell, s->addflag);
  TQ_LAUNCH_CHECK();
  if (int e = scan_u32(s, s->addflag, s->pkey, count)) return e;
  const int o = s->cur;
  launch<256, 0, k_add_append>(s, s->st, dim3((count + 255) / 256), dim3(256), 0, s->stage, count, s->addflag, s->pkey,
                               s->sc, s->seq_next, s->cap, s->pos[o], s->vel[o], s->cm[o], s->ext[o], s->seq[o]);
  TQ_LAUNCH_CHECK();
  launch<64, 0, k_add_finish>(s, s->st, dim3(1), dim3(1), 0, s->sc, s->pkey + count, s->cap);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== emitirTorneira (principal.js:301-321)
// Particle k of an emission takes draws 2k+1 (angle) and 2k+2 (radius) of the LCG stream that
// starts at `seed`; records as the reference builds them (binary64 arithmetic, f32 stores).
__device__ __forceinline__ void k_faucet(Blk B_, float* __restrict__ out, int n, uint32_t seed, double x, double y,
                                         double raio, double vx, double vy, float a0, float a1, float a2) {
#pragma clang fp contract(off)
  const int k = B_.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t s1 = lcg_jump(seed, 2ull * (uint64_t)k + 1);
  const uint32_t s2 = s1 * LCG_A + LCG_C;
  const double ang = (double)s1 / 4294967296.0 * M_PI * 2;
  const double r = sqrt((double)s2 / 4294967296.0) * raio;
  float* o = out + 12 * (size_t)k;
  o[0] = (float)(x + cos(ang) * r);
  o[1] = (float)(y + sin(ang) * r);
  o[2] = (float)vx;
  o[3] = (float)vy;
#pragma unroll
  for (int q = 4; q < 9; q++) o[q] = 0.f;
  o[9] = a0;
  o[10] = a1;
  o[11] = a2;
}

// n particles generated into the staging list (the first `keep` of them are added); returns
// through *seed the stream advanced past all n
int faucet_dev(tq_sim* s, const tq_faucet* f, double vx, double vy, int n, int keep, uint32_t* seed) {
  if (keep > 0) {
    launch<256, 0, k_faucet>(s, s->st, dim3((keep + 255) / 256), dim3(256), 0, s->stage, keep, *seed, f->x, f->y,
                             f->raio, vx, vy, (float)f->absorcao[0], (float)f->absorcao[1], (float)f->absorcao[2]);
    TQ_LAUNCH_CHECK();
    if (int e = add_particles_dev(s, keep, 1)) return e;
  }
  *seed = lcg_jump(*seed, 2ull * (uint64_t)n);
  return 0;
}

// ================================================================== aplicarFerramenta (fisica-cpu.js:129-152)
__device__ __forceinline__ void k_tool(Blk B_, float2* __restrict__ pos, float2* __restrict__ vel, float4* __restrict__ ext,
                                       const DevScalars* __restrict__ sc, int kind, double fx, double fy,
                                       double raio, double fvx, double fvy, float a0, float a1, float a2) {
#pragma clang fp contract(off)
  const uint32_t k = B_.x * blockDim.x + threadIdx.x;
  if (k >= sc->n_slots) return;
  float2 P = pos[k];
  if (P.x != P.x) return;
  const double r2 = raio * raio;
  const double dx = (double)P.x - fx, dy = (double)P.y - fy;
  const double d2 = dx * dx + dy * dy;
  if (kind == 1) {
    if (d2 < r2) pos[k] = make_float2(__builtin_nanf(""), P.y);
    return;
  }
  if (d2 >= r2) return;
  const double peso = 1 - sqrt(d2 / r2);
  if (kind == 2) {
    const double a = fmin(1.0, peso * 0.9);
    float2 V = vel[k];
    V.x = (float)((double)V.x + (fvx - (double)V.x) * a);
    V.y = (float)((double)V.y + (fvy - (double)V.y) * a);
    vel[k] = V;
  } else if (kind == 3) {
    float4 E = ext[k];
    E.y = a0; E.z = a1; E.w = a2;
    ext[k] = E;
  }
}

int apply_tool_dev(tq_sim* s, const tq_tool* t) {
  if (!t || t->kind == 0 || s->n_upper <= 0) return 0;
  const int o = s->cur;
  launch<256, 0, k_tool>(s, s->st, dim3((s->n_upper + 255) / 256), dim3(256), 0, s->pos[o], s->vel[o], s->ext[o], s->sc,
                         t->kind, t->x, t->y, t->raio, t->vx, t->vy, (float)t->absorcao[0], (float)t->absorcao[1],
                         (float)t->absorcao[2]);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== live count
__device__ __forceinline__ void k_count_live(Blk B_, const float2* __restrict__ pos, const DevScalars* __restrict__ sc,
                                             uint32_t* __restrict__ out) {
  // one atomic per block (per-wave atomics on the one counter serialise in the cache)
  __shared__ uint32_t wc[4];
  const uint32_t k = B_.x * blockDim.x + threadIdx.x;
  uint32_t live = (k < sc->n_slots && pos[k].x == pos[k].x) ? 1u : 0u;
  unsigned long long b = __ballot(live);
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) t += wc[q];
    if (t) atomicAdd(out, t);
  }
}

int count_live(tq_sim* s, int* n) {
  uint32_t* tmp = s->addflag;
  TQ_HIP(hipMemsetAsync(tmp, 0, sizeof(uint32_t), s->st));
  if (s->n_upper > 0) {
    launch<256, 0, k_count_live>(s, s->st, dim3((s->n_upper + 255) / 256), dim3(256), 0, s->pos[s->cur], s->sc, tmp);
    TQ_LAUNCH_CHECK();
  }
  uint32_t h = 0;
  TQ_HIP(hipMemcpyAsync(&h, tmp, sizeof(uint32_t), hipMemcpyDeviceToHost, s->st));
  TQ_HIP(hipStreamSynchronize(s->st));
  *n = (int)h;
  return 0;
}

// the status words of a non-blocking readback: the count is cleared and vMax copied (one thread);
// k_count_live adds the live particles behind it
__device__ __forceinline__ void k_status_begin(Blk B_, unsigned long long* __restrict__ st,
                                               const DevScalars* __restrict__ sc) {
  if (threadIdx.x == 0) {
    st[0] = 0ull;
    st[1] = sc->vmax_bits;
  }
}

int status_enqueue(tq_sim* s) {
  launch<64, 0, k_status_begin>(s, s->st, dim3(1), dim3(64), 0, (unsigned long long*)s->status_dev,
                                (const DevScalars*)s->sc);
  if (s->n_upper > 0)
    launch<256, 0, k_count_live>(s, s->st, dim3((s->n_upper + 255) / 256), dim3(256), 0, s->pos[s->cur], s->sc,
                                 (uint32_t*)s->status_dev);
  TQ_LAUNCH_CHECK();
  return 0;
}

// ================================================================== atualizarParedes (fisica-cpu.js:80-85)
// Cena.celulaSolida (cena.js:119-127): the sum of the cell's 4 SDF samples, in binary64, < 0.
__device__ __forceinline__ void k_solid(Blk B_, const float* __restrict__ sdf, int nx, int ny, uint8_t* __restrict__ solid) {
#pragma clang fp contract(off)
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c >= nx * ny) return;
  const int i = c % nx, j = c / nx, w = 2 * nx, a = 2 * i, b = 2 * j;
  double s = (double)sdf[(size_t)b * w + a] + (double)sdf[(size_t)b * w + a + 1];
  s = s + (double)sdf[(size_t)(b + 1) * w + a];
  s = s + (double)sdf[(size_t)(b + 1) * w + a + 1];
  solid[c] = s < 0 ? 1 : 0;
}

int solid_cells(tq_sim* s) {
  launch<256, 0, k_solid>(s, s->st, dim3((s->C + 255) / 256), dim3(256), 0, s->sdf, s->nx, s->ny, s->solid);
  TQ_LAUNCH_CHECK();
  return 0;
}

}  // namespace tq
