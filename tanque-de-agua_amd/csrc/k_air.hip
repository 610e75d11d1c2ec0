// k_air.hip -- the air tracker (RastreadorDeAr.atualizar, js/ar.js:56-153) on the GPU.
//
//   1. union-find over air cells (4-neighbourhood) with lock-free CAS linking of the larger
//      root under the smaller one, so every component's root is its smallest cell index --
//      the root the reference's sequential union-find ends with (ar.js:155-158).  Each 32x32
//      tile is first united in LDS (its components' roots are their smallest cells, which are
//      also the smallest global indices); only the tile-border cells then unite through the
//      global array;
//   2. components are numbered by their root's raster position (an exclusive scan of root
//      flags): the reference numbers them by first raster appearance, which is the root;
//   3. per-component counts (cells, atmosphere cells, top-row contact) and the
//      (component, old label) overlap table (count, first cell) with wavefront-aggregated
//      atomics into an open-addressing hash table;
//   4. one workgroup does the label bookkeeping in the reference's order: overlap totals,
//      per-component volumes summed over old labels in first-appearance order, the stable
//      volume-descending ordering of candidate bubbles, stable label reuse, free-label
//      allocation and the born bubbles' pressures;
//   5. every cell takes its component's label.
#include "tq_device.hpp"

namespace tq {

constexpr unsigned long long HEMPTY = ~0ull;

__device__ inline int ld_rel(const int32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline int uf_find(int32_t* pai, int x) {
  while (true) {
    const int p = ld_rel(&pai[x]);
    if (p == x) return x;
    const int gp = ld_rel(&pai[p]);
    if (gp != p) __hip_atomic_store(&pai[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = p;
  }
}

__device__ inline void uf_unite(int32_t* pai, int a, int b) {
  while (true) {
    a = uf_find(pai, a);
    b = uf_find(pai, b);
    if (a == b) return;
    if (a < b) { int t = a; a = b; b = t; }
    const int old = atomicCAS(&pai[a], a, b);
    if (old == a) return;
    a = old;
  }
}

__device__ __forceinline__ void k_uf_init(Blk B_, int C, uint32_t* __restrict__ cells, uint32_t* __restrict__ atm,
                                          uint32_t* __restrict__ topo, int32_t* __restrict__ dono,
                                          uint32_t* __restrict__ label, float* __restrict__ alvo) {
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  cells[c] = 0; atm[c] = 0; topo[c] = 0; dono[c] = -1; label[c] = 0; alvo[c] = 0.f;
}

// ---- 1a. union-find inside a 32x32 tile (LDS, workgroup-scope atomics)
constexpr int UF_T = 32;
__device__ inline int lds_find(int* lp, int x) {
  while (true) {
    const int p = __hip_atomic_load(&lp[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (p == x) return x;
    const int gp = __hip_atomic_load(&lp[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (gp != p) __hip_atomic_store(&lp[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    x = p;
  }
}

__device__ inline void lds_unite(int* lp, int a, int b) {
  while (true) {
    a = lds_find(lp, a);
    b = lds_find(lp, b);
    if (a == b) return;
    if (a < b) { int t = a; a = b; b = t; }
    int expected = a;
    if (__hip_atomic_compare_exchange_strong(&lp[a], &expected, b, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP))
      return;
    a = expected;
  }
}

// Every cell of the tile gets pai = the global index of its tile-component's root (-1: not air).
// Local indices are row-major in the tile, so the smallest local index is the smallest global one.
__device__ __forceinline__ void k_uf_local(Blk B_, const uint8_t* __restrict__ tipo, int nx, int ny,
                                           int32_t* __restrict__ pai) {
  __shared__ int lp[UF_T * UF_T];
  const int ox = B_.x * UF_T, oy = B_.y * UF_T;
  bool air[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k;
    const int i = ox + (e & (UF_T - 1)), j = oy + e / UF_T;
    air[k] = i < nx && j < ny && tipo[(size_t)i + (size_t)j * nx] == AR;
    lp[e] = air[k] ? e : -1;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k;
    if (!air[k]) continue;
    const int lx = e & (UF_T - 1), ly = e / UF_T;
    if (lx > 0 && __hip_atomic_load(&lp[e - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= 0)
      lds_unite(lp, e, e - 1);
    if (ly > 0 && __hip_atomic_load(&lp[e - UF_T], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= 0)
      lds_unite(lp, e, e - UF_T);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int e = threadIdx.x + 256 * k;
    const int i = ox + (e & (UF_T - 1)), j = oy + e / UF_T;
    if (i >= nx || j >= ny) continue;
    int g = -1;
    if (air[k]) {
      const int r = lds_find(lp, e);
      g = (oy + r / UF_T) * nx + ox + (r & (UF_T - 1));
    }
    pai[(size_t)i + (size_t)j * nx] = g;
  }
}

// ---- 1b. tile-border cells unite with their neighbour across the border (global array):
// first the left borders (i = 32a, a > 0), then the top borders (j = 32b, b > 0)
__device__ __forceinline__ void k_uf_border(Blk B_, const uint8_t* __restrict__ tipo, int nx, int ny,
                                            int32_t* __restrict__ pai) {
  const int t = B_.x * blockDim.x + threadIdx.x;
  const int ntx = (nx + UF_T - 1) / UF_T, nty = (ny + UF_T - 1) / UF_T;
  const int nv = (ntx - 1) * ny;
  int c, d;
  if (t < nv) {
    const int a = 1 + t / ny, j = t - (t / ny) * ny;
    c = j * nx + a * UF_T;
    d = c - 1;
  } else if (t < nv + (nty - 1) * nx) {
    const int u = t - nv;
    const int b = 1 + u / nx, i = u - (u / nx) * nx;
    c = b * UF_T * nx + i;
    d = c - nx;
  } else {
    return;
  }
  if (tipo[c] == AR && tipo[d] == AR) uf_unite(pai, c, d);
}

__device__ __forceinline__ void k_uf_flatten(Blk B_, const uint8_t* __restrict__ tipo, int C, int32_t* __restrict__ pai,
                                             uint32_t* __restrict__ rootFlag) {
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  uint32_t f = 0;
  if (tipo[c] == AR) {
    const int r = uf_find(pai, c);
    f = (r == c) ? 1u : 0u;
  }
  rootFlag[c] = f;
}

__device__ inline uint64_t hmix(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdull; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ull; k ^= k >> 33;
  return k;
}

// one (component, old label) group of `cnt` cells whose first cell in raster order is `first`
// into the overlap hash table (open addressing; new entries listed in hlist)
__device__ inline void overlap_add(unsigned long long key, uint32_t cnt, uint32_t first,
                                   unsigned long long* __restrict__ hkeys, uint32_t* __restrict__ hcount,
                                   uint32_t* __restrict__ hfirst, uint32_t* __restrict__ hlist, int hmask,
                                   DevScalars* __restrict__ sc) {
  uint32_t h = (uint32_t)hmix(key) & (uint32_t)hmask;
  while (true) {
    const unsigned long long prev = atomicCAS(&hkeys[h], HEMPTY, key);
    if (prev == HEMPTY) {
      const uint32_t idx = atomicAdd(&sc->npairs, 1u);
      hlist[idx] = h;
      break;
    }
    if (prev == key) break;
    h = (h + 1) & (uint32_t)hmask;
  }
  atomicAdd(&hcount[h], cnt);
  atomicMin(&hfirst[h], first);
}

// per-cell: component index, per-component statistics, overlap table.  Each wave groups its
// cells by component (and by component and old label); the wave leaders add into block tables
// in LDS (CS_SLOTS entries each), flushed to the global arrays once per block, so a component
// of millions of cells costs a few thousand global atomics instead of several per wave.
// Groups the full tables cannot take go to the global arrays directly.  Every sum is an
// integer count, a flag or a minimum, so the order of the adds does not matter.
constexpr int CS_SLOTS = 32;
constexpr int CS_GRID = 2048;
__device__ __forceinline__ void k_comp_stats(Blk B_, const uint8_t* __restrict__ tipo, const uint32_t* __restrict__ rot,
                                             int nx, int ny, const int32_t* __restrict__ pai,
                                             const uint32_t* __restrict__ compIdx, int32_t* __restrict__ comp,
                                             uint32_t* __restrict__ cells, uint32_t* __restrict__ atm,
                                             uint32_t* __restrict__ topo, unsigned long long* __restrict__ hkeys,
                                             uint32_t* __restrict__ hcount, uint32_t* __restrict__ hfirst,
                                             uint32_t* __restrict__ hlist, int hmask,
                                             DevScalars* __restrict__ sc) {
  __shared__ int ck[CS_SLOTS];
  __shared__ uint32_t cc[CS_SLOTS], ca[CS_SLOTS], ct[CS_SLOTS];
  __shared__ unsigned long long ok[CS_SLOTS];
  __shared__ uint32_t oc[CS_SLOTS], of[CS_SLOTS];
  for (int e = threadIdx.x; e < CS_SLOTS; e += blockDim.x) {
    ck[e] = -1;
    cc[e] = ca[e] = ct[e] = 0;
    ok[e] = HEMPTY;
    oc[e] = 0;
    of[e] = 0xFFFFFFFFu;
  }
  __syncthreads();
  const int C = nx * ny;
  const int lane = threadIdx.x & 63;
  for (int base = B_.x * blockDim.x; base < C; base += B_.gx * blockDim.x) {
    const int c = base + threadIdx.x;
    int k = -1;
    uint32_t old = 0;
    bool istop = false;
    if (c < C) {
      if (tipo[c] == AR) {
        int r = pai[c];
        while (pai[r] != r) r = pai[r];
        k = (int)compIdx[r];
        old = rot[c];
        istop = (c / nx) >= ny - MOLDURA - 1;
      }
      comp[c] = k;
    }
    // statistics grouped by component within the wavefront
    bool pending = k >= 0;
    while (true) {
      unsigned long long b = __ballot(pending);
      if (!b) break;
      const int leader = __ffsll((long long)b) - 1;
      const int lk = __shfl(k, leader, 64);
      const bool m = pending && k == lk;
      const unsigned long long bm = __ballot(m), ba = __ballot(m && old == 1), bt = __ballot(m && istop);
      if (lane == leader) {
        int sl = -1;
        for (int q = 0; q < CS_SLOTS; q++) {
          const int hs = (int)(((uint32_t)lk + (uint32_t)q) % CS_SLOTS);
          const int prev = atomicCAS(&ck[hs], -1, lk);
          if (prev == -1 || prev == lk) { sl = hs; break; }
        }
        if (sl >= 0) {
          atomicAdd(&cc[sl], (uint32_t)__popcll(bm));
          if (ba) atomicAdd(&ca[sl], (uint32_t)__popcll(ba));
          if (bt) atomicOr(&ct[sl], 1u);
        } else {
          atomicAdd(&cells[lk], (uint32_t)__popcll(bm));
          if (ba) atomicAdd(&atm[lk], (uint32_t)__popcll(ba));
          if (bt) atomicOr(&topo[lk], 1u);
        }
      }
      if (m) pending = false;
    }
    // overlap table entries for old bubble labels
    pending = k >= 0 && old >= 2;
    while (true) {
      unsigned long long b = __ballot(pending);
      if (!b) break;
      const int leader = __ffsll((long long)b) - 1;
      const int lk = __shfl(k, leader, 64);
      const uint32_t lo = __shfl(old, leader, 64);
      const bool m = pending && k == lk && old == lo;
      const unsigned long long bm = __ballot(m);
      if (lane == leader) {
        const unsigned long long key = ((unsigned long long)(uint32_t)lk << 12) | lo;
        int sl = -1;
        for (int q = 0; q < CS_SLOTS; q++) {
          const int hs = (int)(((uint32_t)hmix(key) + (uint32_t)q) % CS_SLOTS);
          const unsigned long long prev = atomicCAS(&ok[hs], HEMPTY, key);
          if (prev == HEMPTY || prev == key) { sl = hs; break; }
        }
        // the leader is the group's first cell in raster order
        if (sl >= 0) {
          atomicAdd(&oc[sl], (uint32_t)__popcll(bm));
          atomicMin(&of[sl], (uint32_t)c);
        } else {
          overlap_add(key, (uint32_t)__popcll(bm), (uint32_t)c, hkeys, hcount, hfirst, hlist, hmask, sc);
        }
      }
      if (m) pending = false;
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < CS_SLOTS; e += blockDim.x) {
    const int lk = ck[e];
    if (lk >= 0) {
      atomicAdd(&cells[lk], cc[e]);
      if (ca[e]) atomicAdd(&atm[lk], ca[e]);
      if (ct[e]) atomicOr(&topo[lk], 1u);
    }
    if (ok[e] != HEMPTY) overlap_add(ok[e], oc[e], of[e], hkeys, hcount, hfirst, hlist, hmask, sc);
  }
}

// single-workgroup bitonic sort of n (key, value) pairs in global memory (padded with max keys)
__device__ void block_bitonic(unsigned long long* keys, uint32_t* vals, int n) {
  int P = 1;
  while (P < n) P <<= 1;
  for (int i = n + threadIdx.x; i < P; i += blockDim.x) { keys[i] = ~0ull; vals[i] = 0xFFFFFFFFu; }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < P; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool up = (i & k) == 0;
          const unsigned long long a = keys[i], b = keys[ixj];
          if ((a > b) == up) {
            keys[i] = b; keys[ixj] = a;
            const uint32_t t = vals[i]; vals[i] = vals[ixj]; vals[ixj] = t;
          }
        }
      }
      __syncthreads();
    }
}

// Lists of overlap pairs and of candidate components up to this long are sorted, and labels
// are assigned, in LDS; longer lists use global memory and the reference's sequential pass.
constexpr int BK_LDS = 4096;
static_assert(BK_LDS == MAX_BOLHAS, "the label scans cover BK_LDS entries");

// Exclusive prefix counts of the flags f(i), i in [0, n), n <= 4 * 1024, in a 1024-thread
// block: idx[i] = number of flagged entries before i (written where f(i)); returns the total in
// every thread.  `ws` holds 17 ints of LDS.
template <class F>
__device__ inline int block_scan_flags(int n, F f, uint16_t* idx, int* ws) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  int v[4], s = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = 4 * tid + q;
    v[q] = (i < n && f(i)) ? 1 : 0;
    s += v[q];
  }
  int x = s;  // inclusive scan over the wave
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[w] = x;
  __syncthreads();
  if (tid == 0) {
    int a = 0;
    for (int k = 0; k < 16; k++) {
      const int t = ws[k];
      ws[k] = a;
      a += t;
    }
    ws[16] = a;
  }
  __syncthreads();
  int off = ws[w] + x - s;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (v[q]) idx[4 * tid + q] = (uint16_t)off;
    off += v[q];
  }
  const int total = ws[16];
  __syncthreads();
  return total;
}

__device__ __forceinline__ void k_bookkeep(Blk B_, const uint32_t* __restrict__ compIdxTotal, const uint32_t* __restrict__ cells,
                                           const uint32_t* __restrict__ atm, const uint32_t* __restrict__ topo,
                                           float* __restrict__ alvo, int32_t* __restrict__ dono,
                                           uint32_t* __restrict__ label, unsigned long long* __restrict__ hkeys,
                                           uint32_t* __restrict__ hcount, uint32_t* __restrict__ hfirst,
                                           const uint32_t* __restrict__ hlist, unsigned long long* __restrict__ skeys,
                                           uint32_t* __restrict__ svals, float* __restrict__ volAlvo,
                                           uint8_t* __restrict__ vivo, float* __restrict__ pBolha,
                                           DevScalars* __restrict__ sc) {
  __shared__ uint32_t total[MAX_BOLHAS];
  __shared__ uint8_t vivoNovo[MAX_BOLHAS], vivoV[MAX_BOLHAS];
  __shared__ float pB[MAX_BOLHAS];
  __shared__ unsigned long long lkeys[BK_LDS];
  __shared__ uint32_t lvals[BK_LDS];
  __shared__ int32_t cdono[BK_LDS];
  __shared__ uint32_t clab[BK_LDS];
  __shared__ uint32_t ncand_s;
  const int tid = threadIdx.x;
  const uint32_t ncomp = *compIdxTotal;
  const uint32_t np = sc->npairs;
  const bool primeira = sc->primeira != 0;
  for (int r = tid; r < MAX_BOLHAS; r += 1024) {
    total[r] = 0;
    vivoNovo[r] = 0;
    vivoV[r] = vivo[r];
    pB[r] = pBolha[r];
  }
  if (tid == 0) { sc->ncomp = ncomp; ncand_s = 0; }
  __syncthreads();
  // ---- overlap totals of the components that stay (ar.js:101-106)
  unsigned long long* SK = np <= (uint32_t)BK_LDS ? lkeys : skeys;
  uint32_t* SV = np <= (uint32_t)BK_LDS ? lvals : svals;
  for (uint32_t i = tid; i < np; i += 1024) {
    const uint32_t h = hlist[i];
    const unsigned long long key = hkeys[h];
    const uint32_t k = (uint32_t)(key >> 12), r = (uint32_t)(key & 4095);
    if (topo[k] || cells[k] >= MIN_BOLHA_CELULAS) atomicAdd(&total[r], hcount[h]);
    SK[i] = ((unsigned long long)k << 32) | hfirst[h];
    SV[i] = h;
  }
  __syncthreads();
  // ---- per-component volume over old labels in first-appearance order (ar.js:108-125)
  if (np > 0 && !primeira) {
    block_bitonic(SK, SV, (int)np);
    for (uint32_t i = tid; i < np; i += 1024) {
      const uint32_t k = (uint32_t)(SK[i] >> 32);
      if (i > 0 && (uint32_t)(SK[i - 1] >> 32) == k) continue;  // one thread per component segment
      if (topo[k] || cells[k] < MIN_BOLHA_CELULAS) continue;
      double v = atm[k];
      uint32_t melhor = 0;
      int dn = -1;
      for (uint32_t e = i; e < np && (uint32_t)(SK[e] >> 32) == k; e++) {
        const uint32_t h = SV[e];
        const uint32_t r = (uint32_t)(hkeys[h] & 4095);
        const uint32_t qv = hcount[h];
        v += (double)volAlvo[r] * ((double)qv / (double)total[r]);
        if (qv > melhor) { melhor = qv; dn = (int)r; }
      }
      alvo[k] = (float)v;
      dono[k] = dn;
      label[k] = 0xFFFFFFFFu;  // marks "volume computed from overlaps"
    }
  }
  __syncthreads();  // the pair list is no longer read (its LDS holds the candidates next)
  // ---- volumes of the other components, candidates (ar.js:110-130)
  unsigned long long* CK = ncomp <= (uint32_t)BK_LDS ? lkeys : skeys;
  uint32_t* CV = ncomp <= (uint32_t)BK_LDS ? lvals : svals;
  for (uint32_t k = tid; k < ncomp; k += 1024) {
    float a = 0.f;
    if (topo[k]) { label[k] = 1u; alvo[k] = 0.f; continue; }
    if (cells[k] < MIN_BOLHA_CELULAS) { label[k] = 0; alvo[k] = 0.f; continue; }
    if (primeira) a = (float)cells[k];
    else if (label[k] == 0xFFFFFFFFu) a = alvo[k];
    else a = (float)atm[k];
    alvo[k] = a;
    label[k] = 0;
    if ((double)a >= MIN_BOLHA_CELULAS) {
      const uint32_t slot = atomicAdd(&ncand_s, 1u);
      // descending volume, ascending component index (a stable sort of the raster order)
      CK[slot] = ((unsigned long long)(~__float_as_uint(a)) << 32) | k;
      CV[slot] = k;
    }
  }
  __syncthreads();
  const int ncand = (int)ncand_s;
  if (ncand > 0) block_bitonic(CK, CV, ncand);
  // the sequential pass's inputs in LDS: every candidate's owner (the old label of largest overlap)
  const bool lds = ncand <= BK_LDS;
  if (lds)
    for (int e = tid; e < ncand; e += 1024) {
      cdono[e] = dono[CV[e]];
      clab[e] = 0u;  // label[k] of a candidate is 0 here
    }
  __syncthreads();
  // ---- stable labels, free labels, born bubbles (ar.js:131-145).  The reference walks the sorted
  // candidates twice: (1) a candidate takes its owner label r >= 2 unless an earlier candidate took
  // it; (2) every candidate still unlabelled takes the smallest label that is neither taken nor
  // alive before, in candidate order, until the labels run out, and a born bubble inherits its
  // owner's pressure.  In LDS both passes run in parallel with the same result: (1) r goes to the
  // smallest candidate index owning it (an atomic minimum); (2) the j-th unlabelled candidate
  // takes the j-th free label (two prefix counts).  Pass 2 reads the owners' pressures while it
  // writes the born ones; when an owner is itself a free label (never, while every owner is a
  // label alive before) a later read could see an earlier write, so that case, and lists beyond
  // LDS, keep the sequential pass.
  __shared__ int ws[17];
  __shared__ int odd;
  uint16_t* F = reinterpret_cast<uint16_t*>(lkeys);  // the free labels in ascending order
  uint16_t* rank = F + MAX_BOLHAS;                    // rank of an unlabelled candidate
  uint16_t* fpos = rank + MAX_BOLHAS;                 // position of a free label in F
  if (lds) {
    int* win = reinterpret_cast<int*>(total);  // the overlap totals are no longer read
    for (int r = tid; r < MAX_BOLHAS; r += 1024) win[r] = 0x7FFFFFFF;
    if (tid == 0) odd = 0;
    __syncthreads();
    for (int e = tid; e < ncand; e += 1024)
      if (cdono[e] >= 2) atomicMin(&win[cdono[e]], e);
    __syncthreads();
    for (int e = tid; e < ncand; e += 1024) {
      const int r = cdono[e];
      if (r >= 2 && win[r] == e) {
        clab[e] = (uint32_t)r;
        vivoNovo[r] = 1;
      }
    }
    __syncthreads();
    for (int e = tid; e < ncand; e += 1024) {
      const int dn = cdono[e];
      if (clab[e] == 0u && dn >= 2 && !vivoNovo[dn] && !vivoV[dn]) odd = 1;  // an owner that is a free label
    }
    __syncthreads();
  }
  if (lds && !odd) {
    const int nfree = block_scan_flags(MAX_BOLHAS, [&](int r) { return r >= 2 && !vivoNovo[r] && !vivoV[r]; },
                                       fpos, ws);
    for (int r = 2 + tid; r < MAX_BOLHAS; r += 1024)
      if (!vivoNovo[r] && !vivoV[r]) F[fpos[r]] = (uint16_t)r;
    (void)block_scan_flags(ncand, [&](int e) { return clab[e] == 0u; }, rank, ws);  // its barriers order F
    for (int e = tid; e < ncand; e += 1024) {
      if (clab[e] != 0u || rank[e] >= nfree) continue;
      const int lv = F[rank[e]];
      clab[e] = (uint32_t)lv;
      vivoNovo[lv] = 1;
      const int dn = cdono[e];
      const float p = dn >= 2 ? pB[dn] : 0.f;
      pB[lv] = p;
      pBolha[lv] = p;
    }
  } else if (tid == 0) {
    if (!lds)
      for (int e = 0; e < ncand; e++) {
        const int r = dono[CV[e]];
        if (r >= 2 && !vivoNovo[r]) {
          label[CV[e]] = (uint32_t)r;
          vivoNovo[r] = 1;
        }
      }
    int livre = 2;
    for (int e = 0; e < ncand; e++) {
      if (lds ? clab[e] != 0u : label[CV[e]] != 0u) continue;
      while (livre < MAX_BOLHAS && (vivoNovo[livre] || vivoV[livre])) livre++;
      if (livre >= MAX_BOLHAS) break;
      if (lds) clab[e] = (uint32_t)livre;
      else label[CV[e]] = (uint32_t)livre;
      vivoNovo[livre] = 1;
      const int dn = lds ? cdono[e] : dono[CV[e]];
      pB[livre] = dn >= 2 ? pB[dn] : 0.f;
      pBolha[livre] = pB[livre];
    }
  }
  __syncthreads();
  // ---- volumeAlvo, vivo (ar.js:147-150)
  for (int r = tid; r < MAX_BOLHAS; r += 1024) { volAlvo[r] = 0.f; vivo[r] = vivoNovo[r]; }
  if (lds)
    for (int e = tid; e < ncand; e += 1024) label[CV[e]] = clab[e];
  __syncthreads();
  for (int e = tid; e < ncand; e += 1024) {
    const uint32_t k = CV[e];
    const uint32_t lb = lds ? clab[e] : label[k];
    if (lb >= 2) volAlvo[lb] = alvo[k];
  }
  // ---- clear the overlap table for the next update
  for (uint32_t i = tid; i < np; i += 1024) {
    const uint32_t h = hlist[i];
    hkeys[h] = HEMPTY; hcount[h] = 0; hfirst[h] = 0xFFFFFFFFu;
  }
  __syncthreads();
  if (tid == 0) {
    sc->last_pairs = np;
    sc->last_cand = (uint32_t)ncand;
    sc->npairs = 0;
    sc->primeira = 0;
  }
}

__device__ __forceinline__ void k_label_write(Blk B_, const int32_t* __restrict__ comp, const uint32_t* __restrict__ label,
                                              int C, uint32_t* __restrict__ rot) {
  const int c = B_.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int k = comp[c];
  rot[c] = k >= 0 ? label[k] : 0u;
}

int tracker_update(tq_sim* s) {
  const int C = s->C;
  const dim3 g((C + 255) / 256), b(256);
  int t = time_begin(s, "air_tracker", 16.0 * C);
  launch<256, 0, k_uf_init>(s, s->st, g, b, 0, C, s->compCells, s->compAtm, s->compTopo, s->compDono, s->compLabel,
                            s->compAlvo);
  const int ntx = (s->nx + UF_T - 1) / UF_T, nty = (s->ny + UF_T - 1) / UF_T;
  launch<256, 0, k_uf_local>(s, s->st, dim3(ntx, nty), b, 0, s->tipo, s->nx, s->ny, s->pai);
  const int nborder = (ntx - 1) * s->ny + (nty - 1) * s->nx;
  if (nborder > 0)
    launch<256, 0, k_uf_border>(s, s->st, dim3((nborder + 255) / 256), b, 0, s->tipo, s->nx, s->ny, s->pai);
  launch<256, 0, k_uf_flatten>(s, s->st, g, b, 0, s->tipo, C, s->pai, s->rootFlag);
  TQ_LAUNCH_CHECK();
  if (int e = scan_u32(s, s->rootFlag, s->compIdx, C)) return e;
  launch<256, 0, k_comp_stats>(s, s->st, dim3(std::min((C + 255) / 256, CS_GRID)), b, 0, s->tipo, s->rot, s->nx, s->ny,
      s->pai, s->compIdx, s->comp,
      s->compCells, s->compAtm, s->compTopo, s->hkeys, s->hcount, s->hfirst, s->hlist, s->hsize - 1, s->sc);
  launch<1024, 0, k_bookkeep>(s, s->st, dim3(1), dim3(1024), 0, s->compIdx + C, s->compCells, s->compAtm, s->compTopo,
      s->compAlvo, s->compDono, s->compLabel, s->hkeys, s->hcount, s->hfirst, s->hlist,
      (unsigned long long*)s->ordBuf, s->ordBuf + 2 * (size_t)s->hsize, s->volAlvo, s->vivo, s->pBolha,
      s->sc);
  launch<256, 0, k_label_write>(s, s->st, g, b, 0, s->comp, s->compLabel, C, s->rot);
  time_end(s, t);
  TQ_LAUNCH_CHECK();
  return 0;
}

int tracker_reset(tq_sim* s) {
  TQ_HIP(hipMemsetAsync(s->rot, 0, sizeof(uint32_t) * s->C, s->st));
  TQ_HIP(hipMemsetAsync(s->volAlvo, 0, sizeof(float) * MAX_BOLHAS, s->st));
  TQ_HIP(hipMemsetAsync(s->vivo, 0, MAX_BOLHAS, s->st));
  TQ_HIP(hipMemsetAsync(s->pBolha, 0, sizeof(float) * MAX_BOLHAS, s->st));
  const int one = 1;
  TQ_HIP(hipMemcpyAsync(&s->sc->primeira, &one, sizeof(int), hipMemcpyHostToDevice, s->st));
  return 0;
}

}  // namespace tq
