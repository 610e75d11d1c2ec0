// tq_device.hpp -- device helpers: wavefront reductions, the deterministic last-block
// reduction, and the binary64 scene/number helpers shared by P2G, G2P and particle adds.
#pragma once
#include "tq_internal.hpp"

namespace tq {

// ------------------------------------------------------------------ wavefront (64 lanes) reductions
// Data-parallel-primitive (DPP) moves: a VALU operand taken from another lane of the same row of
// 16 (row_shr:k) or broadcast from lane 15 of the previous row (row_bcast:15, rows 1 and 3) and
// from lane 31 (row_bcast:31, rows 2 and 3).  A lane without a source reads `fill`.  An inclusive
// scan over the rows and then across them leaves the wave total in lane 63 after six steps of one
// or two VALU moves and an add; the order of the additions is fixed, so sums are deterministic.
// (A butterfly over ds_bpermute would cost six LDS-latency round trips per value.)
constexpr int DPP_ROW_SHR1 = 0x111, DPP_ROW_SHR2 = 0x112, DPP_ROW_SHR4 = 0x114, DPP_ROW_SHR8 = 0x118;
constexpr int DPP_ROW_BCAST15 = 0x142, DPP_ROW_BCAST31 = 0x143;
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, CTRL, ROWS, 0xF, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f(float v, float fill) {
  return __int_as_float(dpp_i<CTRL, ROWS>(__float_as_int(v), __float_as_int(fill)));
}
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v, double fill) {
  const int lo = dpp_i<CTRL, ROWS>(__double2loint(v), __double2loint(fill));
  const int hi = dpp_i<CTRL, ROWS>(__double2hiint(v), __double2hiint(fill));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane63_d(double v) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane63_f(float v) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// the total of the wave's 64 lanes, in every lane (every lane must be active)
__device__ inline double wave_sum(double v) {
  v += dpp_d<DPP_ROW_SHR1, 0xF>(v, 0.0);
  v += dpp_d<DPP_ROW_SHR2, 0xF>(v, 0.0);
  v += dpp_d<DPP_ROW_SHR4, 0xF>(v, 0.0);
  v += dpp_d<DPP_ROW_SHR8, 0xF>(v, 0.0);
  v += dpp_d<DPP_ROW_BCAST15, 0xA>(v, 0.0);
  v += dpp_d<DPP_ROW_BCAST31, 0xC>(v, 0.0);
  return lane63_d(v);
}
__device__ inline float wave_sumf(float v) {
  v += dpp_f<DPP_ROW_SHR1, 0xF>(v, 0.f);
  v += dpp_f<DPP_ROW_SHR2, 0xF>(v, 0.f);
  v += dpp_f<DPP_ROW_SHR4, 0xF>(v, 0.f);
  v += dpp_f<DPP_ROW_SHR8, 0xF>(v, 0.f);
  v += dpp_f<DPP_ROW_BCAST15, 0xA>(v, 0.f);
  v += dpp_f<DPP_ROW_BCAST31, 0xC>(v, 0.f);
  return lane63_f(v);
}
__device__ inline float wave_maxf(float v) {
  const float lo = -__builtin_inff();
  v = fmaxf(v, dpp_f<DPP_ROW_SHR1, 0xF>(v, lo));
  v = fmaxf(v, dpp_f<DPP_ROW_SHR2, 0xF>(v, lo));
  v = fmaxf(v, dpp_f<DPP_ROW_SHR4, 0xF>(v, lo));
  v = fmaxf(v, dpp_f<DPP_ROW_SHR8, 0xF>(v, lo));
  v = fmaxf(v, dpp_f<DPP_ROW_BCAST15, 0xA>(v, lo));
  v = fmaxf(v, dpp_f<DPP_ROW_BCAST31, 0xC>(v, lo));
  return lane63_f(v);
}
__device__ inline double wave_maxd(double v) {
  const double lo = -__builtin_inf();
  v = fmax(v, dpp_d<DPP_ROW_SHR1, 0xF>(v, lo));
  v = fmax(v, dpp_d<DPP_ROW_SHR2, 0xF>(v, lo));
  v = fmax(v, dpp_d<DPP_ROW_SHR4, 0xF>(v, lo));
  v = fmax(v, dpp_d<DPP_ROW_SHR8, 0xF>(v, lo));
  v = fmax(v, dpp_d<DPP_ROW_BCAST15, 0xA>(v, lo));
  v = fmax(v, dpp_d<DPP_ROW_BCAST31, 0xC>(v, lo));
  return lane63_d(v);
}

// Workgroups are dealt to the 8 XCDs round-robin by linear id, and each XCD has its own L2.
// Blocks are renumbered so that XCD k runs one contiguous range of logical blocks: neighbouring
// tiles, whose halo windows overlap, then mostly share an L2.  A bijection on [0, n) for any n.
constexpr int NXCD = 8;
__device__ inline int xcd_block(int id, int n) {
  const int q = n / NXCD, r = n % NXCD;
  const int x = id % NXCD, slot = id / NXCD;
  return x * q + (x < r ? x : r) + slot;
}
__device__ inline void xcd_block2(int* bx, int* by) {
  const int gx = gridDim.x;
  const int l = xcd_block(blockIdx.x + blockIdx.y * gx, gx * gridDim.y);
  *by = l / gx;
  *bx = l - *by * gx;
}
// the same over a logical block (bodies launched alone or from a tank table)
__device__ inline void xcd_block2(const Blk& B, int* bx, int* by) {
  const int l = xcd_block(B.x + B.y * B.gx, B.gx * B.gy);
  *by = l / B.gx;
  *bx = l - *by * B.gx;
}

// Block sum of NV doubles per thread; result valid in thread 0.  `sh` holds >= NV*16 doubles.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* sh) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; k++) sh[k * 16 + wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) {
      double t = 0;
      for (int w = 0; w < nw; w++) t += sh[k * 16 + w];
      v[k] = t;
    }
  }
  __syncthreads();
}

// Deterministic grid reduction of NV doubles: each block stores its partials, the last block
// to finish (ticket) sums all partials in block order.  Returns true in every thread of the
// last block; then `tot` (valid in all its threads) holds the grid totals.
template <int NV>
__device__ inline bool grid_sum_last(double (&v)[NV], double* partials, uint32_t* ticket, double (&tot)[NV],
                                     double* sh) {
  block_sum<NV>(v, sh);
  __shared__ bool last;
  const int nb = gridDim.x * gridDim.y;
  const int bid = blockIdx.x + blockIdx.y * gridDim.x;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) partials[(size_t)k * MAX_RED_BLOCKS + bid] = v[k];
    __threadfence();
    unsigned t = atomicAdd(ticket, 1u);
    last = (t == (unsigned)(nb - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = 0;
  // fixed assignment of partials to threads -> deterministic totals
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; k++) acc[k] += __builtin_nontemporal_load(&partials[(size_t)k * MAX_RED_BLOCKS + b]);
  block_sum<NV>(acc, sh);
  __shared__ double res[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) res[k] = acc[k];
    *ticket = 0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++) tot[k] = res[k];
  return true;
}

// ------------------------------------------------------------------ binary64 number helpers
// Math.hypot as V8 computes it (scaled, Kahan-compensated sum of squares).
__host__ __device__ inline double js_hypot(double a, double b) {
#pragma clang fp contract(off)
  double xa = fabs(a), xb = fabs(b);
  if (xa == INFINITY || xb == INFINITY) return INFINITY;
  if (xa != xa || xb != xb) return NAN;
  double mx = fmax(xa, xb);
  if (mx == 0) return 0;
  double sum = 0, comp = 0;
  {
    double n = xa / mx;
    double summand = n * n - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  {
    double n = xb / mx;
    double summand = n * n - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return sqrt(sum) * mx;
}

// Math.round: nearest integer, ties toward +infinity (ECMA-262)
__host__ __device__ inline double js_round(double x) {
  const double f = floor(x);
  return (x - f >= 0.5) ? f + 1 : f;
}

// The seeded LCG of the reference's tests (testes/apoio.js:4-9), s -> (s*1664525 + 1013904223)
// mod 2^32 (exact in binary64 there, so uint32 arithmetic reproduces it), advanced k steps at
// once by composing powers of the affine map.
constexpr uint32_t LCG_A = 1664525u, LCG_C = 1013904223u;
__host__ __device__ inline uint32_t lcg_jump(uint32_t s, uint64_t k) {
  uint32_t A = 1, C = 0, a = LCG_A, c = LCG_C;
  while (k) {
    if (k & 1) {
      A = a * A;
      C = a * C + c;
    }
    c = a * c + c;
    a = a * a;
    k >>= 1;
  }
  return A * s + C;
}

// clampi (fisica-cpu.js:697-699) on a double
__device__ inline double clampd(double v, double a, double b) { return v < a ? a : v > b ? b : v; }

// Cena #bilinear (cena.js:156-165) on the (2nx) x (2ny) SDF
__device__ inline double sdf_bilinear(const float* __restrict__ sdf, int w, int hh, double fx, double fy) {
#pragma clang fp contract(off)
  if (fx != fx || fy != fy) return NAN;  // JS: Math.min(NaN, ..) is NaN and the lookup is NaN
  if (fx < 0) fx = 0; else if (fx > w - 1) fx = w - 1;
  if (fy < 0) fy = 0; else if (fy > hh - 1) fy = hh - 1;
  double x0 = fmin(floor(fx), (double)(w - 2)), y0 = fmin(floor(fy), (double)(hh - 2));
  double tx = fx - x0, ty = fy - y0;
  size_t k = (size_t)y0 * w + (size_t)x0;
  return ((double)sdf[k] * (1 - tx) + (double)sdf[k + 1] * tx) * (1 - ty) +
         ((double)sdf[k + w] * (1 - tx) + (double)sdf[k + w + 1] * tx) * ty;
}
// Cena.distancia (cena.js:102-106)
__device__ inline double sdf_distancia(const float* __restrict__ sdf, int w, int hh, double h, double x, double y) {
#pragma clang fp contract(off)
  double fx = (x / h) * 2 - 0.5, fy = (y / h) * 2 - 0.5;
  return sdf_bilinear(sdf, w, hh, fx, fy);
}
// Cena.normal (cena.js:109-117)
__device__ inline void sdf_normal(const float* __restrict__ sdf, int w, int hh, double h, double x, double y,
                                  double* nxo, double* nyo) {
#pragma clang fp contract(off)
  double fx = (x / h) * 2 - 0.5, fy = (y / h) * 2 - 0.5;
  double gx = sdf_bilinear(sdf, w, hh, fx + 1, fy) - sdf_bilinear(sdf, w, hh, fx - 1, fy);
  double gy = sdf_bilinear(sdf, w, hh, fx, fy + 1) - sdf_bilinear(sdf, w, hh, fx, fy - 1);
  double n = js_hypot(gx, gy);
  if (n > 1e-9) { *nxo = gx / n; *nyo = gy / n; } else { *nxo = 0; *nyo = 1; }
}

// tile-major binning key of a cell (16x16 tiles)
__device__ __host__ inline uint32_t cell_key(int ci, int cj, int ntx) {
  return (uint32_t)((((cj >> 4) * ntx + (ci >> 4)) << 8) | ((cj & 15) << 4) | (ci & 15));
}

// count cell of a particle: clampi(floor(x/h*...)) in binary64 (fisica-cpu.js:258)
__device__ inline void count_cell(float x, float y, double inv, int nx, int ny, int* ci, int* cj) {
#pragma clang fp contract(off)
  double a = clampd(floor((double)x * inv), 0, nx - 1), b = clampd(floor((double)y * inv), 0, ny - 1);
  *ci = (int)a;
  *cj = (int)b;
}

__device__ inline long long to_fx(double v) {
  // fixed-point with 2^-46 resolution; non-finite or absurd contributions saturate
  double s = v * FX_SCALE;
  if (!(s > -9.0e18)) s = s != s ? 0.0 : -9.0e18;
  if (s > 9.0e18) s = 9.0e18;
  return (long long)__builtin_rint(s);
}

}  // namespace tq
