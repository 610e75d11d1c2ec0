// bench_stream.hip -- what HBM bandwidth simple streaming kernels reach on this GPU, the ceiling
// the solver's level-0 kernels are measured against next to the 8 TB/s peak.  Prints one JSON
// line (GB/s, HIP events over 20 launches after one warm-up):
//   * copy and read of 1 GiB arrays (4x the 256 MiB last-level cache, so HBM and not the cache
//     serves them), 16-byte lanes, grid-stride;
//   * the update kernel's streams at 4096x2304 (9.44 M cells): five float arrays and two byte
//     arrays read, five float arrays written, 42 bytes per cell, with no arithmetic to speak of,
//     laid out as the solver lays them out -- a wave per 16x16 tile with the lane's cells two
//     adjacent columns in rows r and r + 8, blocks of 2x2 and of 4x1 tiles, the latter also with
//     the two cells of a row as one 8-byte access -- and, for
//     comparison, one cell per thread in row order and a wave per 32x8 block.  Two sets of these
//     arrays, one allocated before the 1 GiB arrays and one after, since the tile layouts' rate
//     depends on where the arrays lie (the row order's does not).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) {                                                                  \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);    \
      std::exit(1);                                                                          \
    }                                                                                        \
  } while (0)

__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ x, float4* __restrict__ y, size_t n4) {
  for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) y[i] = x[i];
}

__global__ __launch_bounds__(256) void k_read(const float4* __restrict__ x, float* __restrict__ out, size_t n4) {
  float s = 0.f;
  for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n4; i += (size_t)gridDim.x * 256) {
    const float4 v = x[i];
    s += (v.x + v.y) + (v.z + v.w);
  }
  if (s == 1234.5f) out[0] = s;  // never true for the zero-filled input: keeps the loads
}

struct Streams {
  const float *a, *b, *c, *d, *e;
  const unsigned char *m1, *m2;
  float *o1, *o2, *o3, *o4, *o5;
};

__device__ inline void mix_cell(const Streams& S, size_t x) {
  const float av = S.a[x], bv = S.b[x], cv = S.c[x], dv = S.d[x], ev = S.e[x];
  const int k = S.m1[x] + S.m2[x];
  S.o1[x] = av + 0.5f * bv;
  S.o2[x] = cv - 0.5f * dv;
  S.o3[x] = ev * (float)k;
  S.o4[x] = av * cv;
  S.o5[x] = bv + ev;
}

// one cell per thread, row order
__global__ __launch_bounds__(256) void k_mix_linear(Streams S, size_t n) {
  const size_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) mix_cell(S, i);
}

// a wave per 16x16 tile in the smoother's lane layout, cell by cell: the lane's cells of colour
// q in rows r and r + 8 (the column within the pair depends on the lane's row, so every access
// moves 4 bytes); a block holds GW x GH tiles.  All loads of the lane's four cells first, then
// the stores.
template <int GW, int GH>
__global__ __launch_bounds__(256) void k_mix_tiles(Streams S, int nx) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tx = GW * blockIdx.x + (w % GW), ty = GH * blockIdx.y + (w / GW);
  float av[4], bv[4], cv[4], dv[4], ev[4];
  int k[4];
  size_t idx[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = tx * 16 + 2 * (lane & 7) + (((q & 1) + (lane >> 3) + 1) & 1), j = ty * 16 + (lane >> 3) + 8 * (q >> 1);
    const size_t x = (size_t)i + (size_t)j * nx;
    idx[q] = x;
    av[q] = S.a[x];
    bv[q] = S.b[x];
    cv[q] = S.c[x];
    dv[q] = S.d[x];
    ev[q] = S.e[x];
    k[q] = S.m1[x] + S.m2[x];
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const size_t x = idx[q];
    S.o1[x] = av[q] + 0.5f * bv[q];
    S.o2[x] = cv[q] - 0.5f * dv[q];
    S.o3[x] = ev[q] * (float)k[q];
    S.o4[x] = av[q] * cv[q];
    S.o5[x] = bv[q] + ev[q];
  }
}

// the same tiles with the lane's two cells of a row moved as one 8-byte access
template <int GW, int GH>
__global__ __launch_bounds__(256) void k_mix_tiles_pairs(Streams S, int nx) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tx = GW * blockIdx.x + (w % GW), ty = GH * blockIdx.y + (w / GW);
  float2 av[2], bv[2], cv[2], dv[2], ev[2];
  uchar2 k1[2], k2[2];
  size_t idx[2];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int i = tx * 16 + 2 * (lane & 7), j = ty * 16 + (lane >> 3) + 8 * r;
    const size_t x = (size_t)i + (size_t)j * nx;  // even: i and nx are
    idx[r] = x;
    av[r] = *reinterpret_cast<const float2*>(S.a + x);
    bv[r] = *reinterpret_cast<const float2*>(S.b + x);
    cv[r] = *reinterpret_cast<const float2*>(S.c + x);
    dv[r] = *reinterpret_cast<const float2*>(S.d + x);
    ev[r] = *reinterpret_cast<const float2*>(S.e + x);
    k1[r] = *reinterpret_cast<const uchar2*>(S.m1 + x);
    k2[r] = *reinterpret_cast<const uchar2*>(S.m2 + x);
  }
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const size_t x = idx[r];
    const float kx = (float)(k1[r].x + k2[r].x), ky = (float)(k1[r].y + k2[r].y);
    *reinterpret_cast<float2*>(S.o1 + x) = make_float2(av[r].x + 0.5f * bv[r].x, av[r].y + 0.5f * bv[r].y);
    *reinterpret_cast<float2*>(S.o2 + x) = make_float2(cv[r].x - 0.5f * dv[r].x, cv[r].y - 0.5f * dv[r].y);
    *reinterpret_cast<float2*>(S.o3 + x) = make_float2(ev[r].x * kx, ev[r].y * ky);
    *reinterpret_cast<float2*>(S.o4 + x) = make_float2(av[r].x * cv[r].x, av[r].y * cv[r].y);
    *reinterpret_cast<float2*>(S.o5 + x) = make_float2(bv[r].x + ev[r].x, bv[r].y + ev[r].y);
  }
}

// a wave per 32x8 block (a lane's four cells in one column, rows r, r + 2, r + 4, r + 6),
// blocks of 2x2 waves
__global__ __launch_bounds__(256) void k_mix_32x8(Streams S, int nx) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int tx = 2 * blockIdx.x + (w & 1), ty = 2 * blockIdx.y + (w >> 1);
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int i = tx * 32 + (lane & 31), j = ty * 8 + (lane >> 5) + 2 * q;
    mix_cell(S, (size_t)i + (size_t)j * nx);
  }
}

struct MixTimes {
  double lin, t22, t41, t41p, t32;
};

int main() {
  const int nx = 4096, ny = 2304;  // multiples of 64 and 32: every launch below covers the grid exactly
  const size_t n = (size_t)nx * ny;
  const size_t big = 256ull << 20;  // floats: 1 GiB per array
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int R = 20;
  auto us_per = [&](auto&& fn) {
    fn();
    CK(hipDeviceSynchronize());
    CK(hipEventRecord(e0, 0));
    for (int r = 0; r < R; r++) fn();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, e0, e1));
    CK(hipGetLastError());
    return 1000.0 * ms / R;
  };
  auto make_streams = [&]() {
    float* f[10];
    unsigned char* m[2];
    for (int i = 0; i < 10; i++) {
      CK(hipMalloc(&f[i], n * 4));
      CK(hipMemset(f[i], 0, n * 4));
    }
    for (int i = 0; i < 2; i++) {
      CK(hipMalloc(&m[i], n));
      CK(hipMemset(m[i], 1, n));
    }
    return Streams{f[0], f[1], f[2], f[3], f[4], m[0], m[1], f[5], f[6], f[7], f[8], f[9]};
  };
  auto mix = [&](const Streams& S) {
    MixTimes t;
    t.lin = us_per([&] { k_mix_linear<<<(unsigned)(n / 256), 256>>>(S, n); });
    t.t22 = us_per([&] { k_mix_tiles<2, 2><<<dim3(nx / 32, ny / 32), 256>>>(S, nx); });
    t.t41 = us_per([&] { k_mix_tiles<4, 1><<<dim3(nx / 64, ny / 16), 256>>>(S, nx); });
    t.t41p = us_per([&] { k_mix_tiles_pairs<4, 1><<<dim3(nx / 64, ny / 16), 256>>>(S, nx); });
    t.t32 = us_per([&] { k_mix_32x8<<<dim3(nx / 64, ny / 16), 256>>>(S, nx); });
    return t;
  };
  const Streams S1 = make_streams();
  float *x, *y, *out;
  CK(hipMalloc(&x, big * 4));
  CK(hipMalloc(&y, big * 4));
  CK(hipMalloc(&out, 64));
  CK(hipMemset(x, 0, big * 4));
  CK(hipMemset(y, 0, big * 4));
  const Streams S2 = make_streams();
  const double cb = 2.0 * big * 4, rb = 1.0 * big * 4, mb = 42.0 * n;
  const int grid = 4096;
  const double t_copy = us_per([&] { k_copy<<<grid, 256>>>((const float4*)x, (float4*)y, big / 4); });
  const double t_read = us_per([&] { k_read<<<grid, 256>>>((const float4*)x, out, big / 4); });
  const MixTimes a = mix(S1), b = mix(S2);
  std::printf("{\"copy_1GiB_GBs\": %.0f, \"read_1GiB_GBs\": %.0f, \"update_streams_4096x2304\": {\"bytes\": %.0f",
              cb / t_copy / 1e3, rb / t_read / 1e3, mb);
  const char* names[2] = {"first", "second"};
  const MixTimes* ts[2] = {&a, &b};
  for (int k = 0; k < 2; k++) {
    const MixTimes& t = *ts[k];
    std::printf(", \"%s_set_GBs\": {\"row_order\": %.0f, \"tiles16_blocks2x2\": %.0f, \"tiles16_blocks4x1\": %.0f, "
                "\"tiles16_blocks4x1_row_pairs\": %.0f, \"waves32x8_blocks2x2\": %.0f}",
                names[k], mb / t.lin / 1e3, mb / t.t22 / 1e3, mb / t.t41 / 1e3, mb / t.t41p / 1e3, mb / t.t32 / 1e3);
  }
  std::printf("}}\n");
  return 0;
}
