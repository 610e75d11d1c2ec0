// bench_sync.hip -- what a device-wide synchronisation costs on this GPU, the measurement behind
// the solver's launch structure (one kernel per dependent pass, no persistent kernel with grid
// barriers).  Prints one JSON line:
//   * an empty dependent kernel launched back to back (1 and 1024 blocks): the cost of a
//     kernel boundary on one stream;
//   * a 5-point stencil pass over a 896x504 grid (the Ultra grid) as back-to-back kernels;
//   * a persistent cooperative kernel (co-residency guaranteed by hipLaunchCooperativeKernel)
//     running the same number of iterations with a hand-written device-scope barrier after each,
//     with and without the stencil pass, for 256..2048 blocks and two barrier forms: every block
//     adding to one counter, and blocks grouped by blockIdx % 8 (one group per XCD under the
//     round-robin dispatch) whose last arrival adds to the top counter.
// Every spin is bounded by wall-clock time (s_memrealtime, 100 MHz): after 200 ms a block sets an
// abort flag that every block checks, so all waves leave their loops and the kernel ends.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                        \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

struct Bar {
  unsigned* grp;  // 8 group counters, 64 B apart (monotonic)
  unsigned* top;  // top counter (monotonic)
  unsigned* gen;  // generation
  int* abort_;
};

constexpr unsigned long long SPIN_LIMIT = 20000000ull;  // 200 ms at 100 MHz

// the waiting part of both barriers: thread 0 spins on the generation (bounded)
__device__ bool wait_gen(const Bar& B, unsigned epoch) {
  const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
  while (__hip_atomic_load(B.gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= epoch) {
    if (__hip_atomic_load(B.abort_, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
    if (__builtin_amdgcn_s_memrealtime() - t0 > SPIN_LIMIT) {
      __hip_atomic_store(B.abort_, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
  }
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  return true;
}

template <int TREE>
__device__ bool grid_barrier(const Bar& B, unsigned nb, unsigned epoch) {
  __syncthreads();
  __shared__ int ok;
  if (threadIdx.x == 0) {
    ok = 1;
    bool last = false;
    if (TREE) {
      const unsigned g = blockIdx.x & 7;
      const unsigned gsize = nb / 8 + (g < (nb & 7) ? 1u : 0u);
      const unsigned o = __hip_atomic_fetch_add(B.grp + 16 * g, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (o == gsize * (epoch + 1) - 1)
        last = __hip_atomic_fetch_add(B.top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) == 8u * (epoch + 1) - 1;
    } else {
      last = __hip_atomic_fetch_add(B.top, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == nb * (epoch + 1) - 1;
    }
    if (last) __hip_atomic_store(B.gen, epoch + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    else if (!wait_gen(B, epoch)) ok = 0;
  }
  __syncthreads();
  return ok != 0;
}

__device__ inline void stencil_cell(const float* x, const float* b, float* y, int nx, int ny, int c) {
  const int j = c / nx, i = c - j * nx;
  const float l = x[i > 0 ? c - 1 : c], r = x[i < nx - 1 ? c + 1 : c];
  const float d = x[j > 0 ? c - nx : c], u = x[j < ny - 1 ? c + nx : c];
  y[c] = 0.25f * (l + r + d + u + b[c]);
}

__global__ __launch_bounds__(256) void k_empty() {}

__global__ __launch_bounds__(256) void k_stencil(const float* __restrict__ x, const float* __restrict__ b,
                                                 float* __restrict__ y, int nx, int ny) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < nx * ny) stencil_cell(x, b, y, nx, ny, c);
}

template <int TREE>
__global__ __launch_bounds__(256) void k_persist(Bar B, int iters, int work, float* xa, float* xb, const float* b,
                                                 int nx, int ny) {
  const unsigned nb = gridDim.x;
  for (int it = 0; it < iters; it++) {
    if (work) {
      const float* x = (it & 1) ? xb : xa;
      float* y = (it & 1) ? xa : xb;
      for (int c = blockIdx.x * 256 + threadIdx.x; c < nx * ny; c += nb * 256) stencil_cell(x, b, y, nx, ny, c);
    }
    if (!grid_barrier<TREE>(B, nb, (unsigned)it)) break;
  }
}

int main() {
  const int nx = 896, ny = 504, C = nx * ny, N = 400;
  float *xa, *xb, *bb;
  CK(hipMalloc(&xa, C * 4));
  CK(hipMalloc(&xb, C * 4));
  CK(hipMalloc(&bb, C * 4));
  std::vector<float> h(C);
  for (int i = 0; i < C; i++) h[i] = (float)((i * 2654435761u) % 1000) * 1e-3f;
  CK(hipMemcpy(xa, h.data(), C * 4, hipMemcpyHostToDevice));
  CK(hipMemcpy(bb, h.data(), C * 4, hipMemcpyHostToDevice));
  CK(hipMemset(xb, 0, C * 4));
  unsigned* ctr;
  int* abort_;
  CK(hipMalloc(&ctr, 4096));
  CK(hipMalloc(&abort_, 64));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  hipDeviceProp_t prop;
  CK(hipGetDeviceProperties(&prop, 0));
  auto timed = [&](auto&& fn) {
    CK(hipEventRecord(e0, 0));
    fn();
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, e0, e1));
    return 1000.0 * ms / N;
  };
  double empty1 = 0, empty1024 = 0, stencil = 0;
  for (int rep = 0; rep < 2; rep++) {  // the second round is reported (the first warms up)
    empty1 = timed([&] { for (int i = 0; i < N; i++) k_empty<<<1, 256>>>(); });
    empty1024 = timed([&] { for (int i = 0; i < N; i++) k_empty<<<1024, 256>>>(); });
    stencil = timed([&] {
      for (int i = 0; i < N; i++) k_stencil<<<(C + 255) / 256, 256>>>((i & 1) ? xb : xa, bb, (i & 1) ? xa : xb, nx, ny);
    });
  }
  int occ = 0;
  CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_persist<1>, 256, 0));
  const int maxb = occ * prop.multiProcessorCount;
  std::printf("{\"cus\": %d, \"empty_kernel_us\": {\"1\": %.3f, \"1024\": %.3f}, \"stencil_kernel_us\": %.3f, "
              "\"persistent_us_per_iteration\": [",
              prop.multiProcessorCount, empty1, empty1024, stencil);
  bool first = true;
  for (int tree = 0; tree < 2; tree++)
    for (int work = 0; work < 2; work++)
      for (int nb : {256, 512, 1024, 2048}) {
        if (nb > maxb) continue;
        CK(hipMemset(ctr, 0, 4096));
        CK(hipMemset(abort_, 0, 64));
        Bar B{ctr, ctr + 512, ctr + 768, abort_};
        int iters = N;
        int w = work;
        int nxv = nx, nyv = ny;
        void* args[] = {&B, &iters, &w, &xa, &xb, &bb, &nxv, &nyv};
        const void* fn = tree ? (const void*)k_persist<1> : (const void*)k_persist<0>;
        const double us = timed([&] { CK(hipLaunchCooperativeKernel(fn, dim3(nb), dim3(256), args, 0, 0)); });
        int ab = 0;
        CK(hipMemcpy(&ab, abort_, 4, hipMemcpyDeviceToHost));
        std::printf("%s{\"barrier\": \"%s\", \"stencil\": %d, \"blocks\": %d, \"us\": %.3f, \"aborted\": %d}",
                    first ? "" : ", ", tree ? "per-XCD groups" : "one counter", work, nb, us, ab);
        first = false;
      }
  std::printf("]}\n");
  return 0;
}
