// tq_launch.hpp -- kernels as bodies over a logical block, launched for one tank or from a tank
// table.
//
// Every kernel of the step is a __device__ body whose first parameter is its logical block
// (Blk: the block's index in its tank's grid and that grid's shape).  launch<LB, W, Body>(s,
// stream, grid, block, shm, args...) either
//   * launches k_one<LB, Body> on the stream: the hardware block is the logical block, or
//   * while the tank's launches are being recorded (s->lrec set, tq_batch's table mode), appends
//     the launch to the log instead: the table kernel of Body, grid, block, LDS bytes and the
//     arguments converted to the body's parameter types and packed at their natural alignment.
// A table launch (k_tab<LB, Body>) runs one body for up to TAB_MAX tanks at once.  Its blocks are
// the concatenation of the tanks' grids; block b finds its tank from the block offsets in the
// kernel arguments (a binary search over scalar loads), reads that tank's grid shape and
// arguments from the tank's record and runs the body as that tank's logical block.  The body
// sees exactly the logical block, grid and arguments of a launch of its own, so a tank stepped in
// a table computes what it computes alone, bit for bit.
// LB and W are the body's launch bounds (__launch_bounds__(LB) or (LB, W)): both launch forms
// get the same register budget.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <utility>
#include <vector>

namespace tq {

// the logical block: index (x, y) in a grid of gx x gy blocks
struct Blk {
  int x, y, gx, gy;
};
__device__ __forceinline__ Blk hw_blk() {
  return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)gridDim.x, (int)gridDim.y};
}

// byte offsets of the parameters V... packed at their natural alignment; off.o[N] = total bytes
template <class... V>
struct ArgPack {
  static constexpr size_t N = sizeof...(V);
  struct Offs {
    size_t o[N + 1];
  };
  static constexpr Offs make() {
    Offs r{};
    const size_t sz[N + 1] = {sizeof(V)..., 0};
    const size_t al[N + 1] = {alignof(V)..., 1};
    size_t p = 0;
    for (size_t i = 0; i < N; i++) {
      p = (p + al[i] - 1) / al[i] * al[i];
      r.o[i] = p;
      p += sz[i];
    }
    r.o[N] = p;
    return r;
  }
  static constexpr Offs off = make();
};

constexpr int TAB_MAX = 128;     // tanks per table launch (more are split over several launches)
constexpr size_t TAB_HDR = 16;   // a record: int gx, int gy, 8 bytes unused, then the arguments
constexpr size_t TAB_ALIGN = 16; // records and packed arguments start on 16-byte boundaries

struct TabHdr {
  uint32_t start[TAB_MAX + 1];  // first block of every tank; start[ntank] = all blocks
  int ntank;
  int stride;  // bytes per record
};

template <auto Body, class... V, size_t... I>
__device__ __forceinline__ void call_rec(const Blk& b, const unsigned char* __restrict__ a,
                                         std::index_sequence<I...>) {
  using L = ArgPack<V...>;
  Body(b, *reinterpret_cast<const V*>(a + L::off.o[I])...);
}

// the record of this block's tank, and the block's logical index in that tank's grid
__device__ __forceinline__ const unsigned char* tab_block(const TabHdr& T, const unsigned char* __restrict__ rec,
                                                          Blk* b) {
  const uint32_t id = blockIdx.x;
  int lo = 0, hi = T.ntank - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (T.start[mid] <= id) lo = mid;
    else hi = mid - 1;
  }
  const unsigned char* r = rec + (size_t)lo * (size_t)T.stride;
  const int gx = *reinterpret_cast<const int*>(r), gy = *reinterpret_cast<const int*>(r + 4);
  const int lb = (int)(id - T.start[lo]);
  const int y = lb / gx;
  *b = Blk{lb - y * gx, y, gx, gy};
  return r + TAB_HDR;
}

template <int LB, auto Body, class... V>
__global__ __launch_bounds__(LB) void k_one(V... v) {
  Body(hw_blk(), v...);
}
template <int LB, int W, auto Body, class... V>
__global__ __launch_bounds__(LB, W) void k_one_w(V... v) {
  Body(hw_blk(), v...);
}
template <int LB, auto Body, class... V>
__global__ __launch_bounds__(LB) void k_tab(TabHdr T, const unsigned char* __restrict__ rec) {
  Blk b;
  const unsigned char* a = tab_block(T, rec, &b);
  call_rec<Body, V...>(b, a, std::index_sequence_for<V...>{});
}
template <int LB, int W, auto Body, class... V>
__global__ __launch_bounds__(LB, W) void k_tab_w(TabHdr T, const unsigned char* __restrict__ rec) {
  Blk b;
  const unsigned char* a = tab_block(T, rec, &b);
  call_rec<Body, V...>(b, a, std::index_sequence_for<V...>{});
}

// ------------------------------------------------------------------ host: the launch log
struct LaunchOp {
  const void* tab;  // the table kernel of the body
  dim3 grid, block;
  uint32_t shm;
  uint32_t off, bytes;  // the packed arguments in the arena
};
struct LaunchLog {
  std::vector<LaunchOp> ops;
  std::vector<unsigned char> arena;
  std::vector<uint32_t> rounds;  // index of the first op of every round
  void clear() {
    ops.clear();
    arena.clear();
    rounds.clear();
  }
};

template <class V, class A>
inline void put_arg(unsigned char* p, A&& a) {
  const V v = std::forward<A>(a);
  std::memcpy(p, &v, sizeof(V));
}

template <int LB, int W, auto Body, class Sig = decltype(Body)>
struct Kernel;
template <int LB, int W, auto Body, class... P>
struct Kernel<LB, W, Body, void (*)(Blk, P...)> {
  using Pack = ArgPack<std::decay_t<P>...>;
  static const void* table() {
    if constexpr (W > 0) return reinterpret_cast<const void*>(&k_tab_w<LB, W, Body, std::decay_t<P>...>);
    else return reinterpret_cast<const void*>(&k_tab<LB, Body, std::decay_t<P>...>);
  }
  template <size_t... I, class... A>
  static void record(LaunchLog* L, dim3 g, dim3 b, size_t shm, std::index_sequence<I...>, A&&... a) {
    const size_t base = (L->arena.size() + TAB_ALIGN - 1) & ~(TAB_ALIGN - 1);
    L->arena.resize(base + Pack::off.o[Pack::N]);
    unsigned char* p = L->arena.data() + base;
    (put_arg<std::decay_t<P>>(p + Pack::off.o[I], std::forward<A>(a)), ...);
    L->ops.push_back(LaunchOp{table(), g, b, (uint32_t)shm, (uint32_t)base, (uint32_t)Pack::off.o[Pack::N]});
  }
  template <class... A>
  static void go(LaunchLog* log, hipStream_t st, dim3 g, dim3 b, size_t shm, A&&... a) {
    static_assert(sizeof...(A) == sizeof...(P), "launch: argument count differs from the body's");
    if (log) {
      record(log, g, b, shm, std::index_sequence_for<P...>{}, std::forward<A>(a)...);
      return;
    }
    if constexpr (W > 0) k_one_w<LB, W, Body, std::decay_t<P>...><<<g, b, shm, st>>>(std::forward<A>(a)...);
    else k_one<LB, Body, std::decay_t<P>...><<<g, b, shm, st>>>(std::forward<A>(a)...);
  }
};

}  // namespace tq
