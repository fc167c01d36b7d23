// Device-side pieces shared by the pass kernels of the HIP library (one
// kernel family per translation unit, so that they compile in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.h"

namespace p3e {

struct DevBarrier {
  __device__ void operator()() const { __syncthreads(); }
};

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin (block b
// to XCD b % 8), and each XCD has its own L2.  Tiles that are adjacent along
// the tile axis share cache lines when their rows are shorter than a line,
// so consecutive tiles are given to workgroups of the same XCD: block b
// takes tile (b % 8) * (G / 8) + b / 8 (the last G % 8 blocks keep their
// index).  A bijection on [0, G).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t G) {
  const uint32_t per = G >> 3;
  return b < (per << 3) ? (b & 7u) * per + (b >> 3) : b;
}

// One radix pass of the power-of-two path for this thread: read half,
// barrier, write half, barrier; the thread's values live in registers.
template <typename T> struct DevPassRun {
  int tid, nth;
  template <class PS> __device__ __forceinline__ void operator()(PS &ps) const {
    cx<T> v[PS::SLOTS];
    if (tid < nth) ps.read(tid, nth, v);
    __syncthreads();
    if (tid < nth) ps.write(tid, nth, v);
    __syncthreads();
  }
  // one pass of each of two sequences: threads 0 .. na - 1 run pa, the next
  // nb run pb (fft_fast.h, dual_seq)
  template <class PA, class PB> __device__ __forceinline__ void pair(PA &pa, int na, PB &pb, int nb) const {
    constexpr int S = PA::SLOTS > PB::SLOTS ? PA::SLOTS : PB::SLOTS;
    cx<T> v[S];
    const bool in_a = tid < na, in_b = !in_a && tid - na < nb;
    if (in_a) pa.read(tid, na, v);
    else if (in_b) pb.read(tid - na, nb, v);
    __syncthreads();
    if (in_a) pa.write(tid, na, v);
    else if (in_b) pb.write(tid - na, nb, v);
    __syncthreads();
  }
};

// A plain phase of a tile for this thread, then a barrier.
struct DevEach {
  int tid;
  template <class F> __device__ __forceinline__ void operator()(F f) const {
    f(tid);
    __syncthreads();
  }
};

// The item of a grouped launch that owns workgroup b: item i owns
// workgroups start[i] .. start[i+1] - 1 (start[0] = 0, n items); a uniform
// binary search.
__device__ __forceinline__ uint32_t group_item(const uint32_t *start, int n, uint32_t b) {
  uint32_t lo = 0, hi = (uint32_t)n;
  while (hi - lo > 1) {
    const uint32_t m = (lo + hi) >> 1;
    if (start[m] <= b) lo = m;
    else hi = m;
  }
  return lo;
}

}  // namespace p3e
