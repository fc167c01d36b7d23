// Kernels of the compile-time-length path (fft_fast.h) and their
// single-length launchers, instantiated by fft_pow2.hip (powers of two) and
// fft_pow2x.hip (3 and 5 times a power of two).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "fft_fast.h"
#include "pass_dev.h"

namespace p3e {

// Compile-time-length path: complex length M at compile time, one LDS buffer for
// the tile, followed in LDS by the twiddle table (ntw entries) and, for
// R2C/C2R, the split factors (M + 1 entries): the radix passes and the
// split/merge read them from LDS instead of issuing global loads (which
// would compete with the tile's stream for the memory pipeline).  The
// twiddles are first read after the barrier that ends the load phase; the
// split factors are read by the C2R load phase itself, so R2C/C2R passes
// wait at a barrier after copying them.
template <typename T, int M>
__device__ __forceinline__ void pow2_body(const PassArgs &p, uint32_t tile) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *ltw = buf + p.tab_off;
  cx<T> *lwn = ltw + p.ntw;
  const cx<T> *gtw = reinterpret_cast<const cx<T> *>(p.tw);
  const cx<T> *gwn = reinterpret_cast<const cx<T> *>(p.wN);
  const int tid = threadIdx.x, nth = p.nthreads;
  for (int i = tid; i < p.ntw; i += blockDim.x) ltw[i] = gtw[i];
  if (p.cls == CLS_R2C || p.cls == CLS_C2R) {
    for (int i = tid; i <= M; i += blockDim.x) lwn[i] = gwn[i];
    // the C2R load phase merges (X_k, X_{M-k}) with the split factors, so
    // they must be in LDS before any thread's load phase
    __syncthreads();
  }
  DevPassRun<T> run{tid, nth};
  DevEach each{tid};
  fast_tile<T, M>(p, tile, buf, nth, run, each, ltw, lwn);
}

template <typename T, int M>
__global__ void __launch_bounds__(P3E_FAST_MAX_THREADS) pass_kernel_pow2(PassArgs p) {
  pow2_body<T, M>(p, p.xcd_map ? xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x);
}

// The same pass of many transforms in one launch (a batch stage): item i
// owns workgroups start[i] .. start[i+1] - 1 and runs with tab[i]; the
// workgroup size is the largest nthreads of the items (the others idle).
template <typename T, int M>
__global__ void __launch_bounds__(P3E_FAST_MAX_THREADS)
    pass_kernel_pow2_grp(const PassArgs *__restrict__ tab, const uint32_t *__restrict__ start, int n) {
  const uint32_t i = group_item(start, n, blockIdx.x);
  pow2_body<T, M>(tab[i], blockIdx.x - start[i]);
}

template <typename T, int M>
static inline int launch_pow2(const PassArgs *a, size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  if (lds_bytes > 65536)
    (void)hipFuncSetAttribute((const void *)pass_kernel_pow2<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes);
  hipLaunchKernelGGL((pass_kernel_pow2<T, M>), dim3(nblocks), dim3(threads), lds_bytes, st, *a);
  return (int)hipGetLastError();
}

template <typename T, int M>
static inline int launch_pow2_grp(const PassArgs *tab, const uint32_t *start, int n, size_t lds_bytes, int threads,
                           unsigned int nblocks, hipStream_t st) {
  if (lds_bytes > 65536)
    (void)hipFuncSetAttribute((const void *)pass_kernel_pow2_grp<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes);
  hipLaunchKernelGGL((pass_kernel_pow2_grp<T, M>), dim3(nblocks), dim3(threads), lds_bytes, st, tab, start, n);
  return (int)hipGetLastError();
}

}  // namespace p3e
