// Single-buffer mixed-radix pass kernels (fft_mid.h), one per odd prime
// radix, and their launcher.
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "fft_mid.h"
#include "pass_dev.h"

namespace p3e {

// Single-buffer mixed-radix path: one padded LDS buffer, then the twiddles
// when they fit.  P: the odd prime radix of the pass (0: none).
template <typename T, int P>
__device__ __forceinline__ void mid_body(const PassArgs &p, uint32_t tile) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  const cx<T> *tw = reinterpret_cast<const cx<T> *>(p.tw);
  const int tid = threadIdx.x, nth = p.nthreads;
  if (p.tw_in_lds) {
    cx<T> *ltw = buf + (size_t)tile_lines(p, p.B) * p.lpad;
    for (int i = tid; i < p.ntw; i += blockDim.x) ltw[i] = tw[i];
    tw = ltw;
  }
  DevPassRun<T> run{tid, nth};
  DevEach each{tid};
  // one call site (the twiddle reads then compile to flat loads, as tw may
  // point to LDS or to global memory): two inlined copies, each with one
  // address space, measured slower on MI355X
  mid_tile<T, P>(p, tile, buf, nth, run, each, tw);
}

template <typename T, int P>
__global__ void __launch_bounds__(P3E_MID_THREADS(sizeof(T), P)) pass_kernel_mid(PassArgs p) {
  mid_body<T, P>(p, xcd_tile(blockIdx.x, gridDim.x));
}

// The same pass of many transforms in one launch (see pass_kernel_pow2_grp).
template <typename T, int P>
__global__ void __launch_bounds__(P3E_MID_THREADS(sizeof(T), P))
    pass_kernel_mid_grp(const PassArgs *__restrict__ tab, const uint32_t *__restrict__ start, int n) {
  const uint32_t i = group_item(start, n, blockIdx.x);
  mid_body<T, P>(tab[i], blockIdx.x - start[i]);
}

template <typename T, int P>
static int launch_mid(const PassArgs *a, size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  if (lds_bytes > 65536)
    (void)hipFuncSetAttribute((const void *)pass_kernel_mid<T, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes);
  hipLaunchKernelGGL((pass_kernel_mid<T, P>), dim3(nblocks), dim3(threads), lds_bytes, st, *a);
  return (int)hipGetLastError();
}

template <typename T>
static int launch_mid_any(const PassArgs *a, size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  switch (a->mid_prime) {
    case 0: return launch_mid<T, 0>(a, lds_bytes, threads, nblocks, st);
    case 7: return launch_mid<T, 7>(a, lds_bytes, threads, nblocks, st);
    case 11: return launch_mid<T, 11>(a, lds_bytes, threads, nblocks, st);
    case 13: return launch_mid<T, 13>(a, lds_bytes, threads, nblocks, st);
    case 17: return launch_mid<T, 17>(a, lds_bytes, threads, nblocks, st);
    case 19: return launch_mid<T, 19>(a, lds_bytes, threads, nblocks, st);
    case 23: return launch_mid<T, 23>(a, lds_bytes, threads, nblocks, st);
    case 29: return launch_mid<T, 29>(a, lds_bytes, threads, nblocks, st);
  }
  // primes above 29: FP32 only (the planner's limit for FP64)
  if constexpr (sizeof(T) == 4) {
    switch (a->mid_prime) {
      case 31: return launch_mid<T, 31>(a, lds_bytes, threads, nblocks, st);
      case 37: return launch_mid<T, 37>(a, lds_bytes, threads, nblocks, st);
      case 41: return launch_mid<T, 41>(a, lds_bytes, threads, nblocks, st);
      case 43: return launch_mid<T, 43>(a, lds_bytes, threads, nblocks, st);
      case 47: return launch_mid<T, 47>(a, lds_bytes, threads, nblocks, st);
    }
  }
  return (int)hipErrorInvalidValue;
}

template <typename T, int P>
static int launch_mid_grp(const PassArgs *tab, const uint32_t *start, int n, size_t lds_bytes, int threads,
                          unsigned int nblocks, hipStream_t st) {
  if (lds_bytes > 65536)
    (void)hipFuncSetAttribute((const void *)pass_kernel_mid_grp<T, P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes);
  hipLaunchKernelGGL((pass_kernel_mid_grp<T, P>), dim3(nblocks), dim3(threads), lds_bytes, st, tab, start, n);
  return (int)hipGetLastError();
}

// A grouped launch of single-buffer passes without an odd prime radix
// above 5 (the batch items' lengths); tab and start in device memory.
int launch_mid_group(const PassArgs *tab, const uint32_t *start, int n, int prec, size_t lds_bytes, int threads,
                     unsigned int nblocks, hipStream_t st) {
  return prec == 8 ? launch_mid_grp<double, 0>(tab, start, n, lds_bytes, threads, nblocks, st)
                   : launch_mid_grp<float, 0>(tab, start, n, lds_bytes, threads, nblocks, st);
}

int launch_mid_pass(const PassArgs *a, int prec, size_t lds_bytes, int threads, unsigned int nblocks,
                    hipStream_t st) {
  return prec == 8 ? launch_mid_any<double>(a, lds_bytes, threads, nblocks, st)
                   : launch_mid_any<float>(a, lds_bytes, threads, nblocks, st);
}

}  // namespace p3e
