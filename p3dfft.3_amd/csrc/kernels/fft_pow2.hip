// Power-of-two pass kernels (fft_fast.h) and their launcher.
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "fft_fast.h"
#include "pass_dev.h"

namespace p3e {

// Power-of-two path: complex length M at compile time, one LDS buffer for
// the tile, followed in LDS by the twiddle table (M entries) and, for
// R2C/C2R, the split factors (M + 1 entries): the radix passes and the
// split/merge read them from LDS instead of issuing global loads (which
// would compete with the tile's stream for the memory pipeline).  The
// twiddles are first read after the barrier that ends the load phase; the
// split factors are read by the C2R load phase itself, so R2C/C2R passes
// wait at a barrier after copying them.
template <typename T, int M>
__global__ void __launch_bounds__(P3E_FAST_MAX_THREADS) pass_kernel_pow2(PassArgs p0) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *ltw = buf + (size_t)p0.B * p0.lp;
  cx<T> *lwn = ltw + M;
  const cx<T> *gtw = reinterpret_cast<const cx<T> *>(p0.tw);
  const cx<T> *gwn = reinterpret_cast<const cx<T> *>(p0.wN);
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int i = tid; i < p0.ntw; i += nth) ltw[i] = gtw[i];
  if (p0.cls != CLS_C2C) {
    for (int i = tid; i <= M; i += nth) lwn[i] = gwn[i];
    // the C2R load phase merges (X_k, X_{M-k}) with the split factors, so
    // they must be in LDS before any thread's load phase
    __syncthreads();
  }
  PassArgs p = p0;
  p.tw = ltw;
  p.wN = lwn;
  DevPassRun<T> run{tid, nth};
  DevEach each{tid};
  fast_tile<T, M>(p, blockIdx.x, buf, nth, run, each);
}

template <typename T, int M>
static int launch_pow2(const PassArgs *a, size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  if (lds_bytes > 65536)
    (void)hipFuncSetAttribute((const void *)pass_kernel_pow2<T, M>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds_bytes);
  hipLaunchKernelGGL((pass_kernel_pow2<T, M>), dim3(nblocks), dim3(threads), lds_bytes, st, *a);
  return (int)hipGetLastError();
}

template <typename T>
static int launch_pow2_any(const PassArgs *a, size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  switch (a->M) {
    case 16: return launch_pow2<T, 16>(a, lds_bytes, threads, nblocks, st);
    case 32: return launch_pow2<T, 32>(a, lds_bytes, threads, nblocks, st);
    case 64: return launch_pow2<T, 64>(a, lds_bytes, threads, nblocks, st);
    case 128: return launch_pow2<T, 128>(a, lds_bytes, threads, nblocks, st);
    case 256: return launch_pow2<T, 256>(a, lds_bytes, threads, nblocks, st);
    case 512: return launch_pow2<T, 512>(a, lds_bytes, threads, nblocks, st);
    case 1024: return launch_pow2<T, 1024>(a, lds_bytes, threads, nblocks, st);
    case 2048: return launch_pow2<T, 2048>(a, lds_bytes, threads, nblocks, st);
    case 4096: return launch_pow2<T, 4096>(a, lds_bytes, threads, nblocks, st);
  }
  return (int)hipErrorInvalidValue;
}

// Launch of a power-of-two pass (the array alignment checks of the paired
// accesses are made by the caller).
int launch_pow2_pass(const PassArgs *a, int prec, size_t lds_bytes, int threads, unsigned int nblocks,
                     hipStream_t st) {
  return prec == 8 ? launch_pow2_any<double>(a, lds_bytes, threads, nblocks, st)
                   : launch_pow2_any<float>(a, lds_bytes, threads, nblocks, st);
}

}  // namespace p3e
