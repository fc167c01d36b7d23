// Fused reorder + 1D transform pass for CDNA4 (gfx950).
//
// One workgroup owns a tile of B lines along the transform axis d, adjacent
// along a second axis e, at one index of the third axis f.  It streams the
// tile from HBM into LDS in the order that keeps global accesses coalesced
// (along d when d is stride-1, along e otherwise), builds the sequence whose
// DFT is needed (R2C/C2R Hermitian completion, DCT/DST even/odd extensions
// with pre-twiddles, Bluestein chirp), runs a mixed-radix Stockham FFT
// between two LDS buffers, applies the output rule (R2R post-twiddle,
// derivative i*k, scale) and writes the tile with the output strides, which
// may permute the memory order.  This one kernel replaces, per stage, the
// reference's transpose loops (build/exec.C:737-2266) and the FFTW call
// they wrap (build/init.C:1146-1188).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "fft_core.h"
#include "fft_fast.h"
#include "fft_mid.h"
#include "pass_dev.h"

namespace p3e {

// Generic path: two LDS line buffers, then the L twiddles (copied from
// global memory before the first barrier of the tile).
template <typename T>
__global__ void __launch_bounds__(256) pass_kernel(PassArgs p) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *bufA = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *bufB = bufA + (size_t)p.B * p.lpad;
  const cx<T> *tw = reinterpret_cast<const cx<T> *>(p.tw);
  if (p.tw_in_lds) {
    cx<T> *ltw = bufB + (size_t)p.B * p.lpad;
    for (int i = threadIdx.x; i < p.ntw; i += blockDim.x) ltw[i] = tw[i];
    tw = ltw;
  }
  // one call site: a second inlined copy of the tile code makes the
  // compiler keep PassArgs in scratch memory
  run_tile<T>(p, xcd_tile(blockIdx.x, gridDim.x), bufA, bufB, threadIdx.x, blockDim.x, DevBarrier(), tw);
}

// Batch step: a grid-stride loop over a work list of (pass, tile) entries
// drawn from many transforms (and from different stages of each), every
// workgroup with P3E_STEP_THREADS threads and LDS for the largest tile.
// Stages of one transform are ordered by launch boundaries: on MI355X each
// XCD has its own L2, so a kernel boundary is the cheap way to make one
// stage's output visible to the next stage's tiles on other XCDs.
template <typename T, int M>
__device__ void step_pow2(const PassArgs &p, uint32_t tile, cx<T> *buf) {
  DevPassRun<T> run{(int)threadIdx.x, p.nthreads};
  DevEach each{(int)threadIdx.x};
  fast_tile<T, M>(p, tile, buf, p.nthreads, run, each, reinterpret_cast<const cx<T> *>(p.tw),
                  reinterpret_cast<const cx<T> *>(p.wN));
}

template <typename T>
__global__ void __launch_bounds__(P3E_STEP_THREADS)
    step_kernel(const PassArgs *__restrict__ table, const WorkItem *__restrict__ work, int nwork) {
  extern __shared__ unsigned char smem_raw[];
  __shared__ PassArgs sp;
  cx<T> *bufA = reinterpret_cast<cx<T> *>(smem_raw);
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const WorkItem wi = work[w];
    __syncthreads();  // the previous tile is done with sp and the LDS
    {
      const uint32_t *src = reinterpret_cast<const uint32_t *>(table + wi.pass);
      uint32_t *dst = reinterpret_cast<uint32_t *>(&sp);
      for (int i = threadIdx.x; i < (int)(sizeof(PassArgs) / 4); i += blockDim.x) dst[i] = src[i];
    }
    __syncthreads();
    if (sp.fast) {
      switch (sp.M) {
        case 16: step_pow2<T, 16>(sp, wi.tile, bufA); break;
        case 32: step_pow2<T, 32>(sp, wi.tile, bufA); break;
        case 64: step_pow2<T, 64>(sp, wi.tile, bufA); break;
        case 128: step_pow2<T, 128>(sp, wi.tile, bufA); break;
        case 256: step_pow2<T, 256>(sp, wi.tile, bufA); break;
        case 512: step_pow2<T, 512>(sp, wi.tile, bufA); break;
        case 1024: step_pow2<T, 1024>(sp, wi.tile, bufA); break;
        case 2048: step_pow2<T, 2048>(sp, wi.tile, bufA); break;
        default: step_pow2<T, 4096>(sp, wi.tile, bufA); break;
      }
    } else if (sp.mid) {
      DevPassRun<T> run{(int)threadIdx.x, sp.nthreads};
      DevEach each{(int)threadIdx.x};
      // batch plans admit no prime radix to the single-buffer path
      mid_tile<T, 0>(sp, wi.tile, bufA, sp.nthreads, run, each, reinterpret_cast<const cx<T> *>(sp.tw));
    } else {
      cx<T> *bufB = bufA + (size_t)sp.B * sp.lpad;
      run_tile<T>(sp, wi.tile, bufA, bufB, threadIdx.x, blockDim.x, DevBarrier(),
                  reinterpret_cast<const cx<T> *>(sp.tw));
    }
  }
}

// ---------------------------------------------------------------- compute_deriv
// build/deriv.C:85-186: multiply a stored spectrum by i*k along storage rank
// ldir, with wavenumbers from the logical axis (g, mid, glob_start).
template <typename T>
__global__ void deriv_kernel(const cx<T> *in, cx<T> *out, long long n, int s0, int s1, int ldir, int g,
                             int mid, int gstart) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int i;
  if (ldir == 0) i = (int)(p % s0);
  else if (ldir == 1) i = (int)((p / s0) % s1);
  else i = (int)(p / ((long long)s0 * s1));
  cx<T> v = in[p];
  T kk;
  if (i < mid) kk = (T)(i + gstart);
  else if (i == mid) kk = (T)0;
  else kk = (T)(i + gstart - g);
  out[p] = mk(-kk * v.y, kk * v.x);
}

int launch_pow2_pass(const PassArgs *a, int prec, size_t lds_bytes, int threads, unsigned int nblocks,
                     hipStream_t st);  // fft_pow2.hip
int launch_mid_pass(const PassArgs *a, int prec, size_t lds_bytes, int threads, unsigned int nblocks,
                    hipStream_t st);  // fft_midk.hip
int launch_pow2_group(int M, const PassArgs *tab, const uint32_t *start, int n, int prec, size_t lds_bytes,
                      int threads, unsigned int nblocks, hipStream_t st);  // fft_pow2.hip
int launch_mid_group(const PassArgs *tab, const uint32_t *start, int n, int prec, size_t lds_bytes, int threads,
                     unsigned int nblocks, hipStream_t st);  // fft_midk.hip

}  // namespace p3e

// ---------------------------------------------------------------- launchers
extern "C" int p3e_launch_pass(const p3e::PassArgs *a, int prec, size_t lds_bytes, int threads,
                               unsigned int nblocks, hipStream_t st) {
  if (nblocks == 0) return 0;
  if (a->fast) {
    if (threads < 1 || threads > P3E_FAST_MAX_THREADS) return (int)hipErrorInvalidValue;
    p3e::PassArgs b = *a;
    p3e::drop_unaligned_pairs(b, prec);
    return p3e::launch_pow2_pass(&b, prec, lds_bytes, threads, nblocks, st);
  }
  if (a->mid) {
    if (threads < 1 || threads > P3E_MID_THREADS(prec, a->mid_prime)) return (int)hipErrorInvalidValue;
    return p3e::launch_mid_pass(a, prec, lds_bytes, threads, nblocks, st);
  }
  if (prec == 8) {
    if (lds_bytes > 65536)
      (void)hipFuncSetAttribute((const void *)p3e::pass_kernel<double>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    hipLaunchKernelGGL(p3e::pass_kernel<double>, dim3(nblocks), dim3(threads), lds_bytes, st, *a);
  } else {
    if (lds_bytes > 65536)
      (void)hipFuncSetAttribute((const void *)p3e::pass_kernel<float>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    hipLaunchKernelGGL(p3e::pass_kernel<float>, dim3(nblocks), dim3(threads), lds_bytes, st, *a);
  }
  return (int)hipGetLastError();
}

// One launch of the same pass of n transforms (a batch stage group): kind
// 1 power-of-two passes of complex length M, kind 2 single-buffer passes
// without an odd prime radix above 5.  tab (n PassArgs) and start (n + 1
// workgroup offsets, start[n] = nblocks) are device arrays.
extern "C" int p3e_launch_group(int kind, int M, const void *tab, const void *start, int n, int prec, size_t lds,
                                int threads, unsigned int nblocks, hipStream_t st) {
  if (nblocks == 0 || n <= 0) return 0;
  const p3e::PassArgs *t = (const p3e::PassArgs *)tab;
  const uint32_t *s = (const uint32_t *)start;
  if (kind == 1) {
    if (threads < 1 || threads > P3E_FAST_MAX_THREADS) return (int)hipErrorInvalidValue;
    return p3e::launch_pow2_group(M, t, s, n, prec, lds, threads, nblocks, st);
  }
  if (kind == 2) {
    if (threads < 1 || threads > P3E_MID_THREADS(prec, 0)) return (int)hipErrorInvalidValue;
    return p3e::launch_mid_group(t, s, n, prec, lds, threads, nblocks, st);
  }
  return (int)hipErrorInvalidValue;
}

extern "C" int p3e_launch_step(const void *table, const void *work, int nwork, int prec, size_t lds,
                               unsigned int grid, hipStream_t st) {
  if (nwork <= 0) return 0;
  const void *fn = prec == 8 ? (const void *)p3e::step_kernel<double> : (const void *)p3e::step_kernel<float>;
  if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (prec == 8)
    hipLaunchKernelGGL(p3e::step_kernel<double>, dim3(grid), dim3(P3E_STEP_THREADS), lds, st,
                       (const p3e::PassArgs *)table, (const p3e::WorkItem *)work, nwork);
  else
    hipLaunchKernelGGL(p3e::step_kernel<float>, dim3(grid), dim3(P3E_STEP_THREADS), lds, st,
                       (const p3e::PassArgs *)table, (const p3e::WorkItem *)work, nwork);
  return (int)hipGetLastError();
}

// Workgroups of the step kernel resident per CU for a given dynamic LDS size
// (cached per precision and LDS size; the batch and the Poisson service call
// this from different threads).
extern "C" int p3e_step_occupancy(int prec, size_t lds) {
  static std::mutex mu;
  static std::map<std::pair<int, size_t>, int> cache;
  const int k = prec == 8 ? 8 : 4;
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find({k, lds});
  if (it != cache.end()) return it->second;
  const void *fn = prec == 8 ? (const void *)p3e::step_kernel<double> : (const void *)p3e::step_kernel<float>;
  if (lds > 65536) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, P3E_STEP_THREADS, lds) != hipSuccess || n < 1) {
    (void)hipGetLastError();
    n = 1;
  }
  cache[{k, lds}] = n;
  return n;
}

extern "C" int p3e_launch_deriv(const void *in, void *out, long long n, int s0, int s1, int ldir, int g,
                                int mid, int gstart, int prec, hipStream_t st) {
  if (n <= 0) return 0;
  const int th = 256;
  unsigned int nb = (unsigned int)((n + th - 1) / th);
  if (prec == 8)
    hipLaunchKernelGGL(p3e::deriv_kernel<double>, dim3(nb), dim3(th), 0, st,
                       (const p3e::cx<double> *)in, (p3e::cx<double> *)out, n, s0, s1, ldir, g, mid, gstart);
  else
    hipLaunchKernelGGL(p3e::deriv_kernel<float>, dim3(nb), dim3(th), 0, st,
                       (const p3e::cx<float> *)in, (p3e::cx<float> *)out, n, s0, s1, ldir, g, mid, gstart);
  return (int)hipGetLastError();
}
