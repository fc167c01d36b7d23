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
};

// A plain phase of a tile for this thread, then a barrier.
struct DevEach {
  int tid;
  template <class F> __device__ __forceinline__ void operator()(F f) const {
    f(tid);
    __syncthreads();
  }
};

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

// Single-buffer mixed-radix path: one padded LDS buffer, then the twiddles
// when they fit.  P: the odd prime radix of the pass (0: none).
template <typename T, int P>
__global__ void __launch_bounds__(P3E_MID_THREADS(sizeof(T), P)) pass_kernel_mid(PassArgs p) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  const cx<T> *tw = reinterpret_cast<const cx<T> *>(p.tw);
  const int tid = threadIdx.x, nth = blockDim.x;
  if (p.tw_in_lds) {
    cx<T> *ltw = buf + (size_t)tile_lines(p, p.B) * p.lpad;
    for (int i = tid; i < p.ntw; i += nth) ltw[i] = tw[i];
    tw = ltw;
  }
  DevPassRun<T> run{tid, nth};
  DevEach each{tid};
  // one call site (the twiddle reads then compile to flat loads, as tw may
  // point to LDS or to global memory): two inlined copies, each with one
  // address space, measured slower on MI355X
  mid_tile<T, P>(p, xcd_tile(blockIdx.x, gridDim.x), buf, nth, run, each, tw);
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
  fast_tile<T, M>(p, tile, buf, p.nthreads, run, each);
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

}  // namespace p3e

// ---------------------------------------------------------------- launchers
extern "C" int p3e_launch_pass(const p3e::PassArgs *a, int prec, size_t lds_bytes, int threads,
                               unsigned int nblocks, hipStream_t st) {
  if (nblocks == 0) return 0;
  if (a->fast) {
    if (threads < 1 || threads > P3E_FAST_MAX_THREADS) return (int)hipErrorInvalidValue;
    // a pair is two reals (R2C in, C2R out) or two complex values
    const size_t in_pair = (a->cls == p3e::CLS_R2C ? 2 : 4) * (size_t)prec;
    const size_t out_pair = (a->cls == p3e::CLS_C2R ? 2 : 4) * (size_t)prec;
    const bool in_odd = a->vec2 && (reinterpret_cast<uintptr_t>(a->in) % in_pair) != 0;
    const bool out_odd = a->vec2s && (reinterpret_cast<uintptr_t>(a->out) % out_pair) != 0;
    if (in_odd || out_odd) {
      // the paired loads / stores need the array aligned to one pair
      p3e::PassArgs b = *a;
      if (in_odd) b.vec2 = 0;
      if (out_odd) b.vec2s = 0;
      return prec == 8 ? p3e::launch_pow2_any<double>(&b, lds_bytes, threads, nblocks, st)
                       : p3e::launch_pow2_any<float>(&b, lds_bytes, threads, nblocks, st);
    }
    return prec == 8 ? p3e::launch_pow2_any<double>(a, lds_bytes, threads, nblocks, st)
                     : p3e::launch_pow2_any<float>(a, lds_bytes, threads, nblocks, st);
  }
  if (a->mid) {
    if (threads < 1 || threads > P3E_MID_THREADS(prec, a->mid_prime)) return (int)hipErrorInvalidValue;
    return prec == 8 ? p3e::launch_mid_any<double>(a, lds_bytes, threads, nblocks, st)
                     : p3e::launch_mid_any<float>(a, lds_bytes, threads, nblocks, st);
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
