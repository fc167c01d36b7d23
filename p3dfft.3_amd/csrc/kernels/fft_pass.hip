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

#include "fft_core.h"
#include "fft_fast.h"

namespace p3e {

struct DevBarrier {
  __device__ void operator()() const { __syncthreads(); }
};

template <typename T>
__global__ void __launch_bounds__(256) pass_kernel(PassArgs p) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *bufA = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *bufB = bufA + (size_t)p.B * p.lpad;
  run_tile<T>(p, blockIdx.x, bufA, bufB, threadIdx.x, blockDim.x, DevBarrier());
}

// Power-of-two path: one LDS buffer, P3E_FAST_E elements per thread.
template <typename T>
__global__ void __launch_bounds__(1024) pass_kernel_fast(PassArgs p) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *buf = reinterpret_cast<cx<T> *>(smem_raw);
  uint32_t f;
  int e0, nb;
  fast_tile_coords(p, blockIdx.x, &f, &e0, &nb);
  const int tid = threadIdx.x, nth = blockDim.x;
  fast_load<T>(p, f, e0, nb, buf, tid, nth);
  __syncthreads();
  if (p.cls == CLS_C2R) {
    fast_c2r_merge<T>(p, nb, buf, tid, nth);
    __syncthreads();
  }
  const cx<T> *tw = reinterpret_cast<const cx<T> *>(p.tw);
  const int s = (p.cls == CLS_R2C) ? -1 : p.sign;
  cx<T> v[P3E_FAST_E];
  int lines[P3E_FAST_E], js[P3E_FAST_E];
  int Ns = 1, logNs = 0;
  for (int i = 0; i < p.nrad; i++) {
    const int R = p.radix[i];
    fast_phase<T>(0, R, buf, p, nb, Ns, logNs, s, tw, tid, nth, v, lines, js);
    __syncthreads();
    fast_phase<T>(1, R, buf, p, nb, Ns, logNs, s, tw, tid, nth, v, lines, js);
    __syncthreads();
    Ns *= R;
    logNs += (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : 4;
  }
  fast_store<T>(p, f, e0, nb, buf, tid, nth);
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
    if (prec == 8) {
      if (lds_bytes > 65536)
        (void)hipFuncSetAttribute((const void *)p3e::pass_kernel_fast<double>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
      hipLaunchKernelGGL(p3e::pass_kernel_fast<double>, dim3(nblocks), dim3(threads), lds_bytes, st, *a);
    } else {
      if (lds_bytes > 65536)
        (void)hipFuncSetAttribute((const void *)p3e::pass_kernel_fast<float>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
      hipLaunchKernelGGL(p3e::pass_kernel_fast<float>, dim3(nblocks), dim3(threads), lds_bytes, st, *a);
    }
    return (int)hipGetLastError();
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
