// Power-of-two pass kernels of the compile-time-length path and the
// launchers of that path (other lengths: fft_pow2x.hip).
#include "fixed_kernels.h"

namespace p3e {

// fft_pow2x.hip: a pass (a) or a grouped launch (tab, start, n) of length
// 3 or 5 times a power of two.
int launch_fixed_odd(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n, int esz,
                     size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st);

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
  return launch_fixed_odd(a->M, a, nullptr, nullptr, 0, sizeof(T), lds_bytes, threads, nblocks, st);
}

template <typename T>
static int launch_pow2_grp_any(int M, const PassArgs *tab, const uint32_t *start, int n, size_t lds_bytes,
                               int threads, unsigned int nblocks, hipStream_t st) {
  switch (M) {
    case 16: return launch_pow2_grp<T, 16>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 32: return launch_pow2_grp<T, 32>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 64: return launch_pow2_grp<T, 64>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 128: return launch_pow2_grp<T, 128>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 256: return launch_pow2_grp<T, 256>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 512: return launch_pow2_grp<T, 512>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 1024: return launch_pow2_grp<T, 1024>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 2048: return launch_pow2_grp<T, 2048>(tab, start, n, lds_bytes, threads, nblocks, st);
    case 4096: return launch_pow2_grp<T, 4096>(tab, start, n, lds_bytes, threads, nblocks, st);
  }
  return launch_fixed_odd(M, nullptr, tab, start, n, sizeof(T), lds_bytes, threads, nblocks, st);
}

// A grouped launch of power-of-two passes of complex length M (tab and
// start in device memory; nblocks = start[n]).
int launch_pow2_group(int M, const PassArgs *tab, const uint32_t *start, int n, int prec, size_t lds_bytes,
                      int threads, unsigned int nblocks, hipStream_t st) {
  return prec == 8 ? launch_pow2_grp_any<double>(M, tab, start, n, lds_bytes, threads, nblocks, st)
                   : launch_pow2_grp_any<float>(M, tab, start, n, lds_bytes, threads, nblocks, st);
}

// Launch of a power-of-two pass (the array alignment checks of the paired
// accesses are made by the caller).
int launch_pow2_pass(const PassArgs *a, int prec, size_t lds_bytes, int threads, unsigned int nblocks,
                     hipStream_t st) {
  return prec == 8 ? launch_pow2_any<double>(a, lds_bytes, threads, nblocks, st)
                   : launch_pow2_any<float>(a, lds_bytes, threads, nblocks, st);
}

}  // namespace p3e
