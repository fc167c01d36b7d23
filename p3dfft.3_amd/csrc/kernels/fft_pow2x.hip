// Kernels of the compile-time-length path for 3 times a power of two
// (48 ... 768), and the launcher of every length that is not a power of two:
// fft_pow2y.hip and fft_pow2z.hip instantiate the others (split so that the
// units compile in parallel).
#include "fixed_kernels.h"

namespace p3e {

int launch_fixed_set_y(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n, int esz,
                       size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st);
int launch_fixed_set_z(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n, int esz,
                       size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st);

template <typename T>
static int launch_odd(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n,
                      size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
#define P3E_ODD_CASE(L)                                                                \
  case L:                                                                              \
    return a ? launch_pow2<T, L>(a, lds_bytes, threads, nblocks, st)                   \
             : launch_pow2_grp<T, L>(tab, start, n, lds_bytes, threads, nblocks, st);
  switch (M) {
    P3E_ODD_CASE(48)
    P3E_ODD_CASE(96)
    P3E_ODD_CASE(192)
    P3E_ODD_CASE(384)
    P3E_ODD_CASE(768)
  }
#undef P3E_ODD_CASE
  return -1;
}

int launch_fixed_odd(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n, int esz,
                     size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  int rc = esz == 8 ? launch_odd<double>(M, a, tab, start, n, lds_bytes, threads, nblocks, st)
                    : launch_odd<float>(M, a, tab, start, n, lds_bytes, threads, nblocks, st);
  if (rc == -1) rc = launch_fixed_set_y(M, a, tab, start, n, esz, lds_bytes, threads, nblocks, st);
  if (rc == -1) rc = launch_fixed_set_z(M, a, tab, start, n, esz, lds_bytes, threads, nblocks, st);
  return rc == -1 ? (int)hipErrorInvalidValue : rc;
}

}  // namespace p3e
