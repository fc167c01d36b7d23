// Kernels of the compile-time-length path for 3 and 5 times a power of two
// (48 ... 768, 80 ... 640, 120) and their launcher.
#include "fixed_kernels.h"

namespace p3e {

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
    P3E_ODD_CASE(80)
    P3E_ODD_CASE(160)
    P3E_ODD_CASE(320)
    P3E_ODD_CASE(640)
    P3E_ODD_CASE(120)
  }
#undef P3E_ODD_CASE
  return (int)hipErrorInvalidValue;
}

int launch_fixed_odd(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n, int esz,
                     size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  return esz == 8 ? launch_odd<double>(M, a, tab, start, n, lds_bytes, threads, nblocks, st)
                  : launch_odd<float>(M, a, tab, start, n, lds_bytes, threads, nblocks, st);
}

}  // namespace p3e
