// Kernels of the compile-time-length path for 7 times a power of two (112, 224) and 9 times one (36, 72, 144) and their
// launcher; fft_pow2x.hip dispatches the lengths that are not powers of two
// over the units that instantiate them (split so that they compile in
// parallel).
#include "fixed_kernels.h"

namespace p3e {

template <typename T>
static int launch_set(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n,
                      size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
#define P3E_ODD_CASE(L)                                                                \
  case L:                                                                              \
    return a ? launch_pow2<T, L>(a, lds_bytes, threads, nblocks, st)                   \
             : launch_pow2_grp<T, L>(tab, start, n, lds_bytes, threads, nblocks, st);
  switch (M) {
    P3E_ODD_CASE(112)
    P3E_ODD_CASE(224)
    P3E_ODD_CASE(36)
    P3E_ODD_CASE(72)
    P3E_ODD_CASE(144)
  }
#undef P3E_ODD_CASE
  return -1;
}

// -1 when M is not one of this unit's lengths
int launch_fixed_set_z(int M, const PassArgs *a, const PassArgs *tab, const uint32_t *start, int n, int esz,
           size_t lds_bytes, int threads, unsigned int nblocks, hipStream_t st) {
  return esz == 8 ? launch_set<double>(M, a, tab, start, n, lds_bytes, threads, nblocks, st)
                  : launch_set<float>(M, a, tab, start, n, lds_bytes, threads, nblocks, st);
}

}  // namespace p3e
