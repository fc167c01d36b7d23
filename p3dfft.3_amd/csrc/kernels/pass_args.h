// Launch arguments of the fused reorder + 1D transform pass kernel.
// Shared by the host planner (C++) and the HIP kernels; plain data only.
#pragma once
#include <stdint.h>

namespace p3e {

enum Cls { CLS_EMPTY = 0, CLS_R2C = 1, CLS_C2R = 2, CLS_C2C = 3, CLS_R2R = 4 };

// FFTW r2r kinds (same numbering as the oracle)
enum R2RKind {
  REDFT00 = 0, REDFT10 = 1, REDFT01 = 2, REDFT11 = 3,
  RODFT00 = 4, RODFT10 = 5, RODFT01 = 6, RODFT11 = 7
};

// Unsigned division by a runtime constant: q = (umulhi(n, mul) + n) >> shift,
// exact for 0 <= n < 2^31 (Granlund-Montgomery round-up method).
struct FastDiv {
  uint32_t d, mul, shift;
};

static inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  f.shift = s;
  f.mul = (uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1);
  return f;
}

#define P3E_MAX_RADIX_PASSES 24

// Load/store orderings of the tile, chosen from the strides so that
// consecutive threads touch consecutive addresses.
enum Order { ORD_LINE = 0, ORD_TILE = 1 };

struct PassArgs {
  const void *in;
  void *out;
  const void *tw;         // L entries: exp(-2 pi i t / L)
  const void *chirp;      // W entries (Bluestein): exp(sign i pi n^2 / W)
  const void *chirp_hat;  // L entries: DFT_L of the conjugate chirp, scaled 1/L
  const void *pre;        // r2r pre-twiddle, W entries (or null)
  const void *post;       // r2r post-twiddle, nout entries (or null)
  long long is_d, is_e, is_f;  // input element strides (units of input elements)
  long long os_d, os_e, os_f;  // output element strides
  int cls, r2r_kind;
  int in_cplx, out_cplx;
  int N;           // transform length (real length for R2C/C2R)
  int nin, nout;   // elements along the transform axis, input / output
  int W;           // inner complex DFT length
  int L;           // length of the FFT done in LDS (W, or the Bluestein length)
  int sign;        // sign of the inner DFT of length W
  int bluestein;
  int nrad;
  int radix[P3E_MAX_RADIX_PASSES];
  int ne, nf;      // extents of the tile axis e and the outer axis f
  int B, logB;     // lines per workgroup (power of two)
  int ntile_e;
  int lpad;        // LDS pitch of one line, in complex elements
  int zero_fill;   // clear the line before scattering the extended input
  int post_off;    // output k reads work element k + post_off
  int post_mode;   // 0: none, 1: multiply by post[k]
  int load_order, store_order;
  int deriv, dg, dmid;  // spectral derivative (compute_deriv_loc semantics)
  int has_scale;
  double scale;
  FastDiv div_nin, div_nout, div_ntile;
  // power-of-two path (fast != 0): one LDS buffer, register-staged passes,
  // R2C/C2R through a half-length complex FFT of the packed real line
  int fast;
  int M, logM;       // complex FFT length of the fast path (N, or N/2 for R2C/C2R)
  int lp;            // LDS pitch of one line in the fast path (padded)
  const void *wN;    // R2C/C2R: M+1 entries exp(-2 pi i k / N)
  int nthreads;      // threads that work on a tile (a batch launch may have more)
  FastDiv div_pairs; // C2R: division by M/2 + 1, the pairs (X_k, X_{M-k}) per line
  int tw_in_lds;     // generic path: the standalone kernel copies tw to LDS
  int ntw;           // entries of the twiddle table tw
  int vec2;          // power-of-two R2C, tile-order load: one load reads lines j, j+1 (fft_fast.h)
  int vec2s;         // power-of-two C2R, tile-order store: one store writes lines j, j+1
  int mid;           // single-buffer mixed-radix path (fft_mid.h)
  int mid_prime;     // its odd prime radix 7..47 (0: none); selects the kernel instantiation
  int pair;          // R2C/C2R on the single-buffer path: lines 2t, 2t+1 share LDS line t
                    // as real and imaginary parts (one complex FFT for two real lines)
  // generic path, radix pass i: division by L / radix[i] (butterflies per
  // line) and by the pass's input stride Ns = radix[0] ... radix[i-1]
  FastDiv div_nbf[P3E_MAX_RADIX_PASSES], div_ns[P3E_MAX_RADIX_PASSES];
  // Spectral operator on the output (the Poisson service): 0 none; 1
  // multiply by -1/|k|^2, 0 at k = 0.  The wavenumber of output index i along
  // the pass's axis d, tile axis e and outer axis f (roles r = 0, 1, 2) is i
  // if khalf[r] (the half spectrum of an R2C axis) or 2i < kn[r], else
  // i - kn[r], with kn[r] the axis' logical real length.
  int spec_op;
  int kn[3], khalf[3];
  // compile-time-length path: workgroups take tiles in the XCD-grouped
  // order (pass_dev.h, xcd_tile), for tiles whose tile-order rows are
  // shorter than a 128-byte cache line
  int xcd_map;
  // offset of the constant tables in the kernel's LDS, in complex elements
  // (after the tile)
  int tab_off;
  // complex DCT-I of M + 1 points split into an M-point FFT and an
  // M/2-point inverse FFT (fft_fast.h, dct1_tile): the second one's lines
  // have LDS pitch lp2 after the first one's; the tables are the two
  // twiddle tables (ntw_a and ntw_b entries) and M/4 + 1 pre-twiddles
  int dct_split;
  int lp2, ntw_a, ntw_b;
};

// One entry of a batch work list: tile `tile` of the pass table[pass].
struct WorkItem {
  uint32_t pass, tile;
};

// The paired global accesses of the power-of-two path (vec2, vec2s) need
// the array aligned to one pair: two reals (R2C input, C2R output) or two
// complex values.  Clears the flag of a side whose array is not.
static inline void drop_unaligned_pairs(PassArgs &a, int prec) {
  const unsigned long in_pair = (a.cls == CLS_R2C ? 2ul : 4ul) * (unsigned long)prec;
  const unsigned long out_pair = (a.cls == CLS_C2R ? 2ul : 4ul) * (unsigned long)prec;
  if (a.vec2 && reinterpret_cast<unsigned long>(a.in) % in_pair != 0) a.vec2 = 0;
  if (a.vec2s && reinterpret_cast<unsigned long>(a.out) % out_pair != 0) a.vec2s = 0;
}

// Workgroup size of the batch (step) kernel.
#define P3E_STEP_THREADS 256

// elements held in registers per thread by the fast path
#define P3E_FAST_E 16
// Single-buffer mixed-radix path: elements of the tile per thread, at most
#ifndef P3E_MID_E
#define P3E_MID_E 16
#endif
// and its LDS budget per workgroup in bytes (a tile may exceed it up to the
// 160 KiB of a CU when one line alone does)
#ifndef P3E_MID_LDS
#define P3E_MID_LDS (80 * 1024)
#endif
// and its largest workgroup for FP32 passes without an odd prime radix
// (the others, whose butterflies need more registers, use 512)
#ifndef P3E_MID_MAXTH
#define P3E_MID_MAXTH 512
#endif
#define P3E_MID_THREADS(esz, prime) (((esz) == 4 && (prime) == 0) ? P3E_MID_MAXTH : 512)
// threads per workgroup of the fast path, at most
#define P3E_FAST_MAX_THREADS 512

// Radix sequence of the compile-time-length path for each complex length M
// (the kernel's compile-time sequences in fft_fast.h are the same); 0 if M
// is not supported.  Powers of two from 16 to 4096; 3 times 16 to 256, 5
// times 16 to 128, 7 times 16 and 32, 9 times 4 to 16, and 15 times 4, 8
// and 16, with the odd radices last (radices 7 and 9 run as direct odd
// DFTs, bf_prime).
static inline int fast_radices(int M, int *rad) {
  switch (M) {
    case 48: rad[0] = 16; rad[1] = 3; return 2;
    case 96: rad[0] = 8; rad[1] = 4; rad[2] = 3; return 3;
    case 192: rad[0] = 16; rad[1] = 4; rad[2] = 3; return 3;
    case 384: rad[0] = 16; rad[1] = 8; rad[2] = 3; return 3;
    case 768: rad[0] = 16; rad[1] = 16; rad[2] = 3; return 3;
    case 80: rad[0] = 16; rad[1] = 5; return 2;
    case 160: rad[0] = 8; rad[1] = 4; rad[2] = 5; return 3;
    case 320: rad[0] = 16; rad[1] = 4; rad[2] = 5; return 3;
    case 640: rad[0] = 16; rad[1] = 8; rad[2] = 5; return 3;
    case 120: rad[0] = 8; rad[1] = 5; rad[2] = 3; return 3;
    case 60: rad[0] = 4; rad[1] = 5; rad[2] = 3; return 3;
    case 72: rad[0] = 8; rad[1] = 9; return 2;
    case 144: rad[0] = 16; rad[1] = 9; return 2;
    case 112: rad[0] = 16; rad[1] = 7; return 2;
    case 224: rad[0] = 8; rad[1] = 4; rad[2] = 7; return 3;
    case 240: rad[0] = 16; rad[1] = 5; rad[2] = 3; return 3;
    case 36: rad[0] = 4; rad[1] = 9; return 2;

    case 16: rad[0] = 16; return 1;
    case 32: rad[0] = 8; rad[1] = 4; return 2;
    case 64: rad[0] = 8; rad[1] = 8; return 2;
    case 128: rad[0] = 16; rad[1] = 8; return 2;
    case 256: rad[0] = 16; rad[1] = 16; return 2;
    case 512: rad[0] = 8; rad[1] = 8; rad[2] = 8; return 3;
    case 1024: rad[0] = 16; rad[1] = 8; rad[2] = 8; return 3;
    case 2048: rad[0] = 16; rad[1] = 16; rad[2] = 8; return 3;
    case 4096: rad[0] = 16; rad[1] = 16; rad[2] = 16; return 3;
  }
  return 0;
}
// Radix sequences of the split complex DCT-I of P + 1 points (fft_fast.h,
// dct1_tile): the P-point FFT (ra) and the P/2-point inverse FFT (rb), the
// same number of passes; 0 if P has no split form.
static inline int dct1_radices(int P, int *ra, int *rb) {
  switch (P) {
    case 128: ra[0] = 16; ra[1] = 8; rb[0] = 8; rb[1] = 8; return 2;
    case 256: ra[0] = 16; ra[1] = 16; rb[0] = 16; rb[1] = 8; return 2;
    case 512: ra[0] = 8; ra[1] = 8; ra[2] = 8; rb[0] = 8; rb[1] = 8; rb[2] = 4; return 3;
    case 1024: ra[0] = 16; ra[1] = 8; ra[2] = 8; rb[0] = 8; rb[1] = 8; rb[2] = 8; return 3;
  }
  return 0;
}
// padded LDS index within a line (both paths): one spare slot every 16
// elements, so that strided accesses of a radix pass spread over the banks
#define P3E_PAD(i) ((i) + ((i) >> 4))

// Entries per k of a radix-R pass's twiddle table in the power-of-two path:
// R - 1 rounded up to odd, so that consecutive k fall on different LDS
// banks.
constexpr int tw_stride(int R) { return (R - 1) | 1; }

}  // namespace p3e
