// Compile-time-length path of the fused reorder + 1D transform pass
// (powers of two, and 3 or 5 times a power of two).
//
// Same tile as the generic path (B lines along the transform axis, adjacent
// along the tile axis), but:
//   - the complex length M and its radix sequence are compile-time, one
//     kernel per M, so every index is a shift (or a multiply-high by a
//     constant) and the values a thread owns stay in registers;
//   - one LDS buffer: each radix pass reads the P3E_FAST_E elements a thread
//     owns into registers (read half), barrier, twiddles and butterflies and
//     writes them to their Stockham positions (write half), barrier;
//   - R2C and C2R of even length N run one complex FFT of length M = N/2 on
//     the packed line z_n = x_{2n} + i x_{2n+1}, with the split/merge of the
//     even and odd halves done in the store (R2C) or load (C2R) phase.
// Shared by the HIP kernel and the host build used by the CPU tests.
#pragma once
#include "fft_core.h"

namespace p3e {

template <int R> struct Log2 {
  static constexpr int v = (R <= 1) ? 0 : 1 + Log2<R / 2>::v;
};
template <> struct Log2<1> {
  static constexpr int v = 0;
};

// Quotient and remainder by a compile-time divisor D (unsigned, so a power
// of two is a shift and a mask, other divisors a multiply-high).
template <int D> P3E_HD P3E_INL int qdiv(int i) { return (int)((uint32_t)i / (uint32_t)D); }
template <int D> P3E_HD P3E_INL int qmod(int i) { return (int)((uint32_t)i % (uint32_t)D); }

// Radix-R Stockham pass at stride Ns of a length-M FFT with DFT sign S,
// split in halves.  Every thread owns NB butterflies of the B lines of the
// tile (the planner makes nth * E = B * M), including lines past the end
// of a partial tile: their LDS slots hold stale values, transformed and
// never stored, so no butterfly needs a bounds check when R divides E.  An
// odd radix R (M = 3 or 5 times a power of two) gives each thread
// ceil(E / R) butterfly slots, the last ones past the tile's B M / R
// butterflies (skipped).
template <typename T, int R, int M, int Ns, int S> struct FPass {
  static constexpr int E = P3E_FAST_E;
  static constexpr int NB = (E + R - 1) / R;  // butterflies per thread
  static constexpr bool EXACT = E % R == 0;
  static constexpr int SLOTS = NB * R;  // registers per thread between the halves
  static constexpr int nbf = M / R;
  cx<T> *buf;
  int lp;
  const cx<T> *tw;
  int nbt;  // butterflies of the tile: B nbf

  // Padded LDS positions as a per-thread base plus a compile-time offset per
  // r (so the compiler folds the offsets into the LDS instructions):
  // P3E_PAD(x + r STEP) - P3E_PAD(x) = pad_off<STEP>(r) when STEP is a
  // multiple of 16 (any x), or when 16 is a multiple of STEP and x is below
  // STEP (x + r STEP then never crosses more 16-element groups than r STEP).
  template <int STEP> static constexpr int pad_off(int r) {
    return STEP % 16 == 0 ? r * (STEP + STEP / 16) : r * STEP + ((r * STEP) >> 4);
  }

  P3E_HD P3E_INL void read(int tid, int nth, cx<T> *v) const {
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (!EXACT && b >= nbt) continue;
      const int j = qmod<nbf>(b);
      // elements j + r nbf, j < nbf
      if constexpr (nbf % 16 == 0 || 16 % nbf == 0) {
        const cx<T> *sl = buf + qdiv<nbf>(b) * lp + (nbf % 16 == 0 ? P3E_PAD(j) : j);
#pragma unroll
        for (int r = 0; r < R; r++) v[m * R + r] = sl[pad_off<nbf>(r)];
      } else {
        const cx<T> *sl = buf + qdiv<nbf>(b) * lp;
#pragma unroll
        for (int r = 0; r < R; r++) v[m * R + r] = sl[P3E_PAD(j + r * nbf)];
      }
    }
  }

  P3E_HD P3E_INL void write(int tid, int nth, cx<T> *v) const {
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (!EXACT && b >= nbt) continue;
      const int j = qmod<nbf>(b);
      const int k = qmod<Ns>(j);
      cx<T> *w = v + m * R;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = mul(w[r], twid(tw, k * tw_stride(R) + r - 1, S));
      }
      bfly<T, R>(w, S);
      cx<T> *dl = buf + qdiv<nbf>(b) * lp;
      // outputs hi + k + r Ns, hi = (j - k) R a multiple of Ns R, k < Ns
      const int hi = (j - k) * R;
      if constexpr ((Ns * R) % 16 == 0) {
        cx<T> *d = dl + (Ns % 16 == 0 ? P3E_PAD(hi + k) : P3E_PAD(hi) + k);
#pragma unroll
        for (int r = 0; r < R; r++) d[pad_off<Ns>(r)] = w[r];
      } else {
#pragma unroll
        for (int r = 0; r < R; r++) dl[P3E_PAD(hi + k + r * Ns)] = w[r];
      }
    }
  }
};

// Run the radix sequence R, Rest... starting at stride Ns; `run` executes
// one pass (read half, barrier, write half, barrier).  tw: the twiddles of
// this pass onward, pass by pass (pass_twiddles on the host): Ns
// tw_stride(R) entries for a radix-R pass at stride Ns.
template <typename T, int M, int Ns, int S, typename Run, int R, int... Rest>
P3E_HD P3E_INL void fft_seq(Run &run, cx<T> *buf, int lp, const cx<T> *tw, int B) {
  FPass<T, R, M, Ns, S> ps{buf, lp, tw, B * (M / R)};
  run(ps);
  if constexpr (sizeof...(Rest) > 0)
    fft_seq<T, M, Ns * R, S, Run, Rest...>(run, buf, lp, tw + Ns * tw_stride(R), B);
}

// Radix sequence of each supported length (the radices of fast_radices in
// pass_args.h; odd radices come last).
template <typename T, int M, int S, typename Run>
P3E_HD P3E_INL void fft_fixed(Run &run, cx<T> *buf, int lp, const cx<T> *tw, int B) {
  if constexpr (M == 16) fft_seq<T, M, 1, S, Run, 16>(run, buf, lp, tw, B);
  else if constexpr (M == 32) fft_seq<T, M, 1, S, Run, 8, 4>(run, buf, lp, tw, B);
  else if constexpr (M == 64) fft_seq<T, M, 1, S, Run, 8, 8>(run, buf, lp, tw, B);
  else if constexpr (M == 128) fft_seq<T, M, 1, S, Run, 16, 8>(run, buf, lp, tw, B);
  else if constexpr (M == 256) fft_seq<T, M, 1, S, Run, 16, 16>(run, buf, lp, tw, B);
  else if constexpr (M == 512) fft_seq<T, M, 1, S, Run, 8, 8, 8>(run, buf, lp, tw, B);
  else if constexpr (M == 1024) fft_seq<T, M, 1, S, Run, 16, 8, 8>(run, buf, lp, tw, B);
  else if constexpr (M == 2048) fft_seq<T, M, 1, S, Run, 16, 16, 8>(run, buf, lp, tw, B);
  else if constexpr (M == 4096) fft_seq<T, M, 1, S, Run, 16, 16, 16>(run, buf, lp, tw, B);
  else if constexpr (M == 48) fft_seq<T, M, 1, S, Run, 16, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 96) fft_seq<T, M, 1, S, Run, 8, 4, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 192) fft_seq<T, M, 1, S, Run, 16, 4, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 384) fft_seq<T, M, 1, S, Run, 16, 8, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 768) fft_seq<T, M, 1, S, Run, 16, 16, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 80) fft_seq<T, M, 1, S, Run, 16, 5>(run, buf, lp, tw, B);
  else if constexpr (M == 160) fft_seq<T, M, 1, S, Run, 8, 4, 5>(run, buf, lp, tw, B);
  else if constexpr (M == 320) fft_seq<T, M, 1, S, Run, 16, 4, 5>(run, buf, lp, tw, B);
  else if constexpr (M == 120) fft_seq<T, M, 1, S, Run, 8, 5, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 60) fft_seq<T, M, 1, S, Run, 4, 5, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 72) fft_seq<T, M, 1, S, Run, 8, 9>(run, buf, lp, tw, B);
  else if constexpr (M == 144) fft_seq<T, M, 1, S, Run, 16, 9>(run, buf, lp, tw, B);
  else if constexpr (M == 112) fft_seq<T, M, 1, S, Run, 16, 7>(run, buf, lp, tw, B);
  else if constexpr (M == 224) fft_seq<T, M, 1, S, Run, 8, 4, 7>(run, buf, lp, tw, B);
  else if constexpr (M == 240) fft_seq<T, M, 1, S, Run, 16, 5, 3>(run, buf, lp, tw, B);
  else if constexpr (M == 36) fft_seq<T, M, 1, S, Run, 4, 9>(run, buf, lp, tw, B);
  else fft_seq<T, M, 1, S, Run, 16, 8, 5>(run, buf, lp, tw, B);  // 640
}

// Two radix sequences run side by side, pass for pass (the split DCT-I):
// sequence A (length MA, sign SA) on threads 0 .. nthA - 1 over the lines
// at bufA, sequence B (length MB, sign SB) on the next nthB threads over the
// lines at bufB; `run.pair` executes one pass of each between the same
// barriers.  Both sequences have the same number of passes.
template <int... R> struct Rads {};

template <typename T, int MA, int MB, int NsA, int NsB, int SA, int SB, typename Run, int RA, int... RAs, int RB,
          int... RBs>
P3E_HD P3E_INL void dual_seq(Run &run, Rads<RA, RAs...>, Rads<RB, RBs...>, cx<T> *bufA, int lpA, const cx<T> *twA,
                             int nthA, cx<T> *bufB, int lpB, const cx<T> *twB, int nthB, int B) {
  static_assert(sizeof...(RAs) == sizeof...(RBs), "sequences of different lengths");
  FPass<T, RA, MA, NsA, SA> pa{bufA, lpA, twA, B * (MA / RA)};
  FPass<T, RB, MB, NsB, SB> pb{bufB, lpB, twB, B * (MB / RB)};
  run.pair(pa, nthA, pb, nthB);
  if constexpr (sizeof...(RAs) > 0)
    dual_seq<T, MA, MB, NsA * RA, NsB * RB, SA, SB, Run>(run, Rads<RAs...>{}, Rads<RBs...>{}, bufA, lpA,
                                                         twA + NsA * tw_stride(RA), nthA, bufB, lpB,
                                                         twB + NsB * tw_stride(RB), nthB, B);
}

// The radices of dct1_radices (pass_args.h) as compile-time sequences.
template <int P> struct Dct1Rads;
template <> struct Dct1Rads<128> {
  using A = Rads<16, 8>;
  using B = Rads<8, 8>;
};
template <> struct Dct1Rads<256> {
  using A = Rads<16, 16>;
  using B = Rads<16, 8>;
};
template <> struct Dct1Rads<512> {
  using A = Rads<8, 8, 8>;
  using B = Rads<8, 8, 4>;
};
template <> struct Dct1Rads<1024> {
  using A = Rads<16, 8, 8>;
  using B = Rads<8, 8, 8>;
};

// Tile coordinates of workgroup `tile`.
P3E_HD P3E_INL void fast_tile_coords(const PassArgs &p, uint32_t tile, uint32_t *f, int *e0, int *nb) {
  *f = fdiv(tile, p.div_ntile);
  *e0 = (int)(tile - *f * p.ntile_e) * p.B;
  *nb = (p.B < p.ne - *e0) ? p.B : (p.ne - *e0);
}

// Load phase: tile -> LDS (natural order; real lines packed two reals per
// complex slot; C2R lines already merged into the packed half-length
// spectrum: each thread loads the pairs (X_k, X_{M-k}) it merges).
// Split in two halves so that a kernel can issue the global loads of its
// next tile (fast_fetch, into registers) before transforming the current
// one, and write them to LDS (fast_put) afterwards.
//
// Every thread owns up to 2E reals or E complex of the tile (tot <= that *
// nth by construction): entry q of thread tid is tile element
// i = tid + q * nth.  Registers hold scalars (R2C: v[q] = real q; complex:
// v[2q], v[2q+1] = re, im of q): arrays of cx<double> are not split into
// registers by the compiler and would live in scratch.
template <typename T> struct TileRegs {
  T v[2 * P3E_FAST_E];
};

// (line j, element n) of tile entry i in load order.
template <int M>
P3E_HD P3E_INL void fast_load_index(const PassArgs &p, int i, int *j, int *n) {
  if (p.load_order == ORD_LINE) {
    if (p.cls == CLS_R2C) {
      *j = qdiv<2 * M>(i);
      *n = qmod<2 * M>(i);
    } else if (p.cls == CLS_C2R || p.cls == CLS_R2R) {
      *j = (int)fdiv((uint32_t)i, p.div_nin);
      *n = i - *j * p.nin;
    } else {
      *j = qdiv<M>(i);
      *n = qmod<M>(i);
    }
  } else {
    *j = i & (p.B - 1);
    *n = i >> p.logB;
  }
}

// (line j, pair k) of C2R pair entry i: the pair is (X_k, X_{M-k}),
// k = 0..M/2, along the line in load order.
template <int M>
P3E_HD P3E_INL void fast_pair_index(const PassArgs &p, int i, int *j, int *k) {
  if (p.load_order == ORD_LINE) {
    *j = (int)fdiv((uint32_t)i, p.div_pairs);
    *k = i - *j * (M / 2 + 1);
  } else {
    *j = i & (p.B - 1);
    *k = i >> p.logB;
  }
}

// C2R: packed spectrum entries Z_k, Z_{M-k} of one line from X_k, X_{M-k}:
//   Z_k = (X_k + conj X_{M-k}) + i e^{+2 pi i k/N} (X_k - conj X_{M-k});
// the imaginary parts of X_0 and X_M do not contribute (FFTW c2r).
template <typename T, int M>
P3E_HD P3E_INL void c2r_pair_put(const cx<T> *wN, cx<T> *line, int k, cx<T> xk, cx<T> xm) {
  if (k == 0) {
    line[P3E_PAD(0)] = mk(xk.x + xm.x, xk.x - xm.x);
    return;
  }
  line[P3E_PAD(k)] = add(add(xk, conj(xm)), rot(mul(conj(wN[k]), sub(xk, conj(xm))), 1));
  if (2 * k != M) line[P3E_PAD(M - k)] = add(add(xm, conj(xk)), rot(mul(conj(wN[M - k]), sub(xm, conj(xk))), 1));
}

// Two reals in one load (8 or 16 bytes, aligned to its size).
template <typename T> struct alignas(2 * sizeof(T)) RealPair {
  T a, b;
};
// Two complex values in one load (16 or 32 bytes, aligned to 16).
template <typename T> struct alignas(16) CxPair {
  T ar, ai, br, bi;
};

// Global accesses of a tile address the array from the tile's first
// element (a pointer that is the same for the whole workgroup) with 32-bit
// byte offsets: the planner admits a pass to this path only when every
// offset within a tile, on both sides, stays below 2^31 bytes.
template <typename V> P3E_HD P3E_INL V ldg(const char *base, uint32_t off) {
  return *reinterpret_cast<const V *>(base + off);
}
template <typename V> P3E_HD P3E_INL void stg(char *base, uint32_t off, const V &v) {
  *reinterpret_cast<V *>(base + off) = v;
}
// Non-temporal load, for the paired line-order loads (whole lines read
// once): on a 512^3 FP32 round trip 1.306 -> 1.251 ms (the forward C2C z
// pass 0.232 -> 0.194 ms).  The same hint on the other loads (unpaired line
// order, tile order) was slower or no better: 1.266 / 1.254 ms, and the
// 256-item batch 1.82 -> 2.31 ms with tile-order loads included.
template <typename V> P3E_HD P3E_INL V ldg_nt(const char *base, uint32_t off) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(V) == 8) {
    const unsigned long long u = __builtin_nontemporal_load(reinterpret_cast<const unsigned long long *>(base + off));
    return *reinterpret_cast<const V *>(&u);
  } else if constexpr (sizeof(V) == 16) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 u = __builtin_nontemporal_load(reinterpret_cast<const u4 *>(base + off));
    return *reinterpret_cast<const V *>(&u);
  } else {
    return *reinterpret_cast<const V *>(base + off);
  }
#else
  return *reinterpret_cast<const V *>(base + off);
#endif
}
// Non-temporal store (streaming: the line is not kept in the caches), for
// line-order stores, which write whole lines; tile-order stores write
// partial cache lines that the next tile along e completes, and keep stg.
template <typename V> P3E_HD P3E_INL void stg_nt(char *base, uint32_t off, const V &v) {
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (sizeof(V) == 4) {
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned *>(&v), reinterpret_cast<unsigned *>(base + off));
  } else if constexpr (sizeof(V) == 8) {
    __builtin_nontemporal_store(*reinterpret_cast<const unsigned long long *>(&v),
                                reinterpret_cast<unsigned long long *>(base + off));
  } else if constexpr (sizeof(V) == 16) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    __builtin_nontemporal_store(*reinterpret_cast<const u4 *>(&v), reinterpret_cast<u4 *>(base + off));
  } else {
    *reinterpret_cast<V *>(base + off) = v;
  }
#else
  *reinterpret_cast<V *>(base + off) = v;
#endif
}

// FULL: the tile has all B lines (nb == B); every entry a thread owns is
// then inside the tile and the accesses need no bounds checks.
template <typename T, int M, bool FULL>
P3E_HD P3E_INL void fast_fetch(const PassArgs &p, uint32_t f, int e0, int nb, TileRegs<T> &R, int tid, int nth) {
  constexpr int E = P3E_FAST_E;
  constexpr uint32_t RS = sizeof(T), CS = 2 * sizeof(T);
  const uint32_t esz = p.cls == CLS_R2C ? RS : CS;
  const char *tin = reinterpret_cast<const char *>(p.in) + ((long long)e0 * p.is_e + (long long)f * p.is_f) * esz;
  const int isd = (int)p.is_d, ise = (int)p.is_e;
  const int tot = (p.load_order == ORD_LINE ? nb : p.B) * p.nin;
  // out-of-tile entries load the tile's first element (always valid) and
  // are not written: unconditional loads keep the values in registers
  if (p.cls == CLS_R2R) {
    // DCT-I (REDFT00) of complex lines of N = M/2 + 1 points: B N = B M/2 + B
    // entries per tile, covered by E/2 + 1 entries per thread (nth >= B);
    // paired (vec2: tile order, e stride-1, whole tiles), element n of lines
    // j = 2 jp and j + 1 per load, B N / 2 pairs in E/4 + 1 entries
    if (p.vec2) {
      const int lb = p.logB - 1, ptot = (p.B >> 1) * p.nin;
#pragma unroll
      for (int q = 0; q < E / 4 + 1; q++) {
        const int i = tid + q * nth;
        const int j = 2 * (i & ((1 << lb) - 1)), n = i >> lb;
        const CxPair<T> x = ldg<CxPair<T>>(tin, i < ptot ? (uint32_t)(n * isd + j) * CS : 0u);
        R.v[4 * q] = x.ar;
        R.v[4 * q + 1] = x.ai;
        R.v[4 * q + 2] = x.br;
        R.v[4 * q + 3] = x.bi;
      }
    } else {
#pragma unroll
      for (int q = 0; q < E / 2 + 1; q++) {
        const int i = tid + q * nth;
        int j, n;
        fast_load_index<M>(p, i, &j, &n);
        const bool ok = i < tot && (FULL || j < nb);
        const cx<T> x = ldg<cx<T>>(tin, ok ? (uint32_t)(n * isd + j * ise) * CS : 0u);
        R.v[2 * q] = x.x;
        R.v[2 * q + 1] = x.y;
      }
    }
  } else if (p.vec2 && p.cls == CLS_C2C && p.load_order == ORD_LINE) {
    // C2C, line order (d stride-1): elements n = 2 np and n + 1 of line j
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int j = qdiv<M / 2>(i), n = 2 * qmod<M / 2>(i);
      const bool ok = FULL || j < nb;
      const CxPair<T> x = ldg_nt<CxPair<T>>(tin, ok ? (uint32_t)(n + j * ise) * CS : 0u);
      R.v[4 * q] = x.ar;
      R.v[4 * q + 1] = x.ai;
      R.v[4 * q + 2] = x.br;
      R.v[4 * q + 3] = x.bi;
    }
  } else if (p.vec2 && p.cls == CLS_C2C) {
    // C2C, tile order (e stride-1, whole tiles): element n of lines
    // j = 2 jp and j + 1
    const int lb = p.logB - 1;
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int j = 2 * (i & ((1 << lb) - 1)), n = i >> lb;
      const CxPair<T> x = ldg<CxPair<T>>(tin, (uint32_t)(n * isd + j) * CS);
      R.v[4 * q] = x.ar;
      R.v[4 * q + 1] = x.ai;
      R.v[4 * q + 2] = x.br;
      R.v[4 * q + 3] = x.bi;
    }
  } else if (p.vec2 && p.load_order == ORD_LINE) {
    // R2C, line order (d stride-1): entry q is the real pair (x_2m, x_2m+1)
    // of line j, the packed complex element m, read by one load
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = qdiv<M>(i), m = qmod<M>(i);
      const bool ok = FULL || j < nb;
      const RealPair<T> x = ldg_nt<RealPair<T>>(tin, ok ? (uint32_t)(2 * m + j * ise) * RS : 0u);
      R.v[2 * q] = x.a;
      R.v[2 * q + 1] = x.b;
    }
  } else if (p.vec2) {
    // R2C, tile order, e stride-1 and whole tiles (the planner's condition):
    // entry q is element n of lines j = 2 jp and j + 1, adjacent in memory;
    // the launcher checked that base and strides keep the pair aligned
    const int lb = p.logB - 1;
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int jp = i & ((1 << lb) - 1), n = i >> lb;
      const RealPair<T> x = ldg<RealPair<T>>(tin, (uint32_t)(n * isd + 2 * jp) * RS);
      R.v[2 * q] = x.a;
      R.v[2 * q + 1] = x.b;
    }
  } else if (p.cls == CLS_R2C) {
#pragma unroll
    for (int q = 0; q < 2 * E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      const bool ok = FULL ? i < tot : (i < tot && j < nb);
      R.v[q] = ldg<T>(tin, ok ? (uint32_t)(n * isd + j * ise) * RS : 0u);
    }
  } else if (p.cls == CLS_C2R) {
    // pairs (X_k, X_{M-k}) of one line in one thread, merged in fast_put;
    // the first (E/2) nth pairs of (M/2 + 1) B are all inside the tile's
    // pair range
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      int j, k;
      fast_pair_index<M>(p, i, &j, &k);
      const bool ok = FULL || j < nb;
      const int o = j * ise;
      const cx<T> a = ldg<cx<T>>(tin, ok ? (uint32_t)(o + k * isd) * CS : 0u);
      const cx<T> b = ldg<cx<T>>(tin, ok ? (uint32_t)(o + (M - k) * isd) * CS : 0u);
      R.v[4 * q] = a.x;
      R.v[4 * q + 1] = a.y;
      R.v[4 * q + 2] = b.x;
      R.v[4 * q + 3] = b.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      const bool ok = FULL ? i < tot : (i < tot && j < nb);
      const cx<T> x = ldg<cx<T>>(tin, ok ? (uint32_t)(n * isd + j * ise) * CS : 0u);
      R.v[2 * q] = x.x;
      R.v[2 * q + 1] = x.y;
    }
  }
}

template <typename T, int M, bool FULL>
P3E_HD P3E_INL void fast_put(const PassArgs &p, uint32_t f, int e0, int nb, const TileRegs<T> &R, cx<T> *buf,
                             int tid, int nth, const cx<T> *wN) {
  constexpr int E = P3E_FAST_E;
  const int tot = (p.load_order == ORD_LINE ? nb : p.B) * p.nin;
  const int lp = p.lp;
  if (p.cls == CLS_R2R) {
    // even extension: element n at positions n and M - n (0 < n < M/2)
    if (p.vec2) {
      const int lb = p.logB - 1, ptot = (p.B >> 1) * p.nin;
#pragma unroll
      for (int q = 0; q < E / 4 + 1; q++) {
        const int i = tid + q * nth;
        if (i >= ptot) continue;
        const int j = 2 * (i & ((1 << lb) - 1)), n = i >> lb;
        const cx<T> a = mk(R.v[4 * q], R.v[4 * q + 1]), b = mk(R.v[4 * q + 2], R.v[4 * q + 3]);
        cx<T> *d = buf + j * lp;
        d[P3E_PAD(n)] = a;
        d[lp + P3E_PAD(n)] = b;
        if (n >= 1 && n < M / 2) {
          d[P3E_PAD(M - n)] = a;
          d[lp + P3E_PAD(M - n)] = b;
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < E / 2 + 1; q++) {
        const int i = tid + q * nth;
        int j, n;
        fast_load_index<M>(p, i, &j, &n);
        if (i < tot && (FULL || j < nb)) {
          const cx<T> x = mk(R.v[2 * q], R.v[2 * q + 1]);
          buf[j * lp + P3E_PAD(n)] = x;
          if (n >= 1 && n < M / 2) buf[j * lp + P3E_PAD(M - n)] = x;
        }
      }
    }
  } else if (p.vec2 && p.cls == CLS_C2C && p.load_order == ORD_LINE) {
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int j = qdiv<M / 2>(i), n = 2 * qmod<M / 2>(i);
      if (FULL || j < nb) {
        // n is even: n and n + 1 share a padding group
        cx<T> *d = buf + j * lp + P3E_PAD(n);
        d[0] = mk(R.v[4 * q], R.v[4 * q + 1]);
        d[1] = mk(R.v[4 * q + 2], R.v[4 * q + 3]);
      }
    }
  } else if (p.vec2 && p.cls == CLS_C2C) {
    const int lb = p.logB - 1;
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int j = 2 * (i & ((1 << lb) - 1)), n = i >> lb;
      cx<T> *d = buf + j * lp + P3E_PAD(n);
      d[0] = mk(R.v[4 * q], R.v[4 * q + 1]);
      d[lp] = mk(R.v[4 * q + 2], R.v[4 * q + 3]);
    }
  } else if (p.vec2 && p.load_order == ORD_LINE) {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = qdiv<M>(i), m = qmod<M>(i);
      if (FULL || j < nb) buf[j * lp + P3E_PAD(m)] = mk(R.v[2 * q], R.v[2 * q + 1]);
    }
  } else if (p.vec2) {
    T *bufr = reinterpret_cast<T *>(buf);
    const int lb = p.logB - 1;
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = 2 * (i & ((1 << lb) - 1)), n = i >> lb;
      const int o = 2 * P3E_PAD(n >> 1) + (n & 1);
      bufr[2 * j * lp + o] = R.v[2 * q];
      bufr[2 * (j + 1) * lp + o] = R.v[2 * q + 1];
    }
  } else if (p.cls == CLS_R2C) {
    T *bufr = reinterpret_cast<T *>(buf);
#pragma unroll
    for (int q = 0; q < 2 * E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      if (i < tot && (FULL || j < nb)) bufr[2 * (j * lp + P3E_PAD(n >> 1)) + (n & 1)] = R.v[q];
    }
  } else if (p.cls == CLS_C2R) {
    const int ptot = (p.load_order == ORD_LINE ? nb : p.B) * (M / 2 + 1);
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      int j, k;
      fast_pair_index<M>(p, i, &j, &k);
      if (FULL || j < nb)
        c2r_pair_put<T, M>(wN, buf + j * lp, k, mk(R.v[4 * q], R.v[4 * q + 1]), mk(R.v[4 * q + 2], R.v[4 * q + 3]));
    }
    // (M/2 + 1) B pairs are B more than (E/2) nth: loaded here directly
    const char *tin = reinterpret_cast<const char *>(p.in) +
                      ((long long)e0 * p.is_e + (long long)f * p.is_f) * (long long)(2 * sizeof(T));
    const int isd = (int)p.is_d, ise = (int)p.is_e;
    for (int i = tid + (E / 2) * nth; i < ptot; i += nth) {
      int j, k;
      fast_pair_index<M>(p, i, &j, &k);
      if (j >= nb) continue;
      const int o = j * ise;
      c2r_pair_put<T, M>(wN, buf + j * lp, k, ldg<cx<T>>(tin, (uint32_t)(o + k * isd) * (uint32_t)(2 * sizeof(T))),
                         ldg<cx<T>>(tin, (uint32_t)(o + (M - k) * isd) * (uint32_t)(2 * sizeof(T))));
    }
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      if (i < tot && (FULL || j < nb)) buf[j * lp + P3E_PAD(n)] = mk(R.v[2 * q], R.v[2 * q + 1]);
    }
  }
}

template <typename T, int M, bool FULL>
P3E_HD P3E_INL void fast_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *buf, int tid, int nth,
                              const cx<T> *wN) {
  TileRegs<T> R;
  fast_fetch<T, M, FULL>(p, f, e0, nb, R, tid, nth);
  fast_put<T, M, FULL>(p, f, e0, nb, R, buf, tid, nth, wN);
}

// Output rule of element k of line j (line j + e0 of the array): the
// derivative i k, the Poisson operator, the scale.
template <typename T>
P3E_HD P3E_INL cx<T> out_rule(const PassArgs &p, cx<T> y, int k, int je, uint32_t f) {
  if (p.deriv) {
    T kk;
    if (k < p.dmid) kk = (T)k;
    else if (k == p.dmid) kk = (T)0;
    else kk = (T)(k - p.dg);
    y = mk(-kk * y.y, kk * y.x);
  }
  if (p.spec_op) y = scl(y, poisson_factor<T>(p, k, je, (int)f));
  if (p.has_scale) y = scl(y, (T)p.scale);
  return y;
}

// R2C output k of a line from the packed spectrum Z:
//   X_k = (Z_k + conj Z_{M-k})/2 - i w_N^k (Z_k - conj Z_{M-k})/2
template <typename T, int M>
P3E_HD P3E_INL cx<T> r2c_split(const cx<T> *l, const cx<T> *wN, int k) {
  const cx<T> zk = l[P3E_PAD(qmod<M>(k))];
  const cx<T> zm = l[P3E_PAD(qmod<M>(M - k))];
  const cx<T> e = add(zk, conj(zm)), o = sub(zk, conj(zm));
  return scl(add(e, rot(mul(wN[k], o), -1)), (T)0.5);
}

// Tile-order paired store of entry i (e stride-1, whole tiles): element k
// of lines j = 2 jp and j + 1, adjacent in memory, by one store.
template <typename T, int M>
P3E_HD P3E_INL void store_pair_tile(const PassArgs &p, uint32_t f, int e0, const cx<T> *buf, const cx<T> *wN,
                                    char *tout, int i) {
  const int lb = p.logB - 1, lp = p.lp;
  const int j = 2 * (i & ((1 << lb) - 1)), k = i >> lb;
  cx<T> y0, y1;
  if (p.cls == CLS_R2C) {
    y0 = r2c_split<T, M>(buf + j * lp, wN, k);
    y1 = r2c_split<T, M>(buf + (j + 1) * lp, wN, k);
  } else {
    y0 = buf[j * lp + P3E_PAD(k)];
    y1 = buf[(j + 1) * lp + P3E_PAD(k)];
  }
  y0 = out_rule(p, y0, k, e0 + j, f);
  y1 = out_rule(p, y1, k, e0 + j + 1, f);
  CxPair<T> v;
  v.ar = y0.x;
  v.ai = y0.y;
  v.br = y1.x;
  v.bi = y1.y;
  stg(tout, (uint32_t)(k * (int)p.os_d + j) * (uint32_t)(2 * sizeof(T)), v);
}

// R2C outputs k and M - k of one line from the packed pair (Z_k, Z_{M-k}):
// with e = Z_k + conj Z_{M-k} and t = w_N^k (Z_k - conj Z_{M-k}),
// X_k = (e - i t)/2 and X_{M-k} = conj(e + i t)/2 (w_N^{M-k} = -conj w_N^k).
template <typename T, int M>
P3E_HD P3E_INL void r2c_split2(const cx<T> *l, const cx<T> *wN, int k, cx<T> *xk, cx<T> *xm) {
  const cx<T> zk = l[P3E_PAD(k)], zm = l[P3E_PAD(k == 0 ? 0 : M - k)];
  const cx<T> e = add(zk, conj(zm)), t = mul(wN[k], sub(zk, conj(zm)));
  *xk = scl(add(e, rot(t, -1)), (T)0.5);
  *xm = scl(conj(add(e, rot(t, 1))), (T)0.5);
}

// Tile-order paired R2C store of entry i (e stride-1, whole tiles): outputs
// k and M - k (k = 0..M/2) of lines j = 2 jp and j + 1, adjacent in memory,
// by one store each.
template <typename T, int M>
P3E_HD P3E_INL void store_quad_tile(const PassArgs &p, uint32_t f, int e0, const cx<T> *buf, const cx<T> *wN,
                                    char *tout, int i) {
  const int lb = p.logB - 1, lp = p.lp;
  const int j = 2 * (i & ((1 << lb) - 1)), k = i >> lb;
  cx<T> a0, a1, b0, b1;
  r2c_split2<T, M>(buf + j * lp, wN, k, &a0, &b0);
  r2c_split2<T, M>(buf + (j + 1) * lp, wN, k, &a1, &b1);
  a0 = out_rule(p, a0, k, e0 + j, f);
  a1 = out_rule(p, a1, k, e0 + j + 1, f);
  CxPair<T> v;
  v.ar = a0.x;
  v.ai = a0.y;
  v.br = a1.x;
  v.bi = a1.y;
  const int osd = (int)p.os_d;
  stg(tout, (uint32_t)(k * osd + j) * (uint32_t)(2 * sizeof(T)), v);
  if (2 * k != M) {
    b0 = out_rule(p, b0, M - k, e0 + j, f);
    b1 = out_rule(p, b1, M - k, e0 + j + 1, f);
    v.ar = b0.x;
    v.ai = b0.y;
    v.br = b1.x;
    v.bi = b1.y;
    stg(tout, (uint32_t)((M - k) * osd + j) * (uint32_t)(2 * sizeof(T)), v);
  }
}

// Store phase: output rule and LDS -> tile.
template <typename T, int M, bool FULL>
P3E_HD P3E_INL void fast_store(const PassArgs &p, uint32_t f, int e0, int nb, const cx<T> *buf, int tid, int nth,
                               const cx<T> *wN) {
  constexpr int E = P3E_FAST_E;
  constexpr uint32_t RS = sizeof(T), CS = 2 * sizeof(T);
  const uint32_t esz = p.cls == CLS_C2R ? RS : CS;
  char *tout = reinterpret_cast<char *>(p.out) + ((long long)e0 * p.os_e + (long long)f * p.os_f) * esz;
  const int osd = (int)p.os_d, ose = (int)p.os_e;
  const int lp = p.lp;
  const int tot = (p.store_order == ORD_LINE ? nb : p.B) * p.nout;
  if (p.vec2s && p.cls == CLS_C2R && p.store_order == ORD_LINE) {
    // C2R, line order (d stride-1): the packed complex element m of line j
    // is the real pair (x_2m, x_2m+1), written by one store
#pragma unroll 4
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = qdiv<M>(i), m = qmod<M>(i);
      if (!FULL && j >= nb) continue;
      const cx<T> z = buf[j * lp + P3E_PAD(m)];
      RealPair<T> v;
      v.a = p.has_scale ? z.x * (T)p.scale : z.x;
      v.b = p.has_scale ? z.y * (T)p.scale : z.y;
      stg_nt(tout, (uint32_t)(2 * m + j * ose) * RS, v);
    }
  } else if (p.vec2s && p.cls == CLS_C2R) {
    // C2R, tile order, e stride-1 and whole tiles (the planner's condition):
    // entry q is element k of lines j = 2 jp and j + 1, adjacent in memory,
    // written by one store; the launcher checked the pair alignment
    const T *bufr = reinterpret_cast<const T *>(buf);
    const int lb = p.logB - 1;
    T y[2 * E];
    uint32_t off[E];
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = 2 * (i & ((1 << lb) - 1)), k = i >> lb;
      const int o = 2 * P3E_PAD(k >> 1) + (k & 1);
      y[2 * q] = bufr[2 * j * lp + o];
      y[2 * q + 1] = bufr[2 * (j + 1) * lp + o];
      off[q] = (uint32_t)(k * osd + j) * RS;
    }
#pragma unroll
    for (int q = 0; q < E; q++) {
      RealPair<T> v;
      v.a = p.has_scale ? y[2 * q] * (T)p.scale : y[2 * q];
      v.b = p.has_scale ? y[2 * q + 1] * (T)p.scale : y[2 * q + 1];
      stg(tout, off[q], v);
    }
  } else if (p.cls == CLS_C2R) {
    const T *bufr = reinterpret_cast<const T *>(buf);
    // 2E reals per thread: LDS reads first, then the global stores
    T y[2 * E];
    int off[2 * E];
#pragma unroll
    for (int q = 0; q < 2 * E; q++) {
      const int i = tid + q * nth;
      int k, j;
      if (p.store_order == ORD_LINE) {
        j = qdiv<2 * M>(i);
        k = qmod<2 * M>(i);
      } else {
        j = i & (p.B - 1);
        k = i >> p.logB;
      }
      const bool ok = i < tot && (FULL || j < nb);
      y[q] = bufr[ok ? 2 * (j * lp + P3E_PAD(k >> 1)) + (k & 1) : 0];
      off[q] = ok ? k * osd + j * ose : -1;
    }
#pragma unroll
    for (int q = 0; q < 2 * E; q++)
      if (off[q] >= 0) stg(tout, (uint32_t)off[q] * RS, p.has_scale ? y[q] * (T)p.scale : y[q]);
  } else if (p.cls == CLS_R2C && p.store_order == ORD_LINE) {
    // R2C, line order (d stride-1): outputs k and M - k of line j from one
    // pair (Z_k, Z_{M-k}) of the packed spectrum: with e = Z_k + conj Z_{M-k}
    // and t = w_N^k (Z_k - conj Z_{M-k}), X_k = (e - i t)/2 and
    // X_{M-k} = conj(e + i t)/2 (as w_N^{M-k} = -conj w_N^k)
    const int ptot = nb * (M / 2 + 1);
#pragma unroll 2
    for (int i = tid; i < ptot; i += nth) {
      const int j = qdiv<M / 2 + 1>(i), k = qmod<M / 2 + 1>(i);
      cx<T> xk, xm;
      r2c_split2<T, M>(buf + j * lp, wN, k, &xk, &xm);
      const uint32_t lo = (uint32_t)(j * ose);
      stg_nt(tout, (lo + (uint32_t)k) * CS, out_rule(p, xk, k, e0 + j, f));
      if (2 * k != M) stg_nt(tout, (lo + (uint32_t)(M - k)) * CS, out_rule(p, xm, M - k, e0 + j, f));
    }
  } else if (p.vec2s && p.store_order == ORD_LINE) {
    // C2C, line order (d stride-1): elements k = 2 kp and k + 1 of line j
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int j = qdiv<M / 2>(i), k = 2 * qmod<M / 2>(i);
      if (!FULL && j >= nb) continue;
      const cx<T> *l = buf + j * lp + P3E_PAD(k);
      const cx<T> y0 = out_rule(p, l[0], k, e0 + j, f), y1 = out_rule(p, l[1], k + 1, e0 + j, f);
      CxPair<T> v;
      v.ar = y0.x;
      v.ai = y0.y;
      v.br = y1.x;
      v.bi = y1.y;
      stg_nt(tout, (uint32_t)(k + j * ose) * CS, v);
    }
  } else if (p.vec2s) {
    // C2C and R2C, tile order (e stride-1, whole tiles): element k of lines
    // j = 2 jp and j + 1, one store.  (E/2) nth entries are exactly the C2C
    // tile (nout = M); an R2C tile has nout = M + 1, B/2 entries more.
    const int ptot = (p.B >> 1) * p.nout;
    if (p.cls == CLS_R2C) {
      // outputs (k, M - k) of two lines per entry: (B/2)(M/2 + 1) entries
      const int qtot = (p.B >> 1) * (M / 2 + 1);
      for (int i = tid; i < qtot; i += nth) store_quad_tile<T, M>(p, f, e0, buf, wN, tout, i);
    } else if (p.cls == CLS_R2R) {
      // DCT-I: nout = M/2 + 1, B nout / 2 pairs in E/4 + 1 entries
#pragma unroll
      for (int q = 0; q < E / 4 + 1; q++) {
        const int i = tid + q * nth;
        if (i < ptot) store_pair_tile<T, M>(p, f, e0, buf, wN, tout, i);
      }
    } else {
#pragma unroll
      for (int q = 0; q < E / 2; q++) store_pair_tile<T, M>(p, f, e0, buf, wN, tout, tid + q * nth);
      for (int i = tid + (E / 2) * nth; i < ptot; i += nth) store_pair_tile<T, M>(p, f, e0, buf, wN, tout, i);
    }
  } else {
#pragma unroll 4
    for (int i = tid; i < tot; i += nth) {
      int k, j;
      if (p.store_order == ORD_LINE) {
        if (p.cls == CLS_R2C || p.cls == CLS_R2R) {
          j = (int)fdiv((uint32_t)i, p.div_nout);
          k = i - j * p.nout;
        } else {
          j = qdiv<M>(i);
          k = qmod<M>(i);
        }
        if (!FULL && j >= nb) continue;
      } else {
        j = i & (p.B - 1);
        k = i >> p.logB;
        if (!FULL && j >= nb) continue;
      }
      const cx<T> *l = buf + j * lp;
      const cx<T> y = p.cls == CLS_R2C ? r2c_split<T, M>(l, wN, k) : l[P3E_PAD(k)];
      stg(tout, (uint32_t)(k * osd + j * ose) * CS, out_rule(p, y, k, e0 + j, f));
    }
  }
}

// ------------------------------------------------------------ split DCT-I
// Complex DCT-I (REDFT00) of P + 1 points, P = 2L, as two transforms of half
// the size of the 2P-point FFT of the even extension:
//   even outputs  Y_2m = DFT_P(ext a)_m, m = 0..L, where a_n = x_n + x_{P-n}
//     (n = 0..L) and ext a is its even extension (a DCT-I of L + 1 points);
//   odd outputs   Y_2m+1 = REDFT01_L(c)_m, m = 0..L-1, c_n = x_n - x_{P-n},
//     computed as v = IDFT_L(V) with V_n = (c_n - i c_{L-n}) e^{i pi n/2L}
//     (c_L = 0): Y_4j+1 = v_j, Y_4j+3 = v_{L-1-j} (Makhoul's DCT-III by an
//     L-point FFT; every step is complex-linear, so it holds for complex x).
// The load phase reads the quad x_q, x_{L-q}, x_{L+q}, x_{P-q} (q = 0..L/2)
// in one thread and writes a and V at both ends: V_{L-q} = (c_q + i c_{L-q})
// conj(e^{i pi q/2L}), so the pre-twiddle table has L/2 + 1 entries.
// Region A (B lines of pitch lp) holds ext a, region B (B lines of pitch
// lp2) holds V; tab = [twiddles of A (ntw_a) | of B (ntw_b) | pre-twiddles].
// a and V of line j from the quad (x_q, x_{L-q}, x_{L+q}, x_{P-q})
template <typename T, int P>
P3E_HD P3E_INL void dct1_put_quad(const PassArgs &p, cx<T> *bufA, cx<T> *bufB, const cx<T> *pre, int j, int q,
                                  cx<T> x0, cx<T> x1, cx<T> x2, cx<T> x3) {
  constexpr int L = P / 2;
  const cx<T> aq = add(x0, x3), al = add(x1, x2);
  const cx<T> cq = sub(x0, x3), cl = sub(x1, x2);
  cx<T> *A = bufA + j * p.lp;
  A[P3E_PAD(q)] = aq;
  A[P3E_PAD(L - q)] = al;
  A[P3E_PAD(L + q)] = al;
  if (q > 0) A[P3E_PAD(P - q)] = aq;
  const cx<T> t = pre[q];
  cx<T> *V = bufB + j * p.lp2;
  V[P3E_PAD(q)] = mul(add(cq, rot(cl, -1)), t);
  if (q > 0) V[P3E_PAD(L - q)] = mul(add(cq, rot(cl, 1)), conj(t));
}

template <typename T, int P, bool FULL>
P3E_HD P3E_INL void dct1_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *bufA, cx<T> *bufB,
                              const cx<T> *pre, int tid, int nth) {
  constexpr int L = P / 2, Q = L / 2 + 1;
  // quads per thread: B Q quads over (P + L) B / E threads
  constexpr int QN = (Q * P3E_FAST_E + P + L - 1) / (P + L);
  constexpr uint32_t CS = 2 * sizeof(T);
  const char *tin = reinterpret_cast<const char *>(p.in) + ((long long)e0 * p.is_e + (long long)f * p.is_f) * CS;
  const int isd = (int)p.is_d, ise = (int)p.is_e;
  if (p.vec2) {
    // tile order along a stride-1 e, whole tiles: the quads of lines
    // j = 2 jp and j + 1 in 16-byte loads, B Q / 2 entries
    constexpr int QP = (Q * P3E_FAST_E + 2 * (P + L) - 1) / (2 * (P + L));
    const int lb = p.logB - 1, ptot = (p.B >> 1) * Q;
    CxPair<T> y[QP][4];
    int jq2[QP][2];
#pragma unroll
    for (int u = 0; u < QP; u++) {
      const int i = tid + u * nth;
      const bool ok = i < ptot;
      const int j = 2 * (i & ((1 << lb) - 1)), q = i >> lb;
      jq2[u][0] = ok ? j : -1;
      jq2[u][1] = q;
      const uint32_t o = ok ? (uint32_t)j : 0u, qq = ok ? (uint32_t)q : 0u;
      y[u][0] = ldg<CxPair<T>>(tin, (o + qq * isd) * CS);
      y[u][1] = ldg<CxPair<T>>(tin, (o + (L - qq) * isd) * CS);
      y[u][2] = ldg<CxPair<T>>(tin, (o + (L + qq) * isd) * CS);
      y[u][3] = ldg<CxPair<T>>(tin, (o + (P - qq) * isd) * CS);
    }
#pragma unroll
    for (int u = 0; u < QP; u++) {
      const int j = jq2[u][0], q = jq2[u][1];
      if (j < 0) continue;
      dct1_put_quad<T, P>(p, bufA, bufB, pre, j, q, mk(y[u][0].ar, y[u][0].ai), mk(y[u][1].ar, y[u][1].ai),
                          mk(y[u][2].ar, y[u][2].ai), mk(y[u][3].ar, y[u][3].ai));
      dct1_put_quad<T, P>(p, bufA, bufB, pre, j + 1, q, mk(y[u][0].br, y[u][0].bi), mk(y[u][1].br, y[u][1].bi),
                          mk(y[u][2].br, y[u][2].bi), mk(y[u][3].br, y[u][3].bi));
    }
    return;
  }
  const int tot = (p.load_order == ORD_LINE ? nb : p.B) * Q;
  cx<T> x[QN][4];
  int jq[QN][2];
#pragma unroll
  for (int u = 0; u < QN; u++) {
    const int i = tid + u * nth;
    int j, q;
    if (p.load_order == ORD_LINE) {
      j = qdiv<Q>(i);
      q = qmod<Q>(i);
    } else {
      j = i & (p.B - 1);
      q = i >> p.logB;
    }
    const bool ok = i < tot && (FULL || j < nb);
    jq[u][0] = ok ? j : -1;
    jq[u][1] = q;
    const uint32_t o = ok ? (uint32_t)(j * ise) : 0u;
    const uint32_t qq = ok ? (uint32_t)q : 0u;
    x[u][0] = ldg<cx<T>>(tin, (o + qq * isd) * CS);
    x[u][1] = ldg<cx<T>>(tin, (o + (L - qq) * isd) * CS);
    x[u][2] = ldg<cx<T>>(tin, (o + (L + qq) * isd) * CS);
    x[u][3] = ldg<cx<T>>(tin, (o + (P - qq) * isd) * CS);
  }
#pragma unroll
  for (int u = 0; u < QN; u++) {
    const int j = jq[u][0], q = jq[u][1];
    if (j < 0) continue;
    dct1_put_quad<T, P>(p, bufA, bufB, pre, j, q, x[u][0], x[u][1], x[u][2], x[u][3]);
  }
}

template <typename T, int P, bool FULL>
P3E_HD P3E_INL void dct1_store(const PassArgs &p, uint32_t f, int e0, int nb, const cx<T> *bufA, const cx<T> *bufB,
                               int tid, int nth) {
  constexpr int L = P / 2;
  constexpr uint32_t CS = 2 * sizeof(T);
  char *tout = reinterpret_cast<char *>(p.out) + ((long long)e0 * p.os_e + (long long)f * p.os_f) * CS;
  const int osd = (int)p.os_d, ose = (int)p.os_e;
  if (p.vec2s) {
    // tile order along a stride-1 e, whole tiles: output k of lines
    // j = 2 jp and j + 1 in one 16-byte store
    const int lb = p.logB - 1, ptot = (p.B >> 1) * (P + 1);
#pragma unroll 4
    for (int i = tid; i < ptot; i += nth) {
      const int j = 2 * (i & ((1 << lb) - 1)), k = i >> lb;
      const int vi = (k & 2) ? L - 1 - (k >> 2) : (k >> 2);
      const cx<T> y0 = (k & 1) ? bufB[j * p.lp2 + P3E_PAD(vi)] : bufA[j * p.lp + P3E_PAD(k >> 1)];
      const cx<T> y1 = (k & 1) ? bufB[(j + 1) * p.lp2 + P3E_PAD(vi)] : bufA[(j + 1) * p.lp + P3E_PAD(k >> 1)];
      const cx<T> z0 = out_rule(p, y0, k, e0 + j, f), z1 = out_rule(p, y1, k, e0 + j + 1, f);
      CxPair<T> v;
      v.ar = z0.x;
      v.ai = z0.y;
      v.br = z1.x;
      v.bi = z1.y;
      stg(tout, (uint32_t)(k * osd + j) * CS, v);
    }
    return;
  }
  const int tot = (p.store_order == ORD_LINE ? nb : p.B) * (P + 1);
#pragma unroll 4
  for (int i = tid; i < tot; i += nth) {
    int j, k;
    if (p.store_order == ORD_LINE) {
      j = qdiv<P + 1>(i);
      k = qmod<P + 1>(i);
    } else {
      j = i & (p.B - 1);
      k = i >> p.logB;
      if (!FULL && j >= nb) continue;
    }
    // Y_2m from A; Y_4j+1 = v_j, Y_4j+3 = v_{L-1-j} from B
    const int vi = (k & 2) ? L - 1 - (k >> 2) : (k >> 2);
    const cx<T> y = (k & 1) ? bufB[j * p.lp2 + P3E_PAD(vi)] : bufA[j * p.lp + P3E_PAD(k >> 1)];
    stg(tout, (uint32_t)(k * osd + j * ose) * CS, out_rule(p, y, k, e0 + j, f));
  }
}

template <typename T, int P, typename Run, typename Each>
P3E_HD P3E_INL void dct1_tile(const PassArgs &p, uint32_t tile, cx<T> *buf, int nth, Run &run, Each &each,
                              const cx<T> *tab) {
  constexpr int L = P / 2;
  uint32_t f;
  int e0, nb;
  fast_tile_coords(p, tile, &f, &e0, &nb);
  const bool full = nb == p.B;
  cx<T> *bufA = buf, *bufB = buf + p.B * p.lp;
  const cx<T> *twA = tab, *twB = tab + p.ntw_a, *pre = tab + p.ntw_a + p.ntw_b;
  each([&](int tid) P3E_LAMBDA_INL {
    if (tid >= nth) return;
    if (full) dct1_load<T, P, true>(p, f, e0, nb, bufA, bufB, pre, tid, nth);
    else dct1_load<T, P, false>(p, f, e0, nb, bufA, bufB, pre, tid, nth);
  });
  const int nthA = P * p.B / P3E_FAST_E, nthB = L * p.B / P3E_FAST_E;
  dual_seq<T, P, L, 1, 1, -1, 1, Run>(run, typename Dct1Rads<P>::A{}, typename Dct1Rads<P>::B{}, bufA, p.lp, twA,
                                      nthA, bufB, p.lp2, twB, nthB, p.B);
  each([&](int tid) P3E_LAMBDA_INL {
    if (tid >= nth) return;
    if (full) dct1_store<T, P, true>(p, f, e0, nb, bufA, bufB, tid, nth);
    else dct1_store<T, P, false>(p, f, e0, nb, bufA, bufB, tid, nth);
  });
}

// Whole tile.  `each(f)` runs phase f for the calling thread and ends with a
// barrier (device), or runs it for every thread of the workgroup in turn
// (host build); `run` does the same for the two halves of a radix pass.
// Only threads tid < nth work (a batch workgroup may be larger than the
// pass needs); the others take part in the barriers.  The branches on the
// tile being whole and on the DFT sign are uniform over the workgroup.
// tw, wN: the pass's twiddles and R2C/C2R split factors (in LDS or global
// memory).
template <typename T, int M, typename Run, typename Each>
P3E_HD P3E_INL void fast_tile(const PassArgs &p, uint32_t tile, cx<T> *buf, int nth, Run &run, Each &each,
                              const cx<T> *tw, const cx<T> *wN) {
  if constexpr (M == 128 || M == 256 || M == 512 || M == 1024) {
    if (p.dct_split) {
      dct1_tile<T, M>(p, tile, buf, nth, run, each, tw);
      return;
    }
  }
  uint32_t f;
  int e0, nb;
  fast_tile_coords(p, tile, &f, &e0, &nb);
  const bool full = nb == p.B;
  each([&](int tid) P3E_LAMBDA_INL {
    if (tid >= nth) return;
    if (full) fast_load<T, M, true>(p, f, e0, nb, buf, tid, nth, wN);
    else fast_load<T, M, false>(p, f, e0, nb, buf, tid, nth, wN);
  });
  if (p.cls == CLS_R2C || p.sign < 0) fft_fixed<T, M, -1>(run, buf, p.lp, tw, p.B);
  else fft_fixed<T, M, 1>(run, buf, p.lp, tw, p.B);
  each([&](int tid) P3E_LAMBDA_INL {
    if (tid >= nth) return;
    if (full) fast_store<T, M, true>(p, f, e0, nb, buf, tid, nth, wN);
    else fast_store<T, M, false>(p, f, e0, nb, buf, tid, nth, wN);
  });
}

}  // namespace p3e
