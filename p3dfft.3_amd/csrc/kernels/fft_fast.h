// Power-of-two path of the fused reorder + 1D transform pass.
//
// Same tile as the generic path (B lines along the transform axis, adjacent
// along the tile axis), but:
//   - the complex length M and its radix sequence are compile-time, one
//     kernel per M, so every index is a shift and the values a thread owns
//     stay in registers;
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

// Radix-R Stockham pass at stride Ns of a length-M FFT, split in halves.
template <typename T, int R, int M, int Ns> struct FPass {
  static constexpr int E = P3E_FAST_E;
  static constexpr int SLOTS = E;  // registers per thread between the halves
  static constexpr int NB = E / R;  // butterflies per thread
  static constexpr int logR = Log2<R>::v;
  static constexpr int lognbf = Log2<M>::v - logR;
  static constexpr int nbf = M / R;
  cx<T> *buf;
  int lp, nb, s;
  const cx<T> *tw;

  P3E_HD P3E_INL void read(int tid, int nth, cx<T> *v) const {
    const int total = nb << lognbf;
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (b < total) {
        const cx<T> *sl = buf + (b >> lognbf) * lp;
        const int j = b & (nbf - 1);
#pragma unroll
        for (int r = 0; r < R; r++) v[m * R + r] = sl[P3E_PAD(j + r * nbf)];
      }
    }
  }

  P3E_HD P3E_INL void write(int tid, int nth, cx<T> *v) const {
    const int total = nb << lognbf;
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (b < total) {
        const int j = b & (nbf - 1);
        const int k = j & (Ns - 1);
        cx<T> *w = v + m * R;
        if (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; r++) w[r] = mul(w[r], twid(tw, k * tw_stride(R) + r - 1, s));
        }
        bfly<T, R>(w, s);
        cx<T> *dl = buf + (b >> lognbf) * lp;
        const int base = ((j - k) << logR) + k;
#pragma unroll
        for (int r = 0; r < R; r++) dl[P3E_PAD(base + r * Ns)] = w[r];
      }
    }
  }
};

// Run the radix sequence R, Rest... starting at stride Ns; `run` executes
// one pass (read half, barrier, write half, barrier).  tw: the twiddles of
// this pass onward, pass by pass (pass_twiddles on the host): Ns
// tw_stride(R) entries for a radix-R pass at stride Ns.
template <typename T, int M, int Ns, typename Run, int R, int... Rest>
P3E_HD P3E_INL void fft_seq(Run &run, cx<T> *buf, int lp, int nb, int s, const cx<T> *tw) {
  FPass<T, R, M, Ns> ps{buf, lp, nb, s, tw};
  run(ps);
  if constexpr (sizeof...(Rest) > 0) fft_seq<T, M, Ns * R, Run, Rest...>(run, buf, lp, nb, s, tw + Ns * tw_stride(R));
}

// Radix sequence of each supported length (all radices divide P3E_FAST_E).
template <typename T, int M, typename Run>
P3E_HD P3E_INL void fft_pow2(Run &run, cx<T> *buf, int lp, int nb, int s, const cx<T> *tw) {
  if constexpr (M == 16) fft_seq<T, M, 1, Run, 16>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 32) fft_seq<T, M, 1, Run, 8, 4>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 64) fft_seq<T, M, 1, Run, 8, 8>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 128) fft_seq<T, M, 1, Run, 16, 8>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 256) fft_seq<T, M, 1, Run, 16, 16>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 512) fft_seq<T, M, 1, Run, 8, 8, 8>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 1024) fft_seq<T, M, 1, Run, 16, 8, 8>(run, buf, lp, nb, s, tw);
  else if constexpr (M == 2048) fft_seq<T, M, 1, Run, 16, 16, 8>(run, buf, lp, nb, s, tw);
  else fft_seq<T, M, 1, Run, 16, 16, 16>(run, buf, lp, nb, s, tw);
}

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
  constexpr int logM = Log2<M>::v;
  if (p.load_order == ORD_LINE) {
    if (p.cls == CLS_R2C) {
      *j = i >> (logM + 1);
      *n = i & (2 * M - 1);
    } else if (p.cls == CLS_C2R) {
      *j = (int)fdiv((uint32_t)i, p.div_nin);
      *n = i - *j * p.nin;
    } else {
      *j = i >> logM;
      *n = i & (M - 1);
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
P3E_HD P3E_INL void c2r_pair_put(const PassArgs &p, cx<T> *line, int k, cx<T> xk, cx<T> xm) {
  const cx<T> *wN = reinterpret_cast<const cx<T> *>(p.wN);
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

template <typename T, int M>
P3E_HD P3E_INL void fast_fetch(const PassArgs &p, uint32_t f, int e0, int nb, TileRegs<T> &R, int tid, int nth) {
  constexpr int E = P3E_FAST_E;
  const long long base = (long long)e0 * p.is_e + (long long)f * p.is_f;
  const int tot = (p.load_order == ORD_LINE ? nb : p.B) * p.nin;
  // out-of-tile entries load the tile's first element (always valid) and
  // are not written: unconditional loads keep the values in registers
  if (p.vec2 && p.cls == CLS_C2C) {
    // C2C, two adjacent elements per load: in tile order (e stride-1, whole
    // tiles) element n of lines j = 2 jp and j + 1; in line order (d
    // stride-1) elements n = 2 np and n + 1 of line j
    const cx<T> *in = reinterpret_cast<const cx<T> *>(p.in);
    const bool lo = p.load_order == ORD_LINE;
    const int lb = lo ? Log2<M>::v - 1 : p.logB - 1;
    const int ptot = lo ? nb * (M / 2) : (p.B >> 1) * p.nin;
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int lo_i = 2 * (i & ((1 << lb) - 1)), hi_i = i >> lb;
      const int j = lo ? hi_i : lo_i, n = lo ? lo_i : hi_i;
      const bool ok = i < ptot;
      const CxPair<T> x = *reinterpret_cast<const CxPair<T> *>(
          in + (ok ? base + (long long)n * p.is_d + (long long)j * p.is_e : base));
      R.v[4 * q] = x.ar;
      R.v[4 * q + 1] = x.ai;
      R.v[4 * q + 2] = x.br;
      R.v[4 * q + 3] = x.bi;
    }
  } else if (p.vec2 && p.load_order == ORD_LINE) {
    // R2C, line order (d stride-1): entry q is the real pair (x_2m, x_2m+1)
    // of line j, the packed complex element m, read by one load
    const T *in = reinterpret_cast<const T *>(p.in);
    constexpr int logM = Log2<M>::v;
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = i >> logM, m = i & (M - 1);
      const bool ok = j < nb;
      const RealPair<T> x =
          *reinterpret_cast<const RealPair<T> *>(in + (ok ? base + 2 * m + (long long)j * p.is_e : base));
      R.v[2 * q] = x.a;
      R.v[2 * q + 1] = x.b;
    }
  } else if (p.vec2) {
    // R2C, tile order, e stride-1 and whole tiles (the planner's condition):
    // entry q is element n of lines j = 2 jp and j + 1, adjacent in memory;
    // the launcher checked that base and strides keep the pair aligned
    const T *in = reinterpret_cast<const T *>(p.in);
    const int lb = p.logB - 1;
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int jp = i & ((1 << lb) - 1), n = i >> lb;
      const bool ok = i < (p.B >> 1) * p.nin;
      const RealPair<T> x =
          *reinterpret_cast<const RealPair<T> *>(in + (ok ? base + (long long)n * p.is_d + 2 * jp : base));
      R.v[2 * q] = x.a;
      R.v[2 * q + 1] = x.b;
    }
  } else if (p.cls == CLS_R2C) {
    const T *in = reinterpret_cast<const T *>(p.in);
#pragma unroll
    for (int q = 0; q < 2 * E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      const bool ok = i < tot && j < nb;
      R.v[q] = in[ok ? base + (long long)n * p.is_d + (long long)j * p.is_e : base];
    }
  } else if (p.cls == CLS_C2R) {
    // pairs (X_k, X_{M-k}) of one line in one thread, merged in fast_put
    const cx<T> *in = reinterpret_cast<const cx<T> *>(p.in);
    const int ptot = (p.load_order == ORD_LINE ? nb : p.B) * (M / 2 + 1);
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      int j, k;
      fast_pair_index<M>(p, i, &j, &k);
      const bool ok = i < ptot && j < nb;
      const long long o = base + (long long)j * p.is_e;
      const cx<T> a = in[ok ? o + (long long)k * p.is_d : base];
      const cx<T> b = in[ok ? o + (long long)(M - k) * p.is_d : base];
      R.v[4 * q] = a.x;
      R.v[4 * q + 1] = a.y;
      R.v[4 * q + 2] = b.x;
      R.v[4 * q + 3] = b.y;
    }
  } else {
    const cx<T> *in = reinterpret_cast<const cx<T> *>(p.in);
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      const bool ok = i < tot && j < nb;
      const cx<T> x = in[ok ? base + (long long)n * p.is_d + (long long)j * p.is_e : base];
      R.v[2 * q] = x.x;
      R.v[2 * q + 1] = x.y;
    }
  }
}

template <typename T, int M>
P3E_HD P3E_INL void fast_put(const PassArgs &p, uint32_t f, int e0, int nb, const TileRegs<T> &R, cx<T> *buf,
                             int tid, int nth) {
  constexpr int E = P3E_FAST_E;
  const int tot = (p.load_order == ORD_LINE ? nb : p.B) * p.nin;
  if (p.vec2 && p.cls == CLS_C2C) {
    const bool lo = p.load_order == ORD_LINE;
    const int lb = lo ? Log2<M>::v - 1 : p.logB - 1;
    const int ptot = lo ? nb * (M / 2) : (p.B >> 1) * p.nin;
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      const int lo_i = 2 * (i & ((1 << lb) - 1)), hi_i = i >> lb;
      const int j = lo ? hi_i : lo_i, n = lo ? lo_i : hi_i;
      if (i < ptot) {
        // the second element: line j + 1 (tile order) or element n + 1 (line order)
        buf[j * p.lp + P3E_PAD(n)] = mk(R.v[4 * q], R.v[4 * q + 1]);
        buf[lo ? j * p.lp + P3E_PAD(n + 1) : (j + 1) * p.lp + P3E_PAD(n)] = mk(R.v[4 * q + 2], R.v[4 * q + 3]);
      }
    }
  } else if (p.vec2 && p.load_order == ORD_LINE) {
    constexpr int logM = Log2<M>::v;
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = i >> logM, m = i & (M - 1);
      if (j < nb) buf[j * p.lp + P3E_PAD(m)] = mk(R.v[2 * q], R.v[2 * q + 1]);
    }
  } else if (p.vec2) {
    T *bufr = reinterpret_cast<T *>(buf);
    const int lb = p.logB - 1;
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = 2 * (i & ((1 << lb) - 1)), n = i >> lb;
      if (i < (p.B >> 1) * p.nin) {
        const int o = 2 * P3E_PAD(n >> 1) + (n & 1);
        bufr[2 * j * p.lp + o] = R.v[2 * q];
        bufr[2 * (j + 1) * p.lp + o] = R.v[2 * q + 1];
      }
    }
  } else if (p.cls == CLS_R2C) {
    T *bufr = reinterpret_cast<T *>(buf);
#pragma unroll
    for (int q = 0; q < 2 * E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      if (i < tot && j < nb) bufr[2 * (j * p.lp + P3E_PAD(n >> 1)) + (n & 1)] = R.v[q];
    }
  } else if (p.cls == CLS_C2R) {
    const int ptot = (p.load_order == ORD_LINE ? nb : p.B) * (M / 2 + 1);
#pragma unroll
    for (int q = 0; q < E / 2; q++) {
      const int i = tid + q * nth;
      int j, k;
      fast_pair_index<M>(p, i, &j, &k);
      if (i < ptot && j < nb)
        c2r_pair_put<T, M>(p, buf + j * p.lp, k, mk(R.v[4 * q], R.v[4 * q + 1]), mk(R.v[4 * q + 2], R.v[4 * q + 3]));
    }
    // (M/2 + 1) B pairs are B more than (E/2) nth: loaded here directly
    const cx<T> *in = reinterpret_cast<const cx<T> *>(p.in);
    const long long base = (long long)e0 * p.is_e + (long long)f * p.is_f;
    for (int i = tid + (E / 2) * nth; i < ptot; i += nth) {
      int j, k;
      fast_pair_index<M>(p, i, &j, &k);
      if (j >= nb) continue;
      const long long o = base + (long long)j * p.is_e;
      c2r_pair_put<T, M>(p, buf + j * p.lp, k, in[o + (long long)k * p.is_d], in[o + (long long)(M - k) * p.is_d]);
    }
  } else {
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      int j, n;
      fast_load_index<M>(p, i, &j, &n);
      if (i < tot && j < nb) buf[j * p.lp + P3E_PAD(n)] = mk(R.v[2 * q], R.v[2 * q + 1]);
    }
  }
}

template <typename T, int M>
P3E_HD void fast_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *buf, int tid, int nth) {
  TileRegs<T> R;
  fast_fetch<T, M>(p, f, e0, nb, R, tid, nth);
  fast_put<T, M>(p, f, e0, nb, R, buf, tid, nth);
}

// Store phase: output rule and LDS -> tile.
template <typename T, int M>
P3E_HD void fast_store(const PassArgs &p, uint32_t f, int e0, int nb, const cx<T> *buf, int tid, int nth) {
  constexpr int logM = Log2<M>::v;
  const cx<T> *wN = reinterpret_cast<const cx<T> *>(p.wN);
  const long long base = (long long)e0 * p.os_e + (long long)f * p.os_f;
  const int tot = (p.store_order == ORD_LINE ? nb : p.B) * p.nout;
  constexpr int E = P3E_FAST_E;
  if (p.vec2s && p.cls == CLS_C2R && p.store_order == ORD_LINE) {
    // C2R, line order (d stride-1): the packed complex element m of line j
    // is the real pair (x_2m, x_2m+1), written by one store
    T *out = reinterpret_cast<T *>(p.out);
#pragma unroll 4
    for (int i = tid; i < nb * M; i += nth) {
      const int j = i >> logM, m = i & (M - 1);
      const cx<T> z = buf[j * p.lp + P3E_PAD(m)];
      RealPair<T> v;
      v.a = p.has_scale ? z.x * (T)p.scale : z.x;
      v.b = p.has_scale ? z.y * (T)p.scale : z.y;
      *reinterpret_cast<RealPair<T> *>(out + base + 2 * m + (long long)j * p.os_e) = v;
    }
  } else if (p.vec2s && p.cls == CLS_C2R) {
    // C2R, tile order, e stride-1 and whole tiles (the planner's condition):
    // entry q is element k of lines j = 2 jp and j + 1, adjacent in memory,
    // written by one store; the launcher checked the pair alignment
    T *out = reinterpret_cast<T *>(p.out);
    const T *bufr = reinterpret_cast<const T *>(buf);
    const int lb = p.logB - 1;
    T y[2 * E];
    long long off[E];
#pragma unroll
    for (int q = 0; q < E; q++) {
      const int i = tid + q * nth;
      const int j = 2 * (i & ((1 << lb) - 1)), k = i >> lb;
      const bool ok = i < (p.B >> 1) * p.nout;
      const int o = 2 * P3E_PAD(k >> 1) + (k & 1);
      y[2 * q] = bufr[ok ? 2 * j * p.lp + o : 0];
      y[2 * q + 1] = bufr[ok ? 2 * (j + 1) * p.lp + o : 0];
      off[q] = ok ? base + (long long)k * p.os_d + j : -1;
    }
#pragma unroll
    for (int q = 0; q < E; q++) {
      if (off[q] < 0) continue;
      RealPair<T> v;
      v.a = p.has_scale ? y[2 * q] * (T)p.scale : y[2 * q];
      v.b = p.has_scale ? y[2 * q + 1] * (T)p.scale : y[2 * q + 1];
      *reinterpret_cast<RealPair<T> *>(out + off[q]) = v;
    }
  } else if (p.cls == CLS_C2R) {
    T *out = reinterpret_cast<T *>(p.out);
    const T *bufr = reinterpret_cast<const T *>(buf);
    // 2E reals per thread, 2CH at a time: LDS reads first, then the global
    // stores
    constexpr int CH = E;
#pragma unroll
    for (int c = 0; c < 2 * E; c += 2 * CH) {
    T y[2 * CH];
    long long off[2 * CH];
#pragma unroll
    for (int q = 0; q < 2 * CH; q++) {
      const int i = tid + (c + q) * nth;
      int k, j;
      if (p.store_order == ORD_LINE) {
        j = i >> (logM + 1);
        k = i & (2 * M - 1);
      } else {
        j = i & (p.B - 1);
        k = i >> p.logB;
      }
      const bool ok = i < tot && j < nb;
      y[q] = bufr[ok ? 2 * (j * p.lp + P3E_PAD(k >> 1)) + (k & 1) : 0];
      off[q] = ok ? base + (long long)k * p.os_d + (long long)j * p.os_e : -1;
    }
#pragma unroll
    for (int q = 0; q < 2 * CH; q++)
      if (off[q] >= 0) out[off[q]] = p.has_scale ? y[q] * (T)p.scale : y[q];
    }
  } else if (p.vec2s) {
    // C2C and R2C, two adjacent output elements per store: in tile order (e
    // stride-1, whole tiles) element k of lines j = 2 jp and j + 1; in line
    // order (C2C only, d stride-1) elements k = 2 kp and k + 1 of line j
    cx<T> *out = reinterpret_cast<cx<T> *>(p.out);
    const bool lo = p.store_order == ORD_LINE;
    const int lb = lo ? Log2<M>::v - 1 : p.logB - 1;
    const int ptot = lo ? nb * (M / 2) : (p.B >> 1) * p.nout;
#pragma unroll 4
    for (int i = tid; i < ptot; i += nth) {
      const int lo_i = 2 * (i & ((1 << lb) - 1)), hi_i = i >> lb;
      const int j = lo ? hi_i : lo_i, k = lo ? lo_i : hi_i;
      const int j1 = lo ? j : j + 1, k1 = lo ? k + 1 : k;  // the second element
      cx<T> y[2];
      if (p.cls == CLS_R2C) {
        // X_k = (Z_k + conj Z_{M-k})/2 - i w_N^k (Z_k - conj Z_{M-k})/2, per line
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const cx<T> *l = buf + (h ? j1 : j) * p.lp;
          const cx<T> zk = l[P3E_PAD(k & (M - 1))];
          const cx<T> zm = l[P3E_PAD((M - k) & (M - 1))];
          const cx<T> e = add(zk, conj(zm)), o = sub(zk, conj(zm));
          y[h] = scl(add(e, rot(mul(wN[k], o), -1)), (T)0.5);
        }
      } else {
        y[0] = buf[j * p.lp + P3E_PAD(k)];
        y[1] = buf[j1 * p.lp + P3E_PAD(k1)];
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int kh = h ? k1 : k, jh = h ? j1 : j;
        if (p.deriv) {
          T kk;
          if (kh < p.dmid) kk = (T)kh;
          else if (kh == p.dmid) kk = (T)0;
          else kk = (T)(kh - p.dg);
          y[h] = mk(-kk * y[h].y, kk * y[h].x);
        }
        if (p.spec_op) y[h] = scl(y[h], poisson_factor<T>(p, kh, e0 + jh, (int)f));
        if (p.has_scale) y[h] = scl(y[h], (T)p.scale);
      }
      CxPair<T> v;
      v.ar = y[0].x;
      v.ai = y[0].y;
      v.br = y[1].x;
      v.bi = y[1].y;
      *reinterpret_cast<CxPair<T> *>(out + base + (long long)k * p.os_d + (long long)j * p.os_e) = v;
    }
  } else {
    cx<T> *out = reinterpret_cast<cx<T> *>(p.out);
#pragma unroll 4
    for (int i = tid; i < tot; i += nth) {
      int k, j;
      if (p.store_order == ORD_LINE) {
        if (p.cls == CLS_R2C) {
          j = (int)fdiv((uint32_t)i, p.div_nout);
          k = i - j * p.nout;
        } else {
          j = i >> logM;
          k = i & (M - 1);
        }
      } else {
        j = i & (p.B - 1);
        k = i >> p.logB;
        if (j >= nb) continue;
      }
      const cx<T> *l = buf + j * p.lp;
      cx<T> y;  // output element k of line j
      if (p.cls == CLS_R2C) {
        // X_k = (Z_k + conj Z_{M-k})/2 - i w_N^k (Z_k - conj Z_{M-k})/2
        const cx<T> zk = l[P3E_PAD(k & (M - 1))];
        const cx<T> zm = l[P3E_PAD((M - k) & (M - 1))];
        const cx<T> e = add(zk, conj(zm)), o = sub(zk, conj(zm));
        y = scl(add(e, rot(mul(wN[k], o), -1)), (T)0.5);
      } else {
        y = l[P3E_PAD(k)];
      }
      if (p.deriv) {
        T kk;
        if (k < p.dmid) kk = (T)k;
        else if (k == p.dmid) kk = (T)0;
        else kk = (T)(k - p.dg);
        y = mk(-kk * y.y, kk * y.x);
      }
      if (p.spec_op) y = scl(y, poisson_factor<T>(p, k, e0 + j, (int)f));
      if (p.has_scale) y = scl(y, (T)p.scale);
      out[base + (long long)k * p.os_d + (long long)j * p.os_e] = y;
    }
  }
}

// Whole tile.  `each(f)` runs phase f for the calling thread and ends with a
// barrier (device), or runs it for every thread of the workgroup in turn
// (host build); `run` does the same for the two halves of a radix pass.
// Only threads tid < nth work (a batch workgroup may be larger than the
// pass needs); the others take part in the barriers.
template <typename T, int M, typename Run, typename Each>
P3E_HD void fast_tile(const PassArgs &p, uint32_t tile, cx<T> *buf, int nth, Run &run, Each &each) {
  uint32_t f;
  int e0, nb;
  fast_tile_coords(p, tile, &f, &e0, &nb);
  each([&](int tid) {
    if (tid < nth) fast_load<T, M>(p, f, e0, nb, buf, tid, nth);
  });
  fft_pow2<T, M>(run, buf, p.lp, nb, (p.cls == CLS_R2C) ? -1 : p.sign, reinterpret_cast<const cx<T> *>(p.tw));
  each([&](int tid) {
    if (tid < nth) fast_store<T, M>(p, f, e0, nb, buf, tid, nth);
  });
}

}  // namespace p3e
