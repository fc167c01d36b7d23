// Power-of-two path of the fused reorder + 1D transform pass.
//
// Same tile as the generic path (B lines along the transform axis, adjacent
// along the tile axis), but:
//   - one LDS buffer: every radix pass first reads all the elements a thread
//     owns into registers (P3E_FAST_E of them), then writes them back, with a
//     barrier in between, so the tile takes half the LDS of a ping-pong;
//   - radix 16/8/4/2 butterflies with shifts for all index arithmetic;
//   - R2C and C2R of even length N run one complex FFT of length M = N/2 on
//     the packed line z_n = x_{2n} + i x_{2n+1}, with the split/merge of the
//     even and odd halves done in the store (R2C) or load (C2R) phase.
// Shared by the HIP kernel and the host build used by the CPU tests.
#pragma once
#include "fft_core.h"

namespace p3e {

// One register-staged Stockham pass of radix R over nb lines of length M,
// split in its two barrier-separated phases.  Read: each thread loads the
// R inputs of each of its E/R butterflies into registers.  Write: twiddle,
// butterfly and store at the Stockham output positions.
template <typename T, int R>
P3E_HD P3E_INL void fast_read(const cx<T> *buf, const PassArgs &p, int nb, int tid, int nth, cx<T> *v,
                              int *lines, int *js) {
  constexpr int NB = P3E_FAST_E / R;
  constexpr int logR = (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : 4;
  const int lognbf = p.logM - logR;
  const int nbf = 1 << lognbf;
  const int total = nb << lognbf;
#pragma unroll
  for (int m = 0; m < NB; m++) {
    const int b = tid + m * nth;
    lines[m] = b >> lognbf;
    js[m] = b & (nbf - 1);
    if (b < total) {
      const cx<T> *sl = buf + lines[m] * p.lp;
#pragma unroll
      for (int r = 0; r < R; r++) v[m * R + r] = sl[P3E_PAD(js[m] + r * nbf)];
    }
  }
}

template <typename T, int R>
P3E_HD P3E_INL void fast_write(cx<T> *buf, const PassArgs &p, int nb, int Ns, int logNs, int s,
                               const cx<T> *tw, int tid, int nth, cx<T> *v, const int *lines, const int *js) {
  constexpr int NB = P3E_FAST_E / R;
  constexpr int logR = (R == 2) ? 1 : (R == 4) ? 2 : (R == 8) ? 3 : 4;
  const int total = nb << (p.logM - logR);
  const int tstride_log = p.logM - logNs - logR;  // twiddle index stride M / (Ns R)
#pragma unroll
  for (int m = 0; m < NB; m++) {
    const int b = tid + m * nth;
    if (b < total) {
      const int j = js[m];
      const int k = j & (Ns - 1);
      cx<T> *w = v + m * R;
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) w[r] = mul(w[r], twid(tw, (k * r) << tstride_log, s));
      }
      bfly<T, R>(w, s);
      cx<T> *dl = buf + lines[m] * p.lp;
      const int base = ((j - k) << logR) + k;
#pragma unroll
      for (int r = 0; r < R; r++) dl[P3E_PAD(base + r * Ns)] = w[r];
    }
  }
}

// Phase dispatch on the runtime radix: write == 0 reads, 1 writes.
template <typename T>
P3E_HD P3E_INL void fast_phase(int write, int R, cx<T> *buf, const PassArgs &p, int nb, int Ns, int logNs, int s,
                               const cx<T> *tw, int tid, int nth, cx<T> *v, int *lines, int *js) {
  switch (R) {
    case 2:
      if (write) fast_write<T, 2>(buf, p, nb, Ns, logNs, s, tw, tid, nth, v, lines, js);
      else fast_read<T, 2>(buf, p, nb, tid, nth, v, lines, js);
      break;
    case 4:
      if (write) fast_write<T, 4>(buf, p, nb, Ns, logNs, s, tw, tid, nth, v, lines, js);
      else fast_read<T, 4>(buf, p, nb, tid, nth, v, lines, js);
      break;
    case 8:
      if (write) fast_write<T, 8>(buf, p, nb, Ns, logNs, s, tw, tid, nth, v, lines, js);
      else fast_read<T, 8>(buf, p, nb, tid, nth, v, lines, js);
      break;
    default:
      if (write) fast_write<T, 16>(buf, p, nb, Ns, logNs, s, tw, tid, nth, v, lines, js);
      else fast_read<T, 16>(buf, p, nb, tid, nth, v, lines, js);
      break;
  }
}

// Tile coordinates of workgroup `tile`.
P3E_HD P3E_INL void fast_tile_coords(const PassArgs &p, uint32_t tile, uint32_t *f, int *e0, int *nb) {
  *f = fdiv(tile, p.div_ntile);
  *e0 = (int)(tile - *f * p.ntile_e) * p.B;
  *nb = (p.B < p.ne - *e0) ? p.B : (p.ne - *e0);
}

// Load phase of the power-of-two path: tile -> LDS (natural order; real
// lines packed two reals per complex slot).
template <typename T>
P3E_HD void fast_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *buf, int tid, int nth) {
  const int M = p.M;
  cx<T> *side = buf + (size_t)p.B * p.lp;  // C2R: X_M of each line
  // 1. load
  {
    const long long base = (long long)e0 * p.is_e + (long long)f * p.is_f;
    const int tot = (p.load_order == ORD_LINE ? nb : p.B) * p.nin;
    if (p.cls == CLS_R2C) {
      // real input: element n of line j is component n&1 of packed slot n>>1
      const T *in = reinterpret_cast<const T *>(p.in);
      for (int i = tid; i < tot; i += nth) {
        int n, j;
        if (p.load_order == ORD_LINE) {
          j = i >> (p.logM + 1);
          n = i & (2 * M - 1);
        } else {
          j = i & (p.B - 1);
          n = i >> p.logB;
          if (j >= nb) continue;
        }
        T x = in[base + (long long)n * p.is_d + (long long)j * p.is_e];
        T *slot = reinterpret_cast<T *>(buf + j * p.lp + P3E_PAD(n >> 1));
        slot[n & 1] = x;
      }
    } else {
      const cx<T> *in = reinterpret_cast<const cx<T> *>(p.in);
      for (int i = tid; i < tot; i += nth) {
        int n, j;
        if (p.load_order == ORD_LINE) {
          j = (int)fdiv((uint32_t)i, p.div_nin);
          n = i - j * p.nin;
        } else {
          j = i & (p.B - 1);
          n = i >> p.logB;
          if (j >= nb) continue;
        }
        cx<T> x = in[base + (long long)n * p.is_d + (long long)j * p.is_e];
        if (n < M) buf[j * p.lp + P3E_PAD(n)] = x;
        else side[j] = x;  // C2R: X_M
      }
    }
  }
}

// C2R only: merge X_k and X_{M-k} into the packed spectrum in place.
template <typename T>
P3E_HD void fast_c2r_merge(const PassArgs &p, int nb, cx<T> *buf, int tid, int nth) {
  const int M = p.M;
  const cx<T> *wN = reinterpret_cast<const cx<T> *>(p.wN);
  cx<T> *side = buf + (size_t)p.B * p.lp;
  // 2. C2R: merge X_k and X_{M-k} into the packed spectrum
  //    Z_k = (X_k + conj X_{M-k}) + i w_N^{-k} (X_k - conj X_{M-k})
  {
    const int half = M / 2;  // pairs (k, M-k) for k = 0..M/2
    const int tot = nb * (half + 1);
    for (int i = tid; i < tot; i += nth) {
      const int j = i / (half + 1);
      const int k = i - j * (half + 1);
      cx<T> *l = buf + j * p.lp;
      if (k == 0) {
        // imaginary parts of X_0 and X_M do not contribute (FFTW c2r)
        const T a = l[P3E_PAD(0)].x, c = side[j].x;
        l[P3E_PAD(0)] = mk(a + c, a - c);
      } else {
        const cx<T> xk = l[P3E_PAD(k)], xm = l[P3E_PAD(M - k)];
        const cx<T> wk = conj(wN[k]), wm = conj(wN[M - k]);  // exp(+2 pi i k / N)
        // Z_k
        const cx<T> ek = add(xk, conj(xm)), ok = sub(xk, conj(xm));
        const cx<T> zk = add(ek, rot(mul(wk, ok), 1));
        if (k != M - k) {
          const cx<T> em = add(xm, conj(xk)), om = sub(xm, conj(xk));
          const cx<T> zm = add(em, rot(mul(wm, om), 1));
          l[P3E_PAD(M - k)] = zm;
        }
        l[P3E_PAD(k)] = zk;
      }
    }
  }
}

// Store phase of the power-of-two path: output rule and LDS -> tile.
template <typename T>
P3E_HD void fast_store(const PassArgs &p, uint32_t f, int e0, int nb, const cx<T> *buf, int tid, int nth) {
  const int M = p.M;
  const cx<T> *wN = reinterpret_cast<const cx<T> *>(p.wN);
  // 4. store with the output rule
  {
    const long long base = (long long)e0 * p.os_e + (long long)f * p.os_f;
    const int tot = (p.store_order == ORD_LINE ? nb : p.B) * p.nout;
    if (p.cls == CLS_C2R) {
      T *out = reinterpret_cast<T *>(p.out);
      for (int i = tid; i < tot; i += nth) {
        int k, j;
        if (p.store_order == ORD_LINE) {
          j = i >> (p.logM + 1);
          k = i & (2 * M - 1);
        } else {
          j = i & (p.B - 1);
          k = i >> p.logB;
          if (j >= nb) continue;
        }
        const T *slot = reinterpret_cast<const T *>(buf + j * p.lp + P3E_PAD(k >> 1));
        T y = slot[k & 1];
        if (p.has_scale) y *= (T)p.scale;
        out[base + (long long)k * p.os_d + (long long)j * p.os_e] = y;
      }
    } else {
      cx<T> *out = reinterpret_cast<cx<T> *>(p.out);
      for (int i = tid; i < tot; i += nth) {
        int k, j;
        if (p.store_order == ORD_LINE) {
          j = (int)fdiv((uint32_t)i, p.div_nout);
          k = i - j * p.nout;
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
          const cx<T> t = rot(mul(wN[k], o), -1);
          y = scl(add(e, t), (T)0.5);
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
        if (p.has_scale) y = scl(y, (T)p.scale);
        out[base + (long long)k * p.os_d + (long long)j * p.os_e] = y;
      }
    }
  }
}

}  // namespace p3e
