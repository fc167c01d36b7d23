// Single-buffer mixed-radix path of the fused reorder + 1D transform pass.
//
// For work lengths L whose radices are all in {2, 3, 4, 5, 8, 16} plus at
// most one odd prime P in 7..47 (also the power-of-two Bluestein lengths):
// the same load, scatter, output rule and store as the generic path
// (fft_core.h), but the radix passes run in place on ONE padded LDS
// buffer, as in the power-of-two path: every thread reads the butterflies
// it owns into registers, barrier, twiddles + butterflies + writes to the
// Stockham positions, barrier.  Half the LDS of the two-buffer generic path
// per line, so twice the lines per tile (longer contiguous runs on a
// strided axis).  The planner sizes the workgroup so that every thread owns
// at most P3E_MID_E elements of the tile (nb * L <= P3E_MID_E * nth).  The prime P is a
// template parameter (PassArgs::mid_prime selects the instantiation), so
// the registers of its butterfly are only allocated in the kernels that
// need it.
// Shared by the HIP kernels and the host build used by the CPU tests.
#pragma once
#include "fft_core.h"

namespace p3e {

// Radix-R pass at input stride Ns over nb padded lines of length L.
template <typename T, int R, int S> struct MPass {
  static constexpr int NB = (P3E_MID_E + R - 1) / R;  // butterflies per thread, at most
  static constexpr int SLOTS = NB * R;
  cx<T> *buf;
  int L, lpad, nb, Ns;
  const cx<T> *tw;
  FastDiv dnbf, dns;

  P3E_HD P3E_INL void read(int tid, int nth, cx<T> *v) const {
    const int nbf = L / R, total = nb * nbf;
    // P3E_PAD(j + r nbf) = P3E_PAD(j) + r (nbf + nbf / 16) when 16 divides nbf
    const int step = nbf % 16 == 0 ? nbf + (nbf >> 4) : 0;
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (b < total) {
        const int line = (int)fdiv((uint32_t)b, dnbf);
        const int j = b - line * nbf;
        const cx<T> *sl = buf + line * lpad;
        if (step) {
          const cx<T> *q = sl + P3E_PAD(j);
#pragma unroll
          for (int r = 0; r < R; r++) v[m * R + r] = q[r * step];
        } else {
#pragma unroll
          for (int r = 0; r < R; r++) v[m * R + r] = sl[P3E_PAD(j + r * nbf)];
        }
      }
    }
  }

  P3E_HD P3E_INL void write(int tid, int nth, cx<T> *v) const {
    const int nbf = L / R, total = nb * nbf;
    const int tstride = L / (Ns * R);
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (b < total) {
        const int line = (int)fdiv((uint32_t)b, dnbf);
        const int j = b - line * nbf;
        const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
        cx<T> *w = v + m * R;
        if (Ns > 1) {
#pragma unroll
          for (int r = 1; r < R; r++) w[r] = mul(w[r], twid(tw, k * r * tstride, S));
        }
        cx<T> *dl = buf + line * lpad;
        const int base = (j - k) * R + k;
        if constexpr (is_base_radix<R>) {
          bfly<T, R>(w, S);
          if (Ns % 16 == 0) {
            // P3E_PAD(base + r Ns) = P3E_PAD(base) + r (Ns + Ns / 16)
            cx<T> *q = dl + P3E_PAD(base);
            const int step = Ns + (Ns >> 4);
#pragma unroll
            for (int r = 0; r < R; r++) q[r * step] = w[r];
          } else {
#pragma unroll
            for (int r = 0; r < R; r++) dl[P3E_PAD(base + r * Ns)] = w[r];
          }
        } else {
          bf_prime<T, R>(w, S, [&](int q, cx<T> x) { dl[P3E_PAD(base + q * Ns)] = x; });
        }
      }
    }
  }
};

// FFT of length L on nb lines in place; `run` executes one pass (read half,
// barrier, write half, barrier).  tw: one table of length L, w^(k r) of a
// radix-R pass at stride Ns at entry k r L / (Ns R).  (The pass-by-pass
// layout of the power-of-two path measured slower here on MI355X for
// lengths with radices 3 and 5, and faster only for power-of-two L.)
template <typename T, int P, int S, typename Run>
P3E_HD P3E_INL void mid_fft(Run &run, cx<T> *buf, const PassArgs &p, int nb, const cx<T> *tw) {
  int Ns = 1;
  for (int i = 0; i < p.nrad; i++) {
    const int R = p.radix[i];
    const FastDiv dn = p.div_nbf[i], ds = p.div_ns[i];
    switch (R) {
      case 2: { MPass<T, 2, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds}; run(ps); } break;
      case 3: { MPass<T, 3, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds}; run(ps); } break;
      case 4: { MPass<T, 4, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds}; run(ps); } break;
      case 5: { MPass<T, 5, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds}; run(ps); } break;
      case 8: { MPass<T, 8, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds}; run(ps); } break;
      case 16: { MPass<T, 16, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds}; run(ps); } break;
      default:
        // the planner admits no other radix to this instantiation
        if constexpr (P > 0) {
          MPass<T, P, S> ps{buf, p.L, p.lpad, nb, Ns, tw, dn, ds};
          run(ps);
        }
        break;
    }
    Ns *= R;
  }
}

// ---------------------------------------------------------------- load phase
// Load phase of the single-buffer path in groups: every thread issues the
// global loads of U entries into registers, then scatters the U values into
// LDS, so that U loads are in flight together (a loop of load-then-scatter
// keeps about one in flight; all of a thread's entries at once, as in the
// power-of-two path, unrolls the scatter rules too far and spills).  Entry
// u of a group at i0 is tile element i0 + u nth; out-of-tile entries load
// the tile's first element (always valid) and are not scattered.  Values
// are kept as scalars: arrays of cx<double> would not be split into
// registers.

// (line j, element n) of tile element i in load order, over lines of len
// elements (line order) or rows of 2^logw lines (tile order).
P3E_HD P3E_INL void mid_load_index(bool line_order, int i, int len, const FastDiv &dlen, int logw, int *j, int *n) {
  if (line_order) {
    *j = (int)fdiv((uint32_t)i, dlen);
    *n = i - *j * len;
  } else {
    *j = i & ((1 << logw) - 1);
    *n = i >> logw;
  }
}

template <typename T>
P3E_HD P3E_INL void mid_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *buf, int tid, int nth) {
  constexpr int U = 4;  // loads in flight per thread
  const cx<T> *chirp = reinterpret_cast<const cx<T> *>(p.chirp);
  const cx<T> *pre = reinterpret_cast<const cx<T> *>(p.pre);
  // the tile's elements at 32-bit offsets from its first one (the planner
  // admits a pass to this path only when they stay below 2^31 bytes)
  const uint32_t esz = p.in_cplx ? 2 * sizeof(T) : sizeof(T);
  const char *tin = reinterpret_cast<const char *>(p.in) + ((long long)e0 * p.is_e + (long long)f * p.is_f) * esz;
  const int isd = (int)p.is_d, ise = (int)p.is_e;
  const bool lo = p.load_order == ORD_LINE;
  const int tot = (lo ? nb : p.B) * p.nin;
  for (int i0 = tid; i0 < tot; i0 += U * nth) {
    T v[2 * U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * nth;
      int j, n;
      mid_load_index(lo, i, p.nin, p.div_nin, p.logB, &j, &n);
      const uint32_t off = (i < tot && j < nb) ? (uint32_t)(n * isd + j * ise) * esz : 0u;
      if (p.in_cplx) {
        const cx<T> x = *reinterpret_cast<const cx<T> *>(tin + off);
        v[2 * u] = x.x;
        v[2 * u + 1] = x.y;
      } else {
        v[2 * u] = *reinterpret_cast<const T *>(tin + off);
        v[2 * u + 1] = (T)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * nth;
      int j, n;
      mid_load_index(lo, i, p.nin, p.div_nin, p.logB, &j, &n);
      if (i < tot && j < nb) scatter(buf + j * p.lpad, n, mk(v[2 * u], v[2 * u + 1]), p, chirp, pre);
    }
  }
}

// ---------------------------------------------------------------- paired lines
// R2C and C2R with p.pair: one complex FFT of length W = N serves two real
// lines.  LDS line t holds z = x_{2t} + i x_{2t+1} (line 2t+1 absent -> 0).
// R2C:  X_{2t,k} = (Z_k + conj Z_{N-k}) / 2,  X_{2t+1,k} = (Z_k - conj Z_{N-k}) / (2i).
// C2R:  Z_k = Y_{2t,k} + i Y_{2t+1,k} over the Hermitian extension of both
// half spectra; the imaginary parts of the DC and (even N) Nyquist inputs
// are dropped first, as a real inverse transform ignores them; the outputs
// are Re z (line 2t) and Im z (line 2t+1).

// Lines of LDS a tile of nb lines occupies.
P3E_HD P3E_INL int tile_lines(const PassArgs &p, int nb) { return p.pair ? (nb + 1) / 2 : nb; }

template <typename T>
P3E_HD P3E_INL void pair_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *buf, int tid, int nth) {
  constexpr int U = 4;  // entries (line pairs at one n) loaded together, as in mid_load
  const cx<T> *chirp = reinterpret_cast<const cx<T> *>(p.chirp);
  const uint32_t esz = p.cls == CLS_R2C ? sizeof(T) : 2 * sizeof(T);
  const char *tin = reinterpret_cast<const char *>(p.in) + ((long long)e0 * p.is_e + (long long)f * p.is_f) * esz;
  const int isd = (int)p.is_d, ise = (int)p.is_e;
  const int nl = (nb + 1) / 2, N = p.N;
  const int lbl = p.logB > 0 ? p.logB - 1 : 0;  // B / 2 line pairs per row (1 if B = 1)
  const bool lo = p.load_order == ORD_LINE;
  const int tot = (lo ? nl : 1 << lbl) * p.nin;
  for (int i0 = tid; i0 < tot; i0 += U * nth) {
    // entry u: element n of the lines 2t (a) and 2t+1 (b, 0 if absent)
    T v[4 * U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * nth;
      int t, n;
      mid_load_index(lo, i, p.nin, p.div_nin, lbl, &t, &n);
      const bool ok = i < tot && t < nl, two = ok && 2 * t + 1 < nb;
      const uint32_t off = ok ? (uint32_t)(n * isd + 2 * t * ise) * esz : 0u;
      const uint32_t off2 = two ? off + (uint32_t)ise * esz : 0u;
      if (p.cls == CLS_R2C) {
        const T a = *reinterpret_cast<const T *>(tin + off), b = *reinterpret_cast<const T *>(tin + off2);
        v[4 * u] = a;
        v[4 * u + 1] = two ? b : (T)0;
      } else {
        const cx<T> a = *reinterpret_cast<const cx<T> *>(tin + off), b = *reinterpret_cast<const cx<T> *>(tin + off2);
        v[4 * u] = a.x;
        v[4 * u + 1] = a.y;
        v[4 * u + 2] = two ? b.x : (T)0;
        v[4 * u + 3] = two ? b.y : (T)0;
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = i0 + u * nth;
      int t, n;
      mid_load_index(lo, i, p.nin, p.div_nin, lbl, &t, &n);
      if (!(i < tot && t < nl)) continue;
      cx<T> *line = buf + t * p.lpad;
      if (p.cls == CLS_R2C) {
        put(line, n, mk(v[4 * u], v[4 * u + 1]), p, chirp);
      } else {
        cx<T> a = mk(v[4 * u], v[4 * u + 1]), b = mk(v[4 * u + 2], v[4 * u + 3]);
        if (n == 0 || 2 * n == N) {
          a.y = (T)0;
          b.y = (T)0;
        }
        put(line, n, mk(a.x - b.y, a.y + b.x), p, chirp);
        if (n >= 1 && n < (N + 1) / 2) put(line, N - n, mk(a.x + b.y, b.x - a.y), p, chirp);
      }
    }
  }
}

template <typename T>
P3E_HD P3E_INL void pair_store(const PassArgs &p, uint32_t f, int e0, int nb, const cx<T> *res, int tid,
                               int nth) {
  const cx<T> *chirp = reinterpret_cast<const cx<T> *>(p.chirp);
  const uint32_t esz = p.out_cplx ? 2 * sizeof(T) : sizeof(T);
  char *tout = reinterpret_cast<char *>(p.out) + ((long long)e0 * p.os_e + (long long)f * p.os_f) * esz;
  const int osd = (int)p.os_d, ose = (int)p.os_e;
  const int N = p.N;
  const int tot = (p.store_order == ORD_LINE ? nb : p.B) * p.nout;
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
    const cx<T> *line = res + (j >> 1) * p.lpad;
    const uint32_t off = (uint32_t)(k * osd + j * ose) * esz;
    cx<T> z = line[P3E_PAD(k)];
    if (p.bluestein) z = mul(z, chirp[k]);
    cx<T> y;
    if (p.cls == CLS_C2R) {
      y = mk((j & 1) ? z.y : z.x, (T)0);
    } else {
      const int km = k ? N - k : 0;
      cx<T> zm = line[P3E_PAD(km)];
      if (p.bluestein) zm = mul(zm, chirp[km]);
      y = (j & 1) ? mk((T)0.5 * (z.y + zm.y), (T)0.5 * (zm.x - z.x))
                  : mk((T)0.5 * (z.x + zm.x), (T)0.5 * (z.y - zm.y));
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
    if (p.out_cplx) *reinterpret_cast<cx<T> *>(tout + off) = y;
    else *reinterpret_cast<T *>(tout + off) = y.x;
  }
}

// Whole tile.  `each(f)` runs phase f for the calling thread and ends with a
// barrier (device), or runs it for every thread in turn (host build).  Only
// threads tid < nth work; the others take part in the barriers.
template <typename T, int P, typename Run, typename Each>
P3E_HD P3E_INL void mid_tile(const PassArgs &p, uint32_t tile, cx<T> *buf, int nth, Run &run, Each &each,
                     const cx<T> *tw) {
  uint32_t f;
  int e0, nb;
  generic_tile_coords(p, tile, &f, &e0, &nb);
  const int nl = tile_lines(p, nb);
  if (p.zero_fill)
    each([&](int tid) {
      if (tid < nth) generic_zero<T>(p, nl, buf, tid, nth);
    });
  each([&](int tid) {
    if (tid < nth) {
      if (p.pair) pair_load<T>(p, f, e0, nb, buf, tid, nth);
      else mid_load<T>(p, f, e0, nb, buf, tid, nth);
    }
  });
  if (p.cls != CLS_EMPTY) {
    if (!p.bluestein) {
      if (p.sign < 0) mid_fft<T, P, -1>(run, buf, p, nl, tw);
      else mid_fft<T, P, 1>(run, buf, p, nl, tw);
    } else {
      // Bluestein: forward FFT of the chirped sequence, multiply by the
      // spectrum of the conjugate chirp, inverse FFT (the store applies the
      // output chirp)
      mid_fft<T, P, -1>(run, buf, p, nl, tw);
      const cx<T> *chirp_hat = reinterpret_cast<const cx<T> *>(p.chirp_hat);
      each([&](int tid) {
        if (tid < nth)
          for (int i = tid; i < nl * p.L; i += nth) {
            const int line = i / p.L;
            const int m = i - line * p.L;
            buf[line * p.lpad + P3E_PAD(m)] = mul(buf[line * p.lpad + P3E_PAD(m)], chirp_hat[m]);
          }
      });
      mid_fft<T, P, 1>(run, buf, p, nl, tw);
    }
  }
  each([&](int tid) {
    if (tid < nth) {
      if (p.pair) pair_store<T>(p, f, e0, nb, buf, tid, nth);
      else generic_store<T, true>(p, f, e0, nb, buf, tid, nth);
    }
  });
}

}  // namespace p3e
