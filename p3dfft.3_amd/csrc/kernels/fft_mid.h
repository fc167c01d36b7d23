// Single-buffer mixed-radix path of the fused reorder + 1D transform pass.
//
// For work lengths L whose radices are all in {2, 3, 4, 5, 8, 16} (also the
// power-of-two Bluestein lengths): the same load, scatter, output rule and
// store as the generic
// path (fft_core.h), but the radix passes run in place on ONE padded LDS
// buffer, as in the power-of-two path: every thread reads the butterflies
// it owns into registers, barrier, twiddles + butterflies + writes to the
// Stockham positions, barrier.  Half the LDS of the two-buffer generic path
// per line, so twice the lines per tile (longer contiguous runs on a
// strided axis).  The planner sizes the workgroup so that every thread owns
// at most 16 elements of the tile (nb * L <= 16 * nth).
// Shared by the HIP kernels and the host build used by the CPU tests.
#pragma once
#include "fft_core.h"

namespace p3e {

// Radix-R pass at input stride Ns over nb padded lines of length L.
template <typename T, int R> struct MPass {
  static constexpr int NB = (16 + R - 1) / R;  // butterflies per thread, at most
  static constexpr int SLOTS = NB * R;
  cx<T> *buf;
  int L, lpad, nb, Ns, s;
  const cx<T> *tw;
  FastDiv dnbf, dns;

  P3E_HD P3E_INL void read(int tid, int nth, cx<T> *v) const {
    const int nbf = L / R, total = nb * nbf;
#pragma unroll
    for (int m = 0; m < NB; m++) {
      const int b = tid + m * nth;
      if (b < total) {
        const int line = (int)fdiv((uint32_t)b, dnbf);
        const int j = b - line * nbf;
        const cx<T> *sl = buf + line * lpad;
#pragma unroll
        for (int r = 0; r < R; r++) v[m * R + r] = sl[P3E_PAD(j + r * nbf)];
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
          for (int r = 1; r < R; r++) w[r] = mul(w[r], twid(tw, k * r * tstride, s));
        }
        bfly<T, R>(w, s);
        cx<T> *dl = buf + line * lpad;
        const int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) dl[P3E_PAD(base + r * Ns)] = w[r];
      }
    }
  }
};

// FFT of length L on nb lines in place; `run` executes one pass (read half,
// barrier, write half, barrier).
template <typename T, typename Run>
P3E_HD P3E_INL void mid_fft(Run &run, cx<T> *buf, const PassArgs &p, int nb, int s, const cx<T> *tw) {
  int Ns = 1;
  for (int i = 0; i < p.nrad; i++) {
    const int R = p.radix[i];
    const FastDiv dn = p.div_nbf[i], ds = p.div_ns[i];
    switch (R) {
      case 2: { MPass<T, 2> ps{buf, p.L, p.lpad, nb, Ns, s, tw, dn, ds}; run(ps); } break;
      case 3: { MPass<T, 3> ps{buf, p.L, p.lpad, nb, Ns, s, tw, dn, ds}; run(ps); } break;
      case 4: { MPass<T, 4> ps{buf, p.L, p.lpad, nb, Ns, s, tw, dn, ds}; run(ps); } break;
      case 5: { MPass<T, 5> ps{buf, p.L, p.lpad, nb, Ns, s, tw, dn, ds}; run(ps); } break;
      case 8: { MPass<T, 8> ps{buf, p.L, p.lpad, nb, Ns, s, tw, dn, ds}; run(ps); } break;
      default: { MPass<T, 16> ps{buf, p.L, p.lpad, nb, Ns, s, tw, dn, ds}; run(ps); } break;
    }
    Ns *= R;
  }
}

// Whole tile.  `each(f)` runs phase f for the calling thread and ends with a
// barrier (device), or runs it for every thread in turn (host build).  Only
// threads tid < nth work; the others take part in the barriers.
template <typename T, typename Run, typename Each>
P3E_HD P3E_INL void mid_tile(const PassArgs &p, uint32_t tile, cx<T> *buf, int nth, Run &run, Each &each,
                     const cx<T> *tw) {
  uint32_t f;
  int e0, nb;
  generic_tile_coords(p, tile, &f, &e0, &nb);
  if (p.zero_fill)
    each([&](int tid) {
      if (tid < nth) generic_zero<T>(p, nb, buf, tid, nth);
    });
  each([&](int tid) {
    if (tid < nth) generic_load<T>(p, f, e0, nb, buf, tid, nth);
  });
  if (p.cls != CLS_EMPTY) {
    if (!p.bluestein) {
      mid_fft<T>(run, buf, p, nb, p.sign, tw);
    } else {
      // Bluestein: forward FFT of the chirped sequence, multiply by the
      // spectrum of the conjugate chirp, inverse FFT (the store applies the
      // output chirp)
      mid_fft<T>(run, buf, p, nb, -1, tw);
      const cx<T> *chirp_hat = reinterpret_cast<const cx<T> *>(p.chirp_hat);
      each([&](int tid) {
        if (tid < nth)
          for (int i = tid; i < nb * p.L; i += nth) {
            const int line = i / p.L;
            const int m = i - line * p.L;
            buf[line * p.lpad + P3E_PAD(m)] = mul(buf[line * p.lpad + P3E_PAD(m)], chirp_hat[m]);
          }
      });
      mid_fft<T>(run, buf, p, nb, +1, tw);
    }
  }
  each([&](int tid) {
    if (tid < nth) generic_store<T>(p, f, e0, nb, buf, tid, nth);
  });
}

}  // namespace p3e
