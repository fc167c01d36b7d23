// Device code of the fused reorder + 1D transform pass, shared by the HIP
// kernel (fft_pass.hip) and the host build of the same code used by the
// CPU tests (every function is __host__ __device__).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define P3E_HD __host__ __device__
#define P3E_INL __forceinline__
#else
#define P3E_HD
#define P3E_INL inline
#endif
// a lambda whose body must be inlined into its caller (a lambda passed to a
// phase runner that is not inlined would take the kernel's arguments off
// registers into scratch memory)
#define P3E_LAMBDA_INL __attribute__((always_inline))

#include "pass_args.h"

namespace p3e {

// Aligned to its size, so that one complex value moves with one 8-byte
// (FP32) or 16-byte (FP64) LDS or memory instruction, and two values at
// compile-time offsets with one paired LDS instruction.
template <typename T> struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T> P3E_HD P3E_INL cx<T> mk(T a, T b) { return cx<T>{a, b}; }
template <typename T> P3E_HD P3E_INL cx<T> add(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> P3E_HD P3E_INL cx<T> sub(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> P3E_HD P3E_INL cx<T> mul(cx<T> a, cx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> P3E_HD P3E_INL cx<T> scl(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T> P3E_HD P3E_INL cx<T> conj(cx<T> a) { return mk(a.x, -a.y); }
// multiply by -i*sgn: sgn=-1 -> *(i) ; sgn=+1 -> *(-i).  Used as "rotate by the
// forward quarter turn" when sgn is the DFT sign.
template <typename T> P3E_HD P3E_INL cx<T> rot(cx<T> a, int sgn) {
  // a * (sgn * i)
  return sgn > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}

P3E_HD P3E_INL uint32_t fdiv(uint32_t n, const FastDiv &f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return (__umulhi(n, f.mul) + n) >> f.shift;
#else
  return (uint32_t)((((uint64_t)n * f.mul) >> 32) + n) >> f.shift;
#endif
}

// ---------------------------------------------------------------- butterflies
// In-register DFT of size R with sign s: X[q] = sum_r v[r] exp(s 2 pi i r q / R).

template <typename T> P3E_HD P3E_INL void bf2(cx<T> *v) {
  cx<T> a = v[0], b = v[1];
  v[0] = add(a, b);
  v[1] = sub(a, b);
}

template <typename T> P3E_HD P3E_INL void bf4(cx<T> *v, int s) {
  cx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
  cx<T> t2 = add(v[1], v[3]), t3 = rot(sub(v[1], v[3]), s);
  v[0] = add(t0, t2);
  v[2] = sub(t0, t2);
  v[1] = add(t1, t3);
  v[3] = sub(t1, t3);
}

template <typename T> P3E_HD P3E_INL void bf3(cx<T> *v, int s) {
  const T c = (T)-0.5;
  const T sn = (T)0.86602540378443864676372317075294;  // sin(2 pi / 3)
  cx<T> t = add(v[1], v[2]);
  cx<T> d = rot(sub(v[1], v[2]), s);  // s*i*(a1-a2)
  cx<T> m = mk(v[0].x + c * t.x, v[0].y + c * t.y);
  v[0] = add(v[0], t);
  v[1] = mk(m.x + sn * d.x, m.y + sn * d.y);
  v[2] = mk(m.x - sn * d.x, m.y - sn * d.y);
}

template <typename T> P3E_HD P3E_INL void bf5(cx<T> *v, int s) {
  const T c1 = (T)0.30901699437494742410229341718282;   // cos(2 pi/5)
  const T c2 = (T)-0.80901699437494742410229341718282;  // cos(4 pi/5)
  const T s1 = (T)0.95105651629515357211643933337938;   // sin(2 pi/5)
  const T s2 = (T)0.58778525229247312916870595463907;   // sin(4 pi/5)
  cx<T> p1 = add(v[1], v[4]), m1 = sub(v[1], v[4]);
  cx<T> p2 = add(v[2], v[3]), m2 = sub(v[2], v[3]);
  cx<T> a0 = v[0];
  cx<T> r1 = mk(a0.x + c1 * p1.x + c2 * p2.x, a0.y + c1 * p1.y + c2 * p2.y);
  cx<T> r2 = mk(a0.x + c2 * p1.x + c1 * p2.x, a0.y + c2 * p1.y + c1 * p2.y);
  // X1 = r1 + s*i*(s1 m1 + s2 m2), X4 = r1 - ..., X2 = r2 + s*i*(s2 m1 - s1 m2), X3 = r2 - ...
  cx<T> q1 = rot(mk(s1 * m1.x + s2 * m2.x, s1 * m1.y + s2 * m2.y), s);
  cx<T> q2 = rot(mk(s2 * m1.x - s1 * m2.x, s2 * m1.y - s1 * m2.y), s);
  v[0] = add(a0, add(p1, p2));
  v[1] = add(r1, q1);
  v[4] = sub(r1, q1);
  v[2] = add(r2, q2);
  v[3] = sub(r2, q2);
}

// radix 8 = 2 x 4: X[2k] = DFT4(a_r + a_{r+4}), X[2k+1] = DFT4((a_r - a_{r+4}) w8^r)
template <typename T> P3E_HD P3E_INL void bf8(cx<T> *v, int s) {
  const T h = (T)0.70710678118654752440084436210485;
  cx<T> u[4], w[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    u[r] = add(v[r], v[r + 4]);
    w[r] = sub(v[r], v[r + 4]);
  }
  // w8^r = exp(s 2 pi i r / 8)
  w[1] = mk(h * (w[1].x - (T)s * w[1].y), h * (w[1].y + (T)s * w[1].x));
  w[2] = rot(w[2], s);
  w[3] = mk(h * (-w[3].x - (T)s * w[3].y), h * (-w[3].y + (T)s * w[3].x));
  bf4(u, s);
  bf4(w, s);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    v[2 * k] = u[k];
    v[2 * k + 1] = w[k];
  }
}

// radix 16 = 2 x 8
template <typename T> P3E_HD P3E_INL void bf16(cx<T> *v, int s) {
  // cos(pi r / 8) and sin(pi r / 8), r = 0..7
  const T C[8] = {(T)1.0,
                  (T)0.92387953251128675612818318939679,
                  (T)0.70710678118654752440084436210485,
                  (T)0.38268343236508977172845998403040,
                  (T)0.0,
                  (T)-0.38268343236508977172845998403040,
                  (T)-0.70710678118654752440084436210485,
                  (T)-0.92387953251128675612818318939679};
  const T S[8] = {(T)0.0,
                  (T)0.38268343236508977172845998403040,
                  (T)0.70710678118654752440084436210485,
                  (T)0.92387953251128675612818318939679,
                  (T)1.0,
                  (T)0.92387953251128675612818318939679,
                  (T)0.70710678118654752440084436210485,
                  (T)0.38268343236508977172845998403040};
  cx<T> u[8], w[8];
#pragma unroll
  for (int r = 0; r < 8; r++) {
    u[r] = add(v[r], v[r + 8]);
    w[r] = sub(v[r], v[r + 8]);
  }
#pragma unroll
  for (int r = 1; r < 8; r++) {
    // exp(s i pi r / 8) = cos(pi r/8) + s i sin(pi r/8)
    const T c = C[r], sn = (T)s * S[r];
    w[r] = mk(w[r].x * c - sn * w[r].y, w[r].y * c + sn * w[r].x);
  }
  bf8(u, s);
  bf8(w, s);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    v[2 * k] = u[k];
    v[2 * k + 1] = w[k];
  }
}

// cos and sin of 2 pi m / P for m = 0 .. P-1, evaluated at compile time:
// the angle is reduced to an octant exactly in integers (8m = o P + r), the
// remaining angle (pi/4) r/P or (pi/4)(P-r)/P lies in [0, pi/4], and sin/cos
// there are Taylor series summed to x^27 (the next term is below 1e-30).
constexpr double cx_pi4 = 0.78539816339744830961566084581987572;
constexpr void taylor_sincos(double x, double &sn, double &cs) {
  double t = x, s = 0.0;
  for (int k = 1; k < 28; k += 2) {
    s += t;
    t *= -x * x / ((k + 1) * (k + 2));
  }
  double u = 1.0, c = 0.0;
  for (int k = 0; k < 28; k += 2) {
    c += u;
    u *= -x * x / ((k + 1) * (k + 2));
  }
  sn = s;
  cs = c;
}
template <int P> struct PrimeRoots {
  double c[P], s[P];
  constexpr PrimeRoots() : c(), s() {
    for (int m = 0; m < P; m++) {
      const int o = (8 * m) / P, r = 8 * m - o * P;
      const bool odd = o & 1;
      double sx = 0.0, cx = 0.0;
      taylor_sincos(cx_pi4 * (double)(odd ? P - r : r) / (double)P, sx, cx);
      // angle = o pi/4 + x (o even) or (o+1) pi/4 - x (o odd)
      double cc = 0.0, ss = 0.0;
      switch (o) {
        case 0: cc = cx; ss = sx; break;
        case 1: cc = sx; ss = cx; break;
        case 2: cc = -sx; ss = cx; break;
        case 3: cc = -cx; ss = sx; break;
        case 4: cc = -cx; ss = -sx; break;
        case 5: cc = -sx; ss = -cx; break;
        case 6: cc = sx; ss = -cx; break;
        default: cc = cx; ss = -sx; break;
      }
      c[m] = cc;
      s[m] = ss;
    }
  }
};
template <int P> constexpr PrimeRoots<P> prime_roots{};

// Odd prime P: with a_j = v_j + v_{P-j}, b_j = v_j - v_{P-j} (j = 1..H,
// H = (P-1)/2) and theta = 2 pi j k / P,
//   X_k     = v_0 + sum_j a_j cos(theta) + s i sum_j b_j sin(theta)
//   X_{P-k} = v_0 + sum_j a_j cos(theta) - s i sum_j b_j sin(theta)
// so H^2 real-by-complex products for each of the two sums.  put(q, X_q)
// receives each output as soon as it is formed (the caller stores it), so
// the outputs need no registers of their own.
template <typename T, int P, typename Put> P3E_HD P3E_INL void bf_prime(const cx<T> *v, int s, Put put) {
  constexpr int H = (P - 1) / 2;
  cx<T> a[H], b[H];
  const cx<T> x0 = v[0];
  cx<T> y0 = x0;
#pragma unroll
  for (int j = 0; j < H; j++) {
    a[j] = add(v[j + 1], v[P - 1 - j]);
    b[j] = sub(v[j + 1], v[P - 1 - j]);
    y0 = add(y0, a[j]);
  }
  put(0, y0);
#pragma unroll
  for (int k = 1; k <= H; k++) {
    cx<T> A = x0, Bs = mk((T)0, (T)0);
#pragma unroll
    for (int j = 1; j <= H; j++) {
      const int m = (j * k) % P;
      const T c = (T)prime_roots<P>.c[m], sn = (T)prime_roots<P>.s[m];
      A = mk(A.x + a[j - 1].x * c, A.y + a[j - 1].y * c);
      Bs = mk(Bs.x + b[j - 1].x * sn, Bs.y + b[j - 1].y * sn);
    }
    const cx<T> r = rot(Bs, s);
    put(k, add(A, r));
    put(P - k, sub(A, r));
  }
}

template <int R> constexpr bool is_base_radix = R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16;

template <typename T, int R> P3E_HD P3E_INL void bfly(cx<T> *v, int s) {
  if constexpr (R == 2) bf2(v);
  else if constexpr (R == 3) bf3(v, s);
  else if constexpr (R == 4) bf4(v, s);
  else if constexpr (R == 5) bf5(v, s);
  else if constexpr (R == 8) bf8(v, s);
  else if constexpr (R == 16) bf16(v, s);
  else bf_prime<T, R>(v, s, [v](int q, cx<T> x) { v[q] = x; });
}

template <typename T> P3E_HD P3E_INL cx<T> twid(const cx<T> *tw, int idx, int s) {
  cx<T> w = tw[idx];
  return s < 0 ? w : conj(w);
}

// One Stockham radix-R pass over nb lines: src -> dst (padded LDS lines).
// dnbf, dns: division by L / R and by Ns.
template <typename T, int R>
P3E_HD P3E_INL void stockham_pass(const cx<T> *src, cx<T> *dst, int L, int lpad, int nb, int Ns, int s,
                                  const cx<T> *tw, int tid, int nth, const FastDiv &dnbf, const FastDiv &dns) {
  const int nbf = L / R;
  const int tstride = L / (Ns * R);
  const int total = nb * nbf;
  for (int t = tid; t < total; t += nth) {
    const int line = (int)fdiv((uint32_t)t, dnbf);
    const int j = t - line * nbf;
    const cx<T> *sl = src + line * lpad;
    cx<T> *dl = dst + line * lpad;
    const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = sl[P3E_PAD(j + r * nbf)];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; r++) v[r] = mul(v[r], twid(tw, k * r * tstride, s));
    }
    bfly<T, R>(v, s);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) dl[P3E_PAD(base + r * Ns)] = v[r];
  }
}

// Odd prime radix R (7..31): a direct DFT, one thread per output.  Output o
// of the pass is (line, butterfly j = (o mod L) / R, q = o mod R):
// X_q = sum_r x_r W^{rq}, x_r = input r of butterfly j times its Stockham
// twiddle, W = e^{s 2 pi i / R} = the table root at L/R.  Consecutive threads
// share a butterfly, so its inputs are LDS broadcasts.
template <typename T>
P3E_HD void stockham_pass_generic(const cx<T> *src, cx<T> *dst, int L, int lpad, int nb, int Ns, int R, int s,
                                  const cx<T> *tw, int tid, int nth, const FastDiv &dnbf, const FastDiv &dns) {
  const int nbf = L / R;
  const int tstride = L / (Ns * R);
  const int wstride = L / R;
  const int total = nb * L;
  for (int o = tid; o < total; o += nth) {
    const int b = o / R, q = o - b * R;
    const int line = (int)fdiv((uint32_t)b, dnbf);
    const int j = b - line * nbf;
    const int k = j - (int)fdiv((uint32_t)j, dns) * Ns;
    const cx<T> *sl = src + line * lpad;
    cx<T> acc = sl[P3E_PAD(j)];
    int e = 0;
    for (int r = 1; r < R; r++) {
      cx<T> x = sl[P3E_PAD(j + r * nbf)];
      if (Ns > 1) x = mul(x, twid(tw, k * r * tstride, s));
      e += q;
      if (e >= R) e -= R;
      acc = add(acc, mul(x, twid(tw, e * wstride, s)));
    }
    dst[line * lpad + P3E_PAD((j - k) * R + k + q * Ns)] = acc;
  }
}

// Full FFT of length L on nb lines; returns the buffer holding the result.
template <typename T, typename Bar>
P3E_HD cx<T> *lds_fft(cx<T> *a, cx<T> *b, const PassArgs &p, int nb, int s, const cx<T> *tw,
                      int tid, int nth, Bar bar) {
  int Ns = 1;
  for (int i = 0; i < p.nrad; i++) {
    const int R = p.radix[i];
    const FastDiv &dn = p.div_nbf[i], &ds = p.div_ns[i];
    switch (R) {
      case 2: stockham_pass<T, 2>(a, b, p.L, p.lpad, nb, Ns, s, tw, tid, nth, dn, ds); break;
      case 3: stockham_pass<T, 3>(a, b, p.L, p.lpad, nb, Ns, s, tw, tid, nth, dn, ds); break;
      case 4: stockham_pass<T, 4>(a, b, p.L, p.lpad, nb, Ns, s, tw, tid, nth, dn, ds); break;
      case 5: stockham_pass<T, 5>(a, b, p.L, p.lpad, nb, Ns, s, tw, tid, nth, dn, ds); break;
      case 8: stockham_pass<T, 8>(a, b, p.L, p.lpad, nb, Ns, s, tw, tid, nth, dn, ds); break;
      case 16: stockham_pass<T, 16>(a, b, p.L, p.lpad, nb, Ns, s, tw, tid, nth, dn, ds); break;
      default: stockham_pass_generic<T>(a, b, p.L, p.lpad, nb, Ns, R, s, tw, tid, nth, dn, ds); break;
    }
    bar();
    cx<T> *t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// ---------------------------------------------------------------- load rules

template <typename T>
P3E_HD P3E_INL void put(cx<T> *line, int m, cx<T> v, const PassArgs &p, const cx<T> *chirp) {
  if (p.bluestein) v = mul(v, chirp[m]);
  line[P3E_PAD(m)] = v;
}

// Scatter input element n of a line into the work sequence (length W).
template <typename T>
P3E_HD P3E_INL void scatter(cx<T> *line, int n, cx<T> x, const PassArgs &p, const cx<T> *chirp,
                                        const cx<T> *pre) {
  const int N = p.N;
  switch (p.cls) {
    case CLS_EMPTY:
    case CLS_C2C:
    case CLS_R2C:
      put(line, n, x, p, chirp);
      break;
    case CLS_C2R:
      put(line, n, x, p, chirp);
      if (n >= 1 && n < (N + 1) / 2) put(line, N - n, conj(x), p, chirp);
      break;
    case CLS_R2R: {
      const int W = p.W;
      switch (p.r2r_kind) {
        case REDFT00:
          put(line, n, x, p, chirp);
          if (n >= 1 && n <= N - 2) put(line, W - n, x, p, chirp);
          break;
        case RODFT00:
          put(line, n + 1, x, p, chirp);
          put(line, W - n - 1, mk(-x.x, -x.y), p, chirp);
          break;
        case REDFT10:
          put(line, n, x, p, chirp);
          put(line, 2 * N - 1 - n, x, p, chirp);
          break;
        case RODFT10:
          put(line, n, x, p, chirp);
          put(line, 2 * N - 1 - n, mk(-x.x, -x.y), p, chirp);
          break;
        case REDFT01:
          put(line, n, mul(x, pre[n]), p, chirp);
          if (n >= 1) put(line, 2 * N - n, mul(mk(-x.x, -x.y), pre[2 * N - n]), p, chirp);
          break;
        case RODFT01:
          if (n <= N - 2) {
            put(line, n + 1, mul(x, pre[n + 1]), p, chirp);
            put(line, 2 * N - n - 1, mul(x, pre[2 * N - n - 1]), p, chirp);
          } else {
            put(line, N, mul(x, pre[N]), p, chirp);
          }
          break;
        case REDFT11:
          put(line, n, mul(x, pre[n]), p, chirp);
          put(line, 2 * N - 1 - n, mul(mk(-x.x, -x.y), pre[2 * N - 1 - n]), p, chirp);
          break;
        case RODFT11:
          put(line, n, mul(x, pre[n]), p, chirp);
          put(line, 2 * N - 1 - n, mul(x, pre[2 * N - 1 - n]), p, chirp);
          break;
      }
      break;
    }
  }
}

// Poisson factor -1/|k|^2 (0 at k = 0) of output indices (kd, ke, kf).
template <typename T> P3E_HD P3E_INL T poisson_factor(const PassArgs &p, int kd, int ke, int kf) {
  const int idx[3] = {kd, ke, kf};
  T s = (T)0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const int n = p.kn[r], i = idx[r];
    const int k = (p.khalf[r] || 2 * i < n) ? i : i - n;
    s += (T)k * (T)k;
  }
  return s > (T)0 ? (T)-1 / s : (T)0;
}

// Phases of one tile of the generic path, shared with the single-buffer
// mixed-radix path (fft_mid.h).

P3E_HD P3E_INL void generic_tile_coords(const PassArgs &p, uint32_t tile, uint32_t *f, int *e0, int *nb) {
  *f = fdiv(tile, p.div_ntile);
  *e0 = (int)(tile - *f * p.ntile_e) * p.B;
  *nb = (p.B < p.ne - *e0) ? p.B : (p.ne - *e0);
}

// Clear the lines when the scatter does not write every work element.
template <typename T>
P3E_HD P3E_INL void generic_zero(const PassArgs &p, int nb, cx<T> *bufA, int tid, int nth) {
  const int tot = nb * p.L;
  for (int i = tid; i < tot; i += nth) {
    const int line = i / p.L;
    bufA[line * p.lpad + P3E_PAD(i - line * p.L)] = mk((T)0, (T)0);
  }
}

// Load the tile and scatter each element into its line's work sequence.
template <typename T>
P3E_HD P3E_INL void generic_load(const PassArgs &p, uint32_t f, int e0, int nb, cx<T> *bufA, int tid, int nth) {
  const cx<T> *chirp = reinterpret_cast<const cx<T> *>(p.chirp);
  const cx<T> *pre = reinterpret_cast<const cx<T> *>(p.pre);
  const long long base = (long long)e0 * p.is_e + (long long)f * p.is_f;
  const int tot = (p.load_order == ORD_LINE ? nb : p.B) * p.nin;
  for (int i = tid; i < tot; i += nth) {
    int n, j;
    if (p.load_order == ORD_LINE) {
      j = (int)fdiv((uint32_t)i, p.div_nin);
      n = i - j * p.nin;
    } else {
      // nb may be < B on the last tile: lines j >= nb are skipped
      j = i & (p.B - 1);
      n = i >> p.logB;
      if (j >= nb) continue;
    }
    const long long off = base + (long long)n * p.is_d + (long long)j * p.is_e;
    cx<T> x;
    if (p.in_cplx) {
      x = reinterpret_cast<const cx<T> *>(p.in)[off];
    } else {
      x = mk(reinterpret_cast<const T *>(p.in)[off], (T)0);
    }
    scatter(bufA + j * p.lpad, n, x, p, chirp, pre);
  }
}

// Output rule (chirp, post-twiddle, derivative, spectral operator, scale)
// and store of the transformed lines in `res`.
// OFF32: 32-bit offsets within the tile (the single-buffer path, whose
// passes the planner admits only when they stay below 2^31 bytes).
template <typename T, bool OFF32 = false>
P3E_HD P3E_INL void generic_store(const PassArgs &p, uint32_t f, int e0, int nb, const cx<T> *res, int tid,
                                  int nth) {
  const cx<T> *chirp = reinterpret_cast<const cx<T> *>(p.chirp);
  const cx<T> *post = reinterpret_cast<const cx<T> *>(p.post);
  const long long base = (long long)e0 * p.os_e + (long long)f * p.os_f;
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
    const int m = k + p.post_off;
    cx<T> y = res[j * p.lpad + P3E_PAD(m)];
    if (p.bluestein) y = mul(y, chirp[m]);
    if (p.post_mode) y = mul(y, post[k]);
    if (p.deriv) {
      T kk;
      if (k < p.dmid) kk = (T)k;
      else if (k == p.dmid) kk = (T)0;
      else kk = (T)(k - p.dg);
      y = mk(-kk * y.y, kk * y.x);
    }
    if (p.spec_op) y = scl(y, poisson_factor<T>(p, k, e0 + j, (int)f));
    if (p.has_scale) y = scl(y, (T)p.scale);
    if (OFF32) {
      const uint32_t esz = p.out_cplx ? 2 * sizeof(T) : sizeof(T);
      char *t = reinterpret_cast<char *>(p.out) + base * esz;
      const uint32_t off = (uint32_t)(k * (int)p.os_d + j * (int)p.os_e) * esz;
      if (p.out_cplx) *reinterpret_cast<cx<T> *>(t + off) = y;
      else *reinterpret_cast<T *>(t + off) = y.x;
    } else {
      const long long off = base + (long long)k * p.os_d + (long long)j * p.os_e;
      if (p.out_cplx) {
        reinterpret_cast<cx<T> *>(p.out)[off] = y;
      } else {
        reinterpret_cast<T *>(p.out)[off] = y.x;
      }
    }
  }
}

// One workgroup's work on one tile of the generic (two-buffer) path.
// Phases are separated by bar(); the host test build runs each phase with a
// single thread (tid 0, nth 1) and a no-op barrier.  tw_table: the twiddle
// table to use (p.tw, or a copy in LDS).
template <typename T, typename Bar>
P3E_HD void run_tile(const PassArgs &p, uint32_t tile, cx<T> *bufA, cx<T> *bufB, int tid, int nth, Bar bar,
                     const cx<T> *tw_table) {
  uint32_t f;
  int e0, nb;
  generic_tile_coords(p, tile, &f, &e0, &nb);
  const cx<T> *chirp_hat = reinterpret_cast<const cx<T> *>(p.chirp_hat);
  if (p.zero_fill) {
    generic_zero<T>(p, nb, bufA, tid, nth);
    bar();
  }
  generic_load<T>(p, f, e0, nb, bufA, tid, nth);
  bar();
  cx<T> *res = bufA;
  if (p.cls != CLS_EMPTY) {
    if (!p.bluestein) {
      res = lds_fft<T>(bufA, bufB, p, nb, p.sign, tw_table, tid, nth, bar);
    } else {
      res = lds_fft<T>(bufA, bufB, p, nb, -1, tw_table, tid, nth, bar);
      const int tot = nb * p.L;
      for (int i = tid; i < tot; i += nth) {
        const int line = i / p.L;
        const int m = i - line * p.L;
        res[line * p.lpad + P3E_PAD(m)] = mul(res[line * p.lpad + P3E_PAD(m)], chirp_hat[m]);
      }
      bar();
      cx<T> *other = (res == bufA) ? bufB : bufA;
      res = lds_fft<T>(res, other, p, nb, +1, tw_table, tid, nth, bar);
    }
  }
  generic_store<T>(p, f, e0, nb, res, tid, nth);
}

}  // namespace p3e
