// Pass planning (plain C++).  See pass_plan.h.
#include "pass_plan.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>

namespace p3e {

// ------------------------------------------------------------------ errors
// per thread: the Poisson service runs its own thread
static thread_local char g_err[1024];
static thread_local int g_err_code = 0;

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  g_err_code = 1;
  fprintf(stderr, "p3dfft3_hip: %s\n", g_err);
}
const char *last_error() { return g_err; }
int last_error_code() { return g_err_code; }
void clear_error() {
  g_err[0] = 0;
  g_err_code = 0;
}

// ------------------------------------------------------------------ helpers
static int stride1_axis(const int mo[3]) {
  for (int a = 0; a < 3; a++)
    if (mo[a] == 0) return a;
  return 0;
}

// Allocated extent of logical axis a (the pitch for the stride-1 axis).
static int extent(const Layout &l, int a) { return (l.mo[a] == 0 && l.pitch > l.dims[a]) ? l.pitch : l.dims[a]; }

size_t layout_elems(const Layout &l) { return (size_t)extent(l, 0) * extent(l, 1) * extent(l, 2); }
size_t elem_bytes(int prec, int dt) { return (size_t)prec * dt; }

static void strides_of(const Layout &l, long long st[3]) {
  for (int a = 0; a < 3; a++) {
    long long s = 1;
    for (int b = 0; b < 3; b++)
      if (l.mo[b] < l.mo[a]) s *= extent(l, b);
    st[a] = s;
  }
}

// Rows (stride-1 runs) whose length is a multiple of 256 bytes put the
// rows of a large array at strides with many factors of two; a pass that
// reads or writes such an array across rows (a tile with one element of
// each of many rows at once) then concentrates its traffic on few memory
// channels.  Measured on MI355X for a C2C pass over 512^3 FP32 along the
// 4 KiB-strided axis: 4.9 TB/s dense, 5.8 TB/s with rows padded by 128
// bytes; along the 2 MiB-strided axis 3.6 TB/s dense, 4.5 TB/s with the
// middle axis padded.  Intermediate arrays of at least g_pad_min bytes get
// 128 bytes of padding per row.
static size_t g_pad_min = (size_t)32 << 20;
void set_pad_threshold(size_t bytes) { g_pad_min = bytes; }

static int padded_pitch(int prec, const Layout &l) {
  const int a = stride1_axis(l.mo);
  const size_t esz = elem_bytes(prec, l.dt);
  const size_t row = (size_t)l.dims[a] * esz;
  if (row % 256 != 0 || layout_elems(l) * esz < g_pad_min) return 0;
  return l.dims[a] + (int)(128 / esz);
}


// Radix plan for length n: 8s (with one 16 or 4 absorbing the remainder) for
// the power-of-two part, then odd primes up to maxp, largest first.  Returns
// false if a larger prime factor remains (the pass then uses Bluestein).
static bool factor(int n, std::vector<int> &rad, int maxp = 31) {
  rad.clear();
  int m = n, p2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    p2++;
  }
  std::vector<int> twos;
  while (p2 >= 3) {
    twos.push_back(8);
    p2 -= 3;
  }
  if (p2 == 2) twos.push_back(4);
  if (p2 == 1) {
    if (!twos.empty()) twos[0] = 16;
    else twos.push_back(2);
  }
  std::vector<int> odd;
  for (int q : {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47}) {
    if (q > maxp) break;
    while (m % q == 0) {
      odd.push_back(q);
      m /= q;
    }
  }
  if (m != 1) return false;
  std::sort(odd.begin(), odd.end(), [](int a, int b) { return a > b; });
  rad = twos;
  rad.insert(rad.end(), odd.begin(), odd.end());
  return (int)rad.size() <= P3E_MAX_RADIX_PASSES;
}

static int next_pow2(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

static const long double PI_L = 3.141592653589793238462643383279502884L;

// exp(i * pi * num / den), num reduced modulo 2*den in integers first.
static cld expi_pi(long long num, long long den) {
  long long m = num % (2 * den);
  if (m < 0) m += 2 * den;
  long double a = PI_L * (long double)m / (long double)den;
  return cld(cosl(a), sinl(a));
}

// Twiddles of the power-of-two path, pass by pass: for
// the radix sequence rad, pass i (radix R, stride Ns = the product of the
// earlier radices) holds w^(k r), w = e^(-2 pi i / (Ns R)), at entry
// k S + r - 1 (k < Ns, 1 <= r < R), S = tw_stride(R) = R - 1 rounded up to
// odd.  Consecutive threads of a pass read entries an odd distance apart;
// in one table of length L they would read entries k r L / (Ns R) apart, a
// multiple of the LDS bank period for many (Ns, r).  For power-of-two
// radices the passes' sizes sum to L - 1, and one unused entry pads the
// table to L.
static void pass_twiddles(const int *rad, int nrad, std::vector<cld> &tw) {
  tw.clear();
  long long Ns = 1;
  for (int i = 0; i < nrad; i++) {
    const int R = rad[i], S = p3e::tw_stride(R);
    for (long long k = 0; k < Ns; k++)
      for (int r = 1; r <= S; r++) tw.push_back(r < R ? expi_pi(-2LL * k * r, Ns * R) : cld(1, 0));
    Ns *= R;
  }
  tw.push_back(cld(1, 0));
}

// In-place radix-2 DFT (sign -1) in long double, n a power of two.  Plan-time
// only: it computes the spectrum of the Bluestein chirp, a constant table.
static void host_fft_pow2(std::vector<cld> &a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    for (size_t i = 0; i < n; i += len) {
      for (size_t k = 0; k < len / 2; k++) {
        cld w = expi_pi(-(long long)(2 * k), (long long)len);
        cld u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

static const char *cls_name(int c) {
  switch (c) {
    case CLS_EMPTY: return "copy";
    case CLS_R2C: return "r2c";
    case CLS_C2R: return "c2r";
    case CLS_C2C: return "c2c";
    case CLS_R2R: return "r2r";
  }
  return "?";
}

// Divisions of the generic path's radix passes (see PassArgs::div_nbf).
static void set_radix_divs(PassArgs &a) {
  long ns = 1;
  for (int i = 0; i < a.nrad; i++) {
    a.div_nbf[i] = make_fastdiv((uint32_t)(a.L / a.radix[i]));
    a.div_ns[i] = make_fastdiv((uint32_t)ns);
    ns *= a.radix[i];
  }
}

// LDS pitch of a generic-path line of L elements: room for the padded line,
// rounded up to 1 mod 32 elements for long lines (consecutive lines then
// start on different banks for the tile-order load and store).
static int generic_pitch(int L) {
  const int need = P3E_PAD(L - 1) + 1;
  return L >= 32 ? ((need + 31) & ~31) + 1 : need;
}

static const size_t LDS_BUDGET = 80 * 1024;
static const size_t LDS_MAX = 160 * 1024;


// LDS line pitch of the compile-time-length path (setup_fast)
static int fast_pitch(int prec, int M) {
  const int g = prec == 8 ? 16 : 32;
  return ((M + M / 16 + g - 1) / g) * g + 1;
}

static bool g_fast_enabled = true;
static bool g_dct_split = true;
void set_dct_split(bool on) { g_dct_split = on; }
void set_fast_path(bool enable) { g_fast_enabled = enable; }

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static bool radix_supported(int r);

static bool mid_base_radix(int r) { return r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 16; }

// Largest odd prime radix of the single-buffer kernels: 47 in FP32; 29 in
// FP64, where the registers of a larger butterfly (two doubles per complex
// value) exceed the 256 VGPRs of a wave and would spill.
static int mid_max_prime(int prec) { return prec == 8 ? 29 : 47; }

// The odd prime radix of a radix list that the single-buffer path takes: 0
// if every radix is 2, 3, 4, 5, 8 or 16, the prime if there is one distinct
// other radix and it is a prime from 7 to mid_max_prime(prec), -1 otherwise.
static int mid_prime_of(const int *rad, int nrad, int prec) {
  int p = 0;
  for (int i = 0; i < nrad; i++) {
    const int r = rad[i];
    if (mid_base_radix(r)) continue;
    if (r < 7 || r > mid_max_prime(prec) || (p != 0 && r != p)) return -1;
    for (int q = 3; q * q <= r; q += 2)
      if (r % q == 0) return -1;
    p = r;
  }
  return p;
}

// Switch an already described generic pass to the single-buffer mixed-radix
// path (fft_mid.h) when every radix of the work length (the Bluestein length
// for a large prime) is 2, 3, 4, 5, 8 or 16, plus — for a standalone pass
// kernel, not the batch kernel — at most one distinct odd prime 7..47 (its
// own kernel instantiation).  Other passes stay on the two-buffer generic
// path.  Tile: B lines with B * L <= 16 threads' worth of elements, as
// large as LDS and the workgroup size (512 threads, or the batch kernel's)
// allow; the twiddles go to LDS after the lines when they fit.
static void setup_mid(int prec, const OpSpec &op, PassDesc &ps, int batch_threads) {
  PassArgs &a = ps.args;
  if (op.cls == CLS_EMPTY || a.nrad < 1) return;
  const int prime = mid_prime_of(a.radix, a.nrad, prec);
  if (prime < 0 || (prime > 0 && batch_threads > 0)) return;
  const int L = a.L;
  const long maxth = batch_threads > 0 ? batch_threads : P3E_MID_THREADS(prec, prime);
  const size_t esz = 2 * (size_t)prec;
  const size_t per_line = (size_t)a.lpad * esz;
  size_t twl = batch_threads > 0 ? 0 : ps.tw.size() * esz;
  if (per_line + twl > LDS_MAX) twl = 0;
  // R2C / C2R: two real lines per LDS line (fft_mid.h, pair_load)
  const bool pair = op.cls == CLS_R2C || op.cls == CLS_C2R;
  auto lines = [&](int b) { return pair ? (b + 1) / 2 : b; };
  int B = 1;
  const int Bmax = std::min(pair ? 64 : 32, next_pow2(a.ne));
  const long E = P3E_MID_E;
  while (B * 2 <= Bmax && per_line * lines(B * 2) + twl <= P3E_MID_LDS && (long)lines(B * 2) * L <= E * maxth)
    B *= 2;
  if (per_line * lines(B) + twl > LDS_MAX || (long)lines(B) * L > E * maxth) return;
  int th = (int)(((long)lines(B) * L + E - 1) / E);
  th = std::max(64, (th + 63) & ~63);
  if (th > maxth) return;
  // the kernel addresses a tile's elements with 32-bit byte offsets from
  // the tile's first element (below 2^31 on both sides)
  const double in_esz = (double)prec * (a.in_cplx ? 2 : 1), out_esz = (double)prec * (a.out_cplx ? 2 : 1);
  if (((double)a.nin * a.is_d + (double)B * a.is_e + 1) * in_esz >= 2147483648.0 ||
      ((double)a.nout * a.os_d + (double)B * a.os_e + 1) * out_esz >= 2147483648.0)
    return;
  a.mid = 1;
  a.mid_prime = prime;
  a.pair = pair;
  a.tw_in_lds = twl > 0;
  a.ntw = (int)ps.tw.size();
  a.B = B;
  a.logB = 0;
  while ((1 << a.logB) < B) a.logB++;
  a.ntile_e = (a.ne + B - 1) / B;
  a.div_ntile = make_fastdiv((uint32_t)a.ntile_e);
  ps.lds_bytes = per_line * lines(B) + twl;
  ps.threads = th;
  a.nthreads = th;
  ps.nblocks = (unsigned int)((unsigned long long)a.ntile_e * (unsigned long long)a.nf);
}

// Complex DCT-I of P + 1 points on the split form of the compile-time-length
// path (fft_fast.h, dct1_tile): a P-point FFT and a P/2-point inverse FFT
// per line instead of the 2P-point FFT of the even extension.  Tile: B
// lines, (P + P/2) B / P3E_FAST_E threads, LDS for both line sets and the
// tables.  Returns false (the pass unchanged) when P has no split form or
// the tile does not fit.
static bool setup_dct1_split(int prec, PassDesc &ps) {
  PassArgs &a = ps.args;
  const int P = a.N - 1, L = P / 2;
  int ra[P3E_MAX_RADIX_PASSES], rb[P3E_MAX_RADIX_PASSES];
  const int nrad = dct1_radices(P, ra, rb);
  if (nrad == 0) return false;
  const size_t esz = 2 * (size_t)prec;
  const int lp = fast_pitch(prec, P), lp2 = fast_pitch(prec, L);
  std::vector<cld> twa, twb;
  pass_twiddles(ra, nrad, twa);
  pass_twiddles(rb, nrad, twb);
  const size_t ntw = twa.size() + twb.size() + (size_t)(L / 2 + 1);
  auto lds = [&](int b) { return (size_t)b * (lp + lp2) * esz + ntw * esz; };
  auto threads = [&](int b) { return (long)(P + L) * b / P3E_FAST_E; };
  const int Bmax = std::min(64, next_pow2(a.ne));
  int B = 1;
  while (B * 2 <= Bmax && lds(B * 2) <= LDS_BUDGET && threads(B * 2) <= P3E_FAST_MAX_THREADS) B *= 2;
  if (lds(B) > LDS_MAX || threads(B) > P3E_FAST_MAX_THREADS || (long)L * B % P3E_FAST_E != 0) return false;
  const double span_in = ((double)a.nin * a.is_d + (double)B * a.is_e + 1) * esz;
  const double span_out = ((double)a.nout * a.os_d + (double)B * a.os_e + 1) * esz;
  if (span_in >= 2147483648.0 || span_out >= 2147483648.0) return false;
  a.fast = 1;
  a.dct_split = 1;
  a.mid = 0;
  a.mid_prime = 0;
  a.pair = 0;
  // paired lines on a tile-order side along a stride-1 e (whole tiles,
  // strides keeping pairs aligned; the launcher checks the addresses)
  a.vec2 = a.load_order == ORD_TILE && a.is_e == 1 && B >= 2 && a.ne % B == 0 && a.is_d % 2 == 0 && a.is_f % 2 == 0;
  a.vec2s = a.store_order == ORD_TILE && a.os_e == 1 && B >= 2 && a.ne % B == 0 && a.os_d % 2 == 0 && a.os_f % 2 == 0;
  a.xcd_map = (a.load_order == ORD_TILE && B * esz < 128) || (a.store_order == ORD_TILE && B * esz < 128);
  a.tw_in_lds = 0;
  a.M = P;
  a.logM = 0;
  while ((1 << a.logM) < P) a.logM++;
  a.lp = lp;
  a.lp2 = lp2;
  a.nrad = nrad;
  for (int i = 0; i < nrad; i++) a.radix[i] = ra[i];
  a.L = P;
  a.zero_fill = 0;
  a.B = B;
  a.logB = 0;
  while ((1 << a.logB) < B) a.logB++;
  a.ntile_e = (a.ne + B - 1) / B;
  a.div_ntile = make_fastdiv((uint32_t)a.ntile_e);
  a.tab_off = B * (lp + lp2);
  ps.lds_bytes = lds(B);
  ps.threads = (int)threads(B);
  a.nthreads = ps.threads;
  ps.nblocks = (unsigned int)((unsigned long long)a.ntile_e * (unsigned long long)a.nf);
  ps.tw = twa;
  ps.tw.insert(ps.tw.end(), twb.begin(), twb.end());
  for (int q = 0; q <= L / 2; q++) ps.tw.push_back(expi_pi(q, 2LL * L));  // e^{i pi q / 2L}
  a.ntw_a = (int)twa.size();
  a.ntw_b = (int)twb.size();
  a.ntw = (int)ps.tw.size();
  ps.wN.clear();
  return true;
}

// Switch an already described pass to the compile-time-length path when it
// is a C2C, R2C, C2R or complex DCT-I whose complex length M is one of the
// fast lengths (fast_radices: powers of two 16..4096, 3 and 5 times a power
// of two; M = N/2 for the real transforms, 2 (N - 1) for DCT-I).  Tile: B lines, B as large as the LDS budget and
// P3E_FAST_MAX_THREADS threads allow (up to 64), P3E_FAST_E elements per
// thread.
static void setup_fast(int prec, const OpSpec &op, PassDesc &ps, int batch_threads) {
  PassArgs &a = ps.args;
  if (!g_fast_enabled || a.bluestein) return;
  int M;
  if (op.cls == CLS_C2C) M = a.N;
  else if ((op.cls == CLS_R2C || op.cls == CLS_C2R) && a.N % 2 == 0) M = a.N / 2;
  // DCT-I of complex lines: the split form, or the M = 2 (N - 1)-point FFT
  // of the even extension
  else if (op.cls == CLS_R2R && op.r2r_kind == REDFT00 && a.in_cplx && a.out_cplx) {
    if (batch_threads == 0 && g_dct_split && setup_dct1_split(prec, ps)) return;
    M = a.W;
  } else return;
  const bool split = op.cls == CLS_R2C || op.cls == CLS_C2R;
  int rad[P3E_MAX_RADIX_PASSES];
  const int nrad = fast_radices(M, rad);
  if (nrad == 0) return;
  // the batch step kernel has only the power-of-two lengths
  if (batch_threads > 0 && !is_pow2(M)) return;
  const size_t esz = 2 * (size_t)prec;
  // Line pitch in LDS: room for the padded line (P3E_PAD(M-1) < M + M/16),
  // rounded up to 1 mod 32 complex elements (FP32) or 1 mod 16 (FP64).  The
  // load and store phases of tile order touch one element of each of many
  // lines at once; such a pitch puts consecutive lines on consecutive LDS
  // bank pairs (FP32: 32 lines round the 64 banks) or quads (FP64: 16 lines)
  // instead of on the same banks.  The FP64 rounding is the smaller one: it
  // lets a 256-point FP64 tile of 16 lines and its tables fit 80 KiB
  // (128-byte rows of reals for the R2C/C2R passes of a 512^3 grid).
  const int lp = fast_pitch(prec, M);
  const long maxth = batch_threads > 0 ? batch_threads : P3E_FAST_MAX_THREADS;
  int B = 1;
  // Tile width: when the tile axis e is strided on one side (tile-order load
  // or store), the contiguous run per row is B elements, so B is as large as
  // LDS and threads allow.  When both sides run along whole lines, B only
  // sets the workgroup size; about 256 threads (several workgroups per CU,
  // which overlap one another's load, transform and store phases) measured
  // fastest for 512-point FP32 lines on MI355X (0.25 ms vs 0.26 ms at 512
  // threads for a 512^3 pass).
  const bool line_both = a.load_order == ORD_LINE && a.store_order == ORD_LINE;
  int Bcap = batch_threads > 0 ? 256 : 64;
  if (batch_threads == 0 && line_both) Bcap = std::max(1, 256 * P3E_FAST_E / M);
  const int Bmax = std::min(Bcap, next_pow2(a.ne));
  // the pass kernel keeps its twiddles (the pass-by-pass table: M entries
  // for power-of-two radices, more with a radix 3 or 5, whose tw_stride
  // exceeds R - 1) and split factors (M + 1) in LDS after the tile; the
  // batch kernel reads them from global memory
  size_t ntw = 1;
  {
    long long ns = 1;
    for (int i = 0; i < nrad; i++) {
      ntw += (size_t)ns * tw_stride(rad[i]);
      ns *= rad[i];
    }
  }
  const size_t tables = batch_threads > 0 ? 0 : (ntw + (split ? M + 1 : 0)) * esz;
  auto lds = [&](int b) { return (size_t)b * lp * esz + tables; };
  // (a 160 KiB budget, B = 64 lines of 256 or 32 of 512 with one workgroup
  // per CU, measured slower on MI355X: 0.31 ms vs 0.27 ms for the R2C pass
  // of a 512^3 grid, 0.30 ms vs 0.23 ms for a C2C pass)
  while (B * 2 <= Bmax && lds(B * 2) <= LDS_BUDGET && (long)M * B * 2 / P3E_FAST_E <= maxth) B *= 2;
  // B M elements, P3E_FAST_E per thread: an exact number of threads (120 =
  // 7.5 x 16 needs B even)
  while ((long)M * B % P3E_FAST_E != 0 && B * 2 <= Bmax) B *= 2;
  if ((long)M * B % P3E_FAST_E != 0) return;
  if (lds(B) > LDS_MAX || (long)M * B / P3E_FAST_E > maxth) return;
  // the kernel addresses a tile's elements with 32-bit byte offsets from
  // the tile's first element: the largest offset within a tile (one element
  // past it, for paired accesses) must stay below 2^31 on both sides
  const double in_esz = (double)prec * (op.cls == CLS_R2C ? 1 : 2);
  const double out_esz = (double)prec * (op.cls == CLS_C2R ? 1 : 2);
  const double span_in = ((double)a.nin * a.is_d + (double)B * a.is_e + 1) * in_esz;
  const double span_out = ((double)a.nout * a.os_d + (double)B * a.os_e + 1) * out_esz;
  if (span_in >= 2147483648.0 || span_out >= 2147483648.0) return;
  a.fast = 1;
  a.mid = 0;
  a.mid_prime = 0;
  a.pair = 0;
  // R2C, C2C or DCT-I with a tile-order load along a stride-1 e: two adjacent
  // lines' elements per load when every tile is whole and the strides keep
  // pairs aligned (the launcher also checks the array's address);
  // standalone passes only
  a.vec2 = batch_threads == 0 && op.cls != CLS_C2R && a.load_order == ORD_TILE &&
           a.is_e == 1 && B >= 2 && a.ne % B == 0 && a.is_d % 2 == 0 && a.is_f % 2 == 0;
  // and the mirror for all three with a tile-order store along a stride-1 e
  a.vec2s = batch_threads == 0 && a.store_order == ORD_TILE && a.os_e == 1 && B >= 2 && a.ne % B == 0 &&
            a.os_d % 2 == 0 && a.os_f % 2 == 0;
  // R2C loads and C2R stores in line order (d stride-1): a real pair is one
  // packed complex element; aligned when the line strides are even
  if (batch_threads == 0 && op.cls == CLS_R2C && a.load_order == ORD_LINE && a.is_d == 1 && a.is_e % 2 == 0 &&
      a.is_f % 2 == 0)
    a.vec2 = 1;
  if (batch_threads == 0 && op.cls == CLS_C2R && a.store_order == ORD_LINE && a.os_d == 1 && a.os_e % 2 == 0 &&
      a.os_f % 2 == 0)
    a.vec2s = 1;
  // C2C with a line-order side (d stride-1): two adjacent elements of a line
  // per load / store when the line strides keep pairs aligned
  if (batch_threads == 0 && op.cls == CLS_C2C && a.load_order == ORD_LINE && a.is_d == 1 && a.is_e % 2 == 0 &&
      a.is_f % 2 == 0)
    a.vec2 = 1;
  if (batch_threads == 0 && op.cls == CLS_C2C && a.store_order == ORD_LINE && a.os_d == 1 && a.os_e % 2 == 0 &&
      a.os_f % 2 == 0)
    a.vec2s = 1;
  // rows shorter than a cache line: adjacent tiles share lines, so they
  // run on the same XCD (one L2); measured 0.81 -> 0.66 ms for the complex
  // 513-point DCT-I pass (B = 8, 64-byte rows) of a 385x640x513 array
  a.xcd_map = (a.load_order == ORD_TILE && B * in_esz < 128) || (a.store_order == ORD_TILE && B * out_esz < 128);
  a.tw_in_lds = 0;
  a.M = M;
  a.logM = 0;
  while ((1 << a.logM) < M) a.logM++;
  a.lp = lp;
  a.nrad = nrad;
  for (int i = 0; i < nrad; i++) a.radix[i] = rad[i];
  a.L = M;
  a.zero_fill = 0;
  a.B = B;
  a.logB = 0;
  while ((1 << a.logB) < B) a.logB++;
  a.ntile_e = (a.ne + B - 1) / B;
  a.div_ntile = make_fastdiv((uint32_t)a.ntile_e);
  a.div_pairs = make_fastdiv((uint32_t)(M / 2 + 1));
  a.tab_off = B * lp;
  ps.lds_bytes = lds(B);
  ps.threads = M * B / P3E_FAST_E;
  a.nthreads = ps.threads;
  ps.nblocks = (unsigned int)((unsigned long long)a.ntile_e * (unsigned long long)a.nf);
  pass_twiddles(rad, nrad, ps.tw);
  a.ntw = (int)ps.tw.size();  // ntw entries; the kernel puts the split factors after them
  ps.wN.clear();
  if (split) {
    ps.wN.resize(M + 1);
    for (int k = 0; k <= M; k++) ps.wN[k] = expi_pi(-2LL * k, a.N);
  }
}

// ------------------------------------------------------------------ one pass
bool describe_pass(int prec, int axis, const OpSpec &op, const Layout &lin, const Layout &lout, PassDesc &ps,
                   int batch_threads) {
  PassArgs &a = ps.args;
  memset(&a, 0, sizeof(a));
  ps.axis = axis;
  ps.lin = lin;
  ps.lout = lout;
  const int d = axis;
  const int nin = lin.dims[d], nout = lout.dims[d];
  a.cls = op.cls;
  a.r2r_kind = op.r2r_kind;
  a.in_cplx = lin.dt == 2;
  a.out_cplx = lout.dt == 2;
  a.nin = nin;
  a.nout = nout;
  for (int b = 0; b < 3; b++) {
    if (lin.dims[b] < 1 || lout.dims[b] < 1) {
      set_error("array dims must be positive");
      return false;
    }
    if (b != d && lin.dims[b] != lout.dims[b]) {
      set_error("pass along axis %d: dims of other axes differ", d);
      return false;
    }
  }
  int N = nin;
  switch (op.cls) {
    case CLS_EMPTY:
    case CLS_C2C:
    case CLS_R2R:
      if (nin != nout) {
        set_error("transform along axis %d: input length %d != output length %d", d, nin, nout);
        return false;
      }
      break;
    case CLS_R2C:
      if (nout != nin / 2 + 1) {
        set_error("R2C along axis %d: output length %d, expected %d", d, nout, nin / 2 + 1);
        return false;
      }
      break;
    case CLS_C2R:
      N = nout;
      if (nin != nout / 2 + 1) {
        set_error("C2R along axis %d: input length %d, expected %d", d, nin, nout / 2 + 1);
        return false;
      }
      break;
  }
  a.N = N;
  int W = N, sign = -1;
  switch (op.cls) {
    case CLS_C2C: sign = op.sign; break;
    case CLS_C2R: sign = +1; break;
    case CLS_R2R:
      switch (op.r2r_kind) {
        case REDFT00:
          if (N < 2) {
            set_error("DCT-I needs at least 2 points (axis %d has %d)", d, N);
            return false;
          }
          W = 2 * (N - 1);
          break;
        case RODFT00: W = 2 * (N + 1); break;
        case REDFT10:
        case RODFT10: W = 2 * N; break;
        default: W = 2 * N; sign = +1; break;
      }
      break;
    default: break;
  }
  a.W = W;
  a.sign = sign;
  std::vector<int> rad;
  int L = W;
  if (op.cls != CLS_EMPTY) {
    // A standalone FP32 pass takes odd primes up to 47 as radices when the
    // single-buffer path can run the whole length (one distinct such prime,
    // at most 16 elements per thread of 512, one padded line in LDS);
    // otherwise primes up to 31 run as direct DFTs on the generic path, and
    // a length with a larger prime factor uses Bluestein.
    bool ok = false;
    if (batch_threads == 0 && factor(W, rad, mid_max_prime(prec)) &&
        mid_prime_of(rad.data(), (int)rad.size(), prec) > 31 &&
        W <= P3E_MID_E * 512 && (size_t)generic_pitch(W) * 2 * (size_t)prec <= LDS_MAX)
      ok = true;
    if (!ok) ok = factor(W, rad);
    if (!ok) {
      a.bluestein = 1;
      L = next_pow2(2 * W - 1);
      factor(L, rad);
    }
  }
  a.L = L;
  a.nrad = (int)rad.size();
  for (size_t i = 0; i < rad.size(); i++) a.radix[i] = rad[i];
  // constant tables (long double; rounded to the working precision on upload)
  ps.tw.clear();
  ps.chirp.clear();
  ps.chirp_hat.clear();
  ps.pre.clear();
  ps.post.clear();
  if (op.cls != CLS_EMPTY) {
    ps.tw.resize(L);
    for (int t = 0; t < L; t++) ps.tw[t] = expi_pi(-2LL * t, L);
  }
  a.ntw = (int)ps.tw.size();
  if (a.bluestein) {
    ps.chirp.resize(W);
    std::vector<cld> b(L, cld(0, 0));
    for (int n = 0; n < W; n++) {
      long long n2 = ((long long)n * n) % (2LL * W);
      ps.chirp[n] = expi_pi(sign * n2, W);
    }
    for (int m = 0; m < W; m++) b[m] = std::conj(ps.chirp[m]);
    for (int m = 1; m < W; m++) b[L - m] = std::conj(ps.chirp[m]);
    host_fft_pow2(b);
    for (int m = 0; m < L; m++) b[m] /= (long double)L;
    ps.chirp_hat = b;
  }
  a.zero_fill = a.bluestein;
  if (op.cls == CLS_R2R) {
    const int k = op.r2r_kind;
    if (k == RODFT00 || k == REDFT01 || k == RODFT01) a.zero_fill = 1;
    if (k == REDFT01 || k == RODFT01 || k == REDFT11 || k == RODFT11) {
      ps.pre.resize(W);
      for (int m = 0; m < W; m++) ps.pre[m] = expi_pi(m, 2LL * N);  // e^{i pi m / 2N}
    }
    const cld I(0, 1);
    std::vector<cld> post(N);
    bool need = true;
    switch (k) {
      case REDFT00: case REDFT01: need = false; break;
      case RODFT00: a.post_off = 1; for (int q = 0; q < N; q++) post[q] = I; break;
      case REDFT10: for (int q = 0; q < N; q++) post[q] = expi_pi(-q, 2LL * N); break;
      case RODFT10: a.post_off = 1; for (int q = 0; q < N; q++) post[q] = I * expi_pi(-(q + 1), 2LL * N); break;
      case RODFT01: for (int q = 0; q < N; q++) post[q] = -I; break;
      case REDFT11: for (int q = 0; q < N; q++) post[q] = expi_pi(2 * q + 1, 4LL * N); break;
      case RODFT11: for (int q = 0; q < N; q++) post[q] = -I * expi_pi(2 * q + 1, 4LL * N); break;
    }
    if (need) {
      a.post_mode = 1;
      ps.post = post;
    }
  }
  // tile geometry: e is the input's stride-1 axis when d is strided on
  // input, else the output's; f is the remaining axis
  long long is[3], os[3];
  strides_of(lin, is);
  strides_of(lout, os);
  const int o1 = (d + 1) % 3, o2 = (d + 2) % 3;
  int e;
  if (is[d] != 1) e = (is[o1] <= is[o2]) ? o1 : o2;
  else e = (os[o1] <= os[o2]) ? o1 : o2;
  const int f = 3 - d - e;
  ps.axis_e = e;
  ps.axis_f = f;
  a.is_d = is[d]; a.is_e = is[e]; a.is_f = is[f];
  a.os_d = os[d]; a.os_e = os[e]; a.os_f = os[f];
  a.ne = lin.dims[e];
  a.nf = lin.dims[f];
  a.load_order = (is[d] == 1 && nin > 1) ? ORD_LINE : ORD_TILE;
  a.store_order = (os[d] == 1 && nout > 1) ? ORD_LINE : ORD_TILE;
  const size_t esz = 2 * (size_t)prec;
  a.lpad = generic_pitch(L);
  set_radix_divs(a);
  const size_t per_line = 2 * (size_t)a.lpad * esz;
  int B = 1;
  const int Bmax = std::min(32, next_pow2(a.ne));
  // the standalone pass kernel keeps the L twiddles in LDS after the two
  // line buffers; the batch kernel reads them from global memory
  size_t twl = (batch_threads > 0 || op.cls == CLS_EMPTY) ? 0 : (size_t)L * esz;
  if (per_line + twl > LDS_MAX) twl = 0;  // long lines: twiddles stay in global memory
  a.tw_in_lds = twl > 0;
  while (B * 2 <= Bmax && per_line * (B * 2) + twl <= LDS_BUDGET) B *= 2;
  if (per_line * B + twl > LDS_MAX) {
    set_error("transform length %d (work length %d) exceeds the LDS capacity of one workgroup", N, L);
    return false;
  }
  a.B = B;
  a.logB = 0;
  while ((1 << a.logB) < B) a.logB++;
  a.ntile_e = (a.ne + B - 1) / B;
  a.div_nin = make_fastdiv((uint32_t)std::max(nin, 1));
  a.div_nout = make_fastdiv((uint32_t)std::max(nout, 1));
  a.div_ntile = make_fastdiv((uint32_t)a.ntile_e);
  ps.lds_bytes = per_line * B + twl;
  ps.threads = 256;
  a.nthreads = 256;
  const unsigned long long nbk = (unsigned long long)a.ntile_e * (unsigned long long)a.nf;
  if (nbk > 0x7fffffffULL) {
    set_error("grid too large");
    return false;
  }
  ps.nblocks = (unsigned int)nbk;
  ps.in_bytes = layout_elems(lin) * elem_bytes(prec, lin.dt);
  ps.out_bytes = layout_elems(lout) * elem_bytes(prec, lout.dt);
  ps.inplace_ok = (lin.mo[0] == lout.mo[0] && lin.mo[1] == lout.mo[1] && lin.mo[2] == lout.mo[2] &&
                   lin.dt == lout.dt && nin == nout && is[0] == os[0] && is[1] == os[1] && is[2] == os[2]);
  ps.deriv_capable = (op.cls != CLS_EMPTY) && lout.dt == 2;
  // compute_deriv_loc (build/exec.C:228-287): g = 2(n-1) after R2C, else n
  a.dg = (op.cls == CLS_R2C) ? 2 * (nout - 1) : nout;
  a.dmid = a.dg / 2;
  a.scale = 1.0;
  setup_mid(prec, op, ps, batch_threads);
  setup_fast(prec, op, ps, batch_threads);
  return true;
}

// ------------------------------------------------------------------ chain
static const int PERMS[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// 1 when one tile shape coalesces both sides of the pass, 4 when the
// transform axis and both stride-1 axes are all different.
static int pass_cost(int axis, const int moi[3], const int moo[3]) {
  const int ai = stride1_axis(moi), ao = stride1_axis(moo);
  if (ai == axis || ao == axis || ai == ao) return 1;
  return 4;
}

static bool same_mo(const int a[3], const int b[3]) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2]; }

// Secondary cost of a pass along `axis` from layout lin to lout (memory
// orders as given): 1 when a side is tile order (the axis strided there) and
// the tile axis e (chosen as describe_pass chooses it) has an extent that
// is not a multiple of 8.  Such a pass has a partial tile in every row and
// no paired 16-byte accesses; on the 768x640x513 R2C / DCT-I / C2C chain
// (tile axes of 385 complex values) the candidate orders with one such pass
// instead of two ran 0.15-0.17 ms faster in each direction than the
// others of equal primary cost (1.45-1.53 ms against 1.62-1.88 ms for the
// three passes).
static int ragged_tile_cost(int axis, const Layout &lin, const Layout &lout) {
  long long is[3], os[3];
  strides_of(lin, is);
  strides_of(lout, os);
  const bool tile_in = is[axis] != 1, tile_out = os[axis] != 1;
  if (!tile_in && !tile_out) return 0;
  const int o1 = (axis + 1) % 3, o2 = (axis + 2) % 3;
  int e;
  if (tile_in) e = (is[o1] <= is[o2]) ? o1 : o2;
  else e = (os[o1] <= os[o2]) ? o1 : o2;
  return lin.dims[e] % 8 != 0 ? 1 : 0;
}

bool plan_chain(int prec, const OpSpec ops[3], const int order[3], const Layout &in, const Layout &out,
                PassChain &chain, int batch_threads) {
  chain.passes.clear();
  chain.copy_only = false;
  chain.desc.clear();
  for (int a = 0; a < 3; a++) chain.pass_of_axis[a] = -1;
  std::vector<int> axes;
  for (int i = 0; i < 3; i++)
    if (ops[order[i]].cls != CLS_EMPTY) axes.push_back(order[i]);
  const int k = (int)axes.size();
  if (k == 0) {
    for (int b = 0; b < 3; b++)
      if (in.dims[b] != out.dims[b]) {
        set_error("EMPTY transform: input and output dims differ");
        return false;
      }
    if (in.dt != out.dt) {
      set_error("EMPTY transform: input and output element types differ");
      return false;
    }
    if (same_mo(in.mo, out.mo)) {
      chain.copy_only = true;
      chain.desc = "copy";
      return true;
    }
    PassDesc ps;
    OpSpec cp = {CLS_EMPTY, -1, 0, in.dt, in.dt};
    if (!describe_pass(prec, stride1_axis(out.mo), cp, in, out, ps, batch_threads)) return false;
    chain.passes.push_back(ps);
    chain.desc = "reorder";
    return true;
  }
  std::vector<Layout> lay(k + 1);
  lay[0] = in;
  for (int i = 0; i < k; i++) {
    Layout l = lay[i];
    const int ax = axes[i];
    const OpSpec &op = ops[ax];
    if (op.cls == CLS_R2C) l.dims[ax] = l.dims[ax] / 2 + 1;
    else if (op.cls == CLS_C2R) l.dims[ax] = out.dims[ax];
    if (op.in_dt != l.dt) {
      set_error("transform along axis %d expects %s input", ax, op.in_dt == 2 ? "complex" : "real");
      return false;
    }
    l.dt = op.out_dt;
    lay[i + 1] = l;
  }
  for (int b = 0; b < 3; b++)
    if (lay[k].dims[b] != out.dims[b]) {
      set_error("output grid does not match the transform: axis %d has %d, transform gives %d", b,
                out.dims[b], lay[k].dims[b]);
      return false;
    }
  if (lay[k].dt != out.dt) {
    set_error("output element type does not match the transform");
    return false;
  }
  // intermediate memory orders: exhaustive over 6^(k-1); candidates are
  // tried with the input order first so that ties keep it
  std::vector<int> cand;
  for (int q = 0; q < 6; q++)
    if (same_mo(PERMS[q], in.mo)) cand.insert(cand.begin(), q);
    else cand.push_back(q);
  int ncomb = 1;
  for (int i = 1; i < k; i++) ncomb *= 6;
  int best_cost = 1 << 30, best_ragged = 1 << 30;
  std::vector<int> best(k + 1, 0);
  for (int c = 0; c < ncomb; c++) {
    int cc = c;
    const int *mo[4];
    mo[0] = in.mo;
    mo[k] = out.mo;
    std::vector<int> sel(k + 1, 0);
    for (int i = 1; i < k; i++) {
      sel[i] = cand[cc % 6];
      cc /= 6;
      mo[i] = PERMS[sel[i]];
    }
    int cost = 0;
    for (int i = 0; i < k; i++) cost += pass_cost(axes[i], mo[i], mo[i + 1]);
    if (cost > best_cost) continue;
    // ties: fewest passes with a ragged tile axis
    int ragged = 0;
    for (int i = 0; i < k; i++) {
      Layout a = lay[i], b = lay[i + 1];
      memcpy(a.mo, mo[i], sizeof(a.mo));
      memcpy(b.mo, mo[i + 1], sizeof(b.mo));
      a.pitch = b.pitch = 0;
      ragged += ragged_tile_cost(axes[i], a, b);
    }
    if (cost < best_cost || ragged < best_ragged) {
      best_cost = cost;
      best_ragged = ragged;
      best = sel;
    }
  }
  for (int i = 1; i < k; i++) {
    memcpy(lay[i].mo, PERMS[best[i]], sizeof(lay[i].mo));
    lay[i].pitch = 0;
    lay[i].pitch = padded_pitch(prec, lay[i]);
  }
  memcpy(lay[k].mo, out.mo, sizeof(lay[k].mo));
  lay[k].pitch = out.pitch;
  char buf[256];
  for (int i = 0; i < k; i++) {
    PassDesc ps;
    if (!describe_pass(prec, axes[i], ops[axes[i]], lay[i], lay[i + 1], ps, batch_threads)) {
      chain.passes.clear();
      return false;
    }
    chain.pass_of_axis[axes[i]] = (int)chain.passes.size();
    char kind[40];
    if (ps.args.mid_prime) snprintf(kind, sizeof(kind), ", mid p%d", ps.args.mid_prime);
    else
      snprintf(kind, sizeof(kind), "%s",
               ps.args.bluestein ? (ps.args.mid ? ", bluestein mid" : ", bluestein")
                                 : (ps.args.fast ? (ps.args.dct_split ? ", pow2 split" : ", pow2")
                                                 : (ps.args.mid ? ", mid" : "")));
    if (ps.args.pair) strncat(kind, " pair", sizeof(kind) - strlen(kind) - 1);
    char pitch[24] = "";
    if (lay[i + 1].pitch > 0) snprintf(pitch, sizeof(pitch), ", out pitch %d", lay[i + 1].pitch);
    snprintf(buf, sizeof(buf), "%s%s[axis %d, N=%d%s, mo (%d%d%d)->(%d%d%d), B=%d%s]", i ? " | " : "",
             cls_name(ops[axes[i]].cls), axes[i], ps.args.N, kind,
             lay[i].mo[0], lay[i].mo[1], lay[i].mo[2], lay[i + 1].mo[0], lay[i + 1].mo[1], lay[i + 1].mo[2],
             ps.args.B, pitch);
    chain.desc += buf;
    chain.passes.push_back(ps);
  }
  return true;
}

// ------------------------------------------------------------------ tuning knobs
static bool radix_supported(int r) {
  if (r == 2 || r == 3 || r == 4 || r == 5 || r == 8 || r == 16) return true;
  if (r < 7 || r > 31 || r % 2 == 0) return false;
  for (int q = 3; q * q <= r; q += 2)
    if (r % q == 0) return false;
  return true;
}

bool retune_pass(int prec, PassDesc &ps, int B, const int *rad, int nrad) {
  PassArgs &a = ps.args;
  if (a.fast || a.mid_prime || a.cls == CLS_EMPTY || B < 1 || (B & (B - 1)) || nrad < 0 ||
      nrad > P3E_MAX_RADIX_PASSES)
    return false;
  long prod = 1;
  for (int i = 0; i < nrad; i++) {
    if (!radix_supported(rad[i])) return false;
    // the single-buffer path of the batch kernel: radices 2, 3, 4, 5, 8, 16
    if (a.mid && !mid_base_radix(rad[i])) return false;
    prod *= rad[i];
  }
  if (prod != a.L) return false;
  // passes planned for the batch kernel: no tables in LDS; a single-buffer
  // pass keeps at most 16 elements per thread of the batch workgroup
  const int nl = a.pair ? (B + 1) / 2 : B;  // LDS lines of a tile
  const size_t lds = (a.mid ? 1 : 2) * (size_t)a.lpad * 2 * (size_t)prec * nl;
  if (lds > LDS_MAX) return false;
  if (a.mid) {
    if ((long)nl * a.L > (long)P3E_MID_E * P3E_STEP_THREADS) return false;
    const int th = std::max(64, ((int)(((long)nl * a.L + P3E_MID_E - 1) / P3E_MID_E) + 63) & ~63);
    ps.threads = th;
    a.nthreads = th;
  }
  for (int i = 0; i < nrad; i++) a.radix[i] = rad[i];
  a.nrad = nrad;
  set_radix_divs(a);
  a.B = B;
  a.logB = 0;
  while ((1 << a.logB) < B) a.logB++;
  a.ntile_e = (a.ne + B - 1) / B;
  a.div_ntile = make_fastdiv((uint32_t)a.ntile_e);
  ps.lds_bytes = lds;
  ps.nblocks = (unsigned int)((unsigned long long)a.ntile_e * (unsigned long long)a.nf);
  return true;
}

bool plan_poisson(int prec, const int dims[3], const int mo_real[3], const int mo_spec[3], PassChain &fwd,
                  PassChain &bwd, int batch_threads) {
  for (int a = 0; a < 3; a++)
    if (dims[a] < 2) {
      set_error("Poisson grid needs at least 2 points per axis");
      return false;
    }
  const OpSpec f_ops[3] = {{CLS_R2C, -1, 0, 1, 2}, {CLS_C2C, -1, 0, 2, 2}, {CLS_C2C, -1, 0, 2, 2}};
  const OpSpec b_ops[3] = {{CLS_C2R, 1, 0, 2, 1}, {CLS_C2C, 1, 0, 2, 2}, {CLS_C2C, 1, 0, 2, 2}};
  const int f_order[3] = {0, 1, 2}, b_order[3] = {1, 2, 0};
  Layout real, spec;
  for (int a = 0; a < 3; a++) {
    real.dims[a] = dims[a];
    real.mo[a] = mo_real[a];
    spec.dims[a] = a == 0 ? dims[0] / 2 + 1 : dims[a];
    spec.mo[a] = mo_spec[a];
  }
  real.dt = 1;
  spec.dt = 2;
  if (!plan_chain(prec, f_ops, f_order, real, spec, fwd, batch_threads)) return false;
  if (!plan_chain(prec, b_ops, b_order, spec, real, bwd, batch_threads)) return false;
  if (fwd.passes.empty() || bwd.passes.empty()) {
    set_error("Poisson plan has no pass");
    return false;
  }
  PassDesc &pd = fwd.passes.back();
  if (pd.lout.dt != 2) {
    set_error("Poisson plan: last forward pass does not write the spectrum");
    return false;
  }
  PassArgs &a = pd.args;
  const int role_axis[3] = {pd.axis, pd.axis_e, pd.axis_f};
  a.spec_op = 1;
  for (int r = 0; r < 3; r++) {
    a.kn[r] = dims[role_axis[r]];
    a.khalf[r] = role_axis[r] == 0;
  }
  a.has_scale = 1;
  a.scale = 1.0 / ((double)dims[0] * dims[1] * dims[2]);
  return true;
}

bool tables_ready(const PassDesc &ps) {
  const PassArgs &a = ps.args;
  const std::vector<cld> *h[6] = {&ps.tw, &ps.chirp, &ps.chirp_hat, &ps.pre, &ps.post, &ps.wN};
  const void *d[6] = {a.tw, a.chirp, a.chirp_hat, a.pre, a.post, a.wN};
  for (int i = 0; i < 6; i++)
    if (!h[i]->empty() && !d[i]) return false;
  return true;
}

static void factorizations(int n, int maxr, std::vector<int> &cur, std::vector<std::vector<int>> &out) {
  if (n == 1) {
    out.push_back(cur);
    return;
  }
  if (cur.size() >= 6) return;
  static const int R[] = {16, 8, 5, 4, 3, 2, 7, 11, 13, 17, 19, 23, 29, 31};
  for (int r : R) {
    if (r > maxr || n % r) continue;
    cur.push_back(r);
    factorizations(n / r, r, cur, out);  // non-increasing multisets
    cur.pop_back();
  }
}

std::vector<std::vector<int>> radix_candidates(int W, int cap) {
  std::vector<std::vector<int>> sets, out;
  std::vector<int> cur;
  factorizations(W, 1 << 30, cur, sets);
  std::sort(sets.begin(), sets.end(),
            [](const std::vector<int> &x, const std::vector<int> &y) { return x.size() < y.size(); });
  for (const auto &s : sets) {
    std::vector<int> desc = s, asc = s;
    std::sort(desc.begin(), desc.end(), [](int a, int b) { return a > b; });
    std::sort(asc.begin(), asc.end());
    for (const auto &c : {desc, asc}) {
      if (std::find(out.begin(), out.end(), c) == out.end()) out.push_back(c);
      if ((int)out.size() >= cap) return out;
    }
  }
  return out;
}

// ------------------------------------------------------------------ buffers
void assign_buffers(const std::vector<PassDesc> &passes, size_t out_cap, bool alias, int dst[8], size_t need[2],
                    bool *copy_back) {
  const int k = (int)passes.size();
  need[0] = need[1] = 0;
  *copy_back = false;
  if (k == 0) return;
  int ncomb = 1;
  for (int i = 0; i < k - 1; i++) ncomb *= 3;
  size_t best_score = (size_t)-1;
  bool found = false;
  auto phys = [&](int b) { return (alias && b == BUF_OUT) ? BUF_IN : b; };
  for (int c = 0; c < ncomb; c++) {
    int d[8];
    int cc = c;
    for (int i = 0; i < k - 1; i++) {
      d[i] = BUF_OUT + (cc % 3);
      cc /= 3;
    }
    d[k - 1] = BUF_OUT;
    bool ok = true;
    size_t nd[2] = {0, 0};
    int src = BUF_IN;
    for (int i = 0; i < k && ok; i++) {
      const PassDesc &ps = passes[i];
      if (phys(d[i]) == phys(src) && !ps.inplace_ok) ok = false;
      if (d[i] == BUF_OUT && ps.out_bytes > out_cap) ok = false;
      if (d[i] == BUF_S0) nd[0] = std::max(nd[0], ps.out_bytes);
      if (d[i] == BUF_S1) nd[1] = std::max(nd[1], ps.out_bytes);
      src = d[i];
    }
    if (!ok) continue;
    const size_t score = nd[0] + nd[1];
    if (score < best_score) {
      best_score = score;
      found = true;
      for (int i = 0; i < k; i++) dst[i] = d[i];
      need[0] = nd[0];
      need[1] = nd[1];
    }
  }
  if (!found) {
    // e.g. a single pass that cannot run in place with in == out
    for (int i = 0; i < k; i++) dst[i] = (i % 2 == 0) ? BUF_S0 : BUF_S1;
    need[0] = need[1] = 0;
    for (int i = 0; i < k; i++) {
      const int s = dst[i] - BUF_S0;
      need[s] = std::max(need[s], passes[i].out_bytes);
    }
    *copy_back = true;
  }
}

}  // namespace p3e
