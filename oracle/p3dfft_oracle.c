/*
 * p3dfft_oracle.c -- CPU FP64 restatement of P3DFFT++'s transform3D path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker for the HIP product library,
 * never part of it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * The arithmetic follows FFTW's published transform definitions, which is
 * what the reference delegates every 1D stage to (build/init.C:1146-1607).
 * FFTW itself is not vendored in the reference and is absent from this
 * image, so the restatement is pinned against (tests/test_oracle.py):
 *   - FFTW-generated r2r vectors shipped with scipy (14 sizes x DCT/DST 1-4),
 *   - numpy pocketfft FP64 fixtures for c2c/r2c/c2r (tests/golden),
 *   - long-double O(N^2) definitions (or_dft_naive, or_r2r_naive),
 *   - the reference samples' closed-form known answers.
 */
#define _GNU_SOURCE
#include "p3dfft_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cx;

static const long double PI_L = 3.141592653589793238462643383279502884L;

/* ------------------------------------------------------------------ */
/* Type registry: registration order of build/init.C:121-731 (setup()). */
/* DCT4_* IDs are registered with the DCT-I planner plan_dct1_* at     */
/* build/init.C:640,652,664,676 -- reproduced here (REDFT00).          */
/* ------------------------------------------------------------------ */
static const or_type_t OR_TYPES[44] = {
  /* 0-3 EMPTY_TYPE_{SINGLE,DOUBLE,SINGLE_COMPLEX,DOUBLE_COMPLEX}  init.C:131-179 */
  {OR_EMPTY, 1, 1, 4, 0, 0}, {OR_EMPTY, 1, 1, 8, 0, 0},
  {OR_EMPTY, 2, 2, 4, 0, 0}, {OR_EMPTY, 2, 2, 8, 0, 0},
  /* 4-7 R2CFFT_S, R2CFFT_D, C2RFFT_S, C2RFFT_D                    init.C:213-259 */
  {OR_R2C, 1, 2, 4, -1, 0}, {OR_R2C, 1, 2, 8, -1, 0},
  {OR_C2R, 2, 1, 4, 1, 0}, {OR_C2R, 2, 1, 8, 1, 0},
  /* 8-11 CFFT_FORWARD_S/D, CFFT_BACKWARD_S/D                       init.C:261-310 */
  {OR_C2C, 2, 2, 4, -1, 0}, {OR_C2C, 2, 2, 8, -1, 0},
  {OR_C2C, 2, 2, 4, 1, 0}, {OR_C2C, 2, 2, 8, 1, 0},
  /* 12-15 DCT1_{REAL_S,REAL_D,COMPLEX_S,COMPLEX_D} = REDFT00 */
  {OR_R2R, 1, 1, 4, 0, OR_REDFT00}, {OR_R2R, 1, 1, 8, 0, OR_REDFT00},
  {OR_R2R, 2, 2, 4, 0, OR_REDFT00}, {OR_R2R, 2, 2, 8, 0, OR_REDFT00},
  /* 16-19 DST1 = RODFT00 */
  {OR_R2R, 1, 1, 4, 0, OR_RODFT00}, {OR_R2R, 1, 1, 8, 0, OR_RODFT00},
  {OR_R2R, 2, 2, 4, 0, OR_RODFT00}, {OR_R2R, 2, 2, 8, 0, OR_RODFT00},
  /* 20-23 DCT2 = REDFT10 */
  {OR_R2R, 1, 1, 4, 0, OR_REDFT10}, {OR_R2R, 1, 1, 8, 0, OR_REDFT10},
  {OR_R2R, 2, 2, 4, 0, OR_REDFT10}, {OR_R2R, 2, 2, 8, 0, OR_REDFT10},
  /* 24-27 DST2 = RODFT10 */
  {OR_R2R, 1, 1, 4, 0, OR_RODFT10}, {OR_R2R, 1, 1, 8, 0, OR_RODFT10},
  {OR_R2R, 2, 2, 4, 0, OR_RODFT10}, {OR_R2R, 2, 2, 8, 0, OR_RODFT10},
  /* 28-31 DCT3 = REDFT01 */
  {OR_R2R, 1, 1, 4, 0, OR_REDFT01}, {OR_R2R, 1, 1, 8, 0, OR_REDFT01},
  {OR_R2R, 2, 2, 4, 0, OR_REDFT01}, {OR_R2R, 2, 2, 8, 0, OR_REDFT01},
  /* 32-35 DST3 = RODFT01 */
  {OR_R2R, 1, 1, 4, 0, OR_RODFT01}, {OR_R2R, 1, 1, 8, 0, OR_RODFT01},
  {OR_R2R, 2, 2, 4, 0, OR_RODFT01}, {OR_R2R, 2, 2, 8, 0, OR_RODFT01},
  /* 36-39 DCT4 -> REDFT00 (reference quirk, init.C:640-676) */
  {OR_R2R, 1, 1, 4, 0, OR_REDFT00}, {OR_R2R, 1, 1, 8, 0, OR_REDFT00},
  {OR_R2R, 2, 2, 4, 0, OR_REDFT00}, {OR_R2R, 2, 2, 8, 0, OR_REDFT00},
  /* 40-43 DST4 = RODFT11 */
  {OR_R2R, 1, 1, 4, 0, OR_RODFT11}, {OR_R2R, 1, 1, 8, 0, OR_RODFT11},
  {OR_R2R, 2, 2, 4, 0, OR_RODFT11}, {OR_R2R, 2, 2, 8, 0, OR_RODFT11},
};

int or_num_types(void) { return 44; }

int or_type_info(int id, or_type_t *t)
{
  if (id < 0 || id >= 44) return -1;
  *t = OR_TYPES[id];
  return 0;
}

/* ------------------------------------------------------------------ */
/* 1D DFT                                                              */
/* ------------------------------------------------------------------ */

/* exp(sign * 2 pi i * t / n), t reduced exactly in integers first. */
static cx root(long t, long n, int sign)
{
  t %= n;
  if (t < 0) t += n;
  long double a = 2.0L * PI_L * (long double)t / (long double)n;
  cx w;
  w.re = (double)cosl(a);
  w.im = (double)((long double)sign * sinl(a));
  return w;
}

/* Definition, long-double accumulation: X_k = sum_j x_j e^{sign 2 pi i jk/n}. */
void or_dft_naive(int n, const double *in, double *out, int sign)
{
  for (int k = 0; k < n; k++) {
    long double sr = 0, si = 0;
    for (int j = 0; j < n; j++) {
      long t = ((long)j * k) % n;
      long double a = 2.0L * PI_L * (long double)t / (long double)n;
      long double c = cosl(a), s = (long double)sign * sinl(a);
      long double xr = in[2 * j], xi = in[2 * j + 1];
      sr += xr * c - xi * s;
      si += xr * s + xi * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
}

static int smallest_factor(int n)
{
  if (n % 4 == 0) return 4;
  if (n % 2 == 0) return 2;
  for (int p = 3; (long)p * p <= n; p += 2)
    if (n % p == 0) return p;
  return n;
}

/* Recursive mixed-radix decimation in time.  w[t] = root(t, N0, sign);
 * W_n^t = w[t * wstep] with wstep = N0 / n. */
static void fft_rec(int n, const cx *in, long is, cx *out, const cx *w, long wstep, cx *tmp)
{
  if (n == 1) { out[0] = in[0]; return; }
  int p = smallest_factor(n);
  int m = n / p;
  for (int r = 0; r < p; r++)
    fft_rec(m, in + r * is, is * p, out + (long)r * m, w, wstep * p, tmp);
  for (int k2 = 0; k2 < m; k2++) {
    for (int r = 0; r < p; r++) tmp[r] = out[(long)r * m + k2];
    for (int q = 0; q < p; q++) {
      long k = k2 + (long)q * m;
      double sr = 0, si = 0;
      for (int r = 0; r < p; r++) {
        cx tw = w[(((long)r * k) % n) * wstep];
        sr += tmp[r].re * tw.re - tmp[r].im * tw.im;
        si += tmp[r].re * tw.im + tmp[r].im * tw.re;
      }
      out[k].re = sr;
      out[k].im = si;
    }
  }
}

/* The roots of the last (n, sign) used by this thread: a 3D transform calls
 * or_fft line after line with the same length, and the long-double roots
 * dominate the cost otherwise.  Per thread, so concurrent callers are safe. */
static _Thread_local cx *g_w = NULL;
static _Thread_local int g_w_n = 0, g_w_sign = 0;

void or_fft(int n, const double *in, double *out, int sign)
{
  if (n <= 0) return;
  if (g_w_n != n || g_w_sign != sign) {
    free(g_w);
    g_w = (cx *)malloc(sizeof(cx) * n);
    for (int t = 0; t < n; t++) g_w[t] = root(t, n, sign);
    g_w_n = n;
    g_w_sign = sign;
  }
  cx *tmp = (cx *)malloc(sizeof(cx) * n);
  cx *o = (cx *)malloc(sizeof(cx) * n);
  fft_rec(n, (const cx *)in, 1, o, g_w, 1, tmp);
  memcpy(out, o, sizeof(cx) * n);
  free(tmp); free(o);
}

/* FFTW r2c: first n/2+1 outputs of the forward DFT of real data. */
void or_r2c(int n, const double *in, double *out)
{
  double *c = (double *)malloc(sizeof(double) * 2 * n);
  double *f = (double *)malloc(sizeof(double) * 2 * n);
  for (int j = 0; j < n; j++) { c[2 * j] = in[j]; c[2 * j + 1] = 0; }
  or_fft(n, c, f, -1);
  memcpy(out, f, sizeof(double) * 2 * (n / 2 + 1));
  free(c); free(f);
}

/* FFTW c2r: y_j = sum_k X_k e^{+2 pi i jk/n}, X_{n-k} = conj(X_k); the
 * imaginary parts of X_0 (and X_{n/2}, n even) do not contribute. */
void or_c2r(int n, const double *in, double *out)
{
  double *c = (double *)malloc(sizeof(double) * 2 * n);
  double *f = (double *)malloc(sizeof(double) * 2 * n);
  int h = n / 2;
  for (int k = 0; k <= h; k++) { c[2 * k] = in[2 * k]; c[2 * k + 1] = in[2 * k + 1]; }
  for (int k = h + 1; k < n; k++) { c[2 * k] = in[2 * (n - k)]; c[2 * k + 1] = -in[2 * (n - k) + 1]; }
  or_fft(n, c, f, 1);
  for (int j = 0; j < n; j++) out[j] = f[2 * j];
  free(c); free(f);
}

/* ------------------------------------------------------------------ */
/* Real-to-real kinds (FFTW manual "1d Real-even/Real-odd DFTs")      */
/* ------------------------------------------------------------------ */

/* cos / sin of pi * num / den with num reduced modulo 2*den. */
static long double cospi_frac(long num, long den)
{
  long m = num % (2 * den);
  if (m < 0) m += 2 * den;
  return cosl(PI_L * (long double)m / (long double)den);
}
static long double sinpi_frac(long num, long den)
{
  long m = num % (2 * den);
  if (m < 0) m += 2 * den;
  return sinl(PI_L * (long double)m / (long double)den);
}

void or_r2r_naive(int kind, int n, const double *x, double *y)
{
  for (int k = 0; k < n; k++) {
    long double s = 0;
    switch (kind) {
    case OR_REDFT00: /* Y_k = X_0 + (-1)^k X_{n-1} + 2 sum_{j=1}^{n-2} X_j cos(pi jk/(n-1)) */
      if (n == 1) { s = x[0]; break; }
      s = x[0] + ((k & 1) ? -1.0L : 1.0L) * x[n - 1];
      for (int j = 1; j < n - 1; j++) s += 2.0L * x[j] * cospi_frac((long)j * k, n - 1);
      break;
    case OR_REDFT10: /* 2 sum X_j cos(pi (2j+1) k / (2n)) */
      for (int j = 0; j < n; j++) s += 2.0L * x[j] * cospi_frac((long)(2 * j + 1) * k, 2L * n);
      break;
    case OR_REDFT01: /* X_0 + 2 sum_{j>=1} X_j cos(pi j (2k+1)/(2n)) */
      s = x[0];
      for (int j = 1; j < n; j++) s += 2.0L * x[j] * cospi_frac((long)j * (2 * k + 1), 2L * n);
      break;
    case OR_REDFT11: /* 2 sum X_j cos(pi (2j+1)(2k+1)/(4n)) */
      for (int j = 0; j < n; j++) s += 2.0L * x[j] * cospi_frac((long)(2 * j + 1) * (2 * k + 1), 4L * n);
      break;
    case OR_RODFT00: /* 2 sum X_j sin(pi (j+1)(k+1)/(n+1)) */
      for (int j = 0; j < n; j++) s += 2.0L * x[j] * sinpi_frac((long)(j + 1) * (k + 1), n + 1);
      break;
    case OR_RODFT10: /* 2 sum X_j sin(pi (2j+1)(k+1)/(2n)) */
      for (int j = 0; j < n; j++) s += 2.0L * x[j] * sinpi_frac((long)(2 * j + 1) * (k + 1), 2L * n);
      break;
    case OR_RODFT01: /* (-1)^k X_{n-1} + 2 sum_{j=0}^{n-2} X_j sin(pi (j+1)(2k+1)/(2n)) */
      s = ((k & 1) ? -1.0L : 1.0L) * x[n - 1];
      for (int j = 0; j < n - 1; j++) s += 2.0L * x[j] * sinpi_frac((long)(j + 1) * (2 * k + 1), 2L * n);
      break;
    case OR_RODFT11: /* 2 sum X_j sin(pi (2j+1)(2k+1)/(4n)) */
      for (int j = 0; j < n; j++) s += 2.0L * x[j] * sinpi_frac((long)(2 * j + 1) * (2 * k + 1), 4L * n);
      break;
    }
    y[k] = (double)s;
  }
}

/* Fast r2r through one complex DFT of an extended sequence of length M:
 *   w = extension of x (even/odd symmetric), optional pre-twiddle
 *   F = DFT_M(w) (forward or backward)
 *   Y_k = Re(post_k * F_{k+off})
 * The identities are derived from cos/sin = sums of exponentials and the
 * M-periodicity of the exponentials (see the per-kind comments). */
void or_r2r(int kind, int n, const double *x, double *y)
{
  if (kind == OR_REDFT00 && n == 1) { y[0] = x[0]; return; }
  int M;
  switch (kind) {
  case OR_REDFT00: M = 2 * (n - 1); break;
  case OR_RODFT00: M = 2 * (n + 1); break;
  default: M = 2 * n; break;
  }
  double *w = (double *)calloc((size_t)2 * M, sizeof(double));
  double *F = (double *)malloc(sizeof(double) * 2 * M);
  int sign = -1;
  switch (kind) {
  case OR_REDFT00: /* w = even extension about 0 and n-1; Y_k = F_k */
    w[0] = x[0];
    for (int j = 1; j < n - 1; j++) { w[2 * j] = x[j]; w[2 * (M - j)] = x[j]; }
    w[2 * (n - 1)] = x[n - 1];
    break;
  case OR_RODFT00: /* odd extension; Y_k = -Im F_{k+1} */
    for (int j = 0; j < n; j++) { w[2 * (j + 1)] = x[j]; w[2 * (M - j - 1)] = -x[j]; }
    break;
  case OR_REDFT10: /* half-sample even; Y_k = Re(e^{-i pi k/2n} F_k) */
    for (int j = 0; j < n; j++) { w[2 * j] = x[j]; w[2 * (2 * n - 1 - j)] = x[j]; }
    break;
  case OR_RODFT10: /* half-sample odd; Y_{k-1} = Re(i e^{-i pi k/2n} F_k) */
    for (int j = 0; j < n; j++) { w[2 * j] = x[j]; w[2 * (2 * n - 1 - j)] = -x[j]; }
    break;
  case OR_REDFT01: /* v_0=x_0, v_j=x_j, v_{2n-j}=-x_j; pre e^{i pi m/2n}; backward */
  case OR_RODFT01: /* w_j=x_{j-1}, w_{2n-j}=x_{j-1}, w_n=x_{n-1}; pre; backward; x(-i) */
  case OR_REDFT11: /* w_j=x_j, w_{2n-1-j}=-x_j; pre; backward; x e^{i pi (k+1/2)/2n} */
  case OR_RODFT11: /* w_j=x_j, w_{2n-1-j}=+x_j; pre; backward; x -i e^{i pi (k+1/2)/2n} */
    sign = 1;
    {
      double *v = (double *)calloc((size_t)M, sizeof(double));
      if (kind == OR_REDFT01) {
        v[0] = x[0];
        for (int j = 1; j < n; j++) { v[j] = x[j]; v[2 * n - j] = -x[j]; }
      } else if (kind == OR_RODFT01) {
        for (int j = 1; j < n; j++) { v[j] = x[j - 1]; v[2 * n - j] = x[j - 1]; }
        v[n] = x[n - 1];
      } else {
        double s = (kind == OR_REDFT11) ? -1.0 : 1.0;
        for (int j = 0; j < n; j++) { v[j] = x[j]; v[2 * n - 1 - j] = s * x[j]; }
      }
      for (int m = 0; m < M; m++) {
        long double a = PI_L * (long double)m / (long double)(2 * n);
        w[2 * m] = (double)(v[m] * cosl(a));
        w[2 * m + 1] = (double)(v[m] * sinl(a));
      }
      free(v);
    }
    break;
  }
  or_fft(M, w, F, sign);
  for (int k = 0; k < n; k++) {
    switch (kind) {
    case OR_REDFT00: y[k] = F[2 * k]; break;
    case OR_RODFT00: y[k] = -F[2 * (k + 1) + 1]; break;
    case OR_REDFT10: {
      long double a = -PI_L * (long double)k / (long double)(2 * n);
      y[k] = (double)(F[2 * k] * cosl(a) - F[2 * k + 1] * sinl(a));
      break;
    }
    case OR_RODFT10: {
      int kk = k + 1;
      long double a = -PI_L * (long double)kk / (long double)(2 * n);
      /* Re(i * e^{ia} * F) = -Im(e^{ia} F) */
      long double im = F[2 * kk] * sinl(a) + F[2 * kk + 1] * cosl(a);
      y[k] = (double)(-im);
      break;
    }
    case OR_REDFT01: y[k] = F[2 * k]; break;
    case OR_RODFT01: y[k] = F[2 * k + 1]; break; /* Re(-i F) = Im F */
    case OR_REDFT11: {
      long double a = PI_L * (long double)(2 * k + 1) / (long double)(4 * n);
      y[k] = (double)(F[2 * k] * cosl(a) - F[2 * k + 1] * sinl(a));
      break;
    }
    case OR_RODFT11: {
      long double a = PI_L * (long double)(2 * k + 1) / (long double)(4 * n);
      /* Re(-i e^{ia} F) = Im(e^{ia} F) */
      y[k] = (double)(F[2 * k] * sinl(a) + F[2 * k + 1] * cosl(a));
      break;
    }
    }
  }
  free(w); free(F);
}

/* ------------------------------------------------------------------ */
/* Planner restatement (1x1 grid)                                      */
/* ------------------------------------------------------------------ */

/* build/exec.C:2959-2967 */
void or_inv_mo(const int mo[3], int imo[3])
{
  for (int i = 0; i < 3; i++) imo[mo[i]] = i;
}

/* build/exec.C:2270-2296: mc[i] is the position j with mo2[j] == mo1[i]. */
void or_rel_change(const int *mo1, const int *mo2, int *mc)
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (mo1[i] == mo2[j]) { mc[i] = j; break; }
}

/* build/templ.C:816-858 */
void or_swap0(int newmo[3], const int mo[3], int L)
{
  int t[3] = {mo[0], mo[1], mo[2]};
  newmo[0] = t[0]; newmo[1] = t[1]; newmo[2] = t[2];
  if (newmo[L] == 0) return;
  if (newmo[L] == 1) {
    int j;
    for (j = 0; j < 3; j++) if (newmo[j] == 0) break;
    newmo[j] = 1;
    newmo[L] = 0;
    return;
  }
  for (int i = 0; i < 3; i++) newmo[i] = (newmo[i] + 1) % 3;
}

/* build/templ.C:872-882 */
int or_excl(int a, int b)
{
  if (a * b == 0) return (a > b ? a : b) == 1 ? 2 : 1;
  return 0;
}

static int find_zero(const int ar[3])
{
  for (int i = 0; i < 3; i++) if (ar[i] == 0) return i;
  return -1;
}

/* build/templ.C:627-802, 1x1 branch.  On a 1x1 grid every Pdims is 1, so
 * the grid's local dims are L = {0,1,2} and nd = 1 (build/init.C:1637-1642,
 * 1811-1815). */
void or_find_order(const int types[3], const int mo1[3], const int mo2[3], int L[3])
{
  L[0] = L[1] = L[2] = -1;
  for (int i = 0; i < 3; i++) {
    const or_type_t *t = &OR_TYPES[types[i]];
    if (t->dt1 < t->dt2) { if (L[0] < 0) L[0] = i; }
    else if (t->dt1 > t->dt2) { if (L[2] < 0) L[2] = i; }
  }
  const int gl0 = 0, gl1 = 1; /* grid1->L[0], grid1->L[1] */
  if (L[0] < 0) {
    if (L[2] < 0) {
      L[0] = gl0;
      if (mo1[L[0]] > mo1[gl1]) L[0] = gl1;
    } else {
      L[0] = gl0;
      if (L[0] == L[2]) L[0] = gl1;
    }
  }
  if (L[2] < 0) {
    L[2] = find_zero(mo2);
    if (L[2] == L[0]) L[2] = (L[0] + 1) % 3;
  }
  L[1] = or_excl(L[0], L[2]);
}

/* build/templ.C:91-435 restated for TRANS_ONLY stages. */
int or_plan3d(const int types[3], const int gdims1[3], const int mo1[3],
              const int mo2[3], or_stage_t st_out[4])
{
  int L[3];
  or_find_order(types, mo1, mo2, L);
  int gdims[3] = {gdims1[0], gdims1[1], gdims1[2]};
  int dim_conj_sym = -1;
  int mocurr[3], monext[3] = {mo1[0], mo1[1], mo1[2]};
  int ns = 0;
  for (int st = 0; st < 3; st++) {
    memcpy(mocurr, monext, sizeof(mocurr));
    int din[3] = {gdims[0], gdims[1], gdims[2]};
    const or_type_t *t = &OR_TYPES[types[L[st]]];
    if (t->dt1 < t->dt2) {
      dim_conj_sym = L[st];
      gdims[dim_conj_sym] = gdims[dim_conj_sym] / 2 + 1;
    } else if (t->dt2 < t->dt1) {
      int d = dim_conj_sym >= 0 ? dim_conj_sym : L[st];
      gdims[d] = (gdims[d] - 1) * 2;
      dim_conj_sym = -1;
    }
    if (st < 2) {
      if (t->cls != OR_EMPTY) or_swap0(monext, mocurr, L[st]);
    } else {
      memcpy(monext, mo2, sizeof(monext));
      if (monext[L[st]] != 0) or_swap0(monext, mocurr, L[st]);
    }
    or_stage_t *s = &st_out[ns++];
    s->trans_dim = L[st];
    s->type_id = types[L[st]];
    memcpy(s->mo_in, mocurr, sizeof(mocurr));
    memcpy(s->mo_out, monext, sizeof(monext));
    memcpy(s->dims_in, din, sizeof(din));
    memcpy(s->dims_out, gdims, sizeof(gdims));
  }
  if (monext[0] != mo2[0] || monext[1] != mo2[1] || monext[2] != mo2[2]) {
    /* final_seq (build/templ.C:422-435): EMPTY reorder into grid2's order */
    or_stage_t *s = &st_out[ns++];
    int td = 0;
    if (monext[td] != 0 && mo2[td] != 0) td = 1;
    const or_type_t *last = &OR_TYPES[types[L[2]]];
    int empty_id = (last->dt2 == 1) ? (last->prec == 4 ? 0 : 1) : (last->prec == 4 ? 2 : 3);
    s->trans_dim = td;
    s->type_id = empty_id;
    memcpy(s->mo_in, monext, sizeof(monext));
    memcpy(s->mo_out, mo2, sizeof(int) * 3);
    memcpy(s->dims_in, gdims, sizeof(gdims));
    memcpy(s->dims_out, gdims, sizeof(gdims));
  }
  return ns;
}

/* ------------------------------------------------------------------ */
/* Execution                                                           */
/* ------------------------------------------------------------------ */

static long storage_offset(const int idx[3], const int dims[3], const int mo[3])
{
  int imo[3], s[3], sd[3];
  or_inv_mo(mo, imo);
  for (int i = 0; i < 3; i++) { s[mo[i]] = idx[i]; sd[mo[i]] = dims[i]; }
  (void)imo;
  return s[0] + (long)sd[0] * (s[1] + (long)sd[1] * s[2]);
}

/* compute_deriv_loc on one line (build/exec.C:228-287). */
static void deriv_line(cx *v, int len, int r2c)
{
  int g = r2c ? (len - 1) * 2 : len;
  int mid = g / 2;
  for (int i = 0; i < len; i++) {
    double kk;
    if (i < mid) kk = i;
    else if (i == mid) { v[i].re = v[i].im = 0; continue; }
    else kk = i - g;
    double re = v[i].re, im = v[i].im;
    v[i].re = -kk * im;
    v[i].im = kk * re;
  }
}

/* Transform one line (complex work layout) of type t. */
static void line_transform(const or_type_t *t, int nin, int nout, cx *vin, cx *vout)
{
  switch (t->cls) {
  case OR_EMPTY:
    memcpy(vout, vin, sizeof(cx) * nout);
    break;
  case OR_C2C:
    or_fft(nin, (const double *)vin, (double *)vout, t->isign);
    break;
  case OR_R2C: {
    double *r = (double *)malloc(sizeof(double) * nin);
    for (int j = 0; j < nin; j++) r[j] = vin[j].re;
    double *o = (double *)malloc(sizeof(cx) * (nin / 2 + 1));
    or_r2c(nin, r, o);
    memcpy(vout, o, sizeof(cx) * nout);
    free(r); free(o);
    break;
  }
  case OR_C2R: {
    /* the C2R length is the output length; input holds nout/2+1 values */
    double *r = (double *)malloc(sizeof(double) * nout);
    double *ci = (double *)calloc((size_t)(nout / 2 + 1), sizeof(cx));
    int h = nout / 2 + 1 < nin ? nout / 2 + 1 : nin;
    memcpy(ci, vin, sizeof(cx) * h);
    or_c2r(nout, ci, r);
    for (int j = 0; j < nout; j++) { vout[j].re = r[j]; vout[j].im = 0; }
    free(r); free(ci);
    break;
  }
  case OR_R2R: {
    double *a = (double *)malloc(sizeof(double) * nin);
    double *b = (double *)malloc(sizeof(double) * nin);
    for (int j = 0; j < nin; j++) a[j] = vin[j].re;
    or_r2r(t->r2r_kind, nin, a, b);
    for (int j = 0; j < nout; j++) { vout[j].re = b[j]; vout[j].im = 0; }
    if (t->dt1 == 2) { /* complex variant: re and im transformed separately (init.C:1179-1188) */
      for (int j = 0; j < nin; j++) a[j] = vin[j].im;
      or_r2r(t->r2r_kind, nin, a, b);
      for (int j = 0; j < nout; j++) vout[j].im = b[j];
    }
    free(a); free(b);
    break;
  }
  }
}

int or_exec_1d(int type_id, int d, const int dims_in[3], const int mo_in[3],
               const int dims_out[3], const int mo_out[3],
               const double *in, double *out, int deriv)
{
  if (type_id < 0 || type_id >= 44 || d < 0 || d > 2) return -1;
  const or_type_t *t = &OR_TYPES[type_id];
  for (int i = 0; i < 3; i++)
    if (i != d && dims_in[i] != dims_out[i]) return -2;
  int e = (d == 0) ? 1 : 0, f = 3 - d - e;
  int nin = dims_in[d], nout = dims_out[d];
  int nmax = nin > nout ? nin : nout;
  cx *vin = (cx *)calloc((size_t)nmax, sizeof(cx));
  cx *vout = (cx *)calloc((size_t)nmax, sizeof(cx));
  int idx[3];
  for (int jf = 0; jf < dims_in[f]; jf++)
    for (int je = 0; je < dims_in[e]; je++) {
      idx[e] = je; idx[f] = jf;
      for (int k = 0; k < nin; k++) {
        idx[d] = k;
        long o = storage_offset(idx, dims_in, mo_in);
        if (t->dt1 == 1) { vin[k].re = in[o]; vin[k].im = 0; }
        else { vin[k].re = in[2 * o]; vin[k].im = in[2 * o + 1]; }
      }
      line_transform(t, nin, nout, vin, vout);
      if (deriv && t->dt2 == 2 && t->cls != OR_EMPTY) deriv_line(vout, nout, t->cls == OR_R2C);
      for (int k = 0; k < nout; k++) {
        idx[d] = k;
        long o = storage_offset(idx, dims_out, mo_out);
        if (t->dt2 == 1) out[o] = vout[k].re;
        else { out[2 * o] = vout[k].re; out[2 * o + 1] = vout[k].im; }
      }
    }
  free(vin); free(vout);
  return 0;
}

int or_exec3d(const int types[3], const int gdims1[3], const int mo1[3],
              const int mo2[3], const double *in, double *out, int idir)
{
  or_stage_t st[4];
  int ns = or_plan3d(types, gdims1, mo1, mo2, st);
  /* ping-pong over full-volume complex buffers */
  size_t cap = 1;
  for (int s = 0; s < ns; s++) {
    size_t a = 2, b = 2;
    for (int i = 0; i < 3; i++) { a *= st[s].dims_in[i]; b *= st[s].dims_out[i]; }
    if (a > cap) cap = a;
    if (b > cap) cap = b;
  }
  double *bufA = (double *)calloc(cap, sizeof(double));
  double *bufB = (double *)calloc(cap, sizeof(double));
  const double *src = in;
  double *dst = NULL;
  int rc = 0;
  for (int s = 0; s < ns; s++) {
    const or_type_t *t = &OR_TYPES[st[s].type_id];
    dst = (s == ns - 1) ? out : ((s % 2 == 0) ? bufA : bufB);
    int deriv = (idir == st[s].trans_dim) && t->cls != OR_EMPTY;
    rc = or_exec_1d(st[s].type_id, st[s].trans_dim, st[s].dims_in, st[s].mo_in,
                    st[s].dims_out, st[s].mo_out, src, dst, deriv);
    if (rc) break;
    src = dst;
  }
  free(bufA); free(bufB);
  return rc;
}

/* build/deriv.C:85-186, including its choice of storage axis: ldir is the
 * logical index i with MemOrder[i] == idir, then used as a storage rank. */
void or_compute_deriv(const double *in, double *out, const int gdims[3],
                      const int ldims[3], const int glob_start[3],
                      const int mem_order[3], int dim_conj_sym, int idir)
{
  int sdims[3], ldir = 0;
  for (int i = 0; i < 3; i++) {
    sdims[mem_order[i]] = ldims[i];
    if (mem_order[i] == idir) ldir = i;
  }
  int g = (dim_conj_sym == idir) ? (gdims[idir] - 1) * 2 : gdims[idir];
  int mid = g / 2 - glob_start[idir];
  long n = (long)sdims[0] * sdims[1] * sdims[2];
  for (long p = 0; p < n; p++) {
    int s[3];
    s[0] = (int)(p % sdims[0]);
    s[1] = (int)((p / sdims[0]) % sdims[1]);
    s[2] = (int)(p / ((long)sdims[0] * sdims[1]));
    int i = s[ldir];
    double re = in[2 * p], im = in[2 * p + 1];
    double kk;
    if (i < mid) kk = i + glob_start[idir];
    else if (i == mid) { out[2 * p] = 0; out[2 * p + 1] = 0; continue; }
    else kk = i + glob_start[idir] - g;
    out[2 * p] = -kk * im;
    out[2 * p + 1] = kk * re;
  }
}
