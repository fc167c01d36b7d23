// Host build of the pass kernel, for CPU tests only (never part of the
// product library).  It plans with the product's pass_plan.cpp and runs each
// workgroup of each pass through the same run_tile() the HIP kernel runs,
// one workgroup at a time with a single thread and a no-op barrier.  This
// checks the kernel's indexing, radix passes, extension rules and layout
// changes without a GPU.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../p3dfft.3_amd/csrc/host/pass_plan.h"
#include "../../p3dfft.3_amd/csrc/kernels/fft_core.h"
#include "../../p3dfft.3_amd/csrc/kernels/fft_fast.h"
#include "../../p3dfft.3_amd/csrc/kernels/fft_mid.h"

using namespace p3e;

struct HostBarrier {
  void operator()() const {}
};

template <typename T> static std::vector<T> tab(const std::vector<cld> &v) {
  std::vector<T> h(2 * v.size());
  for (size_t i = 0; i < v.size(); i++) {
    h[2 * i] = (T)v[i].real();
    h[2 * i + 1] = (T)v[i].imag();
  }
  return h;
}

// Every thread of the workgroup is emulated for a radix pass: all read
// halves, then all write halves, each thread keeping its own registers.
// set when a split pass asks for more threads than its workgroup has
static bool g_pair_overflow = false;

template <typename T> struct HostPassRun {
  int nth;
  std::vector<cx<T>> v;
  template <class PS> void operator()(PS &ps) {
    v.assign((size_t)nth * PS::SLOTS, cx<T>{(T)1e30, (T)-1e30});
    for (int tid = 0; tid < nth; tid++) ps.read(tid, nth, &v[(size_t)tid * PS::SLOTS]);
    for (int tid = 0; tid < nth; tid++) ps.write(tid, nth, &v[(size_t)tid * PS::SLOTS]);
  }
  template <class PA, class PB> void pair(PA &pa, int na, PB &pb, int nb) {
    if (na + nb > nth) {
      g_pair_overflow = true;
      return;
    }
    std::vector<cx<T>> va((size_t)na * PA::SLOTS, cx<T>{(T)1e30, (T)-1e30});
    std::vector<cx<T>> vb((size_t)nb * PB::SLOTS, cx<T>{(T)1e30, (T)-1e30});
    for (int tid = 0; tid < na; tid++) pa.read(tid, na, &va[(size_t)tid * PA::SLOTS]);
    for (int tid = 0; tid < nb; tid++) pb.read(tid, nb, &vb[(size_t)tid * PB::SLOTS]);
    for (int tid = 0; tid < na; tid++) pa.write(tid, na, &va[(size_t)tid * PA::SLOTS]);
    for (int tid = 0; tid < nb; tid++) pb.write(tid, nb, &vb[(size_t)tid * PB::SLOTS]);
  }
};

// A plain phase: every thread of the workgroup in turn.
struct HostEach {
  int nth;
  template <class F> void operator()(F f) const {
    for (int tid = 0; tid < nth; tid++) f(tid);
  }
};

// The workgroup's LDS as the kernel lays it out (fixed_kernels.h): the tile,
// then the twiddle table, then the R2C/C2R split factors, in exactly the
// planned lds_bytes, followed by a guard region that must stay untouched.
// Passes planned for the batch step kernel read their tables from global
// memory (their LDS holds the tile only).
static bool g_batch_plans = false;

template <typename T, int M> static int run_pow2(const PassDesc &pd, const PassArgs &a) {
  HostPassRun<T> run{pd.threads, std::vector<cx<T>>((size_t)pd.threads * P3E_FAST_E)};
  HostEach each{pd.threads};
  const size_t n_lds = pd.lds_bytes / sizeof(cx<T>), guard = 64;
  const size_t tile = (size_t)a.tab_off;
  const size_t ntab = g_batch_plans ? 0 : (size_t)a.ntw + (a.cls == CLS_R2C || a.cls == CLS_C2R ? M + 1 : 0);
  if (tile + ntab > n_lds) {
    set_error("power-of-two pass: tile and tables need %zu entries, LDS holds %zu",
              tile + ntab, n_lds);
    return -1;
  }
  std::vector<cx<T>> lds(n_lds + guard);
  const cx<T> poison{(T)1e30, (T)-1e30};
  for (size_t i = n_lds; i < n_lds + guard; i++) lds[i] = poison;
  const cx<T> *gtw = reinterpret_cast<const cx<T> *>(a.tw), *gwn = reinterpret_cast<const cx<T> *>(a.wN);
  cx<T> *ltw = lds.data() + tile, *lwn = ltw + a.ntw;
  if (!g_batch_plans) {
    for (int i = 0; i < a.ntw; i++) ltw[i] = gtw[i];
    if (a.cls == CLS_R2C || a.cls == CLS_C2R)
      for (int i = 0; i <= M; i++) lwn[i] = gwn[i];
  }
  for (unsigned int t = 0; t < pd.nblocks; t++) {
    for (size_t i = 0; i < tile; i++) lds[i] = poison;
    fast_tile<T, M>(a, t, lds.data(), pd.threads, run, each, g_batch_plans ? gtw : ltw, g_batch_plans ? gwn : lwn);
  }
  if (g_pair_overflow) {
    g_pair_overflow = false;
    set_error("split pass: its two sequences need more threads than the workgroup has");
    return -1;
  }
  for (size_t i = n_lds; i < n_lds + guard; i++)
    if (lds[i].x != poison.x || lds[i].y != poison.y) {
      set_error("power-of-two pass wrote past its LDS");
      return -1;
    }
  return 0;
}

// The single-buffer path for one instantiation of its odd prime radix.
template <typename T, int P> static int run_mid(const PassDesc &pd, const PassArgs &a) {
  HostPassRun<T> run{pd.threads, {}};
  HostEach each{pd.threads};
  std::vector<cx<T>> buf((size_t)a.B * a.lpad);
  for (unsigned int t = 0; t < pd.nblocks; t++) {
    for (auto &x : buf) x = cx<T>{(T)1e30, (T)-1e30};
    mid_tile<T, P>(a, t, buf.data(), pd.threads, run, each, reinterpret_cast<const cx<T> *>(a.tw));
  }
  return 0;
}

template <typename T> static int run_pass(const PassDesc &pd, const void *in, void *out, int deriv) {
  std::vector<T> tw = tab<T>(pd.tw), ch = tab<T>(pd.chirp), chh = tab<T>(pd.chirp_hat), pre = tab<T>(pd.pre),
                 post = tab<T>(pd.post), wn = tab<T>(pd.wN);
  PassArgs a = pd.args;
  a.in = in;
  a.out = out;
  a.tw = tw.empty() ? nullptr : tw.data();
  a.chirp = ch.empty() ? nullptr : ch.data();
  a.chirp_hat = chh.empty() ? nullptr : chh.data();
  a.pre = pre.empty() ? nullptr : pre.data();
  a.post = post.empty() ? nullptr : post.data();
  a.wN = wn.empty() ? nullptr : wn.data();
  a.deriv = deriv;
  if (a.mid) {
    switch (a.mid_prime) {
      case 0: return run_mid<T, 0>(pd, a);
      case 7: return run_mid<T, 7>(pd, a);
      case 11: return run_mid<T, 11>(pd, a);
      case 13: return run_mid<T, 13>(pd, a);
      case 17: return run_mid<T, 17>(pd, a);
      case 19: return run_mid<T, 19>(pd, a);
      case 23: return run_mid<T, 23>(pd, a);
      case 29: return run_mid<T, 29>(pd, a);
      case 31: return run_mid<T, 31>(pd, a);
      case 37: return run_mid<T, 37>(pd, a);
      case 41: return run_mid<T, 41>(pd, a);
      case 43: return run_mid<T, 43>(pd, a);
      case 47: return run_mid<T, 47>(pd, a);
    }
    return -1;
  }
  if (a.fast) {
    switch (a.M) {
      case 16: return run_pow2<T, 16>(pd, a);
      case 32: return run_pow2<T, 32>(pd, a);
      case 64: return run_pow2<T, 64>(pd, a);
      case 128: return run_pow2<T, 128>(pd, a);
      case 256: return run_pow2<T, 256>(pd, a);
      case 512: return run_pow2<T, 512>(pd, a);
      case 1024: return run_pow2<T, 1024>(pd, a);
      case 2048: return run_pow2<T, 2048>(pd, a);
      case 4096: return run_pow2<T, 4096>(pd, a);
      case 48: return run_pow2<T, 48>(pd, a);
      case 96: return run_pow2<T, 96>(pd, a);
      case 192: return run_pow2<T, 192>(pd, a);
      case 384: return run_pow2<T, 384>(pd, a);
      case 768: return run_pow2<T, 768>(pd, a);
      case 80: return run_pow2<T, 80>(pd, a);
      case 160: return run_pow2<T, 160>(pd, a);
      case 320: return run_pow2<T, 320>(pd, a);
      case 640: return run_pow2<T, 640>(pd, a);
      case 120: return run_pow2<T, 120>(pd, a);
      case 60: return run_pow2<T, 60>(pd, a);
      case 72: return run_pow2<T, 72>(pd, a);
      case 144: return run_pow2<T, 144>(pd, a);
      case 112: return run_pow2<T, 112>(pd, a);
      case 224: return run_pow2<T, 224>(pd, a);
      case 240: return run_pow2<T, 240>(pd, a);
      case 36: return run_pow2<T, 36>(pd, a);
    }
    return -1;
  }
  std::vector<cx<T>> A((size_t)a.B * a.lpad), B((size_t)a.B * a.lpad);
  for (unsigned int t = 0; t < pd.nblocks; t++) {
    // poison LDS between tiles so stale data cannot hide a missing write
    for (auto &v : A) v = cx<T>{(T)1e30, (T)-1e30};
    for (auto &v : B) v = cx<T>{(T)-1e30, (T)1e30};
    run_tile<T>(a, t, A.data(), B.data(), 0, 1, HostBarrier(), reinterpret_cast<const cx<T> *>(a.tw));
  }
  return 0;
}

static int run_chain(int prec, const PassChain &ch, const void *in, void *out, int idir) {
  std::vector<std::vector<char>> bufs(ch.passes.size());
  const void *src = in;
  for (size_t i = 0; i < ch.passes.size(); i++) {
    const PassDesc &pd = ch.passes[i];
    void *dst;
    if (i + 1 == ch.passes.size()) dst = out;
    else {
      bufs[i].assign(pd.out_bytes, 0);
      dst = bufs[i].data();
    }
    const int deriv = (idir >= 0 && idir == pd.axis && pd.deriv_capable) ? 1 : 0;
    const int rc = prec == 8 ? run_pass<double>(pd, src, dst, deriv) : run_pass<float>(pd, src, dst, deriv);
    if (rc) return rc;
    src = dst;
  }
  return 0;
}

extern "C" {

// Plans and runs a transform on host arrays.  ops_*[a]: class, sign, r2r
// kind, input/output element type of the operation on logical axis a.
// Returns 0, or -1 with the planner's message in emu_error().
int emu_exec(int prec, const int *cls, const int *sign, const int *kind, const int *idt, const int *odt,
             const int *order, const int *in_dims, const int *in_mo, int in_dt, const int *out_dims,
             const int *out_mo, int out_dt, const void *in, void *out, int idir, char *desc, int desc_len,
             int batch_threads) {
  OpSpec ops[3];
  for (int a = 0; a < 3; a++) ops[a] = OpSpec{cls[a], sign[a], kind[a], idt[a], odt[a]};
  Layout li, lo;
  for (int i = 0; i < 3; i++) {
    li.dims[i] = in_dims[i];
    li.mo[i] = in_mo[i];
    lo.dims[i] = out_dims[i];
    lo.mo[i] = out_mo[i];
  }
  li.dt = in_dt;
  lo.dt = out_dt;
  PassChain ch;
  if (!plan_chain(prec, ops, order, li, lo, ch, batch_threads)) return -1;
  if (desc && desc_len > 0) {
    strncpy(desc, ch.desc.c_str(), desc_len - 1);
    desc[desc_len - 1] = 0;
  }
  const size_t out_bytes = layout_elems(lo) * elem_bytes(prec, lo.dt);
  if (ch.copy_only) {
    memcpy(out, in, out_bytes);
    return 0;
  }
  g_batch_plans = batch_threads > 0;
  return run_chain(prec, ch, in, out, idir);
}

// As emu_exec in batch mode, with every pass that accepts it re-tiled by
// the autotuner's retune_pass to B lines and the radix sequence that is the
// first of radix_candidates(L) in descending order (`pick` selects another
// candidate).  Returns the number of re-tiled passes, or -1.
int emu_exec_retuned(int prec, const int *cls, const int *sign, const int *kind, const int *idt, const int *odt,
                     const int *order, const int *in_dims, const int *in_mo, int in_dt, const int *out_dims,
                     const int *out_mo, int out_dt, const void *in, void *out, int B, int pick, char *desc,
                     int desc_len) {
  OpSpec ops[3];
  for (int a = 0; a < 3; a++) ops[a] = OpSpec{cls[a], sign[a], kind[a], idt[a], odt[a]};
  Layout li, lo;
  for (int i = 0; i < 3; i++) {
    li.dims[i] = in_dims[i];
    li.mo[i] = in_mo[i];
    lo.dims[i] = out_dims[i];
    lo.mo[i] = out_mo[i];
  }
  li.dt = in_dt;
  lo.dt = out_dt;
  PassChain ch;
  if (!plan_chain(prec, ops, order, li, lo, ch, P3E_STEP_THREADS)) return -1;
  int n = 0;
  std::string d;
  for (PassDesc &pd : ch.passes) {
    std::vector<std::vector<int>> c = radix_candidates(pd.args.L, 4);
    if (c.empty()) continue;
    const std::vector<int> &r = c[(size_t)pick % c.size()];
    PassDesc t = pd;
    if (retune_pass(prec, t, B, r.data(), (int)r.size())) {
      pd = t;
      n++;
      char buf[96];
      snprintf(buf, sizeof(buf), "[%s B=%d nrad=%d th=%d] ", pd.args.mid ? "mid" : "gen", pd.args.B, pd.args.nrad,
               pd.threads);
      d += buf;
    }
  }
  if (desc && desc_len > 0) {
    strncpy(desc, d.c_str(), desc_len - 1);
    desc[desc_len - 1] = 0;
  }
  g_batch_plans = true;  // retuned passes are batch-kernel plans
  if (run_chain(prec, ch, in, out, -1)) return -1;
  return n;
}

// Periodic Poisson solve of a real grid through the product's Poisson plan
// (forward chain with the fused -1/|k|^2 and 1/N multiply, backward chain).
int emu_poisson(int prec, const int *dims, const int *mo_real, const int *mo_spec, const void *rhs, void *sol,
                void *spec, char *desc, int desc_len, int batch_threads) {
  PassChain f, b;
  if (!plan_poisson(prec, dims, mo_real, mo_spec, f, b, batch_threads)) return -1;
  if (desc && desc_len > 0) {
    strncpy(desc, (f.desc + " || " + b.desc).c_str(), desc_len - 1);
    desc[desc_len - 1] = 0;
  }
  g_batch_plans = batch_threads > 0;
  if (run_chain(prec, f, rhs, spec, -1) || run_chain(prec, b, spec, sol, -1)) return -1;
  return 0;
}

const char *emu_error(void) { return last_error(); }

// 1: plan with the power-of-two path where eligible (the product default);
// 0: generic path only.
void emu_set_fast(int enable) { set_fast_path(enable != 0); }
void emu_set_dct_split(int on) { set_dct_split(on != 0); }
void emu_set_pad_threshold(long long bytes) { set_pad_threshold((size_t)bytes); }

// Buffer assignment of a plan, for tests: dst[i] per pass, scratch sizes.
int emu_assign(int prec, const int *cls, const int *sign, const int *kind, const int *idt, const int *odt,
               const int *order, const int *in_dims, const int *in_mo, int in_dt, const int *out_dims,
               const int *out_mo, int out_dt, long long out_cap, int alias, int *dst, long long *need,
               int *copy_back) {
  OpSpec ops[3];
  for (int a = 0; a < 3; a++) ops[a] = OpSpec{cls[a], sign[a], kind[a], idt[a], odt[a]};
  Layout li, lo;
  for (int i = 0; i < 3; i++) {
    li.dims[i] = in_dims[i];
    li.mo[i] = in_mo[i];
    lo.dims[i] = out_dims[i];
    lo.mo[i] = out_mo[i];
  }
  li.dt = in_dt;
  lo.dt = out_dt;
  PassChain ch;
  if (!plan_chain(prec, ops, order, li, lo, ch)) return -1;
  size_t nd[2];
  bool cb;
  assign_buffers(ch.passes, (size_t)out_cap, alias != 0, dst, nd, &cb);
  need[0] = (long long)nd[0];
  need[1] = (long long)nd[1];
  *copy_back = cb;
  return (int)ch.passes.size();
}

// The launch guard: number of passes whose tables are not ready with the
// planner's null table addresses (*before) and with every built table given
// an address (*after).
int emu_tables_ready(int prec, const int *cls, const int *sign, const int *kind, const int *idt, const int *odt,
                     const int *order, const int *in_dims, const int *in_mo, int in_dt, const int *out_dims,
                     const int *out_mo, int out_dt, int *before, int *after) {
  OpSpec ops[3];
  for (int a = 0; a < 3; a++) ops[a] = OpSpec{cls[a], sign[a], kind[a], idt[a], odt[a]};
  Layout li, lo;
  for (int i = 0; i < 3; i++) {
    li.dims[i] = in_dims[i];
    li.mo[i] = in_mo[i];
    lo.dims[i] = out_dims[i];
    lo.mo[i] = out_mo[i];
  }
  li.dt = in_dt;
  lo.dt = out_dt;
  PassChain ch;
  if (!plan_chain(prec, ops, order, li, lo, ch)) return -1;
  *before = *after = 0;
  static const char dummy[16] = {0};
  for (PassDesc &pd : ch.passes) {
    if (!tables_ready(pd)) ++*before;
    PassArgs &a = pd.args;
    if (!pd.tw.empty()) a.tw = dummy;
    if (!pd.chirp.empty()) a.chirp = dummy;
    if (!pd.chirp_hat.empty()) a.chirp_hat = dummy;
    if (!pd.pre.empty()) a.pre = dummy;
    if (!pd.post.empty()) a.post = dummy;
    if (!pd.wN.empty()) a.wN = dummy;
    if (!tables_ready(pd)) ++*after;
  }
  return (int)ch.passes.size();
}

}  // extern "C"
