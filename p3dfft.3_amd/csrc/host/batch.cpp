// Batched execution: many independent transforms of different sizes.  When
// every item's passes run on the power-of-two or single-buffer kernels,
// each stage is one grouped launch per pass kernel over all items' tiles
// (the dedicated kernels, with their paired accesses and LDS twiddles);
// otherwise one launch of the persistent step kernel per stage, with tile
// shapes and radix orders of its generic-path passes chosen by an
// autotuner.  Either way the launches can be replayed from a hipGraph.
#include "batch.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>

extern "C" int p3e_launch_step(const void *table, const void *work, int nwork, int prec, size_t lds,
                               unsigned int grid, hipStream_t st);
extern "C" int p3e_step_occupancy(int prec, size_t lds);
extern "C" int p3e_launch_group(int kind, int M, const void *tab, const void *start, int n, int prec, size_t lds,
                                int threads, unsigned int nblocks, hipStream_t st);

namespace p3e {

#define BCHK(x)                                                  \
  do {                                                           \
    hipError_t e_ = (x);                                         \
    if (e_ != hipSuccess) {                                      \
      set_error("%s failed: %s", #x, hipGetErrorString(e_));     \
      return -1;                                                 \
    }                                                            \
  } while (0)

template <typename T> static void *upload_t(const std::vector<cld> &v, std::vector<void *> &owned) {
  std::vector<T> h(2 * v.size());
  for (size_t i = 0; i < v.size(); i++) {
    h[2 * i] = (T)v[i].real();
    h[2 * i + 1] = (T)v[i].imag();
  }
  void *d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(T)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  owned.push_back(d);
  return d;
}

bool upload_pass_tables(int prec, PassDesc &pd, std::vector<void *> &owned) {
  PassArgs &a = pd.args;
  const std::vector<cld> *src[6] = {&pd.tw, &pd.chirp, &pd.chirp_hat, &pd.pre, &pd.post, &pd.wN};
  const void **dst[6] = {&a.tw, &a.chirp, &a.chirp_hat, &a.pre, &a.post, &a.wN};
  for (int i = 0; i < 6; i++) {
    *dst[i] = nullptr;
    if (src[i]->empty()) continue;
    void *d = prec == 8 ? upload_t<double>(*src[i], owned) : upload_t<float>(*src[i], owned);
    if (!d) {
      set_error("hipMalloc/hipMemcpy failed for a constant table");
      return false;
    }
    *dst[i] = d;
  }
  return true;
}

unsigned int persistent_grid(int prec, size_t lds, int nwork) {
  // CU count of the current device, looked up once (the batch and the
  // Poisson service call this from different threads)
  static std::once_flag once;
  static int cus = 256;
  std::call_once(once, [] {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
  });
  int per_cu = p3e_step_occupancy(prec, lds);
  if (per_cu < 1) per_cu = 1;
  long g = (long)cus * per_cu;
  if (g > nwork) g = nwork;
  return (unsigned int)std::max(1L, g);
}

// ------------------------------------------------------------------ autotuner
struct TuneChoice {
  int B;
  std::vector<int> rad;
};
static std::map<std::string, TuneChoice> g_tune;
static bool g_tune_loaded = false;

static std::string cache_path() {
  const char *p = getenv("P3DFFT_HIP_TUNE_CACHE");
  if (p && *p) return p;
  const char *h = getenv("HOME");
  return std::string(h ? h : "/tmp") + "/.cache/p3dfft3_hip_tune.txt";
}

static void load_cache() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  FILE *f = fopen(cache_path().c_str(), "r");
  if (!f) return;
  char key[256], rads[256];
  int B;
  while (fscanf(f, "%255s %d %255s", key, &B, rads) == 3) {
    TuneChoice c;
    c.B = B;
    for (char *t = strtok(rads, ","); t; t = strtok(nullptr, ",")) c.rad.push_back(atoi(t));
    g_tune[key] = c;
  }
  fclose(f);
}

static void save_cache() {
  std::string path = cache_path();
  const size_t slash = path.find_last_of('/');
  if (slash != std::string::npos && slash > 0) (void)mkdir(path.substr(0, slash).c_str(), 0755);
  FILE *f = fopen(path.c_str(), "w");
  if (!f) return;
  for (auto &kv : g_tune) {
    fprintf(f, "%s %d ", kv.first.c_str(), kv.second.B);
    for (size_t i = 0; i < kv.second.rad.size(); i++) fprintf(f, "%s%d", i ? "," : "", kv.second.rad[i]);
    fprintf(f, "\n");
  }
  fclose(f);
}

int tune_cache_size() {
  load_cache();
  return (int)g_tune.size();
}

static std::string tune_key(int prec, const PassDesc &pd) {
  const PassArgs &a = pd.args;
  char k[256];
  snprintf(k, sizeof(k), "gfx950:p%d:c%d:k%d:N%d:W%d:L%d:o%d%d:e%d:m%d", prec, a.cls, a.r2r_kind, a.N, a.W, a.L,
           a.load_order, a.store_order, a.ne, a.mid);
  return k;
}

// Time one candidate: all tiles of the pass through the step kernel.
static float time_candidate(int prec, PassDesc &pd, void *din, void *dout, PassArgs *d_tab, WorkItem *d_work,
                            int maxwork, hipStream_t st) {
  PassArgs a = pd.args;
  a.in = din;
  a.out = dout;
  a.deriv = 0;
  const int nw = std::min<int>((int)pd.nblocks, maxwork);
  std::vector<WorkItem> w(nw);
  for (int t = 0; t < nw; t++) w[t] = WorkItem{0, (uint32_t)t};
  if (hipMemcpy(d_tab, &a, sizeof(a), hipMemcpyHostToDevice) != hipSuccess) return 1e30f;
  if (hipMemcpy(d_work, w.data(), sizeof(WorkItem) * nw, hipMemcpyHostToDevice) != hipSuccess) return 1e30f;
  const unsigned int grid = persistent_grid(prec, pd.lds_bytes, nw);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  float best = 1e30f;
  for (int rep = 0; rep < 4; rep++) {
    (void)hipEventRecord(e0, st);
    int rc = p3e_launch_step(d_tab, d_work, nw, prec, pd.lds_bytes, grid, st);
    (void)hipEventRecord(e1, st);
    (void)hipEventSynchronize(e1);
    if (rc) {
      best = 1e30f;
      break;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (rep > 0) best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// Re-tile pd (tile size B, radix order rad) and give it device tables: a
// single-buffer pass's twiddles follow the radix order, so retune_pass
// leaves it without a device copy, and a new one goes to `owned`.
static bool apply_choice(int prec, PassDesc &pd, int B, const std::vector<int> &rad, std::vector<void *> &owned) {
  PassDesc t = pd;
  if (!retune_pass(prec, t, B, rad.data(), (int)rad.size())) return false;
  if (!tables_ready(t) && !upload_pass_tables(prec, t, owned)) return false;
  if (!tables_ready(t)) return false;
  pd = t;
  return true;
}

static void tune_pass(int prec, PassDesc &pd, hipStream_t st, std::vector<void *> &owned) {
  PassArgs &a = pd.args;
  if (a.fast || a.cls == CLS_EMPTY) return;
  if (!tables_ready(pd)) return;  // never launch a pass whose tables are not on the device
  load_cache();
  const std::string key = tune_key(prec, pd);
  auto it = g_tune.find(key);
  if (it != g_tune.end()) {
    (void)apply_choice(prec, pd, it->second.B, it->second.rad, owned);
    return;
  }
  void *din = nullptr, *dout = nullptr;
  PassArgs *d_tab = nullptr;
  WorkItem *d_work = nullptr;
  const int maxwork = 4096;
  if (hipMalloc(&din, std::max(pd.in_bytes, (size_t)16)) != hipSuccess ||
      hipMalloc(&dout, std::max(pd.out_bytes, (size_t)16)) != hipSuccess ||
      hipMalloc(&d_tab, sizeof(PassArgs)) != hipSuccess ||
      hipMalloc(&d_work, sizeof(WorkItem) * maxwork) != hipSuccess) {
    (void)hipGetLastError();
  } else {
    (void)hipMemsetAsync(din, 0, pd.in_bytes, st);
    std::vector<std::vector<int>> rads = radix_candidates(a.L, 4);
    TuneChoice best{a.B, std::vector<int>(a.radix, a.radix + a.nrad)};
    float best_ms = time_candidate(prec, pd, din, dout, d_tab, d_work, maxwork, st);
    for (int B : {2, 4, 8, 16, 32}) {
      for (const auto &r : rads) {
        // a candidate's own tables (its twiddles follow its radix order)
        // live only while it is timed
        std::vector<void *> tmp;
        PassDesc t = pd;
        if (!apply_choice(prec, t, B, r, tmp) || t.lds_bytes > 96 * 1024) {
          for (void *q : tmp) (void)hipFree(q);
          continue;
        }
        const float ms = time_candidate(prec, t, din, dout, d_tab, d_work, maxwork, st);
        (void)hipStreamSynchronize(st);
        for (void *q : tmp) (void)hipFree(q);
        if (ms < best_ms) {
          best_ms = ms;
          best = TuneChoice{B, r};
        }
      }
    }
    g_tune[key] = best;
    save_cache();
    (void)apply_choice(prec, pd, best.B, best.rad, owned);
  }
  if (din) (void)hipFree(din);
  if (dout) (void)hipFree(dout);
  if (d_tab) (void)hipFree(d_tab);
  if (d_work) (void)hipFree(d_work);
}

// ------------------------------------------------------------------ batch
// Plans every item's chain; false (with the batch plan untouched) if they
// differ in their number of stages or one has none.
static bool plan_items(BatchPlan *b, const OpSpec ops[3], const int order[3], int batch_threads) {
  std::vector<PassChain> chains(b->nitems);
  for (int i = 0; i < b->nitems; i++) {
    if (!plan_chain(b->prec, ops, order, b->lin[i], b->lout[i], chains[i], batch_threads)) return false;
    if (chains[i].copy_only || chains[i].passes.empty()) {
      set_error("batch: item %d has no transform stage", i);
      return false;
    }
    if (i > 0 && chains[i].passes.size() != chains[0].passes.size()) {
      set_error("batch: items differ in their number of stages");
      return false;
    }
  }
  b->chains.swap(chains);
  b->nstages = (int)b->chains[0].passes.size();
  return true;
}

// Grouped kernel of a standalone pass: 1 power-of-two, 2 single-buffer
// without an odd prime above 5, 0 none.
static int group_kind(const PassDesc &pd) {
  if (pd.args.fast) return 1;
  if (pd.args.mid && pd.args.mid_prime == 0) return 2;
  return 0;
}

static bool build_groups(BatchPlan *b) {
  b->groups.assign(b->nstages, {});
  b->slot.assign((size_t)b->nstages * b->nitems, 0);
  b->table.clear();
  char buf[160];
  for (int s = 0; s < b->nstages; s++) {
    std::map<std::pair<int, int>, std::vector<int>> by_kernel;
    for (int i = 0; i < b->nitems; i++) {
      const PassDesc &pd = b->chains[i].passes[s];
      const int kind = group_kind(pd);
      by_kernel[{kind, kind == 1 ? pd.args.M : 0}].push_back(i);
    }
    for (auto &kv : by_kernel) {
      BatchGroup g;
      g.kind = kv.first.first;
      g.M = kv.first.second;
      g.threads = 0;
      g.lds = 0;
      g.tab0 = b->table.size();
      g.n = (int)kv.second.size();
      g.d_start = nullptr;
      g.work = 0;
      g.q = 0;
      std::vector<uint32_t> start(1, 0);
      unsigned long long tiles = 0;
      for (int i : kv.second) {
        const PassDesc &pd = b->chains[i].passes[s];
        b->slot[(size_t)s * b->nitems + i] = b->table.size();
        b->table.push_back(pd.args);
        g.threads = std::max(g.threads, pd.threads);
        g.lds = std::max(g.lds, pd.lds_bytes);
        tiles += pd.nblocks;
        g.work += (double)pd.in_bytes + (double)pd.out_bytes;
        if (tiles >= (1ull << 31)) {
          set_error("batch: too many tiles in one stage");
          return false;
        }
        start.push_back((uint32_t)tiles);
      }
      g.nblocks = (unsigned int)tiles;
      if (hipMalloc((void **)&g.d_start, sizeof(uint32_t) * start.size()) != hipSuccess ||
          hipMemcpy(g.d_start, start.data(), sizeof(uint32_t) * start.size(), hipMemcpyHostToDevice) != hipSuccess) {
        if (g.d_start) (void)hipFree(g.d_start);
        set_error("batch: upload of a group's workgroup offsets failed");
        return false;
      }
      b->owned.push_back(g.d_start);
      b->groups[s].push_back(g);
      snprintf(buf, sizeof(buf), "%sstage %d %s%d: %d items, %u tiles, %d threads, LDS %zu B",
               b->desc.empty() ? "" : " | ", s, g.kind == 1 ? "pow2 M=" : "mid", g.kind == 1 ? g.M : 0, g.n,
               g.nblocks, g.threads, g.lds);
      b->desc += buf;
    }
    // The stage's launches, largest first, each on the stream with the least
    // work so far (longest-processing-time first over BATCH_STREAMS
    // streams): the device runs about that many kernels at once, and a
    // stage ends with its last stream.
    std::vector<BatchGroup> &gs = b->groups[s];
    std::stable_sort(gs.begin(), gs.end(), [](const BatchGroup &x, const BatchGroup &y) { return x.work > y.work; });
    double load[BATCH_STREAMS] = {0};
    for (BatchGroup &g : gs) {
      int q = 0;
      for (int k = 1; k < BATCH_STREAMS; k++)
        if (load[k] < load[q]) q = k;
      g.q = q;
      load[q] += g.work;
    }
  }
  return true;
}

static bool build_step_lists(BatchPlan *b) {
  b->slot.assign((size_t)b->nstages * b->nitems, 0);
  b->table.resize((size_t)b->nstages * b->nitems);
  for (int s = 0; s < b->nstages; s++)
    for (int i = 0; i < b->nitems; i++) {
      b->slot[(size_t)s * b->nitems + i] = (size_t)s * b->nitems + i;
      b->table[(size_t)s * b->nitems + i] = b->chains[i].passes[s].args;
    }
  b->nwork.resize(b->nstages);
  b->d_work.resize(b->nstages, nullptr);
  char buf[128];
  for (int s = 0; s < b->nstages; s++) {
    std::vector<WorkItem> w;
    size_t lds = 0;
    for (int i = 0; i < b->nitems; i++) {
      const PassDesc &pd = b->chains[i].passes[s];
      for (unsigned int t = 0; t < pd.nblocks; t++) w.push_back(WorkItem{(uint32_t)(s * b->nitems + i), t});
      lds = std::max(lds, pd.lds_bytes);
    }
    b->nwork[s] = (int)w.size();
    b->lds = std::max(b->lds, lds);
    if (hipMalloc((void **)&b->d_work[s], sizeof(WorkItem) * std::max<size_t>(1, w.size())) != hipSuccess ||
        hipMemcpy(b->d_work[s], w.data(), sizeof(WorkItem) * w.size(), hipMemcpyHostToDevice) != hipSuccess) {
      set_error("batch: work list upload failed");
      return false;
    }
    snprintf(buf, sizeof(buf), "%sstage %d: %d tiles, LDS %zu B", s ? " | " : "", s, b->nwork[s], lds);
    b->desc += buf;
  }
  return true;
}

BatchPlan *batch_create(int prec, int nitems, const OpSpec ops[3], const int order[3], const Layout *in,
                        const Layout *out, bool tune) {
  if (nitems < 1) {
    set_error("batch: no items");
    return nullptr;
  }
  BatchPlan *b = new BatchPlan();
  b->prec = prec;
  b->nitems = nitems;
  b->grouped = false;
  b->h_table[0] = b->h_table[1] = nullptr;
  b->d_table = nullptr;
  b->scratch = nullptr;
  b->scratch_bytes = 0;
  b->use_graph = false;
  b->parity = 0;
  for (int k = 0; k < 2; k++) {
    b->graph_ok[k] = false;
    b->graph[k] = nullptr;
    b->gexec[k] = nullptr;
    b->table_pending[k] = false;
    (void)hipEventCreateWithFlags(&b->table_done[k], hipEventDisableTiming);
  }
  b->lds = 0;
  b->grid = 1;
  b->lin.assign(in, in + nitems);
  b->lout.assign(out, out + nitems);
  // grouped launches when every pass of every item, planned as a standalone
  // pass, runs on a kernel that has a grouped form; else the step kernel
  if (plan_items(b, ops, order, 0)) {
    b->grouped = true;
    for (const PassChain &c : b->chains)
      for (const PassDesc &pd : c.passes)
        if (!group_kind(pd)) b->grouped = false;
  }
  if (!b->grouped && !plan_items(b, ops, order, P3E_STEP_THREADS)) {
    batch_destroy(b);
    return nullptr;
  }
  hipStream_t st = nullptr;
  // constant tables first (the tuner launches the pass, which reads them;
  // re-tiling keeps the work length and so the tables), then tuning (of the
  // step kernel's passes)
  for (int i = 0; i < nitems; i++)
    for (PassDesc &pd : b->chains[i].passes) {
      const bool up = upload_pass_tables(prec, pd, b->owned);
      if (!up || !tables_ready(pd)) {
        if (up) set_error("batch: a constant table has no device copy");
        batch_destroy(b);
        return nullptr;
      }
      if (tune && !b->grouped) tune_pass(prec, pd, st, b->owned);
      if (!tables_ready(pd)) {
        set_error("batch: tuning lost a constant table");
        batch_destroy(b);
        return nullptr;
      }
    }
  // scratch: worst case over the in != out and in == out assignments
  b->scratch_off.resize(2 * nitems);
  b->scratch_sz.resize(2 * nitems);
  size_t off = 0;
  for (int i = 0; i < nitems; i++) {
    size_t need_a[2], need_b[2];
    int dst[8];
    bool cb;
    const size_t outb = layout_elems(out[i]) * elem_bytes(prec, out[i].dt);
    assign_buffers(b->chains[i].passes, outb, false, dst, need_a, &cb);
    assign_buffers(b->chains[i].passes, outb, true, dst, need_b, &cb);
    for (int s = 0; s < 2; s++) {
      const size_t n = (std::max(need_a[s], need_b[s]) + 255) & ~(size_t)255;
      b->scratch_off[2 * i + s] = off;
      b->scratch_sz[2 * i + s] = n;
      off += n;
    }
  }
  b->scratch_bytes = off;
  if (off && hipMalloc(&b->scratch, off) != hipSuccess) {
    set_error("batch: hipMalloc of %zu scratch bytes failed", off);
    batch_destroy(b);
    return nullptr;
  }
  if (!(b->grouped ? build_groups(b) : build_step_lists(b))) {
    batch_destroy(b);
    return nullptr;
  }
  if (b->grouped) {
    int most = 0;
    for (const auto &gs : b->groups)
      for (const BatchGroup &g : gs) most = std::max(most, g.q + 1);
    b->fork_ev.assign(b->nstages, nullptr);
    for (int s = 0; s < b->nstages; s++) (void)hipEventCreateWithFlags(&b->fork_ev[s], hipEventDisableTiming);
    for (int q = 0; q + 1 < most; q++) {
      hipStream_t hs = nullptr;
      hipEvent_t he = nullptr;
      if (hipStreamCreateWithFlags(&hs, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&he, hipEventDisableTiming) != hipSuccess) {
        if (hs) (void)hipStreamDestroy(hs);
        set_error("batch: creation of a side stream failed");
        batch_destroy(b);
        return nullptr;
      }
      b->side.push_back(hs);
      b->join_ev.push_back(he);
    }
  }
  const size_t ntab = b->table.size();
  if (hipHostMalloc((void **)&b->h_table[0], sizeof(PassArgs) * ntab, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void **)&b->h_table[1], sizeof(PassArgs) * ntab, hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void **)&b->d_table, sizeof(PassArgs) * ntab) != hipSuccess) {
    set_error("batch: allocation of the pass table failed");
    batch_destroy(b);
    return nullptr;
  }
  b->grid = 0;
  return b;
}

// The launches of one exec (recorded into a graph when capturing).
static int batch_launch(BatchPlan *b, hipStream_t st) {
  for (int s = 0; s < b->nstages; s++) {
    if (b->grouped) {
      const std::vector<BatchGroup> &gs = b->groups[s];
      // the groups of a stage are independent: they run concurrently on the
      // exec stream and side streams forked from it (their stream planned in
      // build_groups), so one group's tail overlaps the others (1.99 -> 1.80
      // ms for 256 items of 32..128 points, six groups a stage)
      bool used[BATCH_STREAMS] = {false};
      for (const BatchGroup &g : gs) used[g.q] = true;
      for (int q = 1; q < BATCH_STREAMS; q++)
        if (used[q] && q - 1 >= (int)b->side.size()) {
          set_error("batch: group planned on a missing side stream");
          return -1;
        }
      bool forked = false;
      for (int q = 1; q < BATCH_STREAMS; q++) {
        if (!used[q]) continue;
        if (!forked) BCHK(hipEventRecord(b->fork_ev[s], st));
        forked = true;
        BCHK(hipStreamWaitEvent(b->side[q - 1], b->fork_ev[s], 0));
      }
      for (const BatchGroup &g : gs) {
        hipStream_t gst = g.q == 0 ? st : b->side[g.q - 1];
        const int rc = p3e_launch_group(g.kind, g.M, b->d_table + g.tab0, g.d_start, g.n, b->prec, g.lds, g.threads,
                                        g.nblocks, gst);
        if (rc) {
          set_error("batch group launch failed: %s", hipGetErrorString((hipError_t)rc));
          return -1;
        }
      }
      for (int q = 1; q < BATCH_STREAMS; q++) {
        if (!used[q]) continue;
        BCHK(hipEventRecord(b->join_ev[q - 1], b->side[q - 1]));
        BCHK(hipStreamWaitEvent(st, b->join_ev[q - 1], 0));
      }
    } else {
      size_t lds = 0;
      for (int i = 0; i < b->nitems; i++) lds = std::max(lds, b->chains[i].passes[s].lds_bytes);
      const unsigned int grid = persistent_grid(b->prec, lds, b->nwork[s]);
      const int rc = p3e_launch_step(b->d_table, b->d_work[s], b->nwork[s], b->prec, lds, grid, st);
      if (rc) {
        set_error("batch step launch failed: %s", hipGetErrorString((hipError_t)rc));
        return -1;
      }
    }
  }
  return 0;
}

int batch_exec(BatchPlan *b, void *const *in, void *const *out, hipStream_t st, bool sync) {
  if (!b) {
    set_error("batch exec on an invalid plan");
    return -1;
  }
  const int k = b->parity;
  // staging copy k was last read by the exec two calls ago
  if (b->table_pending[k]) BCHK(hipEventSynchronize(b->table_done[k]));
  // buffers of every stage of every item
  for (int i = 0; i < b->nitems; i++) {
    const std::vector<PassDesc> &ps = b->chains[i].passes;
    const size_t outb = layout_elems(b->lout[i]) * elem_bytes(b->prec, b->lout[i].dt);
    int dst[8];
    size_t need[2];
    bool copy_back = false;
    assign_buffers(ps, outb, in[i] == out[i], dst, need, &copy_back);
    if (copy_back) {
      set_error("batch: item %d cannot run in place", i);
      return -1;
    }
    const void *src = in[i];
    for (int s = 0; s < b->nstages; s++) {
      void *d = out[i];
      if (dst[s] == BUF_S0 || dst[s] == BUF_S1) {
        const int q = 2 * i + (dst[s] - BUF_S0);
        d = (char *)b->scratch + b->scratch_off[q];
      }
      PassArgs &a = b->table[b->slot[(size_t)s * b->nitems + i]];
      a.in = src;
      a.out = d;
      if (b->grouped) {
        a.vec2 = ps[s].args.vec2;
        a.vec2s = ps[s].args.vec2s;
        drop_unaligned_pairs(a, b->prec);
      }
      src = d;
    }
  }
  memcpy(b->h_table[k], b->table.data(), sizeof(PassArgs) * b->table.size());
  if (b->use_graph && b->graph_ok[k]) {
    BCHK(hipGraphLaunch(b->gexec[k], st));
  } else {
    const bool capture = b->use_graph;
    if (capture) BCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(b->d_table, b->h_table[k], sizeof(PassArgs) * b->table.size(),
                                  hipMemcpyHostToDevice, st);
    int rc = e == hipSuccess ? batch_launch(b, st) : -1;
    if (e != hipSuccess) set_error("batch table upload failed: %s", hipGetErrorString(e));
    if (capture) {
      hipGraph_t g = nullptr;
      const hipError_t ee = hipStreamEndCapture(st, &g);
      if (rc || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (!rc) set_error("hipStreamEndCapture failed: %s", hipGetErrorString(ee));
        return -1;
      }
      b->graph[k] = g;
      BCHK(hipGraphInstantiate(&b->gexec[k], b->graph[k], nullptr, nullptr, 0));
      b->graph_ok[k] = true;
      BCHK(hipGraphLaunch(b->gexec[k], st));
    } else if (rc) {
      return -1;
    }
  }
  BCHK(hipEventRecord(b->table_done[k], st));
  b->table_pending[k] = true;
  b->parity = 1 - k;
  if (sync) BCHK(hipStreamSynchronize(st));
  return 0;
}

void batch_destroy(BatchPlan *b) {
  if (!b) return;
  for (int k = 0; k < 2; k++) {
    if (b->gexec[k]) (void)hipGraphExecDestroy(b->gexec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    if (b->h_table[k]) (void)hipHostFree(b->h_table[k]);
    if (b->table_done[k]) (void)hipEventDestroy(b->table_done[k]);
  }
  for (hipStream_t hs : b->side) (void)hipStreamDestroy(hs);
  for (hipEvent_t he : b->join_ev) (void)hipEventDestroy(he);
  for (hipEvent_t he : b->fork_ev)
    if (he) (void)hipEventDestroy(he);
  for (void *p : b->owned) (void)hipFree(p);
  for (WorkItem *w : b->d_work)
    if (w) (void)hipFree(w);
  if (b->d_table) (void)hipFree(b->d_table);
  if (b->scratch) (void)hipFree(b->scratch);
  delete b;
}

}  // namespace p3e
