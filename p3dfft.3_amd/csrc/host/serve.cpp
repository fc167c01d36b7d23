// Poisson service: see serve.h.
#include "serve.h"

#include <string.h>

#include <algorithm>
#include <chrono>

#include "batch.h"

extern "C" int p3e_launch_step(const void *table, const void *work, int nwork, int prec, size_t lds,
                               unsigned int grid, hipStream_t st);
extern "C" int p3e_launch_pass(const p3e::PassArgs *a, int prec, size_t lds_bytes, int threads,
                               unsigned int nblocks, hipStream_t st);

namespace p3e {

static const int MO_REAL[3] = {0, 1, 2};
static const int MO_SPEC[3] = {1, 2, 0};

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static long long dims_key(const int d[3]) { return ((long long)d[0] << 40) | ((long long)d[1] << 20) | d[2]; }

static size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

static PoissonPlan *make_plan(int prec, const int dims[3]) {
  const bool dedicated = (long long)dims[0] * dims[1] * dims[2] >= DEDICATED_POINTS;
  PassChain f, b;
  if (!plan_poisson(prec, dims, MO_REAL, MO_SPEC, f, b, dedicated ? 0 : P3E_STEP_THREADS)) return nullptr;
  PoissonPlan *p = new PoissonPlan();
  p->dedicated = dedicated;
  for (int a = 0; a < 3; a++) p->dims[a] = dims[a];
  p->passes = f.passes;
  p->passes.insert(p->passes.end(), b.passes.begin(), b.passes.end());
  p->work_bytes = 0;
  for (size_t i = 0; i + 1 < p->passes.size(); i++) p->work_bytes = std::max(p->work_bytes, p->passes[i].out_bytes);
  p->work_bytes = align256(p->work_bytes);
  p->real_bytes = (size_t)dims[0] * dims[1] * dims[2] * (size_t)prec;
  for (PassDesc &pd : p->passes) {
    if (!upload_pass_tables(prec, pd, p->owned) || !tables_ready(pd)) {
      for (void *t : p->owned) (void)hipFree(t);
      delete p;
      set_error("Poisson plan: table upload failed");
      return nullptr;
    }
  }
  return p;
}

static void free_plan(PoissonPlan *p) {
  for (void *t : p->owned) (void)hipFree(t);
  delete p;
}

// ------------------------------------------------------------------ pool
static bool pool_alloc(PoissonServer *s, size_t n, size_t *off) {
  for (auto it = s->free_blocks.begin(); it != s->free_blocks.end(); ++it) {
    if (it->second < n) continue;
    *off = it->first;
    const size_t rest = it->second - n, o = it->first + n;
    s->free_blocks.erase(it);
    if (rest) s->free_blocks[o] = rest;
    return true;
  }
  return false;
}

static void pool_free(PoissonServer *s, size_t off, size_t n) {
  auto it = s->free_blocks.emplace(off, n).first;
  auto nx = std::next(it);
  if (nx != s->free_blocks.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    s->free_blocks.erase(nx);
  }
  if (it != s->free_blocks.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      s->free_blocks.erase(it);
    }
  }
}

// ------------------------------------------------------------------ steps
static bool ensure_cap(PoissonServer::Step &b, size_t ntab, size_t nwork) {
  if (ntab > b.cap_tab) {
    if (b.h_tab) (void)hipHostFree(b.h_tab);
    if (b.d_tab) (void)hipFree(b.d_tab);
    b.h_tab = nullptr;
    b.d_tab = nullptr;
    const size_t c = std::max(ntab, 2 * b.cap_tab);
    if (hipHostMalloc((void **)&b.h_tab, c * sizeof(PassArgs), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&b.d_tab, c * sizeof(PassArgs)) != hipSuccess) {
      b.cap_tab = 0;
      return false;
    }
    b.cap_tab = c;
  }
  if (nwork > b.cap_work) {
    if (b.h_work) (void)hipHostFree(b.h_work);
    if (b.d_work) (void)hipFree(b.d_work);
    b.h_work = nullptr;
    b.d_work = nullptr;
    const size_t c = std::max(nwork, 2 * b.cap_work);
    if (hipHostMalloc((void **)&b.h_work, c * sizeof(WorkItem), hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&b.d_work, c * sizeof(WorkItem)) != hipSuccess) {
      b.cap_work = 0;
      return false;
    }
    b.cap_work = c;
  }
  return true;
}

// Called with s->mu held: the requests of a completed step are done.
static void retire(PoissonServer *s, PoissonServer::Step &b) {
  const double t = now_s() - s->t0;
  for (PoissonRequest *r : b.finishing) {
    r->t_done = t;
    pool_free(s, r->off, 2 * r->plan->work_bytes);
    s->holding--;
    if (r->plan->dedicated) s->holding_large--;
    s->completed.push_back(r);
  }
  b.finishing.clear();
  b.busy = false;
  s->done_cv.notify_all();
}

static void fail_all(PoissonServer *s, const char *msg) {
  s->failed = true;
  s->error = msg;
  const double t = now_s() - s->t0;
  auto finish = [&](PoissonRequest *r) {
    r->failed = true;
    r->t_done = t;
    s->completed.push_back(r);
  };
  for (PoissonRequest *r : s->active) finish(r);
  for (PoissonRequest *r : s->pending) finish(r);
  for (auto &b : s->steps) {
    for (PoissonRequest *r : b.finishing) finish(r);
    b.finishing.clear();
  }
  s->active.clear();
  s->pending.clear();
  s->done_cv.notify_all();
}

// Builds and launches one step over the current pass of every active
// request (s->mu not held; only the service thread touches `active`): the
// passes of large grids as their own launches, then one step-kernel launch
// over the tiles of all the others.
static int launch_step(PoissonServer *s, PoissonServer::Step &b) {
  const int prec = s->prec;
  size_t ntab = 0, nwork = 0, lds = 0;
  for (PoissonRequest *r : s->active) {
    if (r->plan->dedicated) continue;
    const PassDesc &pd = r->plan->passes[r->phase];
    ntab++;
    nwork += pd.nblocks;
    lds = std::max(lds, pd.lds_bytes);
  }
  if (!ensure_cap(b, std::max<size_t>(ntab, 1), std::max<size_t>(nwork, 1))) return -1;
  size_t w = 0, t_i = 0;
  for (PoissonRequest *r : s->active) {
    const PoissonPlan *pl = r->plan;
    const int ph = r->phase, last = (int)pl->passes.size() - 1;
    char *w0 = s->pool + r->off, *w1 = w0 + pl->work_bytes;
    const PassDesc &pd = pl->passes[ph];
    PassArgs a = pd.args;
    a.in = ph == 0 ? r->rhs : ((ph - 1) % 2 ? w1 : w0);
    a.out = ph == last ? r->sol : (ph % 2 ? w1 : w0);
    a.deriv = 0;
    if (pl->dedicated) {
      if (p3e_launch_pass(&a, prec, pd.lds_bytes, pd.threads, pd.nblocks, s->st)) return -1;
      s->ntiles += pd.nblocks;
      continue;
    }
    b.h_tab[t_i] = a;
    for (unsigned t = 0; t < pd.nblocks; t++) b.h_work[w++] = WorkItem{(uint32_t)t_i, t};
    t_i++;
  }
  if (nwork) {
    if (hipMemcpyAsync(b.d_tab, b.h_tab, ntab * sizeof(PassArgs), hipMemcpyHostToDevice, s->st) != hipSuccess ||
        hipMemcpyAsync(b.d_work, b.h_work, nwork * sizeof(WorkItem), hipMemcpyHostToDevice, s->st) != hipSuccess)
      return -1;
    const unsigned grid = persistent_grid(prec, lds, (int)nwork);
    if (p3e_launch_step(b.d_tab, b.d_work, (int)nwork, prec, lds, grid, s->st)) return -1;
  }
  if (hipEventRecord(b.done, s->st) != hipSuccess) return -1;
  s->nsteps++;
  s->ntiles += (long long)nwork;
  return 0;
}

static void service_loop(PoissonServer *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->failed) {
    // completed steps
    for (int k = 0; k < SERVE_DEPTH; k++) {
      PoissonServer::Step &b = s->steps[(s->cur + k) % SERVE_DEPTH];
      if (b.busy && hipEventQuery(b.done) == hipSuccess) retire(s, b);
    }
    // admission: arrival order, a slot, working memory, and for a large
    // grid one of LARGE_INFLIGHT places
    while (!s->pending.empty() && s->holding < s->max_inflight) {
      PoissonRequest *r = s->pending.front();
      if (r->plan->dedicated && s->holding_large >= LARGE_INFLIGHT) break;
      size_t off;
      if (!pool_alloc(s, 2 * r->plan->work_bytes, &off)) break;
      r->off = off;
      r->phase = 0;
      r->t_admit = now_s() - s->t0;
      s->pending.pop_front();
      s->active.push_back(r);
      s->holding++;
      if (r->plan->dedicated) s->holding_large++;
    }
    if (s->active.empty()) {
      PoissonServer::Step *oldest = nullptr;
      for (int k = 0; k < SERVE_DEPTH; k++)
        if (s->steps[(s->cur + k) % SERVE_DEPTH].busy) {
          oldest = &s->steps[(s->cur + k) % SERVE_DEPTH];
          break;
        }
      if (oldest) {
        lk.unlock();
        const hipError_t e = hipEventSynchronize(oldest->done);
        lk.lock();
        if (e != hipSuccess) fail_all(s, "Poisson service: a step failed on the device");
        else retire(s, *oldest);
        continue;
      }
      if (s->stop && s->pending.empty()) break;
      s->cv.wait_for(lk, std::chrono::milliseconds(50));
      continue;
    }
    PoissonServer::Step &b = s->steps[s->cur];
    if (b.busy) {
      lk.unlock();
      const hipError_t e = hipEventSynchronize(b.done);
      lk.lock();
      if (e != hipSuccess) {
        fail_all(s, "Poisson service: a step failed on the device");
        break;
      }
      retire(s, b);
    }
    lk.unlock();
    const int rc = launch_step(s, b);
    lk.lock();
    if (rc) {
      fail_all(s, "Poisson service: step launch failed");
      break;
    }
    b.busy = true;
    std::vector<PoissonRequest *> still;
    for (PoissonRequest *r : s->active) {
      r->phase++;
      if (r->phase == (int)r->plan->passes.size()) b.finishing.push_back(r);
      else still.push_back(r);
    }
    s->active.swap(still);
    s->cur = (s->cur + 1) % SERVE_DEPTH;
  }
}

PoissonServer *poisson_server_create(int prec, size_t pool_bytes, int max_inflight) {
  if ((prec != 4 && prec != 8) || pool_bytes < 4096 || max_inflight < 1) {
    set_error("Poisson service: invalid arguments");
    return nullptr;
  }
  PoissonServer *s = new PoissonServer();
  s->prec = prec;
  s->max_inflight = max_inflight;
  s->pool = nullptr;
  s->pool_bytes = pool_bytes & ~(size_t)255;
  s->st = nullptr;
  s->cur = 0;
  s->holding = 0;
  s->holding_large = 0;
  s->next_id = 0;
  s->nsteps = s->ntiles = 0;
  s->stop = s->failed = false;
  s->t0 = now_s();
  for (auto &b : s->steps) {
    b.h_tab = b.d_tab = nullptr;
    b.h_work = b.d_work = nullptr;
    b.cap_tab = b.cap_work = 0;
    b.busy = false;
    b.done = nullptr;
  }
  bool ok = hipGetDevice(&s->device) == hipSuccess &&
            hipMalloc((void **)&s->pool, s->pool_bytes) == hipSuccess &&
            hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) == hipSuccess;
  for (auto &b : s->steps) ok = ok && hipEventCreateWithFlags(&b.done, hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    set_error("Poisson service: device allocation of %zu bytes failed", s->pool_bytes);
    poisson_server_destroy(s);
    return nullptr;
  }
  s->free_blocks[0] = s->pool_bytes;
  s->th = std::thread(service_loop, s);
  return s;
}

static PoissonPlan *plan_of(PoissonServer *s, const int dims[3]) {
  const long long k = dims_key(dims);
  auto it = s->plans.find(k);
  if (it != s->plans.end()) return it->second;
  PoissonPlan *p = make_plan(s->prec, dims);
  if (!p) return nullptr;
  if (2 * p->work_bytes > s->pool_bytes) {
    set_error("Poisson service: a %dx%dx%d solve needs %zu bytes of working memory, the pool has %zu", dims[0],
              dims[1], dims[2], 2 * p->work_bytes, s->pool_bytes);
    free_plan(p);
    return nullptr;
  }
  s->plans[k] = p;
  return p;
}

int poisson_server_prepare(PoissonServer *s, const int dims[3]) {
  std::lock_guard<std::mutex> g(s->mu);
  return plan_of(s, dims) ? 0 : -1;
}

long long poisson_submit(PoissonServer *s, const int dims[3], const void *rhs, void *sol) {
  std::lock_guard<std::mutex> g(s->mu);
  if (s->failed) {
    set_error("%s", s->error.c_str());
    return -1;
  }
  if (!rhs || !sol) {
    set_error("Poisson service: null array");
    return -1;
  }
  PoissonPlan *p = plan_of(s, dims);
  if (!p) return -1;
  PoissonRequest *r = new PoissonRequest();
  r->id = s->next_id++;
  r->plan = p;
  r->rhs = rhs;
  r->sol = sol;
  r->off = 0;
  r->phase = 0;
  r->t_submit = now_s() - s->t0;
  r->t_admit = r->t_done = -1;
  r->failed = false;
  s->pending.push_back(r);
  s->cv.notify_all();
  return r->id;
}

int poisson_poll(PoissonServer *s, long long *ids, double *t_submit, double *t_admit, double *t_done, int *ok,
                 int cap, int wait_ms) {
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->completed.empty() && wait_ms > 0)
    s->done_cv.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->completed.empty(); });
  int n = 0;
  while (n < cap && !s->completed.empty()) {
    PoissonRequest *r = s->completed.front();
    s->completed.erase(s->completed.begin());
    ids[n] = r->id;
    if (t_submit) t_submit[n] = r->t_submit;
    if (t_admit) t_admit[n] = r->t_admit;
    if (t_done) t_done[n] = r->t_done;
    if (ok) ok[n] = r->failed ? 0 : 1;
    delete r;
    n++;
  }
  return n;
}

void poisson_server_destroy(PoissonServer *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  if (s->st) (void)hipStreamSynchronize(s->st);
  for (PoissonRequest *r : s->completed) delete r;
  for (PoissonRequest *r : s->pending) delete r;
  for (PoissonRequest *r : s->active) delete r;
  for (auto &b : s->steps) {
    for (PoissonRequest *r : b.finishing) delete r;
    if (b.h_tab) (void)hipHostFree(b.h_tab);
    if (b.d_tab) (void)hipFree(b.d_tab);
    if (b.h_work) (void)hipHostFree(b.h_work);
    if (b.d_work) (void)hipFree(b.d_work);
    if (b.done) (void)hipEventDestroy(b.done);
  }
  for (auto &kv : s->plans) free_plan(kv.second);
  if (s->pool) (void)hipFree(s->pool);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

// ------------------------------------------------------------------ per-request path
struct SolvePlans {
  Plan *fwd, *bwd;
  void *spec;
  size_t spec_bytes;
};
static std::map<long long, SolvePlans> g_solve[2];

int poisson_solve(int prec, const int dims[3], const void *rhs, void *sol, hipStream_t st, bool sync) {
  if (prec != 4 && prec != 8) {
    set_error("Poisson solve: precision must be 4 or 8");
    return -1;
  }
  auto &cache = g_solve[prec == 8];
  const long long k = dims_key(dims);
  auto it = cache.find(k);
  if (it == cache.end()) {
    PassChain f, b;
    if (!plan_poisson(prec, dims, MO_REAL, MO_SPEC, f, b)) return -1;
    Layout real, spec;
    for (int a = 0; a < 3; a++) {
      real.dims[a] = dims[a];
      real.mo[a] = MO_REAL[a];
      spec.dims[a] = a == 0 ? dims[0] / 2 + 1 : dims[a];
      spec.mo[a] = MO_SPEC[a];
    }
    real.dt = 1;
    spec.dt = 2;
    SolvePlans sp;
    sp.fwd = plan_from_chain(prec, f, real, spec);
    sp.bwd = sp.fwd ? plan_from_chain(prec, b, spec, real) : nullptr;
    sp.spec_bytes = layout_elems(spec) * elem_bytes(prec, 2);
    sp.spec = nullptr;
    if (!sp.fwd || !sp.bwd || hipMalloc(&sp.spec, sp.spec_bytes) != hipSuccess) {
      if (sp.fwd) plan_destroy(sp.fwd);
      if (sp.bwd) plan_destroy(sp.bwd);
      set_error("Poisson solve: plan or spectrum allocation failed");
      return -1;
    }
    it = cache.emplace(k, sp).first;
  }
  SolvePlans &sp = it->second;
  if (plan_exec(sp.fwd, rhs, sp.spec, -1, false, st, false)) return -1;
  if (plan_exec(sp.bwd, sp.spec, sol, -1, true, st, sync)) return -1;
  return 0;
}

}  // namespace p3e
