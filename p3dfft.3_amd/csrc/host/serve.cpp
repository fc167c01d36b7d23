// Poisson service: see serve.h.
#include "serve.h"

#include <stdlib.h>
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

// ------------------------------------------------------------------ launch checks
bool check_launch(const PassArgs *tab, const PassDesc *const *descs, size_t ntab, const WorkItem *work,
                  size_t nwork) {
  for (size_t i = 0; i < ntab; i++) {
    const PassArgs &a = tab[i];
    // every tile loop of the kernels strides by nthreads
    if (a.nthreads < 1 || a.B < 1 || a.nrad > P3E_MAX_RADIX_PASSES) {
      set_error("Poisson service: pass %zu has nthreads %d, B %d, %d radix passes", i, a.nthreads, a.B, a.nrad);
      return false;
    }
    if (descs && !tables_ready(*descs[i])) {
      set_error("Poisson service: pass %zu has a table without a device copy", i);
      return false;
    }
  }
  for (size_t w = 0; w < nwork; w++) {
    if (work[w].pass >= ntab) {
      set_error("Poisson service: work item %zu names pass %u of %zu", w, work[w].pass, ntab);
      return false;
    }
    if (descs && work[w].tile >= descs[work[w].pass]->nblocks) {
      set_error("Poisson service: work item %zu names tile %u of %u", w, work[w].tile,
                descs[work[w].pass]->nblocks);
      return false;
    }
  }
  return true;
}

// ------------------------------------------------------------------ groups
static bool grow(void **h, void **d, size_t *cap, size_t need, size_t esz) {
  if (need <= *cap) return true;
  // the group is not in flight: its previous copies have completed
  if (*h) (void)hipHostFree(*h);
  if (*d) (void)hipFree(*d);
  *h = *d = nullptr;
  const size_t c = std::max(need, 2 * *cap);
  if (hipHostMalloc(h, c * esz, hipHostMallocDefault) != hipSuccess || hipMalloc(d, c * esz) != hipSuccess) {
    *cap = 0;
    return false;
  }
  *cap = c;
  return true;
}

// Working buffers of pass ph of request r: the rhs, the two halves of its
// pool block in turn, and the solution.
static void bind(const PoissonServer *s, const PoissonRequest *r, int ph, PassArgs &a) {
  const PoissonPlan *pl = r->plan;
  const int last = (int)pl->passes.size() - 1;
  char *w0 = s->pool + r->off, *w1 = w0 + pl->work_bytes;
  a.in = ph == 0 ? r->rhs : ((ph - 1) % 2 ? w1 : w0);
  a.out = ph == last ? r->sol : (ph % 2 ? w1 : w0);
  a.deriv = 0;
}

// Launches a group on its lane (s->mu not held; the group is reserved for
// this thread): a large request's passes as their own launches, or the
// passes of a batch of small requests as one step-kernel launch per pass
// index; then the group's event.
static int launch_group(PoissonServer *s, PoissonGroup &g) {
  const int prec = s->prec;
  hipStream_t st = s->lanes[g.lane];
  if (g.reqs.size() == 1 && g.reqs[0]->plan->dedicated) {
    const PoissonRequest *r = g.reqs[0];
    const PoissonPlan *pl = r->plan;
    for (int ph = 0; ph < (int)pl->passes.size(); ph++) {
      const PassDesc &pd = pl->passes[ph];
      PassArgs a = pd.args;
      bind(s, r, ph, a);
      const PassDesc *dp = &pd;
      if (!check_launch(&a, &dp, 1, nullptr, 0)) return -1;
      if (p3e_launch_pass(&a, prec, pd.lds_bytes, pd.threads, pd.nblocks, st)) {
        set_error("Poisson service: pass launch failed");
        return -1;
      }
      s->nlaunch++;
      s->ntiles += pd.nblocks;
    }
  } else {
    // passes of every request, pass index major: entry ph * n + i
    const size_t n = g.reqs.size();
    const int npass = (int)g.reqs[0]->plan->passes.size();
    size_t nwork = 0;
    for (const PoissonRequest *r : g.reqs) {
      if ((int)r->plan->passes.size() != npass) {
        set_error("Poisson service: requests of one group differ in their number of passes");
        return -1;
      }
      for (const PassDesc &pd : r->plan->passes) nwork += pd.nblocks;
    }
    const size_t ntab = n * npass;
    if (!grow((void **)&g.h_tab, (void **)&g.d_tab, &g.cap_tab, ntab, sizeof(PassArgs)) ||
        !grow((void **)&g.h_work, (void **)&g.d_work, &g.cap_work, nwork, sizeof(WorkItem))) {
      set_error("Poisson service: allocation of launch tables failed");
      return -1;
    }
    std::vector<const PassDesc *> descs(ntab);
    std::vector<size_t> w_off(npass + 1, 0);
    std::vector<size_t> lds(npass, 0);
    size_t w = 0;
    for (int ph = 0; ph < npass; ph++) {
      w_off[ph] = w;
      for (size_t i = 0; i < n; i++) {
        const PassDesc &pd = g.reqs[i]->plan->passes[ph];
        const size_t t = ph * n + i;
        PassArgs a = pd.args;
        bind(s, g.reqs[i], ph, a);
        g.h_tab[t] = a;
        descs[t] = &pd;
        lds[ph] = std::max(lds[ph], pd.lds_bytes);
        for (unsigned k = 0; k < pd.nblocks; k++) g.h_work[w++] = WorkItem{(uint32_t)t, k};
      }
    }
    w_off[npass] = w;
    if (!check_launch(g.h_tab, descs.data(), ntab, g.h_work, nwork)) return -1;
    if (hipMemcpyAsync(g.d_tab, g.h_tab, ntab * sizeof(PassArgs), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(g.d_work, g.h_work, nwork * sizeof(WorkItem), hipMemcpyHostToDevice, st) != hipSuccess) {
      set_error("Poisson service: table upload failed");
      return -1;
    }
    for (int ph = 0; ph < npass; ph++) {
      const int nw = (int)(w_off[ph + 1] - w_off[ph]);
      const unsigned grid = persistent_grid(prec, lds[ph], nw);
      if (p3e_launch_step(g.d_tab, g.d_work + w_off[ph], nw, prec, lds[ph], grid, st)) {
        set_error("Poisson service: step launch failed");
        return -1;
      }
      s->nlaunch++;
      s->ntiles += nw;
    }
  }
  if (hipEventRecord(g.done, st) != hipSuccess) {
    set_error("Poisson service: event record failed");
    return -1;
  }
  s->ngroups++;
  return 0;
}

// Called with s->mu held: the requests of a completed group are done.
static void retire(PoissonServer *s, PoissonGroup &g) {
  const double t = now_s() - s->t0;
  for (PoissonRequest *r : g.reqs) {
    r->t_done = t;
    pool_free(s, r->off, 2 * r->plan->work_bytes);
    s->holding--;
    s->completed.push_back(r);
  }
  g.reqs.clear();
  g.busy = false;
  s->done_cv.notify_all();
}

static void fail_all(PoissonServer *s, const std::string &msg) {
  s->failed = true;
  s->error = msg;
  const double t = now_s() - s->t0;
  auto finish = [&](PoissonRequest *r) {
    r->failed = true;
    r->t_done = t;
    s->completed.push_back(r);
  };
  for (PoissonRequest *r : s->pending) finish(r);
  s->pending.clear();
  // groups in flight keep their requests until the device is done with
  // them (the service thread drains them before it exits)
  s->done_cv.notify_all();
}

// A free group on the lane with the fewest groups in flight, or null.
static PoissonGroup *free_group(PoissonServer *s) {
  std::vector<int> busy(s->nlanes, 0);
  for (const PoissonGroup &g : s->groups)
    if (g.busy) busy[g.lane]++;
  PoissonGroup *best = nullptr;
  for (PoissonGroup &g : s->groups)
    if (!g.busy && (!best || busy[g.lane] < busy[best->lane])) best = &g;
  return best;
}

// Called with s->mu held: takes the next requests that can start into g
// (one large request, or consecutive small ones up to SMALL_GROUP_MAX), with
// their working memory; false if the request at the head cannot start yet.
static bool admit(PoissonServer *s, PoissonGroup &g) {
  const double t = now_s() - s->t0;
  while (!s->pending.empty() && s->holding < s->max_inflight && (int)g.reqs.size() < SMALL_GROUP_MAX) {
    PoissonRequest *r = s->pending.front();
    if (!g.reqs.empty() && (r->plan->dedicated || g.reqs[0]->plan->dedicated)) break;
    if (!g.reqs.empty() && r->plan->passes.size() != g.reqs[0]->plan->passes.size()) break;
    size_t off;
    if (!pool_alloc(s, 2 * r->plan->work_bytes, &off)) break;
    r->off = off;
    r->t_admit = t;
    s->pending.pop_front();
    s->holding++;
    g.reqs.push_back(r);
    if (r->plan->dedicated) break;
  }
  return !g.reqs.empty();
}

static void service_loop(PoissonServer *s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool any_busy = false;
    for (PoissonGroup &g : s->groups) {
      if (!g.busy) continue;
      const hipError_t e = hipEventQuery(g.done);
      if (e == hipSuccess) {
        retire(s, g);
      } else if (e != hipErrorNotReady) {
        // the device reported an error: every request of the group fails
        for (PoissonRequest *r : g.reqs) r->failed = true;
        retire(s, g);
        if (!s->failed) fail_all(s, "Poisson service: a launch group failed on the device");
      } else {
        any_busy = true;
      }
    }
    while (!s->failed && !s->pending.empty()) {
      PoissonGroup *g = free_group(s);
      if (!g || !admit(s, *g)) break;
      g->busy = true;  // reserved: only this thread launches and retires groups
      lk.unlock();
      const int rc = launch_group(s, *g);
      lk.lock();
      if (rc) {
        // nothing of a failed group reached the device past its failing
        // launch; its requests fail with the rest
        const std::string msg = last_error();
        s->ngroups++;
        for (PoissonRequest *r : g->reqs) r->failed = true;
        (void)hipStreamSynchronize(s->lanes[g->lane]);
        retire(s, *g);
        fail_all(s, msg);
        break;
      }
      any_busy = true;
    }
    if (!any_busy && s->pending.empty() && (s->stop || s->failed)) break;
    if (any_busy) {
      // groups in flight: poll their events (a submission is seen at the
      // next turn of the loop)
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    } else {
      s->cv.wait_for(lk, std::chrono::milliseconds(50));
    }
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
  s->nlanes = SERVE_LANES;
  if (const char *e = getenv("P3DFFT_HIP_SERVE_LANES")) {
    const int v = atoi(e);
    if (v >= 1 && v <= 8) s->nlanes = v;
  }
  s->pool = nullptr;
  s->pool_bytes = pool_bytes & ~(size_t)255;
  s->holding = 0;
  s->next_id = 0;
  s->ngroups = s->nlaunch = s->ntiles = 0;
  s->stop = s->failed = false;
  s->t0 = now_s();
  s->lanes.assign(s->nlanes, nullptr);
  s->groups.resize((size_t)s->nlanes * LANE_DEPTH);
  for (size_t i = 0; i < s->groups.size(); i++) {
    PoissonGroup &g = s->groups[i];
    g.lane = (int)(i % s->nlanes);
    g.busy = false;
    g.done = nullptr;
    g.h_tab = g.d_tab = nullptr;
    g.h_work = g.d_work = nullptr;
    g.cap_tab = g.cap_work = 0;
  }
  bool ok = hipGetDevice(&s->device) == hipSuccess && hipMalloc((void **)&s->pool, s->pool_bytes) == hipSuccess;
  for (hipStream_t &st : s->lanes) ok = ok && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
  for (PoissonGroup &g : s->groups)
    ok = ok && hipEventCreateWithFlags(&g.done, hipEventDisableTiming) == hipSuccess;
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

void poisson_state(PoissonServer *s, int *pending, int *holding, int *groups_busy, int *lanes) {
  std::lock_guard<std::mutex> g(s->mu);
  int busy = 0;
  for (const PoissonGroup &gr : s->groups) busy += gr.busy ? 1 : 0;
  if (pending) *pending = (int)s->pending.size();
  if (holding) *holding = s->holding;
  if (groups_busy) *groups_busy = busy;
  if (lanes) *lanes = s->nlanes;
}

void poisson_server_destroy(PoissonServer *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  for (hipStream_t st : s->lanes)
    if (st) (void)hipStreamSynchronize(st);
  for (PoissonRequest *r : s->completed) delete r;
  for (PoissonRequest *r : s->pending) delete r;
  for (PoissonGroup &g : s->groups) {
    for (PoissonRequest *r : g.reqs) delete r;
    if (g.h_tab) (void)hipHostFree(g.h_tab);
    if (g.d_tab) (void)hipFree(g.d_tab);
    if (g.h_work) (void)hipHostFree(g.h_work);
    if (g.d_work) (void)hipFree(g.d_work);
    if (g.done) (void)hipEventDestroy(g.done);
  }
  for (auto &kv : s->plans) free_plan(kv.second);
  if (s->pool) (void)hipFree(s->pool);
  for (hipStream_t st : s->lanes)
    if (st) (void)hipStreamDestroy(st);
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
