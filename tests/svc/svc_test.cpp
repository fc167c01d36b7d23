// CPU test of the Poisson service's scheduling (test infrastructure): the
// product's serve.cpp and planner, the mock HIP runtime (one concurrent
// queue per stream; frees and pinned allocations synchronize the device),
// and kernels that record which buffers each pass read and wrote.  Checks,
// for several lane counts, in-flight limits, pool sizes and device speeds:
// every request completes (no hang); each request's passes ran in order and
// form the chain rhs -> ... -> sol; no pass wrote a buffer between its
// writer and its reader in a chain; no more than max_inflight requests held
// working memory at once.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <thread>
#include <vector>

#include "../../p3dfft.3_amd/csrc/host/batch.h"
#include "../../p3dfft.3_amd/csrc/host/serve.h"

using namespace p3e;

void mock_enqueue(hipStream_t st, std::function<void()> f);
void mock_drain();
void mock_set_delay(int us);
void mock_shutdown();

// ---------------------------------------------------------------- recorded passes
struct Exec {
  const void *in;
  void *out;
};
static std::mutex g_rec_mu;
static std::vector<Exec> g_execs;  // in device order

static void record(const void *in, void *out) {
  std::lock_guard<std::mutex> g(g_rec_mu);
  g_execs.push_back(Exec{in, out});
}

extern "C" int p3e_launch_pass(const PassArgs *a, int, size_t, int, unsigned int, hipStream_t st) {
  const void *in = a->in;
  void *out = a->out;
  mock_enqueue(st, [in, out] { record(in, out); });
  return 0;
}

extern "C" int p3e_launch_step(const void *table, const void *work, int nwork, int, size_t, unsigned int,
                               hipStream_t st) {
  mock_enqueue(st, [table, work, nwork] {
    const PassArgs *t = (const PassArgs *)table;
    const WorkItem *w = (const WorkItem *)work;
    std::set<uint32_t> seen;
    for (int i = 0; i < nwork; i++)
      if (seen.insert(w[i].pass).second) record(t[w[i].pass].in, t[w[i].pass].out);
  });
  return 0;
}

// ---------------------------------------------------------------- stubs
namespace p3e {
static char g_dummy[64];
bool upload_pass_tables(int, PassDesc &pd, std::vector<void *> &) {
  PassArgs &a = pd.args;
  if (!pd.tw.empty()) a.tw = g_dummy;
  if (!pd.chirp.empty()) a.chirp = g_dummy;
  if (!pd.chirp_hat.empty()) a.chirp_hat = g_dummy;
  if (!pd.pre.empty()) a.pre = g_dummy;
  if (!pd.post.empty()) a.post = g_dummy;
  if (!pd.wN.empty()) a.wN = g_dummy;
  return true;
}
unsigned int persistent_grid(int, size_t, int nwork) { return nwork > 0 ? 1 : 1; }
Plan *plan_from_chain(int, PassChain &, const Layout &, const Layout &) { return nullptr; }
int plan_exec(Plan *, const void *, void *, int, bool, hipStream_t, bool) { return -1; }
void plan_destroy(Plan *) {}
hipStream_t default_stream() { return nullptr; }
}  // namespace p3e

// ---------------------------------------------------------------- scenario
static int scenario(int lanes, int max_inflight, size_t pool, int nreq, unsigned seed) {
  {
    std::lock_guard<std::mutex> g(g_rec_mu);
    g_execs.clear();
  }
  std::mt19937 rng(seed);
  const int edges[] = {16, 24, 32, 40, 48, 64, 96, 128};
  char lv[8];
  snprintf(lv, sizeof(lv), "%d", lanes);
  setenv("P3DFFT_HIP_SERVE_LANES", lv, 1);
  PoissonServer *s = poisson_server_create(4, pool, max_inflight);
  if (!s) {
    printf("create failed: %s\n", last_error());
    return 1;
  }
  struct Req {
    int dims[3];
    std::vector<float> rhs, sol;
    long long id;
    int npass;
  };
  std::vector<Req> reqs(nreq);
  for (int i = 0; i < nreq; i++) {
    Req &r = reqs[i];
    for (int a = 0; a < 3; a++) r.dims[a] = edges[rng() % 8];
    if (i % 7 == 3) r.dims[0] = r.dims[1] = 128, r.dims[2] = 64;  // 2^20 points: dedicated launches
    const size_t n = (size_t)r.dims[0] * r.dims[1] * r.dims[2];
    r.rhs.assign(4, 0.f);  // the mock kernels never touch the data
    r.sol.assign(4, 0.f);
    (void)n;
  }
  // a watchdog that reports the service state if the scenario stalls
  bool finished = false;
  std::thread dog([&] {
    for (int t = 0; t < 100 && !finished; t++) std::this_thread::sleep_for(std::chrono::milliseconds(100));
    if (finished) return;
    int pend, hold, busy, nl;
    poisson_state(s, &pend, &hold, &busy, &nl);
    printf("stall: pending %d holding %d groups in flight %d lanes %d\n", pend, hold, busy, nl);
    fflush(stdout);
  });
  // submit in bursts, polling between, as a client would
  std::map<long long, int> done;
  std::map<long long, std::pair<double, double>> span;  // id -> (admitted, done)
  auto t0 = std::chrono::steady_clock::now();
  int next = 0;
  std::vector<long long> ids(nreq, -1);
  while ((int)done.size() < nreq) {
    for (int k = 0; k < 5 && next < nreq; k++, next++) {
      ids[next] = poisson_submit(s, reqs[next].dims, reqs[next].rhs.data(), reqs[next].sol.data());
      if (ids[next] < 0) {
        printf("submit failed: %s\n", last_error());
        return 1;
      }
    }
    long long got[64];
    double ts[64], ta[64], td[64];
    int ok[64];
    const int n = poisson_poll(s, got, ts, ta, td, ok, 64, 2);
    for (int i = 0; i < n; i++) {
      if (!ok[i]) {
        printf("request %lld failed\n", got[i]);
        return 1;
      }
      done[got[i]] = 1;
      span[got[i]] = {ta[i], td[i]};
    }
    if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60)) {
      printf("HANG: %zu of %d requests done (max_inflight %d)\n", done.size(), nreq, max_inflight);
      finished = true;
      dog.join();
      return 2;
    }
  }
  finished = true;
  dog.join();
  poisson_server_destroy(s);
  mock_drain();
  // admitted requests held working memory at most max_inflight at once
  {
    std::vector<std::pair<double, int>> ev;  // (time, +1 admit / -1 done), done first at equal times
    for (int i = 0; i < nreq; i++) {
      ev.push_back({span[ids[i]].first, +1});
      ev.push_back({span[ids[i]].second, -1});
    }
    std::sort(ev.begin(), ev.end());
    int cur = 0, peak = 0;
    for (auto &e : ev) peak = std::max(peak, cur += e.second);
    if (peak > max_inflight) {
      printf("%d requests held at once (limit %d)\n", peak, max_inflight);
      return 6;
    }
  }
  // chains: every request's passes, in device order, rhs -> ... -> sol
  std::vector<Exec> ex;
  {
    std::lock_guard<std::mutex> g(g_rec_mu);
    ex = g_execs;
  }
  for (int i = 0; i < nreq; i++) {
    const void *cur = reqs[i].rhs.data();
    size_t pos = 0;
    int steps = 0;
    while (cur != (const void *)reqs[i].sol.data()) {
      // the next pass that reads `cur` after position pos
      size_t p = pos;
      while (p < ex.size() && ex[p].in != cur) p++;
      if (p == ex.size()) {
        printf("request %d: chain broken after %d passes\n", i, steps);
        return 3;
      }
      // nobody else wrote `cur` between the writer (pos - 1) and this reader
      for (size_t q = pos; q < p; q++)
        if (ex[q].out == cur && cur != (const void *)reqs[i].rhs.data()) {
          printf("request %d: buffer overwritten between passes %d and %d\n", i, steps, steps + 1);
          return 4;
        }
      cur = ex[p].out;
      pos = p + 1;
      if (++steps > 64) {
        printf("request %d: chain does not reach the solution\n", i);
        return 5;
      }
    }
  }
  printf("ok: lanes %d max_inflight %d pool %zu requests %d passes %zu\n", lanes, max_inflight, pool, nreq,
         ex.size());
  return 0;
}

int main() {
  setvbuf(stdout, NULL, _IONBF, 0);
  int rc = 0;
  const size_t big = 1ull << 30, tight = 40ull << 20;  // tight: one 2^20-point request plus a few small ones
  for (int lanes : {1, 4})
    for (int delay : {30, 0, 2})
      for (int mi : {1, 2, 4, 64})
        for (size_t pool : {big, tight}) {
          mock_set_delay(delay);
          printf("device delay %d us: ", delay);
          rc = scenario(lanes, mi, pool, 40, 1234u + mi + delay + 7 * lanes);
          if (rc) {
            mock_shutdown();
            return rc;
          }
        }
  mock_shutdown();
  return 0;
}
