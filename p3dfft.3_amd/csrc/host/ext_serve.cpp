// Poisson service and per-request Poisson solve: C entry points declared in
// include/p3dfft_hip.h.
#include <vector>

#include "../../../include/p3dfft_hip.h"
#include "engine.h"
#include "serve.h"

static std::vector<p3e::PoissonServer *> g_servers;

static p3e::PoissonServer *server_of(int h) {
  if (h < 0 || h >= (int)g_servers.size() || !g_servers[h]) {
    p3e::set_error("invalid Poisson server handle %d", h);
    return nullptr;
  }
  return g_servers[h];
}

extern "C" {

int p3dfft_hip_poisson_server_create(int prec, long long pool_bytes, int max_inflight) {
  if (pool_bytes <= 0) {
    p3e::set_error("p3dfft_hip_poisson_server_create: pool_bytes must be positive");
    return -1;
  }
  p3e::PoissonServer *s = p3e::poisson_server_create(prec, (size_t)pool_bytes, max_inflight);
  if (!s) return -1;
  g_servers.push_back(s);
  return (int)g_servers.size() - 1;
}

int p3dfft_hip_poisson_server_prepare(int h, const int dims[3]) {
  p3e::PoissonServer *s = server_of(h);
  return s ? p3e::poisson_server_prepare(s, dims) : -1;
}

long long p3dfft_hip_poisson_submit(int h, const int dims[3], const void *rhs, void *sol) {
  p3e::PoissonServer *s = server_of(h);
  return s ? p3e::poisson_submit(s, dims, rhs, sol) : -1;
}

int p3dfft_hip_poisson_poll(int h, long long *ids, double *t_submit, double *t_admit, double *t_done, int *ok,
                            int cap, int wait_ms) {
  p3e::PoissonServer *s = server_of(h);
  if (!s || !ids || cap < 1) return -1;
  return p3e::poisson_poll(s, ids, t_submit, t_admit, t_done, ok, cap, wait_ms);
}

int p3dfft_hip_poisson_server_stats(int h, long long *launches, long long *tiles) {
  p3e::PoissonServer *s = server_of(h);
  if (!s) return -1;
  std::lock_guard<std::mutex> g(s->mu);
  if (launches) *launches = s->nlaunch;
  if (tiles) *tiles = s->ntiles;
  return s->failed ? 1 : 0;
}

int p3dfft_hip_poisson_server_state(int h, int *pending, int *holding, int *groups_busy, int *lanes) {
  p3e::PoissonServer *s = server_of(h);
  if (!s) return -1;
  p3e::poisson_state(s, pending, holding, groups_busy, lanes);
  return 0;
}

int p3dfft_hip_poisson_server_destroy(int h) {
  p3e::PoissonServer *s = server_of(h);
  if (!s) return -1;
  p3e::poisson_server_destroy(s);
  g_servers[h] = nullptr;
  return 0;
}

int p3dfft_hip_poisson_solve(int prec, const int dims[3], const void *rhs, void *sol, void *stream, int sync) {
  hipStream_t st = stream ? (hipStream_t)stream : p3e::default_stream();
  return p3e::poisson_solve(prec, dims, rhs, sol, st, sync != 0);
}

}  // extern "C"
