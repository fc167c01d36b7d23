// Poisson service (internal): periodic Poisson solves of many grid sizes,
// continuously batched at pass granularity on one GPU.
//
// Each admitted request advances one pass per step; a step is one launch of
// the step kernel over the current pass of every small request in flight
// (plus one launch of the per-pass kernels for each large one), so requests
// that arrive while others run join at their first pass in the next step
// and their work shares launches with the others'.  Requests are admitted in
// arrival order when a slot and their working memory (two buffers of the
// largest pass output, from one pool allocated up front) are free.  Plans
// and constant tables stay resident per grid size.  Up to SERVE_DEPTH steps
// may be in flight; a request's memory returns to the pool when the step
// with its last pass has completed.
#pragma once
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "engine.h"
#include "pass_plan.h"

namespace p3e {

struct PoissonPlan {
  int dims[3];
  // Grids of at least DEDICATED_POINTS points run each pass as its own
  // launch of the per-pass kernels (one such pass fills the GPU); smaller
  // grids share step-kernel launches.  The choice depends on the grid size
  // only, so a request's arithmetic is the same alone and in any mix.
  bool dedicated;
  std::vector<PassDesc> passes;  // forward then backward, tables on the device
  std::vector<void *> owned;
  size_t work_bytes;             // one of the two working buffers
  size_t real_bytes;
};

// 2^18 points: a 64^3 grid's pass kernels already have about 1000
// workgroups; grids above it ran faster with their own launches on MI355X
// (4230 against 3870 solves/s for 64..384-point edges, where the step
// kernel's batch tiles lack the in-register prime radices).
static const long long DEDICATED_POINTS = 1LL << 18;
// Large (dedicated) requests in flight at most, whatever max_inflight
// allows: the passes of the requests in flight are interleaved step by
// step, and a request's two working arrays stay in the 256 MiB Infinity
// Cache between its passes only when few others are interleaved with it
// (4500 solves/s at 2-4 against 4220 at 64 on MI355X).
static const int LARGE_INFLIGHT = 4;
// Steps in flight on the device: the service thread builds the next step
// while the earlier one runs.
static const int SERVE_DEPTH = 2;

struct PoissonRequest {
  long long id;
  PoissonPlan *plan;
  const void *rhs;
  void *sol;
  size_t off;  // working memory: pool + off, 2 * plan->work_bytes
  int phase;   // next pass
  double t_submit, t_admit, t_done;
  bool failed;
};

struct PoissonServer {
  int prec;
  int device;
  int max_inflight;
  char *pool;
  size_t pool_bytes;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  std::map<long long, PoissonPlan *> plans;
  hipStream_t st;
  struct Step {
    PassArgs *h_tab, *d_tab;
    WorkItem *h_work, *d_work;
    size_t cap_tab, cap_work;
    hipEvent_t done;
    bool busy;
    std::vector<PoissonRequest *> finishing;
  } steps[SERVE_DEPTH];
  int cur;
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<PoissonRequest *> pending;
  std::vector<PoissonRequest *> active;
  std::vector<PoissonRequest *> completed;  // not yet polled
  int holding;        // admitted and not yet retired
  int holding_large;  // of which dedicated (large-grid) requests
  long long next_id;
  long long nsteps, ntiles;
  bool stop, failed;
  std::string error;
  std::thread th;
  double t0;
};

PoissonServer *poisson_server_create(int prec, size_t pool_bytes, int max_inflight);
// Plan (and upload) a grid size ahead of its first request.
int poisson_server_prepare(PoissonServer *s, const int dims[3]);
// rhs, sol: device arrays of dims[0]*dims[1]*dims[2] reals in memory order
// (0,1,2) (x stride-1); rhs must be complete when submitted.  Returns the
// request id, or -1.
long long poisson_submit(PoissonServer *s, const int dims[3], const void *rhs, void *sol);
// Completed requests since the last poll (waits up to wait_ms for one).
int poisson_poll(PoissonServer *s, long long *ids, double *t_submit, double *t_admit, double *t_done, int *ok,
                 int cap, int wait_ms);
void poisson_server_destroy(PoissonServer *s);

// The per-request path: the same Poisson plan on the regular pass kernels,
// one solve at a time (plans cached per size and precision).
int poisson_solve(int prec, const int dims[3], const void *rhs, void *sol, hipStream_t st, bool sync);

}  // namespace p3e
