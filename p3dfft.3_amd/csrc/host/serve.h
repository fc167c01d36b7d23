// Poisson service (internal): periodic Poisson solves of many grid sizes on
// one GPU, admitted continuously and run concurrently on several streams.
//
// Requests are admitted in arrival order when an in-flight slot
// (max_inflight) and their working memory (two buffers of the largest pass
// output, from one pool allocated up front) are free.  An admitted request
// is launched at once as a group: all six passes of a large grid on one
// lane (a stream of the service), or, for grids below DEDICATED_POINTS, the
// waiting small requests together as six step-kernel launches (one per pass,
// each over the tiles of every request in the group).  Groups on different
// lanes run concurrently on the device, so a request that arrives while
// others run starts as soon as a lane has room, and its passes overlap
// theirs.  Plans and constant tables stay resident per grid size.  A
// request's memory returns to the pool when its group has completed.
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
  // launch of the per-pass kernels (one such pass fills most of the GPU);
  // smaller grids share step-kernel launches.  The choice depends on the
  // grid size only, so a request's arithmetic is the same alone and in any
  // mix.
  bool dedicated;
  std::vector<PassDesc> passes;  // forward then backward, tables on the device
  std::vector<void *> owned;
  size_t work_bytes;             // one of the two working buffers
  size_t real_bytes;
};

// 2^18 points: a 64^3 grid's pass kernels already have about 1000
// workgroups.
static const long long DEDICATED_POINTS = 1LL << 18;
// Lanes (streams) of the service unless P3DFFT_HIP_SERVE_LANES (1..8) says
// otherwise, and launch groups queued per lane: the service thread enqueues
// a lane's next group while its current one runs.  Closed loop of 256
// mixed-size requests: 7,666 / 8,597 / 9,034 solves/s with 2 / 4 / 8 lanes.
static const int SERVE_LANES = 8;
static const int LANE_DEPTH = 2;
// Small requests in one step-kernel group, at most.
static const int SMALL_GROUP_MAX = 64;

struct PoissonRequest {
  long long id;
  PoissonPlan *plan;
  const void *rhs;
  void *sol;
  size_t off;  // working memory: pool + off, 2 * plan->work_bytes
  double t_submit, t_admit, t_done;
  bool failed;
};

// One launch group: a large request's passes, or a batch of small requests'
// step launches, on one lane, followed by an event.
struct PoissonGroup {
  int lane;
  bool busy;
  hipEvent_t done;
  std::vector<PoissonRequest *> reqs;
  // step-kernel tables of a small batch (pinned host copy and device copy)
  PassArgs *h_tab, *d_tab;
  WorkItem *h_work, *d_work;
  size_t cap_tab, cap_work;
};

struct PoissonServer {
  int prec;
  int device;
  int max_inflight;
  int nlanes;
  char *pool;
  size_t pool_bytes;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  std::map<long long, PoissonPlan *> plans;
  std::vector<hipStream_t> lanes;
  std::vector<PoissonGroup> groups;  // nlanes * LANE_DEPTH
  std::mutex mu;
  std::condition_variable cv, done_cv;
  std::deque<PoissonRequest *> pending;
  std::vector<PoissonRequest *> completed;  // not yet polled
  int holding;  // admitted and not yet retired
  long long next_id;
  long long ngroups, nlaunch, ntiles;
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
// Requests waiting for admission, admitted and not yet done, launch groups
// in flight, lanes.
void poisson_state(PoissonServer *s, int *pending, int *holding, int *groups_busy, int *lanes);
void poisson_server_destroy(PoissonServer *s);

// The per-request path: the same Poisson plan on the regular pass kernels,
// one solve at a time (plans cached per size and precision).
int poisson_solve(int prec, const int dims[3], const void *rhs, void *sol, hipStream_t st, bool sync);

// Host-side check of a launch before it is made: every pass has threads to
// run its tiles and its tables; every work item names a pass of the table
// and a tile of that pass.  Returns false (and records why) otherwise.
bool check_launch(const PassArgs *tab, const PassDesc *const *descs, size_t ntab, const WorkItem *work,
                  size_t nwork);

}  // namespace p3e
