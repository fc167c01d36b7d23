// TIMERS phase report of libp3dfft3_hip.so: the reference's timer class
// (include/p3dfft.h:958-992, build/init.C:755-825) over per-pass HIP events.
//
// timers.init() zeroes the phases and turns the report on; from then on
// every synchronous exec (the drop-in C and C++ calls) records a pair of
// HIP events around each pass it launches, and after the exec has finished
// adds each pass's GPU time to its phase (the engine classifies the pass
// the way the reference's transplan::exec picks its regime).  With
// libroctx64 loadable, each pass launch is also wrapped in a roctx range
// named after its phase, so that rocprofv3 --marker-trace shows the
// phases on the timeline.  timers.print(comm) writes the reference's report.
#define TIMERS
#include <dlfcn.h>
#include <stdio.h>

#include <atomic>
#include <mutex>

#include "../../../include/p3dfft.h"
#include "engine.h"

// MPI is reached only when the host program links it (the library does
// not link libmpi).
extern "C" int MPI_Initialized(int *flag) __attribute__((weak));
extern "C" int MPI_Comm_rank(MPI_Comm comm, int *rank) __attribute__((weak));

namespace p3e {

static std::atomic<bool> g_on{false};
static std::mutex g_mu;
static double g_sec[PH_COUNT];

typedef int (*roctx_push_fn)(const char *);
typedef int (*roctx_pop_fn)();
static roctx_push_fn g_push = nullptr;
static roctx_pop_fn g_pop = nullptr;

void timers_enable(bool on) {
  if (on) {
    static std::once_flag once;
    std::call_once(once, [] {
      void *h = dlopen("libroctx64.so.4", RTLD_NOW | RTLD_LOCAL);
      if (!h) h = dlopen("libroctx64.so", RTLD_NOW | RTLD_LOCAL);
      if (h) {
        g_push = (roctx_push_fn)dlsym(h, "roctxRangePushA");
        g_pop = (roctx_pop_fn)dlsym(h, "roctxRangePop");
        if (!g_push || !g_pop) g_push = nullptr, g_pop = nullptr;
      }
    });
  }
  g_on = on;
}

bool timers_enabled() { return g_on; }

void timers_reset() {
  std::lock_guard<std::mutex> g(g_mu);
  for (double &v : g_sec) v = 0.0;
}

void timers_add(int phase, double seconds) {
  if (phase < 0 || phase >= PH_COUNT) return;
  std::lock_guard<std::mutex> g(g_mu);
  g_sec[phase] += seconds;
}

void timers_read(double seconds[PH_COUNT]) {
  std::lock_guard<std::mutex> g(g_mu);
  for (int i = 0; i < PH_COUNT; i++) seconds[i] = g_sec[i];
}

void timers_range_push(const char *name) {
  if (g_push) g_push(name);
}

void timers_range_pop() {
  if (g_pop) g_pop();
}

}  // namespace p3e

namespace p3dfft {

timer timers;

// build/init.C:755-769
void timer::init() {
  reorder_deriv = reorder_trans = reorder_out = reorder_in = 0.0;
  trans_exec = trans_deriv = 0.0;
  packsend = packsend_trans = packsend_deriv = unpackrecv = alltoall = 0.0;
  p3e::timers_reset();
  p3e::timers_enable(true);
}

// build/init.C:771-825.  One rank runs the whole grid, so the average,
// minimum and maximum over ranks are the rank's own values.
void timer::print(MPI_Comm comm) {
  double s[p3e::PH_COUNT];
  p3e::timers_read(s);
  reorder_deriv = s[p3e::PH_REORDER_DERIV];
  reorder_trans = s[p3e::PH_REORDER_TRANS];
  reorder_out = s[p3e::PH_REORDER_OUT];
  reorder_in = s[p3e::PH_REORDER_IN];
  trans_exec = s[p3e::PH_TRANS_EXEC];
  trans_deriv = s[p3e::PH_TRANS_DERIV];
  int taskid = 0, flag = 0;
  if (MPI_Initialized && MPI_Comm_rank && MPI_Initialized(&flag) == 0 && flag) (void)MPI_Comm_rank(comm, &taskid);
  if (taskid != 0) return;
  printf("TIMERS (avg/min/max): \nReorder_trans (%lf %lf %lf)\n", reorder_trans, reorder_trans, reorder_trans);
  printf("Reorder_out   (%lf %lf %lf)\n", reorder_out, reorder_out, reorder_out);
  printf("Reorder_in    (%lf %lf %lf)\n", reorder_in, reorder_in, reorder_in);
  printf("Trans_exec    (%lf %lf %lf)\n", trans_exec, trans_exec, trans_exec);
  printf("Packsend      (%lf %lf %lf)\n", packsend, packsend, packsend);
  printf("Packsend_trans(%lf %lf %lf)\n", packsend_trans, packsend_trans, packsend_trans);
  printf("Unpackrecv    (%lf %lf %lf)\n", unpackrecv, unpackrecv, unpackrecv);
  printf("Alltoall      (%lf %lf %lf)\n", alltoall, alltoall, alltoall);
  // the derivative phases (the reference accumulates them but leaves them
  // out of its report)
  printf("Reorder_deriv (%lf %lf %lf)\n", reorder_deriv, reorder_deriv, reorder_deriv);
  printf("Trans_deriv   (%lf %lf %lf)\n", trans_deriv, trans_deriv, trans_deriv);
  fflush(stdout);
}

}  // namespace p3dfft
