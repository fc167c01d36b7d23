// Internal engine of libp3dfft3_hip.so: runs a planned chain of fused
// reorder + 1D transform passes on a HIP stream.  Not a public header.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pass_plan.h"

namespace p3e {

struct Pass {
  PassDesc d;                  // geometry and host tables
  std::vector<void *> tables;  // device copies owned by the pass
};

// Phases of the reference's TIMERS report (fields of p3dfft::timer), which
// the synchronous exec calls accumulate while the report is enabled.
enum Phase {
  PH_REORDER_DERIV = 0, PH_REORDER_TRANS, PH_REORDER_OUT, PH_REORDER_IN, PH_TRANS_EXEC, PH_TRANS_DERIV, PH_COUNT
};
void timers_enable(bool on);
bool timers_enabled();
void timers_reset();
void timers_add(int phase, double seconds);
void timers_read(double seconds[PH_COUNT]);
// roctx range around a pass launch while the report is enabled (no-ops
// when libroctx64 is not loadable)
void timers_range_push(const char *name);
void timers_range_pop();

struct Plan {
  int prec;  // 4 or 8
  Layout lin, lout;
  size_t in_bytes, out_bytes;
  std::vector<Pass> passes;
  int pass_of_axis[3];
  bool copy_only;
  // destination buffer of each pass, scratch sizes and the final copy,
  // for separate (index 0) and aliased (index 1) in and out arrays
  int buf_dst[2][8];
  size_t buf_need[2][2];
  bool buf_copy_back[2];
  void *scratch[2];
  size_t scratch_cap[2];
  void *stage_in, *stage_out;
  size_t stage_in_cap, stage_out_cap;
  std::string desc;
  // per-pass timing with HIP events (off unless enabled)
  bool prof;
  std::vector<hipEvent_t> ev_pool;
  std::vector<int> ev_pass;  // pass index of each recorded (start, stop) pair
  size_t ev_used;
  // TIMERS: one (start, stop) pair per pass of the current synchronous exec
  std::vector<hipEvent_t> tm_ev;
};

int plan_profile(Plan *p, bool enable);
// Average duration of each pass over the launches recorded since enabling;
// returns the number of recorded launches of pass 0.
int plan_profile_read(Plan *p, double *avg_ms, int cap, double *total_ms);

// ops[a]: operation on logical axis a; order: axes in execution order (R2C
// first, C2R last).  Returns nullptr on error (message in last_error()).
Plan *plan_create(int prec, const OpSpec ops[3], const int order[3], const Layout &in, const Layout &out);
// A plan over an already planned chain from layout `in` to `out` (uploads
// its tables).
Plan *plan_from_chain(int prec, PassChain &chain, const Layout &in, const Layout &out);
// idir: logical axis of the spectral derivative, or -1.  in/out may be host
// or device pointers; host arrays are staged and imply a synchronous call.
int plan_exec(Plan *p, const void *in, void *out, int idir, bool OW, hipStream_t st, bool sync);
// The passes of p from device array din to device array dout on st, with
// the intermediate arrays at scr[0] and scr[1] (at least p->buf_need bytes
// each for the aliasing case of din and dout); asynchronous, allocates
// nothing, records no timing.
int plan_launch(Plan *p, const void *din, void *dout, void *const scr[2], int idir, hipStream_t st);
void plan_destroy(Plan *p);

// Standalone derivative of a stored spectrum (compute_deriv semantics).
int deriv_exec(int prec, const void *in, void *out, const int sdims[3], int ldir, int g, int mid, int gstart,
               hipStream_t st);

bool is_device_ptr(const void *ptr);
hipStream_t default_stream();
void release_all();

}  // namespace p3e
