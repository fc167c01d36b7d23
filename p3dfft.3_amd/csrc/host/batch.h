// Batched execution of many independent 1D/3D transforms (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pass_plan.h"

namespace p3e {

// A grouped launch of a batch stage: the same pass kernel (power-of-two of
// complex length M, kind 1; single-buffer without an odd prime above 5,
// kind 2) over n items whose PassArgs are table entries tab0 .. tab0+n-1.
struct BatchGroup {
  int kind, M;
  int threads;             // the largest of the items'
  size_t lds;              // the largest of the items'
  unsigned int nblocks;    // all tiles of the n items
  size_t tab0;
  int n;
  uint32_t *d_start;       // n + 1 workgroup offsets (device)
  double work;             // bytes the n items' passes read and write
  int q;                   // stream: 0 the exec stream, q > 0 side stream q - 1
};

// Streams a batch stage's grouped launches are spread over (the device's
// default number of hardware queues per process).
static const int BATCH_STREAMS = 4;

struct BatchPlan {
  int prec;
  int nitems;
  int nstages;                    // passes per item (the same for all items)
  // grouped: items planned as standalone passes, each stage one grouped
  // launch per pass kernel; else every stage one step-kernel launch
  bool grouped;
  std::vector<Layout> lin, lout;  // per item
  std::vector<PassChain> chains;  // per item
  std::vector<void *> owned;      // device constant tables
  std::vector<PassArgs> table;    // the launches' PassArgs
  std::vector<size_t> slot;       // table index of stage s of item i: slot[s * nitems + i]
  std::vector<std::vector<BatchGroup>> groups;  // per stage (grouped)
  // two pinned staging copies of `table`, used by alternate execs so that
  // the host fills one while the device may still read the other
  PassArgs *h_table[2];
  PassArgs *d_table;
  std::vector<WorkItem *> d_work;  // per stage (step kernel)
  std::vector<int> nwork;          // per stage (step kernel)
  size_t lds;                      // dynamic LDS of the step kernel
  unsigned int grid;               // workgroups per step launch
  std::vector<size_t> scratch_off;  // per item: two scratch slots
  std::vector<size_t> scratch_sz;
  void *scratch;
  size_t scratch_bytes;
  bool use_graph;
  hipGraph_t graph[2];  // one per staging copy
  hipGraphExec_t gexec[2];
  bool graph_ok[2];
  hipEvent_t table_done[2];  // the upload from staging copy k has been made
  bool table_pending[2];
  int parity;                // staging copy of the next exec
  // grouped: the launches of one stage run on the exec stream and side
  // streams, forked from and joined back into the exec stream by events
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> fork_ev;  // per stage
  std::vector<hipEvent_t> join_ev;  // per side stream
  std::string desc;
};

// ops/order as for plan_create; in[i]/out[i] are the layouts of item i.
// tune: run the autotuner for the generic-path passes (results cached).
BatchPlan *batch_create(int prec, int nitems, const OpSpec ops[3], const int order[3], const Layout *in,
                        const Layout *out, bool tune);
// in[i]/out[i]: device pointers of item i.  sync: wait for completion.
int batch_exec(BatchPlan *b, void *const *in, void *const *out, hipStream_t st, bool sync);
void batch_destroy(BatchPlan *b);

// Device copies of a pass's constant tables, owned by `owned`; sets the
// table addresses in pd.args.
bool upload_pass_tables(int prec, PassDesc &pd, std::vector<void *> &owned);
// Workgroups of a persistent step-kernel launch over nwork tiles with the
// given dynamic LDS: every CU filled to its occupancy, at most nwork.
unsigned int persistent_grid(int prec, size_t lds, int nwork);

// Autotuner cache file (P3DFFT_HIP_TUNE_CACHE, default
// $HOME/.cache/p3dfft3_hip_tune.txt); returns the number of tuned keys.
int tune_cache_size();

}  // namespace p3e
