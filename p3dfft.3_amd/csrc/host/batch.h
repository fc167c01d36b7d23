// Batched execution of many independent 1D/3D transforms (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "pass_plan.h"

namespace p3e {

struct BatchPlan {
  int prec;
  int nitems;
  int nstages;                    // passes per item (the same for all items)
  std::vector<Layout> lin, lout;  // per item
  std::vector<PassChain> chains;  // per item
  std::vector<void *> owned;      // device constant tables
  std::vector<PassArgs> table;    // stage-major: table[s * nitems + i]
  PassArgs *h_table;              // pinned staging copy of `table`
  PassArgs *d_table;
  std::vector<WorkItem *> d_work;  // per stage
  std::vector<int> nwork;          // per stage
  size_t lds;                      // dynamic LDS of the step kernel
  unsigned int grid;               // workgroups per step launch
  std::vector<size_t> scratch_off;  // per item: two scratch slots
  std::vector<size_t> scratch_sz;
  void *scratch;
  size_t scratch_bytes;
  bool use_graph;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  bool graph_ok;
  hipEvent_t table_done;  // the last table upload has been consumed
  bool table_pending;
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
