// Batched 3D transforms: C entry points declared in include/p3dfft_hip.h.
#include <stdio.h>
#include <string.h>

#include <vector>

#include "../../../include/p3dfft.h"
#include "../../../include/p3dfft_hip.h"
#include "batch.h"
#include "engine.h"

using namespace p3dfft;

static std::vector<p3e::BatchPlan *> g_batches;

static p3e::BatchPlan *batch_of(int h) {
  if (h < 0 || h >= (int)g_batches.size() || !g_batches[h]) {
    p3e::set_error("invalid batch plan handle %d", h);
    return nullptr;
  }
  return g_batches[h];
}

extern "C" {

int p3dfft_hip_batch_plan(int nitems, const int *dims_in, const int *dims_out, const int types[3],
                          const int mo_in[3], const int mo_out[3], int tune) {
  if (nitems < 1 || !dims_in || !dims_out || !types || !mo_in || !mo_out) {
    p3e::set_error("p3dfft_hip_batch_plan: invalid arguments");
    return -1;
  }
  for (int a = 0; a < 3; a++)
    if (types[a] < 0 || types[a] >= (int)types1D.size()) {
      p3e::set_error("p3dfft_hip_batch_plan: undefined 1D type ID %d (was setup() called?)", types[a]);
      return -1;
    }
  // stage order exactly as for a single transform3D (find_order)
  int pd[3] = {1, 1, 1}, dmap[3] = {0, 1, 2};
  int t3[3] = {types[0], types[1], types[2]};
  int m1[3] = {mo_in[0], mo_in[1], mo_in[2]}, m2[3] = {mo_out[0], mo_out[1], mo_out[2]};
  int g1d[3] = {dims_in[0], dims_in[1], dims_in[2]}, g2d[3] = {dims_out[0], dims_out[1], dims_out[2]};
  ProcGrid pg(pd, (MPI_Comm)0);
  DataGrid g1(g1d, -1, &pg, dmap, m1), g2(g2d, -1, &pg, dmap, m2);
  trans_type3D tt(t3);
  int L[3];
  find_order(L, &tt, &g1, &g2, nullptr);
  const int prec = tt.prec;
  p3e::OpSpec ops[3];
  for (int a = 0; a < 3; a++) {
    const gen_trans_type *t = types1D[types[a]];
    ops[a] = p3e::OpSpec{t->cls, t->isign < 0 ? -1 : 1, t->r2r_kind, t->dt1, t->dt2};
  }
  const int dt_in = types1D[types[L[0]]]->dt1, dt_out = types1D[types[L[2]]]->dt2;
  std::vector<p3e::Layout> lin(nitems), lout(nitems);
  for (int i = 0; i < nitems; i++) {
    for (int a = 0; a < 3; a++) {
      lin[i].dims[a] = dims_in[3 * i + a];
      lin[i].mo[a] = mo_in[a];
      lout[i].dims[a] = dims_out[3 * i + a];
      lout[i].mo[a] = mo_out[a];
    }
    lin[i].dt = dt_in;
    lout[i].dt = dt_out;
  }
  p3e::BatchPlan *b = p3e::batch_create(prec, nitems, ops, L, lin.data(), lout.data(), tune != 0);
  if (!b) return -1;
  g_batches.push_back(b);
  return (int)g_batches.size() - 1;
}

int p3dfft_hip_batch_exec(int h, void *const *in, void *const *out, void *stream, int sync) {
  p3e::BatchPlan *b = batch_of(h);
  if (!b) return -1;
  hipStream_t st = stream ? (hipStream_t)stream : p3e::default_stream();
  return p3e::batch_exec(b, in, out, st, sync != 0);
}

int p3dfft_hip_batch_set_graph(int h, int enable) {
  p3e::BatchPlan *b = batch_of(h);
  if (!b) return -1;
  b->use_graph = enable != 0;
  return 0;
}

int p3dfft_hip_batch_describe(int h, char *buf, int len) {
  p3e::BatchPlan *b = batch_of(h);
  if (!b) return -1;
  if (buf && len > 0) {
    strncpy(buf, b->desc.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return (int)b->desc.size();
}

int p3dfft_hip_batch_destroy(int h) {
  p3e::BatchPlan *b = batch_of(h);
  if (!b) return -1;
  p3e::batch_destroy(b);
  g_batches[h] = nullptr;
  return 0;
}

}  // extern "C"
