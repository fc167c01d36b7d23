// Extensions declared in include/p3dfft_hip.h.
#include <stdio.h>
#include <string.h>

#include "../../../include/p3dfft.h"
#include "../../../include/p3dfft_hip.h"
#include "engine.h"

using namespace p3dfft;

static p3e::Plan *engine_of(int plan) {
  if (plan < 0 || plan >= (int)stored_trans3D.size() || !stored_trans3D[plan]) {
    p3e::set_error("invalid plan handle %d", plan);
    return nullptr;
  }
  return (p3e::Plan *)stored_trans3D[plan]->engine();
}

extern "C" {

int p3dfft_hip_last_error(char *buf, int len) {
  if (buf && len > 0) {
    strncpy(buf, p3e::last_error(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return p3e::last_error_code();
}

void p3dfft_hip_clear_error(void) { p3e::clear_error(); }

void *p3dfft_hip_stream(void) { return (void *)p3e::default_stream(); }

static int exec_async(int plan, int prec, void *in, void *out, int idir, int OW, void *stream) {
  p3e::Plan *p = engine_of(plan);
  if (!p) return -1;
  if (p->prec != prec) {
    p3e::set_error("async exec: plan precision %d does not match the call", p->prec);
    return -1;
  }
  return p3e::plan_exec(p, in, out, idir, OW != 0, (hipStream_t)stream, false);
}

int p3dfft_exec_3Dtrans_double_async(int plan, double *in, double *out, int idir, int OW, void *stream) {
  return exec_async(plan, 8, in, out, idir, OW, stream);
}

int p3dfft_exec_3Dtrans_single_async(int plan, float *in, float *out, int idir, int OW, void *stream) {
  return exec_async(plan, 4, in, out, idir, OW, stream);
}

int p3dfft_hip_plan_npasses(int plan) {
  p3e::Plan *p = engine_of(plan);
  return p ? (int)p->passes.size() : -1;
}

int p3dfft_hip_plan_describe(int plan, char *buf, int len) {
  p3e::Plan *p = engine_of(plan);
  if (!p) return -1;
  if (buf && len > 0) {
    strncpy(buf, p->desc.c_str(), (size_t)len - 1);
    buf[len - 1] = 0;
  }
  return (int)p->desc.size();
}

int p3dfft_hip_pass_bytes(int plan, int i, long long *bytes_in, long long *bytes_out) {
  p3e::Plan *p = engine_of(plan);
  if (!p || i < 0 || i >= (int)p->passes.size()) return -1;
  if (bytes_in) *bytes_in = (long long)p->passes[i].d.in_bytes;
  if (bytes_out) *bytes_out = (long long)p->passes[i].d.out_bytes;
  return 0;
}

void p3dfft_hip_set_fast_path(int enable) { p3e::set_fast_path(enable != 0); }

int p3dfft_hip_profile(int plan, int enable) {
  p3e::Plan *p = engine_of(plan);
  return p ? p3e::plan_profile(p, enable != 0) : -1;
}

int p3dfft_hip_profile_read(int plan, double *avg_ms, int cap, double *total_ms) {
  p3e::Plan *p = engine_of(plan);
  return p ? p3e::plan_profile_read(p, avg_ms, cap, total_ms) : -1;
}

}  // extern "C"
