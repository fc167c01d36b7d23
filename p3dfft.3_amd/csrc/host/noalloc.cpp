// 3D exec that allocates nothing: the intermediate arrays of a plan's pass
// chain live in a fixed device array of this library instead of plan-owned
// hipMalloc memory, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../../include/p3dfft.h"
#include "../../../include/p3dfft_hip.h"
#include "engine.h"

// Intermediate arrays of every no-alloc exec on a device.  Each device that
// runs such an exec holds its own copy until the process ends.  The passes
// write every element of an intermediate array before the next pass reads
// it, so nothing needs clearing between calls; calls that overlap on one
// device share it and race.  Host code never uses this name as an address:
// it passes the device address that hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[P3DFFT_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

static const int MAX_DEVICES = 64;
static std::mutex g_mu;
static void *g_scratch[MAX_DEVICES];  // device address of `scratch`, per device

// Address of this device's copy of `scratch` (looked up once per device,
// with that device current); nullptr on error.
void *scratch_of_current_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
    p3e::set_error("no-alloc exec: no current HIP device (or more than %d devices)", MAX_DEVICES);
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_mu);
  if (!g_scratch[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      p3e::set_error("no-alloc exec: hipGetSymbolAddress of the scratch array failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[dev] = ptr;
  }
  return g_scratch[dev];
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

extern "C" int p3dfft_hip_exec_3Dtrans_noalloc(int plan, void *in, void *out, int idir, int OW, void *stream) {
  (void)OW;  // the input is never written, as in every exec of this library
  using p3dfft::stored_trans3D;
  if (plan < 0 || plan >= (int)stored_trans3D.size() || !stored_trans3D[plan]) {
    p3e::set_error("no-alloc exec: invalid plan handle %d", plan);
    return -1;
  }
  p3e::Plan *p = (p3e::Plan *)stored_trans3D[plan]->engine();
  if (!p3e::is_device_ptr(in) || !p3e::is_device_ptr(out)) {
    p3e::set_error("no-alloc exec: in and out must be device arrays (host arrays would need staging buffers)");
    return -1;
  }
  if (idir > 2) {
    p3e::set_error("no-alloc exec: idir %d is not an axis (0..2, or < 0 for none)", idir);
    return -1;
  }
  const bool alias = in == out;
  const size_t n0 = align256(p->buf_need[alias][0]), n1 = p->buf_need[alias][1];
  if (n0 + n1 > (size_t)P3DFFT_HIP_NOALLOC_SCRATCH_BYTES) {
    p3e::set_error("no-alloc exec: the plan needs %zu bytes of intermediate arrays, the limit is %zu", n0 + n1,
                   (size_t)P3DFFT_HIP_NOALLOC_SCRATCH_BYTES);
    return -1;
  }
  void *base = nullptr;
  if (n0 + n1 > 0 && !(base = scratch_of_current_device())) return -1;
  void *const scr[2] = {base, base ? (char *)base + n0 : nullptr};
  return p3e::plan_launch(p, in, out, scr, idir, (hipStream_t)stream);
}
