// Pass planning for libp3dfft3_hip.so, in plain C++ (no HIP calls): radix
// plans, constant tables in long double, tile geometry, the memory order of
// every intermediate array, and the assignment of passes to buffers.  The
// engine uploads the tables and launches; the CPU tests run the same plans
// through the host build of the kernel code.
#pragma once
#include <complex>
#include <string>
#include <vector>

#include "../kernels/pass_args.h"

namespace p3e {

typedef std::complex<long double> cld;

// A 1D operation along one logical axis.
struct OpSpec {
  int cls;       // CLS_*
  int sign;      // C2C: -1 forward, +1 backward
  int r2r_kind;  // for CLS_R2R
  int in_dt;     // 1 real, 2 complex
  int out_dt;
};

// An array: logical dims, MemOrder (storage rank of each logical axis; 0 is
// stride-1), element type (1 real, 2 complex), and the allocated extent of
// the stride-1 axis (pitch; 0 or at most its dim: dense).  Caller arrays are
// dense; the planner pads the intermediate arrays between passes.
struct Layout {
  int dims[3];
  int mo[3];
  int dt;
  int pitch = 0;
};

struct PassDesc {
  PassArgs args;  // table pointers and in/out are filled by the executor
  int axis;              // logical axis d of the transform
  int axis_e, axis_f;    // logical axes of the tile (e) and outer (f) roles
  Layout lin, lout;
  size_t in_bytes, out_bytes;
  size_t lds_bytes;
  int threads;
  unsigned int nblocks;
  bool inplace_ok;
  bool deriv_capable;  // complex output along a transformed axis
  std::vector<cld> tw, chirp, chirp_hat, pre, post, wN;
};

// Disable the power-of-two path (tests compare both paths).
void set_fast_path(bool enable);
// Test hook: plan complex DCT-I passes of 129, 257, 513 and 1025 points in
// the split form (default) or as the FFT of the even extension.
void set_dct_split(bool on);
// Intermediate arrays of at least this many bytes whose stride-1 rows are a
// multiple of 256 bytes get padded rows (tests lower it to reach small
// grids).
void set_pad_threshold(size_t bytes);

struct PassChain {
  std::vector<PassDesc> passes;
  int pass_of_axis[3];
  bool copy_only;  // no transform and identical layouts
  std::string desc;
};

size_t layout_elems(const Layout &l);
size_t elem_bytes(int prec, int dt);

// batch_threads > 0 plans the pass for the batch (step) kernel, whose
// workgroups have that many threads; 0 plans it for its own launch.
bool describe_pass(int prec, int axis, const OpSpec &op, const Layout &lin, const Layout &lout, PassDesc &ps,
                   int batch_threads = 0);
// ops[a]: operation on logical axis a; order: axes in execution order.
bool plan_chain(int prec, const OpSpec ops[3], const int order[3], const Layout &in, const Layout &out,
                PassChain &chain, int batch_threads = 0);
// The two chains of a periodic Poisson solve on a real grid of logical dims
// (nx, ny, nz) in memory order mo_real: forward R2C x, C2C y, C2C z into the
// spectrum (nx/2+1, ny, nz) in memory order mo_spec, whose last pass also
// multiplies by -1/|k|^2 (0 at k = 0) and 1/(nx ny nz); backward C2C, C2C,
// C2R x back to mo_real.  The domain is [0, 2 pi)^3, so wavenumbers are
// integers.
bool plan_poisson(int prec, const int dims[3], const int mo_real[3], const int mo_spec[3], PassChain &fwd,
                  PassChain &bwd, int batch_threads = 0);
// Re-tile a generic-path or single-buffer mixed-radix pass planned for the
// batch kernel with B lines per tile and the given radix sequence (the
// autotuner's knobs).  Returns false if the choice is invalid.
bool retune_pass(int prec, PassDesc &ps, int B, const int *rad, int nrad);
// True when every constant table the planner built for this pass (twiddles,
// chirps, pre/post factors, R2C/C2R split factors) has a device address in
// its args: the kernels read those tables whenever the planner built them.
// Checked before every launch; a pass that fails it is never launched.
bool tables_ready(const PassDesc &ps);
// Radix sequences worth trying for a generic-path length W (at most `cap`).
std::vector<std::vector<int>> radix_candidates(int W, int cap);

enum { BUF_IN = 0, BUF_OUT = 1, BUF_S0 = 2, BUF_S1 = 3 };
// Destination buffer of each pass (BUF_OUT or scratch), and scratch sizes.
// alias: caller's in and out are the same array.  copy_back: the last pass
// writes scratch and the result must be copied to out.
void assign_buffers(const std::vector<PassDesc> &passes, size_t out_cap, bool alias, int dst[8],
                    size_t need[2], bool *copy_back);

// errors (shared with the engine)
void set_error(const char *fmt, ...);
const char *last_error();
int last_error_code();
void clear_error();

}  // namespace p3e
