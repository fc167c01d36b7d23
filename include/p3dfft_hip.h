/*
 * p3dfft_hip.h -- extensions of libp3dfft3_hip.so beyond the reference ABI.
 * None of these has a counterpart in the reference's Cwrap.h; they serve
 * callers that keep arrays on the GPU and drive their own streams
 * (benchmarks, the Poisson service, profilers).
 *
 * Stream arguments are hipStream_t passed as void* (NULL = the library's
 * stream).  Array pointers passed to the *_async functions must be device
 * pointers; those functions return before the work completes.
 */
#ifndef P3DFFT_HIP_H
#define P3DFFT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Last error recorded by the library: returns 0 if none, else nonzero, and
 * copies the message into buf (NUL-terminated, at most len bytes). */
int p3dfft_hip_last_error(char *buf, int len);
void p3dfft_hip_clear_error(void);

/* Asynchronous exec of a 3D plan on a stream (device pointers only).
 * idir < 0: no derivative.  Returns 0 on success. */
int p3dfft_exec_3Dtrans_double_async(int plan, double *in, double *out, int idir, int OW, void *stream);
int p3dfft_exec_3Dtrans_single_async(int plan, float *in, float *out, int idir, int OW, void *stream);

/* Exec of a 3D plan that allocates nothing (no hipMalloc, hipFree, stream
 * or event creation, no host heap): a variant of the *_async calls above
 * with no counterpart in the reference, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The precision is the plan's; in and out must be device
 * arrays; stream NULL is the null stream.  The intermediate arrays of the
 * plan's passes live in one fixed device array of the library,
 * P3DFFT_HIP_NOALLOC_SCRATCH_BYTES long, held by each device that runs this
 * call until the process ends; a plan needing more is refused with an
 * error.  Every call on one device uses the same array, so calls that
 * overlap on a device (other streams, other threads) race.  No timing or
 * TIMERS phases are recorded.  Returns 0, or -1 with the reason in
 * p3dfft_hip_last_error(). */
#define P3DFFT_HIP_NOALLOC_SCRATCH_BYTES (64 << 20)
int p3dfft_hip_exec_3Dtrans_noalloc(int plan, void *in, void *out, int idir, int OW, void *stream);

/* The library's own stream (created on first use). */
void *p3dfft_hip_stream(void);

/* Introspection of a 3D plan: number of passes, and a text description. */
int p3dfft_hip_plan_npasses(int plan);
int p3dfft_hip_plan_describe(int plan, char *buf, int len);
/* Bytes read and written by pass i of a 3D plan (for roofline accounting). */
int p3dfft_hip_pass_bytes(int plan, int i, long long *bytes_in, long long *bytes_out);

/* Per-pass timing of a 3D plan with HIP events recorded on the exec stream
 * around every pass launch.  p3dfft_hip_profile(plan, 1) starts (and
 * resets) collection; p3dfft_hip_profile_read waits for the recorded events
 * and writes the average duration in ms of each pass (up to cap entries)
 * and the summed duration of all recorded launches.  Returns the number of
 * recorded executions, or -1. */
int p3dfft_hip_profile(int plan, int enable);
int p3dfft_hip_profile_read(int plan, double *avg_ms, int cap, double *total_ms);

/* Plans made after this call use (1, default) or do not use (0) the
 * power-of-two pass kernel where it applies; 0 forces the generic kernel.
 * For testing both kernels on the same transforms. */
void p3dfft_hip_set_fast_path(int enable);

/* Batches of independent 3D transforms of possibly different sizes, all of
 * the same three 1D types and memory orders.  Item i has logical input dims
 * dims_in[3i..3i+2] and output dims dims_out[3i..3i+2].  Every stage of all
 * items runs in one launch of a grid-stride kernel over their tiles.
 * tune != 0 lets an autotuner pick the tile size and radix order of the
 * non-power-of-two stages (results cached in the file named by
 * P3DFFT_HIP_TUNE_CACHE, default $HOME/.cache/p3dfft3_hip_tune.txt).
 * Returns a batch handle, or -1. */
int p3dfft_hip_batch_plan(int nitems, const int *dims_in, const int *dims_out, const int types[3],
                          const int mo_in[3], const int mo_out[3], int tune);
/* in[i]/out[i]: device pointers of item i (in[i] == out[i] runs in place when
 * the stages allow it).  sync != 0 waits for completion. */
int p3dfft_hip_batch_exec(int bplan, void *const *in, void *const *out, void *stream, int sync);
/* enable != 0: the first exec captures the launches in a hipGraph, later
 * execs replay it (pointers may change between calls). */
int p3dfft_hip_batch_set_graph(int bplan, int enable);
int p3dfft_hip_batch_describe(int bplan, char *buf, int len);
int p3dfft_hip_batch_destroy(int bplan);

/* Periodic Poisson solves (domain [0, 2 pi)^3, grid nx x ny x nz of reals
 * in memory order (0,1,2), x stride-1): u = inverse FFT of -f_k / |k|^2, with
 * the k = 0 mode set to 0.  prec: 4 (FP32) or 8 (FP64).
 *
 * The service runs its own thread and several streams ("lanes", 4 unless
 * the environment variable P3DFFT_HIP_SERVE_LANES gives 1..8).  Requests
 * are admitted in arrival order when a slot (max_inflight) and their
 * working memory (from a device pool of pool_bytes allocated at creation)
 * are free.  An admitted request is launched at once on the least busy
 * lane: a grid of 2^18 points or more as its six pass launches, smaller
 * grids waiting together as one group of six step-kernel launches over all
 * their tiles.  Lanes run concurrently, so requests that arrive while others
 * run start as soon as a lane has room and their work overlaps the others'
 * on the device.  rhs and sol are device arrays that the caller keeps alive
 * until the request is polled as done; rhs must be complete when submitted
 * (e.g. the caller synchronized the stream that wrote it).  A request's
 * result does not depend on the other requests in flight.  Times are
 * seconds since the service was created. */
int p3dfft_hip_poisson_server_create(int prec, long long pool_bytes, int max_inflight);
/* Plan a grid size ahead of its first request (plans and tables stay
 * resident per size). */
int p3dfft_hip_poisson_server_prepare(int server, const int dims[3]);
/* Returns the request id, or -1 (e.g. a grid whose working memory exceeds
 * the pool). */
long long p3dfft_hip_poisson_submit(int server, const int dims[3], const void *rhs, void *sol);
/* Writes up to cap completed requests (id, submit/admit/done times, ok flag)
 * and returns their number; waits up to wait_ms for one if none is ready. */
int p3dfft_hip_poisson_poll(int server, long long *ids, double *t_submit, double *t_admit, double *t_done,
                            int *ok, int cap, int wait_ms);
/* Kernel launches made and tiles processed so far; returns 1 if the
 * service has failed, 0 if not, -1 for a bad handle. */
int p3dfft_hip_poisson_server_stats(int server, long long *launches, long long *tiles);
/* Requests waiting for admission, requests admitted and not yet done,
 * launch groups in flight and lanes (for diagnostics).  Returns 0, or -1
 * for a bad handle. */
int p3dfft_hip_poisson_server_state(int server, int *pending, int *holding, int *groups_busy, int *lanes);
/* Waits for every submitted request, then frees the service. */
int p3dfft_hip_poisson_server_destroy(int server);
/* One solve on the regular per-transform path (same plan math, one request
 * at a time); plans cached per size.  sync != 0 waits for the result. */
int p3dfft_hip_poisson_solve(int prec, const int dims[3], const void *rhs, void *sol, void *stream, int sync);

#ifdef __cplusplus
}
#endif

#endif /* P3DFFT_HIP_H */
