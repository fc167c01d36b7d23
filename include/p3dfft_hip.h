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

#ifdef __cplusplus
}
#endif

#endif /* P3DFFT_HIP_H */
