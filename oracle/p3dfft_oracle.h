/*
 * p3dfft_oracle.h -- CPU FP64 restatement of the P3DFFT++ transform3D path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product library (libp3dfft3_hip.so).  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load it.  The product never links,
 * loads or calls it.
 *
 * What it restates (reference files cited per function in the .c file):
 *   - the 44-entry 1D transform type registry   build/init.C:121-753
 *   - the 1x1 planner (find_order, swap0, excl, final_seq, stage chain)
 *                                               build/templ.C:91-435,627-882
 *   - FFTW's DFT / R2C / C2R / R2R conventions   (FFTW 3 manual definitions,
 *     used via build/init.C:1146-1607)
 *   - the stage-wise derivative multiply         build/exec.C:228-287
 *   - the standalone compute_deriv               build/deriv.C:85-186
 *   - memory-order helpers inv_mo / rel_change   build/exec.C:2270-2296,2959-2967
 *
 * Pinning (see tests/test_oracle.py): FFTW reference r2r vectors shipped with
 * scipy, numpy pocketfft fixtures, long-double O(N^2) definitions, and the
 * reference samples' closed-form known answers.
 */
#ifndef P3DFFT_ORACLE_H
#define P3DFFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* r2r kinds, FFTW names */
enum {
  OR_REDFT00 = 0, OR_REDFT10 = 1, OR_REDFT01 = 2, OR_REDFT11 = 3,
  OR_RODFT00 = 4, OR_RODFT10 = 5, OR_RODFT01 = 6, OR_RODFT11 = 7
};

/* transform classes of the type registry */
enum {
  OR_EMPTY = 0, OR_R2C = 1, OR_C2R = 2, OR_C2C = 3, OR_R2R = 4
};

typedef struct {
  int cls;      /* OR_EMPTY .. OR_R2R */
  int dt1, dt2; /* 1 = real, 2 = complex */
  int prec;     /* 4 or 8 */
  int isign;    /* -1 forward, +1 backward (C2C) */
  int r2r_kind; /* OR_REDFT00.. for OR_R2R */
} or_type_t;

int or_num_types(void);
/* Fill *t for type ID id (registration order).  Returns 0, or -1 if invalid. */
int or_type_info(int id, or_type_t *t);

/* -------- 1D kernels (interleaved complex as double[2]) -------- */
void or_dft_naive(int n, const double *in, double *out, int sign);
void or_fft(int n, const double *in, double *out, int sign);
void or_r2c(int n, const double *in, double *out);        /* out: n/2+1 complex */
void or_c2r(int n, const double *in, double *out);        /* in: n/2+1 complex */
void or_r2r_naive(int kind, int n, const double *in, double *out);
void or_r2r(int kind, int n, const double *in, double *out);

/* -------- planner restatement (1x1 processor grid) -------- */
void or_inv_mo(const int mo[3], int imo[3]);
void or_rel_change(const int *mo1, const int *mo2, int *mc);
void or_swap0(int newmo[3], const int mo[3], int L);
int  or_excl(int a, int b);
/* find_order for a 1x1 grid.  mo1/mo2: grid1/grid2 MemOrder. */
void or_find_order(const int types[3], const int mo1[3], const int mo2[3], int L[3]);

typedef struct {
  int trans_dim;  /* logical axis transformed (or carried, for EMPTY) */
  int type_id;
  int mo_in[3], mo_out[3];
  int dims_in[3], dims_out[3];
} or_stage_t;

/* Stage chain of transform3D on a 1x1 grid.  gdims1: grid1 global dims.
 * Returns the number of stages (3 or 4). */
int or_plan3d(const int types[3], const int gdims1[3], const int mo1[3],
              const int mo2[3], or_stage_t stages[4]);

/* -------- execution -------- */
/* One stage / one 1D transplan: transform along logical axis d, reading a
 * (dims_in, mo_in) array and writing a (dims_out, mo_out) array.  Data is
 * double for real, interleaved double pairs for complex.  deriv != 0
 * multiplies by i*k after the transform (compute_deriv_loc semantics). */
int or_exec_1d(int type_id, int d, const int dims_in[3], const int mo_in[3],
               const int dims_out[3], const int mo_out[3],
               const double *in, double *out, int deriv);

/* Full transform3D::exec_deriv on a 1x1 grid (idir < 0: plain exec).
 * Output dims follow the chain (R2C halves the conj-sym axis). */
int or_exec3d(const int types[3], const int gdims1[3], const int mo1[3],
              const int mo2[3], const double *in, double *out, int idir);

/* Standalone compute_deriv on a stored complex spectrum (build/deriv.C). */
void or_compute_deriv(const double *in, double *out, const int gdims[3],
                      const int ldims[3], const int glob_start[3],
                      const int mem_order[3], int dim_conj_sym, int idir);

#ifdef __cplusplus
}
#endif
#endif
