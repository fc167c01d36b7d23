/*
 * Cwrap.h -- C ABI of the MI355X-native P3DFFT++ transform path
 * (libp3dfft3_hip.so).  Included by p3dfft.h inside extern "C"; C programs
 * include p3dfft.h directly, as with the reference.
 *
 * Every declaration below replaces the same-named one of the reference's
 * include/Cwrap.h (line numbers cite that file; implementations in the
 * reference are in build/wrap.C and build/init.C):
 *
 *   P3DFFT_* type-ID globals          include/p3dfft.h:105-120, build/init.C:84-89
 *   struct CDataGrid_struct (Grid)    Cwrap.h:82-98
 *   p3dfft_setup / p3dfft_cleanup     Cwrap.h:101-102, build/wrap.C:90-142
 *   p3dfft_init_3Dtype                Cwrap.h:103, build/wrap.C:148-154
 *   p3dfft_plan_1Dtrans               Cwrap.h:104, build/wrap.C:235-277
 *   p3dfft_plan_3Dtrans               Cwrap.h:105, build/wrap.C:157-233
 *   p3dfft_init_proc_grid             Cwrap.h:107, build/wrap.C:296-301
 *   p3dfft_init_data_grid             Cwrap.h:108, build/wrap.C:303-324
 *   p3dfft_free_data_grid             Cwrap.h:110, build/wrap.C:334-337
 *   p3dfft_free_proc_grid             Cwrap.h:111, build/wrap.C:339-344
 *   p3dfft_inv_mo                     Cwrap.h:112, build/wrap.C:346-348
 *   p3dfft_exec_1Dtrans_{double,single}   Cwrap.h:114-115, build/wrap.C:484-549
 *   p3dfft_exec_3Dtrans_{double,single}   Cwrap.h:116-117, build/wrap.C:354-417
 *   p3dfft_exec_3Dderiv_{double,single}   Cwrap.h:118-119, build/wrap.C:419-482
 *   p3dfft_compute_deriv_{single,double}  Cwrap.h:120-121, build/wrap.C:551-569
 *
 * Conventions kept from the reference: handles are int indices; Grid* is
 * allocated by the library and released by p3dfft_free_data_grid; complex
 * data is interleaved (re,im); idir is 0-based; nothing is normalized;
 * forward sign is -1.  Differences: exec returns after the result is in
 * `out` (synchronous), arrays may be device or host pointers, and errors
 * are also recorded for p3dfft_hip_last_error() (p3dfft_hip.h).
 * The Fortran-side *_f entry points are declared in Fwrap.h.
 *
 * Ordering: the synchronous calls run on the library's own stream after
 * everything already queued on the null (legacy default) stream, so a
 * kernel that fills `in` there needs no synchronization first; work on
 * other streams must be complete (or ordered by the caller) before the call.
 */
#ifndef P3DFFT_CWRAP_H
#define P3DFFT_CWRAP_H

extern int P3DFFT_EMPTY_TYPE_SINGLE, P3DFFT_EMPTY_TYPE_DOUBLE, P3DFFT_EMPTY_TYPE_SINGLE_COMPLEX,
    P3DFFT_EMPTY_TYPE_DOUBLE_COMPLEX;
extern int P3DFFT_R2CFFT_S, P3DFFT_R2CFFT_D, P3DFFT_C2RFFT_S, P3DFFT_C2RFFT_D,
    P3DFFT_CFFT_FORWARD_S, P3DFFT_CFFT_FORWARD_D, P3DFFT_CFFT_BACKWARD_S, P3DFFT_CFFT_BACKWARD_D;
extern int P3DFFT_DCT1_REAL_S, P3DFFT_DCT1_REAL_D, P3DFFT_DST1_REAL_S, P3DFFT_DST1_REAL_D,
    P3DFFT_DCT1_COMPLEX_S, P3DFFT_DCT1_COMPLEX_D, P3DFFT_DST1_COMPLEX_S, P3DFFT_DST1_COMPLEX_D;
extern int P3DFFT_DCT2_REAL_S, P3DFFT_DCT2_REAL_D, P3DFFT_DST2_REAL_S, P3DFFT_DST2_REAL_D,
    P3DFFT_DCT2_COMPLEX_S, P3DFFT_DCT2_COMPLEX_D, P3DFFT_DST2_COMPLEX_S, P3DFFT_DST2_COMPLEX_D;
extern int P3DFFT_DCT3_REAL_S, P3DFFT_DCT3_REAL_D, P3DFFT_DST3_REAL_S, P3DFFT_DST3_REAL_D,
    P3DFFT_DCT3_COMPLEX_S, P3DFFT_DCT3_COMPLEX_D, P3DFFT_DST3_COMPLEX_S, P3DFFT_DST3_COMPLEX_D;
extern int P3DFFT_DCT4_REAL_S, P3DFFT_DCT4_REAL_D, P3DFFT_DST4_REAL_S, P3DFFT_DST4_REAL_D,
    P3DFFT_DCT4_COMPLEX_S, P3DFFT_DCT4_COMPLEX_D, P3DFFT_DST4_COMPLEX_S, P3DFFT_DST4_COMPLEX_D;

struct CDataGrid_struct {
  int nd;
  int Gdims[3];
  int dim_conj_sym;
  int MemOrder[3];
  int Ldims[3];
  int Dmap[3];
  int pgrid;
  int grid_id[3];
  int GlobStart[3];
  int taskid, numtasks;
  int ProcDims[3];
  MPI_Comm mpi_comm_glob;
};

typedef struct CDataGrid_struct CDataGrid;
typedef struct CDataGrid_struct Grid;

void p3dfft_setup();
void p3dfft_cleanup();
Type3D p3dfft_init_3Dtype(int[3]);
int p3dfft_plan_1Dtrans(Grid *, Grid *, int, int);
Plan3D p3dfft_plan_3Dtrans(Grid *, Grid *, Type3D);
/* Cwrap.h:106, build/wrap.C:279-294: handle of the data grid made by
 * p3dfft_init_data_grid_f with the same global dims, processor grid, Dmap
 * and MemOrder, or -1. */
int find_grid(int[3], int, int *, int[3]);
int p3dfft_init_proc_grid(int pdims[3], MPI_Comm comm);
Grid *p3dfft_init_data_grid(int[3], int, int, int[3], int[3]);
void p3dfft_free_data_grid(Grid *gr);
void p3dfft_free_proc_grid(int pgrid);
void p3dfft_inv_mo(int[3], int[3]);
void p3dfft_exec_1Dtrans_double(int, double *, double *, int);
void p3dfft_exec_1Dtrans_single(int, float *, float *, int);
void p3dfft_exec_3Dtrans_double(Plan3D, double *, double *, int);
void p3dfft_exec_3Dtrans_single(Plan3D, float *, float *, int);
void p3dfft_exec_3Dderiv_double(Plan3D, double *, double *, int, int);
void p3dfft_exec_3Dderiv_single(Plan3D, float *, float *, int, int);
void p3dfft_compute_deriv_single(float *, float *, Grid *, int);
void p3dfft_compute_deriv_double(double *, double *, Grid *, int);

#endif /* P3DFFT_CWRAP_H */
