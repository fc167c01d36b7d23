/*
 * Fwrap.h -- the Fortran-side C entry points of libp3dfft3_hip.so.
 * Included by p3dfft.h inside extern "C", after Cwrap.h, as in the
 * reference.
 *
 * Each declaration replaces the same-named one of the reference's
 * include/Fwrap.h (line numbers cite that file; the reference implements
 * them in build/wrap.C:575-792).  They are the functions the reference's
 * Fortran module binds with BIND(C): every argument is passed by address,
 * data grids are int handles (p3dfft_init_data_grid_f returns the handle of
 * an equal grid made earlier, see find_grid in Cwrap.h), and idir is
 * 1-based.  MPI communicators are Fortran handles; with MPICH (the MPI the
 * library's headers are built against) they equal the C handles.
 *
 *   p3dfft_init_3Dtype_f              Fwrap.h:81, build/wrap.C:575-582
 *   p3dfft_plan_1Dtrans_f             Fwrap.h:82, build/wrap.C:584-621
 *   p3dfft_plan_3Dtrans_f             Fwrap.h:83, build/wrap.C:636-716
 *   p3dfft_init_proc_grid_f           Fwrap.h:84, build/wrap.C:718-723
 *   p3dfft_init_data_grid_f           Fwrap.h:85, build/wrap.C:725-762
 *   p3dfft_exec_1Dtrans_{double,single}_f   Fwrap.h:86-87, build/wrap.C:783-791
 *   p3dfft_exec_3Dtrans_{double,single}_f   Fwrap.h:88-89, build/wrap.C:764-772
 *   p3dfft_exec_3Dderiv_{double,single}_f   Fwrap.h:90-91, build/wrap.C:773-781
 *   p3dfft_compute_deriv_{single,double}_f  Fwrap.h:92-93, build/wrap.C:624-634
 *
 * The Fortran module itself (build/fp3dfft++mod.f90) needs a Fortran
 * compiler and is not part of this library.
 */
#ifndef P3DFFT_FWRAP_H
#define P3DFFT_FWRAP_H

void p3dfft_init_3Dtype_f(int *, int[3]);
void p3dfft_plan_1Dtrans_f(int *, int *, int *, int *, int *);
void p3dfft_plan_3Dtrans_f(int *, int *, int *, Type3D *);
int p3dfft_init_proc_grid_f(int *pdims, int *mpicomm);
void p3dfft_init_data_grid_f(int *mygrid, int *ldims, int *glob_start, int *gdims, int *dim_conj_sym, int *pgrid_id,
                             int *dmap, int *mem_order);
void p3dfft_exec_1Dtrans_double_f(int *, double *, double *, int *);
void p3dfft_exec_1Dtrans_single_f(int *, float *, float *, int *);
void p3dfft_exec_3Dtrans_double_f(Plan3D *, double *, double *, int *);
void p3dfft_exec_3Dtrans_single_f(Plan3D *, float *, float *, int *);
void p3dfft_exec_3Dderiv_double_f(Plan3D *, double *, double *, int *, int *);
void p3dfft_exec_3Dderiv_single_f(Plan3D *, float *, float *, int *, int *);
void p3dfft_compute_deriv_single_f(float *, float *, int *, int *);
void p3dfft_compute_deriv_double_f(double *, double *, int *, int *);

#endif /* P3DFFT_FWRAP_H */
