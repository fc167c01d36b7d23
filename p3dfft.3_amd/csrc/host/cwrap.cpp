// C ABI of libp3dfft3_hip.so (include/Cwrap.h).  Mirrors build/wrap.C:
// integer handles into library-owned tables, Grid structs allocated by the
// library, dispatch on (precision, input type, output type).
#include <stdio.h>
#include <string.h>

#include "../../../include/p3dfft.h"
#include "engine.h"

int P3DFFT_EMPTY_TYPE_SINGLE, P3DFFT_EMPTY_TYPE_DOUBLE, P3DFFT_EMPTY_TYPE_SINGLE_COMPLEX,
    P3DFFT_EMPTY_TYPE_DOUBLE_COMPLEX;
int P3DFFT_R2CFFT_S, P3DFFT_R2CFFT_D, P3DFFT_C2RFFT_S, P3DFFT_C2RFFT_D, P3DFFT_CFFT_FORWARD_S,
    P3DFFT_CFFT_FORWARD_D, P3DFFT_CFFT_BACKWARD_S, P3DFFT_CFFT_BACKWARD_D;
int P3DFFT_DCT1_REAL_S, P3DFFT_DCT1_REAL_D, P3DFFT_DST1_REAL_S, P3DFFT_DST1_REAL_D, P3DFFT_DCT1_COMPLEX_S,
    P3DFFT_DCT1_COMPLEX_D, P3DFFT_DST1_COMPLEX_S, P3DFFT_DST1_COMPLEX_D;
int P3DFFT_DCT2_REAL_S, P3DFFT_DCT2_REAL_D, P3DFFT_DST2_REAL_S, P3DFFT_DST2_REAL_D, P3DFFT_DCT2_COMPLEX_S,
    P3DFFT_DCT2_COMPLEX_D, P3DFFT_DST2_COMPLEX_S, P3DFFT_DST2_COMPLEX_D;
int P3DFFT_DCT3_REAL_S, P3DFFT_DCT3_REAL_D, P3DFFT_DST3_REAL_S, P3DFFT_DST3_REAL_D, P3DFFT_DCT3_COMPLEX_S,
    P3DFFT_DCT3_COMPLEX_D, P3DFFT_DST3_COMPLEX_S, P3DFFT_DST3_COMPLEX_D;
int P3DFFT_DCT4_REAL_S, P3DFFT_DCT4_REAL_D, P3DFFT_DST4_REAL_S, P3DFFT_DST4_REAL_D, P3DFFT_DCT4_COMPLEX_S,
    P3DFFT_DCT4_COMPLEX_D, P3DFFT_DST4_COMPLEX_S, P3DFFT_DST4_COMPLEX_D;

using namespace p3dfft;

static DataGrid *grid_from_c(const Grid *C) {
  if (!C || C->pgrid < 0 || C->pgrid >= (int)stored_proc_grids.size()) {
    p3e::set_error("invalid Grid or processor grid handle");
    return nullptr;
  }
  ProcGrid *pg = stored_proc_grids[C->pgrid];
  return new DataGrid(const_cast<int *>(C->Gdims), C->dim_conj_sym, pg, const_cast<int *>(C->Dmap),
                      const_cast<int *>(C->MemOrder));
}

template <class R, class C>
static void exec3d(gen_transform3D *g, R *in, R *out, int idir, int OW) {
  if (g->dt1 == 1) {
    if (g->dt2 == 1) ((transform3D<R, R> *)g)->exec_deriv(in, out, idir, OW);
    else ((transform3D<R, C> *)g)->exec_deriv(in, (C *)out, idir, OW);
  } else {
    if (g->dt2 == 1) ((transform3D<C, R> *)g)->exec_deriv((C *)in, out, idir, OW);
    else ((transform3D<C, C> *)g)->exec_deriv((C *)in, (C *)out, idir, OW);
  }
}

template <class R, class C> static void exec1d(stage *s, R *in, R *out, int OW) {
  if (s->dt1 == 1) {
    if (s->dt2 == 1) ((transplan<R, R> *)s)->exec((char *)in, (char *)out, OW);
    else ((transplan<R, C> *)s)->exec((char *)in, (char *)out, OW);
  } else {
    if (s->dt2 == 1) ((transplan<C, R> *)s)->exec((char *)in, (char *)out, OW);
    else ((transplan<C, C> *)s)->exec((char *)in, (char *)out, OW);
  }
}

// transform3D<T1,T2> for a 3D type, T1/T2 from the precision and the data
// types of the first and last stage (build/wrap.C:157-233); returns the new
// handle in stored_trans3D, or -1.
static int new_plan3d(DataGrid *gr1, DataGrid *gr2, int tp, const char *who) {
  trans_type3D *type3D = &types3D[tp];
  if (!type3D->is_set) {
    p3e::set_error("%s: 3D type %d is not set", who, tp);
    return -1;
  }
  int L[3];
  find_order(L, type3D, gr1, gr2, nullptr);
  int dt1 = types1D[type3D->types[L[0]]]->dt1;
  int dt2 = types1D[type3D->types[L[2]]]->dt2;
  gen_transform3D *tr;
  if (type3D->prec == 4) {
    if (dt1 == 1)
      tr = (dt2 == 1) ? (gen_transform3D *)new transform3D<float, float>(*gr1, *gr2, type3D)
                      : (gen_transform3D *)new transform3D<float, mycomplex>(*gr1, *gr2, type3D);
    else
      tr = (dt2 == 1) ? (gen_transform3D *)new transform3D<mycomplex, float>(*gr1, *gr2, type3D)
                      : (gen_transform3D *)new transform3D<mycomplex, mycomplex>(*gr1, *gr2, type3D);
  } else {
    if (dt1 == 1)
      tr = (dt2 == 1) ? (gen_transform3D *)new transform3D<double, double>(*gr1, *gr2, type3D)
                      : (gen_transform3D *)new transform3D<double, complex_double>(*gr1, *gr2, type3D);
    else
      tr = (dt2 == 1) ? (gen_transform3D *)new transform3D<complex_double, double>(*gr1, *gr2, type3D)
                      : (gen_transform3D *)new transform3D<complex_double, complex_double>(*gr1, *gr2, type3D);
  }
  int count = (int)stored_trans3D.size();
  stored_trans3D.push_back(tr);
  return count;
}

// transplan<T1,T2> for a 1D type (build/wrap.C:235-277); returns the new
// handle in stored_trans1D.
static int new_plan1d(DataGrid *gr1, DataGrid *gr2, int type_ID, int d) {
  gen_trans_type *tp = types1D[type_ID];
  stage *tr;
  if (tp->prec == 4) {
    if (tp->dt1 == 1)
      tr = (tp->dt2 == 1) ? (stage *)new transplan<float, float>(*gr1, *gr2, tp, d)
                          : (stage *)new transplan<float, mycomplex>(*gr1, *gr2, tp, d);
    else
      tr = (tp->dt2 == 1) ? (stage *)new transplan<mycomplex, float>(*gr1, *gr2, tp, d)
                          : (stage *)new transplan<mycomplex, mycomplex>(*gr1, *gr2, tp, d);
  } else {
    if (tp->dt1 == 1)
      tr = (tp->dt2 == 1) ? (stage *)new transplan<double, double>(*gr1, *gr2, tp, d)
                          : (stage *)new transplan<double, complex_double>(*gr1, *gr2, tp, d);
    else
      tr = (tp->dt2 == 1) ? (stage *)new transplan<complex_double, double>(*gr1, *gr2, tp, d)
                          : (stage *)new transplan<complex_double, complex_double>(*gr1, *gr2, tp, d);
  }
  int count = (int)stored_trans1D.size();
  stored_trans1D.push_back(tr);
  return count;
}

extern "C" {

void p3dfft_setup() {
  p3dfft::setup();
  P3DFFT_EMPTY_TYPE_SINGLE = EMPTY_TYPE_SINGLE;
  P3DFFT_EMPTY_TYPE_DOUBLE = EMPTY_TYPE_DOUBLE;
  P3DFFT_EMPTY_TYPE_SINGLE_COMPLEX = EMPTY_TYPE_SINGLE_COMPLEX;
  P3DFFT_EMPTY_TYPE_DOUBLE_COMPLEX = EMPTY_TYPE_DOUBLE_COMPLEX;
  P3DFFT_R2CFFT_S = R2CFFT_S;
  P3DFFT_R2CFFT_D = R2CFFT_D;
  P3DFFT_C2RFFT_S = C2RFFT_S;
  P3DFFT_C2RFFT_D = C2RFFT_D;
  P3DFFT_CFFT_FORWARD_S = CFFT_FORWARD_S;
  P3DFFT_CFFT_FORWARD_D = CFFT_FORWARD_D;
  P3DFFT_CFFT_BACKWARD_S = CFFT_BACKWARD_S;
  P3DFFT_CFFT_BACKWARD_D = CFFT_BACKWARD_D;
  P3DFFT_DCT1_REAL_S = DCT1_REAL_S;
  P3DFFT_DCT1_REAL_D = DCT1_REAL_D;
  P3DFFT_DST1_REAL_S = DST1_REAL_S;
  P3DFFT_DST1_REAL_D = DST1_REAL_D;
  P3DFFT_DCT2_REAL_S = DCT2_REAL_S;
  P3DFFT_DCT2_REAL_D = DCT2_REAL_D;
  P3DFFT_DST2_REAL_S = DST2_REAL_S;
  P3DFFT_DST2_REAL_D = DST2_REAL_D;
  P3DFFT_DCT3_REAL_S = DCT3_REAL_S;
  P3DFFT_DCT3_REAL_D = DCT3_REAL_D;
  P3DFFT_DST3_REAL_S = DST3_REAL_S;
  P3DFFT_DST3_REAL_D = DST3_REAL_D;
  P3DFFT_DCT4_REAL_S = DCT4_REAL_S;
  P3DFFT_DCT4_REAL_D = DCT4_REAL_D;
  P3DFFT_DST4_REAL_S = DST4_REAL_S;
  P3DFFT_DST4_REAL_D = DST4_REAL_D;
  P3DFFT_DCT1_COMPLEX_S = DCT1_COMPLEX_S;
  P3DFFT_DCT1_COMPLEX_D = DCT1_COMPLEX_D;
  P3DFFT_DST1_COMPLEX_S = DST1_COMPLEX_S;
  P3DFFT_DST1_COMPLEX_D = DST1_COMPLEX_D;
  P3DFFT_DCT2_COMPLEX_S = DCT2_COMPLEX_S;
  P3DFFT_DCT2_COMPLEX_D = DCT2_COMPLEX_D;
  P3DFFT_DST2_COMPLEX_S = DST2_COMPLEX_S;
  P3DFFT_DST2_COMPLEX_D = DST2_COMPLEX_D;
  P3DFFT_DCT3_COMPLEX_S = DCT3_COMPLEX_S;
  P3DFFT_DCT3_COMPLEX_D = DCT3_COMPLEX_D;
  P3DFFT_DST3_COMPLEX_S = DST3_COMPLEX_S;
  P3DFFT_DST3_COMPLEX_D = DST3_COMPLEX_D;
  P3DFFT_DCT4_COMPLEX_S = DCT4_COMPLEX_S;
  P3DFFT_DCT4_COMPLEX_D = DCT4_COMPLEX_D;
  P3DFFT_DST4_COMPLEX_S = DST4_COMPLEX_S;
  P3DFFT_DST4_COMPLEX_D = DST4_COMPLEX_D;
}

void p3dfft_cleanup() { p3dfft::cleanup(); }

Type3D p3dfft_init_3Dtype(int types[3]) {
  trans_type3D tp(types);
  int count = (int)types3D.size();
  types3D.push_back(tp);
  return count;
}

Plan3D p3dfft_plan_3Dtrans(Grid *Cgr1, Grid *Cgr2, Type3D tp) {
  if (tp < 0 || tp >= (int)types3D.size()) {
    p3e::set_error("p3dfft_plan_3Dtrans: invalid 3D type handle %d", tp);
    return -1;
  }
  DataGrid *gr1 = grid_from_c(Cgr1);
  DataGrid *gr2 = grid_from_c(Cgr2);
  const int h = (gr1 && gr2) ? new_plan3d(gr1, gr2, tp, "p3dfft_plan_3Dtrans") : -1;
  delete gr1;
  delete gr2;
  return h;
}

int p3dfft_plan_1Dtrans(Grid *Cgr1, Grid *Cgr2, int type_ID, int d) {
  if (type_ID < 0 || type_ID >= (int)types1D.size()) {
    p3e::set_error("p3dfft_plan_1Dtrans: invalid type ID %d", type_ID);
    return -1;
  }
  DataGrid *gr1 = grid_from_c(Cgr1);
  DataGrid *gr2 = grid_from_c(Cgr2);
  const int h = (gr1 && gr2) ? new_plan1d(gr1, gr2, type_ID, d) : -1;
  delete gr1;
  delete gr2;
  return h;
}

int p3dfft_init_proc_grid(int pdims[3], MPI_Comm mpicomm) {
  ProcGrid *pgrid = new ProcGrid(pdims, mpicomm);
  stored_proc_grids.push_back(pgrid);
  return (int)stored_proc_grids.size() - 1;
}

Grid *p3dfft_init_data_grid(int gdims[3], int dim_conj_sym, int pgrid_id, int dmap[3], int mem_order[3]) {
  if (pgrid_id < 0 || pgrid_id >= (int)stored_proc_grids.size()) {
    p3e::set_error("p3dfft_init_data_grid: invalid processor grid %d", pgrid_id);
    return nullptr;
  }
  ProcGrid *pgrid = stored_proc_grids[pgrid_id];
  DataGrid gr(gdims, dim_conj_sym, pgrid, dmap, mem_order);
  Grid *C = new Grid;
  memset(C, 0, sizeof(Grid));
  C->numtasks = pgrid->numtasks;
  C->taskid = pgrid->taskid;
  C->nd = gr.nd;
  C->pgrid = pgrid_id;
  C->dim_conj_sym = dim_conj_sym;
  memcpy(C->Gdims, gdims, 3 * sizeof(int));
  memcpy(C->Ldims, gr.Ldims, 3 * sizeof(int));
  memcpy(C->MemOrder, mem_order, 3 * sizeof(int));
  memcpy(C->Dmap, dmap, 3 * sizeof(int));
  memcpy(C->ProcDims, pgrid->ProcDims, 3 * sizeof(int));
  memcpy(C->grid_id, gr.grid_id, 3 * sizeof(int));
  memcpy(C->GlobStart, gr.GlobStart, 3 * sizeof(int));
  C->mpi_comm_glob = pgrid->mpi_comm_glob;
  return C;
}

void p3dfft_free_data_grid(Grid *gr) { delete gr; }

// Erases the entry, which shifts later handles, as build/wrap.C:339-344 does.
void p3dfft_free_proc_grid(int pgrid_id) {
  if (pgrid_id < 0 || pgrid_id >= (int)stored_proc_grids.size()) return;
  delete stored_proc_grids[pgrid_id];
  stored_proc_grids.erase(stored_proc_grids.begin() + pgrid_id);
}

void p3dfft_inv_mo(int mo[3], int imo[3]) { p3dfft::inv_mo(mo, imo); }

static gen_transform3D *get3d(Plan3D plan, int prec, const char *who) {
  if (plan < 0 || plan >= (int)stored_trans3D.size() || !stored_trans3D[plan]) {
    p3e::set_error("%s: invalid plan handle %d", who, plan);
    return nullptr;
  }
  gen_transform3D *t = stored_trans3D[plan];
  if (t->prec != prec) {
    p3e::set_error("%s: expecting %s precision data", who, prec == 8 ? "double" : "single");
    return nullptr;
  }
  return t;
}

void p3dfft_exec_3Dtrans_single(Plan3D plan, float *in, float *out, int OW) {
  gen_transform3D *g = get3d(plan, 4, "p3dfft_exec_3Dtrans_single");
  if (g) exec3d<float, mycomplex>(g, in, out, -1, OW);
}

void p3dfft_exec_3Dtrans_double(Plan3D plan, double *in, double *out, int OW) {
  gen_transform3D *g = get3d(plan, 8, "p3dfft_exec_3Dtrans_double");
  if (g) exec3d<double, complex_double>(g, in, out, -1, OW);
}

void p3dfft_exec_3Dderiv_single(Plan3D plan, float *in, float *out, int idir, int OW) {
  gen_transform3D *g = get3d(plan, 4, "p3dfft_exec_3Dderiv_single");
  if (g) exec3d<float, mycomplex>(g, in, out, idir, OW);
}

void p3dfft_exec_3Dderiv_double(Plan3D plan, double *in, double *out, int idir, int OW) {
  gen_transform3D *g = get3d(plan, 8, "p3dfft_exec_3Dderiv_double");
  if (g) exec3d<double, complex_double>(g, in, out, idir, OW);
}

static stage *get1d(int plan, int prec, const char *who) {
  if (plan < 0 || plan >= (int)stored_trans1D.size() || !stored_trans1D[plan]) {
    p3e::set_error("%s: invalid plan handle %d", who, plan);
    return nullptr;
  }
  stage *s = stored_trans1D[plan];
  if (s->stage_prec != prec) {
    p3e::set_error("%s: expecting %s precision data", who, prec == 8 ? "double" : "single");
    return nullptr;
  }
  return s;
}

void p3dfft_exec_1Dtrans_double(int plan, double *in, double *out, int OW) {
  stage *s = get1d(plan, 8, "p3dfft_exec_1Dtrans_double");
  if (s) exec1d<double, complex_double>(s, in, out, OW);
}

void p3dfft_exec_1Dtrans_single(int plan, float *in, float *out, int OW) {
  stage *s = get1d(plan, 4, "p3dfft_exec_1Dtrans_single");
  if (s) exec1d<float, mycomplex>(s, in, out, OW);
}

void p3dfft_compute_deriv_single(float *in, float *out, Grid *Cgrid, int idir) {
  DataGrid *g = grid_from_c(Cgrid);
  if (!g) return;
  compute_deriv<mycomplex>((mycomplex *)in, (mycomplex *)out, g, idir);
  delete g;
}

void p3dfft_compute_deriv_double(double *in, double *out, Grid *Cgrid, int idir) {
  DataGrid *g = grid_from_c(Cgrid);
  if (!g) return;
  compute_deriv<complex_double>((complex_double *)in, (complex_double *)out, g, idir);
  delete g;
}

// ------------------------------------------------------------------ Fortran-side entry points
// build/wrap.C:575-791 (declared in the reference's include/Fwrap.h:81-93).
// Every argument is passed by address; data grids are int handles into
// stored_data_grids (made by p3dfft_init_data_grid_f, which reuses an equal
// grid: find_grid); idir is 1-based.  MPI communicators are Fortran
// handles, which for MPICH equal the C handles (MPI_Comm_f2c is the
// identity), so they are carried as they are.

// build/wrap.C:279-294: the stored data grid with these global dims, an
// equal processor grid, Dmap and MemOrder, or -1.  dim_conj_sym is not part
// of the key, as in the reference.
int find_grid(int gdims[3], int pgrid, int *dmap, int mem_order[3]) {
  if (pgrid < 0 || pgrid >= (int)stored_proc_grids.size()) return -1;
  const ProcGrid *pg = stored_proc_grids[pgrid];
  for (size_t i = 0; i < stored_data_grids.size(); i++) {
    const DataGrid *g = stored_data_grids[i];
    bool same = g->Pgrid && *g->Pgrid == *pg;
    for (int a = 0; a < 3 && same; a++)
      same = g->Gdims[a] == gdims[a] && g->Dmap[a] == dmap[a] && g->MemOrder[a] == mem_order[a];
    if (same) return (int)i;
  }
  return -1;
}

static DataGrid *stored_grid(const int *h, const char *who) {
  if (!h || *h < 0 || *h >= (int)stored_data_grids.size() || !stored_data_grids[*h]) {
    p3e::set_error("%s: invalid data grid handle %d", who, h ? *h : -1);
    return nullptr;
  }
  return stored_data_grids[*h];
}

void p3dfft_init_3Dtype_f(int *type, int types[3]) { *type = p3dfft_init_3Dtype(types); }

void p3dfft_plan_1Dtrans_f(int *plan, int *Fgr1, int *Fgr2, int *type_ID, int *d) {
  *plan = -1;
  DataGrid *gr1 = stored_grid(Fgr1, "p3dfft_plan_1Dtrans_f");
  DataGrid *gr2 = stored_grid(Fgr2, "p3dfft_plan_1Dtrans_f");
  if (!gr1 || !gr2) return;
  if (*type_ID < 0 || *type_ID >= (int)types1D.size()) {
    p3e::set_error("p3dfft_plan_1Dtrans_f: invalid type ID %d", *type_ID);
    return;
  }
  *plan = new_plan1d(gr1, gr2, *type_ID, *d);
}

void p3dfft_plan_3Dtrans_f(Plan3D *plan, int *Fgr1, int *Fgr2, Type3D *tp) {
  *plan = -1;
  DataGrid *gr1 = stored_grid(Fgr1, "p3dfft_plan_3Dtrans_f");
  DataGrid *gr2 = stored_grid(Fgr2, "p3dfft_plan_3Dtrans_f");
  if (!gr1 || !gr2) return;
  if (*tp < 0 || *tp >= (int)types3D.size()) {
    p3e::set_error("p3dfft_plan_3Dtrans_f: invalid 3D type handle %d", *tp);
    return;
  }
  *plan = new_plan3d(gr1, gr2, *tp, "p3dfft_plan_3Dtrans_f");
}

int p3dfft_init_proc_grid_f(int *pdims, int *mpicomm) { return p3dfft_init_proc_grid(pdims, (MPI_Comm)*mpicomm); }

void p3dfft_init_data_grid_f(int *mygrid, int *ldims, int *glob_start, int *gdims, int *dim_conj_sym, int *pgrid_id,
                             int *dmap, int *mem_order) {
  *mygrid = -1;
  if (*pgrid_id < 0 || *pgrid_id >= (int)stored_proc_grids.size()) {
    p3e::set_error("p3dfft_init_data_grid_f: invalid processor grid %d", *pgrid_id);
    return;
  }
  int num = find_grid(gdims, *pgrid_id, dmap, mem_order);
  if (num < 0) {
    stored_data_grids.push_back(new DataGrid(gdims, *dim_conj_sym, stored_proc_grids[*pgrid_id], dmap, mem_order));
    num = (int)stored_data_grids.size() - 1;
  }
  const DataGrid *g = stored_data_grids[num];
  memcpy(ldims, g->Ldims, 3 * sizeof(int));
  memcpy(glob_start, g->GlobStart, 3 * sizeof(int));
  *mygrid = num;
}

void p3dfft_exec_1Dtrans_double_f(int *plan, double *in, double *out, int *OW) {
  p3dfft_exec_1Dtrans_double(*plan, in, out, *OW);
}

void p3dfft_exec_1Dtrans_single_f(int *plan, float *in, float *out, int *OW) {
  p3dfft_exec_1Dtrans_single(*plan, in, out, *OW);
}

void p3dfft_exec_3Dtrans_double_f(Plan3D *plan, double *in, double *out, int *OW) {
  p3dfft_exec_3Dtrans_double(*plan, in, out, *OW);
}

void p3dfft_exec_3Dtrans_single_f(Plan3D *plan, float *in, float *out, int *OW) {
  p3dfft_exec_3Dtrans_single(*plan, in, out, *OW);
}

void p3dfft_exec_3Dderiv_double_f(Plan3D *plan, double *in, double *out, int *idir, int *OW) {
  p3dfft_exec_3Dderiv_double(*plan, in, out, *idir - 1, *OW);
}

void p3dfft_exec_3Dderiv_single_f(Plan3D *plan, float *in, float *out, int *idir, int *OW) {
  p3dfft_exec_3Dderiv_single(*plan, in, out, *idir - 1, *OW);
}

void p3dfft_compute_deriv_single_f(float *in, float *out, int *igrid, int *idir) {
  DataGrid *g = stored_grid(igrid, "p3dfft_compute_deriv_single_f");
  if (g) compute_deriv<mycomplex>((mycomplex *)in, (mycomplex *)out, g, *idir - 1);
}

void p3dfft_compute_deriv_double_f(double *in, double *out, int *igrid, int *idir) {
  DataGrid *g = stored_grid(igrid, "p3dfft_compute_deriv_double_f");
  if (g) compute_deriv<complex_double>((complex_double *)in, (complex_double *)out, g, *idir - 1);
}

}  // extern "C"
