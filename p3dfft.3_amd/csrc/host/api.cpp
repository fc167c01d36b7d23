// C++ API of libp3dfft3_hip.so (namespace p3dfft): type registry, grids,
// transform3D / transplan over the HIP engine, compute_deriv.
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "../../../include/p3dfft.h"
#include "engine.h"

// MPI is used only for an optional check that a 1x1 grid runs on one rank.
// The library does not link libmpi: when the host program does, these weak
// references resolve; otherwise they stay null and the check is skipped.
extern "C" int MPI_Initialized(int *flag) __attribute__((weak));
extern "C" int MPI_Comm_size(MPI_Comm comm, int *size) __attribute__((weak));

namespace p3dfft {

static inline p3e::Plan *EP(engine_plan *e) { return reinterpret_cast<p3e::Plan *>(e); }
static inline engine_plan *PE(p3e::Plan *p) { return reinterpret_cast<engine_plan *>(p); }

int EMPTY_TYPE_SINGLE, EMPTY_TYPE_DOUBLE, EMPTY_TYPE_SINGLE_COMPLEX, EMPTY_TYPE_DOUBLE_COMPLEX;
int R2CFFT_S, R2CFFT_D, C2RFFT_S, C2RFFT_D, CFFT_FORWARD_S, CFFT_FORWARD_D, CFFT_BACKWARD_S, CFFT_BACKWARD_D;
int DCT1_REAL_S, DCT1_REAL_D, DST1_REAL_S, DST1_REAL_D, DCT1_COMPLEX_S, DCT1_COMPLEX_D, DST1_COMPLEX_S,
    DST1_COMPLEX_D;
int DCT2_REAL_S, DCT2_REAL_D, DST2_REAL_S, DST2_REAL_D, DCT2_COMPLEX_S, DCT2_COMPLEX_D, DST2_COMPLEX_S,
    DST2_COMPLEX_D;
int DCT3_REAL_S, DCT3_REAL_D, DST3_REAL_S, DST3_REAL_D, DCT3_COMPLEX_S, DCT3_COMPLEX_D, DST3_COMPLEX_S,
    DST3_COMPLEX_D;
int DCT4_REAL_S, DCT4_REAL_D, DST4_REAL_S, DST4_REAL_D, DCT4_COMPLEX_S, DCT4_COMPLEX_D, DST4_COMPLEX_S,
    DST4_COMPLEX_D;

vector<gen_trans_type *> types1D;
vector<trans_type3D> types3D;
vector<gen_transform3D *> stored_trans3D;
vector<stage *> stored_trans1D;
vector<ProcGrid *> stored_proc_grids;
vector<DataGrid *> stored_data_grids;

template <class Type1, class Type2>
trans_type1D<Type1, Type2>::trans_type1D(const char *name_, int cls_, int r2r_kind_, int isign_, bool empty_)
    : gen_trans_type(isign_) {
  ID = -1;
  name = name_;
  cls = cls_;
  r2r_kind = r2r_kind_;
  is_empty = empty_;
  is_set = true;
  prec = (typeid(Type1) == typeid(float) || typeid(Type1) == typeid(mycomplex)) ? 4 : 8;
  dt1 = (typeid(Type1) == typeid(mycomplex) || typeid(Type1) == typeid(complex_double)) ? 2 : 1;
  dt2 = (typeid(Type2) == typeid(mycomplex) || typeid(Type2) == typeid(complex_double)) ? 2 : 1;
}

template class trans_type1D<float, float>;
template class trans_type1D<double, double>;
template class trans_type1D<mycomplex, mycomplex>;
template class trans_type1D<complex_double, complex_double>;
template class trans_type1D<float, mycomplex>;
template class trans_type1D<double, complex_double>;
template class trans_type1D<mycomplex, float>;
template class trans_type1D<complex_double, double>;

static int reg(gen_trans_type *t) {
  int id = (int)types1D.size();
  types1D.push_back(t);
  return id;
}

// Registration order of build/init.C:121-731; the IDs are part of the ABI.
// The DCT4 IDs compute REDFT00 (DCT-I), as the reference registers them
// with its DCT-I planner (build/init.C:640,652,664,676).
void setup() {
  using namespace p3e;
  if (!types1D.empty()) return;
  EMPTY_TYPE_SINGLE = reg(new trans_type1D<float, float>("Empty Type Single", CLS_EMPTY, 0, 1, true));
  EMPTY_TYPE_DOUBLE = reg(new trans_type1D<double, double>("Empty Type Double", CLS_EMPTY, 0, 1, true));
  EMPTY_TYPE_SINGLE_COMPLEX =
      reg(new trans_type1D<mycomplex, mycomplex>("Empty Type Single Complex", CLS_EMPTY, 0, 1, true));
  EMPTY_TYPE_DOUBLE_COMPLEX = reg(
      new trans_type1D<complex_double, complex_double>("Empty Type Double Complex", CLS_EMPTY, 0, 1, true));
  R2CFFT_S = reg(new trans_type1D<float, mycomplex>("Real-to-complex FFT, single precision", CLS_R2C, 0, -1, false));
  R2CFFT_D = reg(new trans_type1D<double, complex_double>("Real-to-complex FFT, double precision", CLS_R2C, 0, -1, false));
  C2RFFT_S = reg(new trans_type1D<mycomplex, float>("Complex-to-real FFT, single precision", CLS_C2R, 0, 1, false));
  C2RFFT_D = reg(new trans_type1D<complex_double, double>("Complex-to-real FFT, double precision", CLS_C2R, 0, 1, false));
  CFFT_FORWARD_S = reg(new trans_type1D<mycomplex, mycomplex>("Complex forward FFT, single precision", CLS_C2C, 0, -1, false));
  CFFT_FORWARD_D = reg(new trans_type1D<complex_double, complex_double>("Complex forward FFT, double precision", CLS_C2C, 0, -1, false));
  CFFT_BACKWARD_S = reg(new trans_type1D<mycomplex, mycomplex>("Complex backward FFT, single precision", CLS_C2C, 0, 1, false));
  CFFT_BACKWARD_D = reg(new trans_type1D<complex_double, complex_double>("Complex backward FFT, double precision", CLS_C2C, 0, 1, false));

  struct R2R {
    int *rs, *rd, *cs, *cd;
    int kind;
    const char *nm;
  };
  const R2R r2r[] = {
      {&DCT1_REAL_S, &DCT1_REAL_D, &DCT1_COMPLEX_S, &DCT1_COMPLEX_D, REDFT00, "DCT1"},
      {&DST1_REAL_S, &DST1_REAL_D, &DST1_COMPLEX_S, &DST1_COMPLEX_D, RODFT00, "DST1"},
      {&DCT2_REAL_S, &DCT2_REAL_D, &DCT2_COMPLEX_S, &DCT2_COMPLEX_D, REDFT10, "DCT2"},
      {&DST2_REAL_S, &DST2_REAL_D, &DST2_COMPLEX_S, &DST2_COMPLEX_D, RODFT10, "DST2"},
      {&DCT3_REAL_S, &DCT3_REAL_D, &DCT3_COMPLEX_S, &DCT3_COMPLEX_D, REDFT01, "DCT3"},
      {&DST3_REAL_S, &DST3_REAL_D, &DST3_COMPLEX_S, &DST3_COMPLEX_D, RODFT01, "DST3"},
      {&DCT4_REAL_S, &DCT4_REAL_D, &DCT4_COMPLEX_S, &DCT4_COMPLEX_D, REDFT00, "DCT4"},
      {&DST4_REAL_S, &DST4_REAL_D, &DST4_COMPLEX_S, &DST4_COMPLEX_D, RODFT11, "DST4"},
  };
  for (const R2R &r : r2r) {
    *r.rs = reg(new trans_type1D<float, float>(r.nm, CLS_R2R, r.kind, 0, false));
    *r.rd = reg(new trans_type1D<double, double>(r.nm, CLS_R2R, r.kind, 0, false));
    *r.cs = reg(new trans_type1D<mycomplex, mycomplex>(r.nm, CLS_R2R, r.kind, 0, false));
    *r.cd = reg(new trans_type1D<complex_double, complex_double>(r.nm, CLS_R2R, r.kind, 0, false));
  }
}

void cleanup() {
  for (gen_transform3D *t : stored_trans3D) delete t;
  stored_trans3D.clear();
  for (stage *s : stored_trans1D) delete s;
  stored_trans1D.clear();
  for (ProcGrid *g : stored_proc_grids) delete g;
  stored_proc_grids.clear();
  for (DataGrid *g : stored_data_grids) delete g;
  stored_data_grids.clear();
  types3D.clear();
  for (gen_trans_type *t : types1D) delete t;
  types1D.clear();
  p3e::release_all();
}

int arcmp(int *A, int *B, int N) {
  for (int i = 0; i < N; i++)
    if (A[i] != B[i]) return 1;
  return 0;
}

// build/exec.C:2959-2967
void inv_mo(int mo[3], int imo[3]) {
  for (int i = 0; i < 3; i++) imo[mo[i]] = i;
}

// build/exec.C:2270-2296: mc[i] is the position of mo1[i] in mo2.
void rel_change(int *mo1, int *mo2, int *mc) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      if (mo2[j] == mo1[i]) mc[i] = j;
}

// ------------------------------------------------------------------ grids
ProcGrid::ProcGrid(int procdims[3], MPI_Comm mpi_comm_init) {
  nd = 0;
  for (int i = 0; i < 3; i++) {
    ProcDims[i] = procdims[i];
    if (procdims[i] > 1) nd++;
    grid_id_cart[i] = 0;
  }
  if (nd == 0) nd = 1;
  mpi_comm_glob = mpi_comm_init;
  mpi_comm_cart = mpi_comm_init;
  for (int i = 0; i < 3; i++) mpicomm[i] = mpi_comm_init;
  taskid = 0;
  numtasks = 1;
  int flag = 0;
  if (MPI_Initialized && MPI_Comm_size && MPI_Initialized(&flag) == 0 && flag) {
    int n = 1;
    if (MPI_Comm_size(mpi_comm_init, &n) == 0) numtasks = n;
  }
  if (ProcDims[0] * ProcDims[1] * ProcDims[2] != 1 || numtasks != 1)
    p3e::set_error(
        "ProcGrid: %d x %d x %d grid on %d rank(s); this library runs the 1x1 (single-GPU) path only",
        ProcDims[0], ProcDims[1], ProcDims[2], numtasks);
}

ProcGrid::ProcGrid(const ProcGrid &rhs) { *this = rhs; }
ProcGrid::~ProcGrid() {}

bool ProcGrid::operator==(const ProcGrid &P) const {
  if (nd != P.nd || taskid != P.taskid || numtasks != P.numtasks) return false;
  if (mpi_comm_glob != P.mpi_comm_glob) return false;
  for (int i = 0; i < 3; i++)
    if (ProcDims[i] != P.ProcDims[i] || grid_id_cart[i] != P.grid_id_cart[i]) return false;
  return true;
}

// build/init.C:1699-1729 and InitPencil (build/init.C:1806-1901), where the
// block distribution of each axis reduces to the whole axis on one rank.
DataGrid::DataGrid(int *gdims_, int dim_conj_sym_, ProcGrid *pgrid, int *dmap, int *mem_order) {
  dim_conj_sym = dim_conj_sym_;
  Pgrid = pgrid;
  nd = pgrid ? pgrid->nd : 1;
  for (int i = 0; i < 3; i++) {
    Gdims[i] = gdims_[i];
    Dmap[i] = dmap[i];
    Pdims[i] = pgrid ? pgrid->ProcDims[dmap[i]] : 1;
    MemOrder[i] = mem_order[i];
    grid_id[i] = pgrid ? pgrid->grid_id_cart[dmap[i]] : 0;
  }
  InitPencil();
  is_set = true;
}

DataGrid::DataGrid(const DataGrid &rhs) { *this = rhs; }

void DataGrid::set_gdims(int gdims[3]) {
  for (int i = 0; i < 3; i++) Gdims[i] = gdims[i];
  InitPencil();
}

void DataGrid::InitPencil() {
  int nl = 0, n = 0;
  for (int k = 0; k < 3; k++) {
    L[k] = D[k] = -1;
  }
  for (int k = 0; k < 3; k++) {
    if (Pdims[k] == 1) L[nl++] = k;
    else D[n++] = k;
  }
  for (int i = 0; i < 3; i++) {
    const int data = Gdims[i], proc = Pdims[i] > 0 ? Pdims[i] : 1;
    const int size = data / proc, nu = data % proc, nlo = proc - nu;
    const int p = grid_id[i];
    int st = (p < nlo) ? p * size : nlo * size + (p - nlo) * (size + 1);
    int sz = (p < nlo) ? size : size + 1;
    Ldims[i] = sz;
    GlobStart[i] = st;
  }
}

// ------------------------------------------------------------------ type3D
trans_type3D::trans_type3D(int types_IDs[3]) {
  name = nullptr;
  is_set = true;
  for (int i = 0; i < 3; i++) types[i] = types_IDs[i];
  if (types[0] >= 0 && types[0] < (int)types1D.size()) prec = types1D[types[0]]->prec;
  else {
    prec = 0;
    is_set = false;
    p3e::set_error("trans_type3D: undefined 1D type ID %d (was setup() called?)", types[0]);
  }
  for (int i = 0; i < 3; i++)
    if (types[i] < 0 || types[i] >= (int)types1D.size()) {
      is_set = false;
      p3e::set_error("trans_type3D: undefined 1D type ID %d", types[i]);
    } else if (types1D[types[i]]->prec != prec) {
      p3e::set_error("trans_type3D: mixed precisions in the three 1D types");
    }
}

trans_type3D::trans_type3D(gen_trans_type *types_[3]) {
  name = nullptr;
  is_set = true;
  prec = types_[0]->prec;
  for (int i = 0; i < 3; i++) {
    types[i] = -1;
    for (size_t j = 0; j < types1D.size(); j++)
      if (types1D[j] == types_[i]) types[i] = (int)j;
    if (types[i] < 0) is_set = false;
  }
}

trans_type3D::trans_type3D(const trans_type3D &rhs) {
  name = nullptr;
  prec = rhs.prec;
  is_set = rhs.is_set;
  for (int i = 0; i < 3; i++) types[i] = rhs.types[i];
}

trans_type3D &trans_type3D::operator=(const trans_type3D &rhs) {
  prec = rhs.prec;
  is_set = rhs.is_set;
  for (int i = 0; i < 3; i++) types[i] = rhs.types[i];
  return *this;
}

trans_type3D::~trans_type3D() {}

// ------------------------------------------------------------------ planning
// Order of the three 1D stages, as the reference's find_order decides it on
// a 1x1 grid (build/templ.C:627-802): R2C first, C2R last; without R2C,
// start from the first local axis, preferring the one stored faster; on a
// single task the last axis is the output's stride-1 axis unless that is
// the first one.  On one rank every axis is local: grid L = {0,1,2}.
static int excl(int a, int b) {
  if (a * b == 0) return (std::max(a, b) == 1) ? 2 : 1;
  return 0;
}

bool find_order(int L[3], const trans_type3D *tp, const DataGrid *gr1, const DataGrid *gr2, bool *reverse_steps) {
  if (reverse_steps) *reverse_steps = false;
  L[0] = L[1] = L[2] = -1;
  for (int i = 0; i < 3; i++) {
    gen_trans_type *t = types1D[tp->types[i]];
    if (t->dt1 < t->dt2) {
      if (L[0] >= 0) p3e::set_error("transform3D: more than one real-to-complex 1D transform");
      else L[0] = i;
    } else if (t->dt1 > t->dt2) {
      if (L[2] >= 0) p3e::set_error("transform3D: more than one complex-to-real 1D transform");
      else L[2] = i;
    }
  }
  if (L[0] >= 0 && L[2] >= 0) p3e::set_error("transform3D: can't have both R2C and C2R transforms");
  const int l0 = gr1->L[0] >= 0 ? gr1->L[0] : 0, l1 = gr1->L[1] >= 0 ? gr1->L[1] : 1;
  if (L[0] < 0) {
    if (L[2] < 0) {
      L[0] = l0;
      if (gr1->nd == 1 && gr1->MemOrder[L[0]] > gr1->MemOrder[l1]) L[0] = l1;
    } else {
      L[0] = l0;
      if (L[0] == L[2]) L[0] = l1;
    }
  }
  if (L[2] < 0) {
    int z = 0;
    for (int i = 0; i < 3; i++)
      if (gr2->MemOrder[i] == 0) z = i;
    L[2] = z;
    if (L[2] == L[0]) L[2] = (L[0] + 1) % 3;
  }
  L[1] = excl(L[0], L[2]);
  return false;
}

template <class T> static int prec_of() {
  return (typeid(T) == typeid(float) || typeid(T) == typeid(mycomplex)) ? 4 : 8;
}
template <class T> static int dt_of() {
  return (typeid(T) == typeid(mycomplex) || typeid(T) == typeid(complex_double)) ? 2 : 1;
}

static p3e::OpSpec op_of(const gen_trans_type *t) {
  p3e::OpSpec o;
  o.cls = t->cls;
  o.sign = t->isign < 0 ? -1 : 1;
  o.r2r_kind = t->r2r_kind;
  o.in_dt = t->dt1;
  o.out_dt = t->dt2;
  return o;
}

template <class Type1, class Type2>
transform3D<Type1, Type2>::transform3D(const DataGrid &grid1_, const DataGrid &grid2_, const trans_type3D *type) {
  eng = nullptr;
  is_set = false;
  prec = prec_of<Type1>();
  dt1 = dt_of<Type1>();
  dt2 = dt_of<Type2>();
  OW = false;
  grid1 = new DataGrid(grid1_);
  grid2 = new DataGrid(grid2_);
  if (prec != prec_of<Type2>()) {
    p3e::set_error("transform3D: precisions of input and output types don't match");
    return;
  }
  if (!type || !type->is_set) {
    p3e::set_error("transform3D: 3D transform type not set");
    return;
  }
  if (type->prec != prec) {
    p3e::set_error("transform3D: transform type precision %d does not match array precision %d", type->prec,
                   prec);
    return;
  }
  if (grid1_.Pgrid && grid2_.Pgrid && !(*grid1_.Pgrid == *grid2_.Pgrid)) {
    p3e::set_error("transform3D: processor grids dont match");
    return;
  }
  int L[3];
  find_order(L, type, &grid1_, &grid2_, nullptr);
  if (types1D[type->types[L[0]]]->dt1 != dt1) {
    p3e::set_error("transform3D: input datatypes don't match");
    return;
  }
  if (types1D[type->types[L[2]]]->dt2 != dt2) {
    p3e::set_error("transform3D: output datatypes don't match");
    return;
  }
  p3e::OpSpec ops[3];
  for (int a = 0; a < 3; a++) ops[a] = op_of(types1D[type->types[a]]);
  p3e::Layout in, out;
  for (int i = 0; i < 3; i++) {
    in.dims[i] = grid1_.Ldims[i];
    in.mo[i] = grid1_.MemOrder[i];
    out.dims[i] = grid2_.Ldims[i];
    out.mo[i] = grid2_.MemOrder[i];
  }
  in.dt = dt1;
  out.dt = dt2;
  eng = PE(p3e::plan_create(prec, ops, L, in, out));
  is_set = eng != nullptr;
}

template <class Type1, class Type2> transform3D<Type1, Type2>::~transform3D() {
  p3e::plan_destroy(EP(eng));
  delete grid1;
  delete grid2;
}

template <class Type1, class Type2> void transform3D<Type1, Type2>::exec(Type1 *in, Type2 *out, bool OW_) {
  exec_deriv(in, out, -1, OW_);
}

template <class Type1, class Type2>
void transform3D<Type1, Type2>::exec_deriv(Type1 *in, Type2 *out, int idir, bool OW_) {
  if (!is_set || !eng) {
    p3e::set_error("transform3D::exec: plan is not set");
    return;
  }
  if ((void *)in == (void *)out && !OW_)
    fprintf(stderr, "Warning in transform3D:exec_deriv: input and output are the same, but overwrite "
                    "priviledge is not set\n");
  if (idir >= 0 && idir < 3) {
    int pi = EP(eng)->pass_of_axis[idir];
    if (pi < 0 || !EP(eng)->passes[pi].d.deriv_capable)
      p3e::set_error("exec_deriv: axis %d has no complex-output transform stage; derivative skipped", idir);
  }
  p3e::plan_exec(EP(eng), in, out, idir, OW_, nullptr, true);
}

// ------------------------------------------------------------------ transplan
template <class Type1, class Type2>
transplan<Type1, Type2>::transplan(const DataGrid &gr1, const DataGrid &gr2, const gen_trans_type *type, int d) {
  init_tr(gr1, gr2, type, d);
}

template <class Type1, class Type2>
transplan<Type1, Type2>::transplan(const DataGrid &gr1, const DataGrid &gr2, int type_ID, int d) {
  if (type_ID < 0 || type_ID >= (int)types1D.size()) {
    eng = nullptr;
    is_set = false;
    grid1 = grid2 = nullptr;
    trans_type = nullptr;
    p3e::set_error("trans_plan: 1D transform type %d not set", type_ID);
    return;
  }
  init_tr(gr1, gr2, types1D[type_ID], d);
}

// build/templ.C:900-992 (argument checks and the per-stage fields)
template <class Type1, class Type2>
void transplan<Type1, Type2>::init_tr(const DataGrid &gr1, const DataGrid &gr2, const gen_trans_type *type, int d) {
  eng = nullptr;
  is_set = false;
  grid1 = new DataGrid(gr1);
  grid2 = new DataGrid(gr2);
  trans_type = const_cast<gen_trans_type *>(type);
  trans_dim = d;
  kind = TRANS_ONLY;
  is_empty = type && type->is_empty;
  N = m = 0;
  isign = type ? type->isign : 0;
  prec = stage_prec = prec_of<Type1>();
  dt1 = dt_of<Type1>();
  dt2 = dt_of<Type2>();
  for (int i = 0; i < 3; i++) {
    dims1[i] = gr1.Ldims[i];
    dims2[i] = gr2.Ldims[i];
    mo1[i] = gr1.MemOrder[i];
    mo2[i] = gr2.MemOrder[i];
  }
  if (!type || !type->is_set) {
    p3e::set_error("trans_plan: 1D transform type not set");
    return;
  }
  if (d < 0 || d > 2) {
    p3e::set_error("transplan: invalid dimension %d", d);
    return;
  }
  if (type->dt1 != dt1 || type->dt2 != dt2 || type->prec != prec) {
    p3e::set_error("transplan: transform type does not match the template data types");
    return;
  }
  if (gr1.Ldims[d] != gr1.Gdims[d] || gr2.Ldims[d] != gr2.Gdims[d]) {
    p3e::set_error("transplan: dimensions dont match %d, %d, %d", gr1.Ldims[d], gr2.Ldims[d], d);
    return;
  }
  N = (type->dt1 > type->dt2) ? gr2.Gdims[d] : gr1.Gdims[d];
  p3e::OpSpec ops[3];
  for (int a = 0; a < 3; a++) ops[a] = p3e::OpSpec{p3e::CLS_EMPTY, -1, 0, dt1, dt1};
  ops[d] = op_of(type);
  int order[3] = {d, (d + 1) % 3, (d + 2) % 3};
  p3e::Layout in, out;
  for (int i = 0; i < 3; i++) {
    in.dims[i] = gr1.Ldims[i];
    in.mo[i] = gr1.MemOrder[i];
    out.dims[i] = gr2.Ldims[i];
    out.mo[i] = gr2.MemOrder[i];
  }
  in.dt = dt1;
  out.dt = dt2;
  // an EMPTY type leaves every axis EMPTY: the plan is a pure reorder
  eng = PE(p3e::plan_create(prec, ops, order, in, out));
  is_set = eng != nullptr;
}

template <class Type1, class Type2> transplan<Type1, Type2>::~transplan() {
  p3e::plan_destroy(EP(eng));
  delete grid1;
  delete grid2;
}

template <class Type1, class Type2> void transplan<Type1, Type2>::exec(char *in, char *out, bool OW) {
  if (!is_set || !eng) {
    p3e::set_error("transplan::exec: plan is not set");
    return;
  }
  p3e::plan_exec(EP(eng), in, out, -1, OW, nullptr, true);
}

template <class Type1, class Type2> void transplan<Type1, Type2>::exec_deriv(char *in, char *out, bool OW) {
  if (!is_set || !eng) {
    p3e::set_error("transplan::exec_deriv: plan is not set");
    return;
  }
  if (dt2 != 2) p3e::set_error("Error in exec_deriv: expected complex output type");
  p3e::plan_exec(EP(eng), in, out, trans_dim, OW, nullptr, true);
}

#define P3_INSTANTIATE(T1, T2) \
  template class transform3D<T1, T2>; \
  template class transplan<T1, T2>;
P3_INSTANTIATE(float, float)
P3_INSTANTIATE(double, double)
P3_INSTANTIATE(mycomplex, mycomplex)
P3_INSTANTIATE(complex_double, complex_double)
P3_INSTANTIATE(float, mycomplex)
P3_INSTANTIATE(double, complex_double)
P3_INSTANTIATE(mycomplex, float)
P3_INSTANTIATE(complex_double, double)

// ------------------------------------------------------------------ compute_deriv
// build/deriv.C:85-186.  The storage axis differentiated is the logical
// index i with MemOrder[i] == idir, used as a storage rank, exactly as the
// reference computes it (the two coincide for memory orders that are their
// own inverse); wavenumbers come from logical axis idir.
template <class Type> void compute_deriv(Type *in, Type *out, DataGrid *gr, int idir) {
  if (!gr || idir < 0 || idir > 2) {
    p3e::set_error("compute_deriv: invalid grid or direction %d", idir);
    return;
  }
  int sdims[3], ldir = 0;
  for (int i = 0; i < 3; i++) {
    sdims[gr->MemOrder[i]] = gr->Ldims[i];
    if (gr->MemOrder[i] == idir) ldir = i;
  }
  int g = (gr->dim_conj_sym == idir) ? (gr->Gdims[idir] - 1) * 2 : gr->Gdims[idir];
  int mid = g / 2 - gr->GlobStart[idir];
  p3e::deriv_exec(prec_of<Type>(), in, out, sdims, ldir, g, mid, gr->GlobStart[idir], nullptr);
}

template void compute_deriv<mycomplex>(mycomplex *, mycomplex *, DataGrid *, int);
template void compute_deriv<complex_double>(complex_double *, complex_double *, DataGrid *, int);

}  // namespace p3dfft
