"""ctypes binding of libp3dfft3_hip.so's C ABI (include/Cwrap.h and
include/p3dfft_hip.h), for Python callers such as the tests and bench.py.

The names follow the reference's C interface (p3dfft_setup,
p3dfft_init_proc_grid, p3dfft_init_data_grid, p3dfft_plan_3Dtrans,
p3dfft_exec_3Dtrans_double, ...).  Array arguments are raw addresses: device
pointers (e.g. torch.Tensor.data_ptr() of a cuda tensor) are used in place,
host pointers are staged by the library.  There is no CPU fallback: if the
HIP library is missing this module raises on load.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libp3dfft3_hip.so")

MPI_COMM_WORLD = 0x44000000  # MPICH's handle; the library never calls MPI with it

TYPE_NAMES = [
    "EMPTY_TYPE_SINGLE", "EMPTY_TYPE_DOUBLE", "EMPTY_TYPE_SINGLE_COMPLEX", "EMPTY_TYPE_DOUBLE_COMPLEX",
    "R2CFFT_S", "R2CFFT_D", "C2RFFT_S", "C2RFFT_D",
    "CFFT_FORWARD_S", "CFFT_FORWARD_D", "CFFT_BACKWARD_S", "CFFT_BACKWARD_D",
] + [f"{k}{i}_{c}_{p}" for i in (1, 2, 3, 4) for k in ("DCT", "DST") for c in ("REAL", "COMPLEX")
     for p in ("S", "D")]


class Grid(ctypes.Structure):
    """struct CDataGrid_struct (include/Cwrap.h)."""
    _fields_ = [("nd", ctypes.c_int), ("Gdims", ctypes.c_int * 3), ("dim_conj_sym", ctypes.c_int),
                ("MemOrder", ctypes.c_int * 3), ("Ldims", ctypes.c_int * 3), ("Dmap", ctypes.c_int * 3),
                ("pgrid", ctypes.c_int), ("grid_id", ctypes.c_int * 3), ("GlobStart", ctypes.c_int * 3),
                ("taskid", ctypes.c_int), ("numtasks", ctypes.c_int), ("ProcDims", ctypes.c_int * 3),
                ("mpi_comm_glob", ctypes.c_int)]


GridP = ctypes.POINTER(Grid)
I3 = ctypes.c_int * 3
VP = ctypes.c_void_p

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.p3dfft_init_data_grid.restype = GridP
        L.p3dfft_init_data_grid.argtypes = [I3, ctypes.c_int, ctypes.c_int, I3, I3]
        L.p3dfft_init_proc_grid.argtypes = [I3, ctypes.c_int]
        L.p3dfft_plan_3Dtrans.argtypes = [GridP, GridP, ctypes.c_int]
        L.p3dfft_plan_1Dtrans.argtypes = [GridP, GridP, ctypes.c_int, ctypes.c_int]
        L.p3dfft_free_data_grid.argtypes = [GridP]
        for n in ("double", "single"):
            getattr(L, f"p3dfft_exec_3Dtrans_{n}").argtypes = [ctypes.c_int, VP, VP, ctypes.c_int]
            getattr(L, f"p3dfft_exec_3Dderiv_{n}").argtypes = [ctypes.c_int, VP, VP, ctypes.c_int, ctypes.c_int]
            getattr(L, f"p3dfft_exec_1Dtrans_{n}").argtypes = [ctypes.c_int, VP, VP, ctypes.c_int]
            getattr(L, f"p3dfft_compute_deriv_{n}").argtypes = [VP, VP, GridP, ctypes.c_int]
            getattr(L, f"p3dfft_exec_3Dtrans_{n}_async").argtypes = [ctypes.c_int, VP, VP, ctypes.c_int,
                                                                     ctypes.c_int, VP]
        L.p3dfft_hip_exec_3Dtrans_noalloc.argtypes = [ctypes.c_int, VP, VP, ctypes.c_int, ctypes.c_int, VP]
        L.p3dfft_hip_last_error.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.p3dfft_hip_plan_describe.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
        L.p3dfft_hip_pass_bytes.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                            ctypes.POINTER(ctypes.c_longlong)]
        L.p3dfft_hip_stream.restype = VP
        _lib = L
    return _lib


def setup():
    lib().p3dfft_setup()


def cleanup():
    lib().p3dfft_cleanup()


def type_id(name):
    """Value of the P3DFFT_<name> global after setup()."""
    return ctypes.c_int.in_dll(lib(), "P3DFFT_" + name).value


def last_error(clear=True):
    buf = ctypes.create_string_buffer(1024)
    rc = lib().p3dfft_hip_last_error(buf, 1024)
    if clear:
        lib().p3dfft_hip_clear_error()
    return rc, buf.value.decode()


def check():
    rc, msg = last_error()
    if rc:
        raise RuntimeError(msg)


def init_3Dtype(types):
    return lib().p3dfft_init_3Dtype(I3(*types))


def init_proc_grid(pdims=(1, 1, 1), comm=MPI_COMM_WORLD):
    return lib().p3dfft_init_proc_grid(I3(*pdims), comm)


def init_data_grid(gdims, dim_conj_sym, pgrid, dmap, mem_order):
    g = lib().p3dfft_init_data_grid(I3(*gdims), dim_conj_sym, pgrid, I3(*dmap), I3(*mem_order))
    if not g:
        check()
        raise RuntimeError("p3dfft_init_data_grid failed")
    return g


def free_data_grid(g):
    lib().p3dfft_free_data_grid(g)


def free_proc_grid(p):
    lib().p3dfft_free_proc_grid(p)


def plan_3Dtrans(g1, g2, type3d):
    h = lib().p3dfft_plan_3Dtrans(g1, g2, type3d)
    check()
    if h < 0:
        raise RuntimeError("p3dfft_plan_3Dtrans failed")
    return h


def plan_1Dtrans(g1, g2, type_id_, d):
    h = lib().p3dfft_plan_1Dtrans(g1, g2, type_id_, d)
    check()
    if h < 0:
        raise RuntimeError("p3dfft_plan_1Dtrans failed")
    return h


def exec_3Dtrans(plan, inp, out, OW=0, prec=8):
    f = lib().p3dfft_exec_3Dtrans_double if prec == 8 else lib().p3dfft_exec_3Dtrans_single
    f(plan, inp, out, OW)
    check()


def exec_3Dderiv(plan, inp, out, idir, OW=0, prec=8):
    f = lib().p3dfft_exec_3Dderiv_double if prec == 8 else lib().p3dfft_exec_3Dderiv_single
    f(plan, inp, out, idir, OW)
    check()


def exec_1Dtrans(plan, inp, out, OW=0, prec=8):
    f = lib().p3dfft_exec_1Dtrans_double if prec == 8 else lib().p3dfft_exec_1Dtrans_single
    f(plan, inp, out, OW)
    check()


def compute_deriv(inp, out, grid, idir, prec=8):
    f = lib().p3dfft_compute_deriv_double if prec == 8 else lib().p3dfft_compute_deriv_single
    f(inp, out, grid, idir)
    check()


def exec_async(plan, inp, out, idir=-1, OW=0, stream=None, prec=8):
    f = lib().p3dfft_exec_3Dtrans_double_async if prec == 8 else lib().p3dfft_exec_3Dtrans_single_async
    rc = f(plan, inp, out, idir, OW, stream)
    if rc:
        check()
        raise RuntimeError("async exec failed")


def exec_noalloc(plan, inp, out, idir=-1, OW=0, stream=None):
    """Exec of a 3D plan on device arrays whose intermediate arrays live in
    the library's fixed device scratch array (allocates nothing)."""
    rc = lib().p3dfft_hip_exec_3Dtrans_noalloc(plan, inp, out, idir, OW, stream)
    if rc:
        _, msg = last_error()
        raise RuntimeError("no-alloc exec failed: " + msg)


def describe(plan):
    buf = ctypes.create_string_buffer(2048)
    lib().p3dfft_hip_plan_describe(plan, buf, 2048)
    return buf.value.decode()


def npasses(plan):
    return lib().p3dfft_hip_plan_npasses(plan)


def pass_bytes(plan, i):
    a, b = ctypes.c_longlong(), ctypes.c_longlong()
    rc = lib().p3dfft_hip_pass_bytes(plan, i, ctypes.byref(a), ctypes.byref(b))
    if rc:
        raise IndexError(i)
    return a.value, b.value


def stream():
    return lib().p3dfft_hip_stream()


def profile(plan, enable=True):
    """Start (and reset) per-pass HIP-event timing of a 3D plan."""
    L = lib()
    L.p3dfft_hip_profile.argtypes = [ctypes.c_int, ctypes.c_int]
    return L.p3dfft_hip_profile(plan, int(enable))


def profile_read(plan, npass=8):
    """(number of recorded executions, [avg ms per pass], total ms)."""
    L = lib()
    L.p3dfft_hip_profile_read.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)]
    avg = (ctypes.c_double * npass)()
    tot = ctypes.c_double()
    n = L.p3dfft_hip_profile_read(plan, avg, npass, ctypes.byref(tot))
    return n, list(avg)[:max(0, npasses(plan))], tot.value


def set_fast_path(enable=True):
    """Plans made afterwards use the power-of-two kernel where it applies
    (default) or always the generic kernel."""
    lib().p3dfft_hip_set_fast_path(int(bool(enable)))


# ------------------------------------------------------------ batches

def batch_plan(dims_in, dims_out, types, mo_in=(0, 1, 2), mo_out=(0, 1, 2), tune=False):
    """Batch of independent 3D transforms (include/p3dfft_hip.h): dims_in and
    dims_out are lists of per-item logical dims; types are three 1D type IDs."""
    L = lib()
    n = len(dims_in)
    if len(dims_out) != n:
        raise ValueError("dims_in and dims_out differ in length")
    IN = ctypes.c_int * (3 * n)
    di = IN(*[int(v) for d in dims_in for v in d])
    do = IN(*[int(v) for d in dims_out for v in d])
    L.p3dfft_hip_batch_plan.argtypes = [ctypes.c_int, IN, IN, I3, I3, I3, ctypes.c_int]
    h = L.p3dfft_hip_batch_plan(n, di, do, I3(*types), I3(*mo_in), I3(*mo_out), int(bool(tune)))
    if h < 0:
        raise RuntimeError(last_error())
    return h


def batch_exec(h, ins, outs, stream=None, sync=True):
    """ins/outs: per-item device addresses."""
    L = lib()
    n = len(ins)
    A = VP * n
    L.p3dfft_hip_batch_exec.argtypes = [ctypes.c_int, A, A, VP, ctypes.c_int]
    rc = L.p3dfft_hip_batch_exec(h, A(*ins), A(*outs), stream, int(bool(sync)))
    if rc != 0:
        raise RuntimeError(last_error())


def batch_set_graph(h, enable=True):
    L = lib()
    L.p3dfft_hip_batch_set_graph.argtypes = [ctypes.c_int, ctypes.c_int]
    if L.p3dfft_hip_batch_set_graph(h, int(bool(enable))) != 0:
        raise RuntimeError(last_error())


def batch_describe(h):
    L = lib()
    L.p3dfft_hip_batch_describe.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    buf = ctypes.create_string_buffer(1 << 16)
    L.p3dfft_hip_batch_describe(h, buf, len(buf))
    return buf.value.decode()


def batch_destroy(h):
    L = lib()
    L.p3dfft_hip_batch_destroy.argtypes = [ctypes.c_int]
    return L.p3dfft_hip_batch_destroy(h)


# ------------------------------------------------------------ Poisson service

def poisson_server_create(prec=4, pool_bytes=8 << 30, max_inflight=64):
    L = lib()
    L.p3dfft_hip_poisson_server_create.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int]
    h = L.p3dfft_hip_poisson_server_create(int(prec), int(pool_bytes), int(max_inflight))
    if h < 0:
        raise RuntimeError(last_error())
    return h


def poisson_server_prepare(h, dims):
    L = lib()
    L.p3dfft_hip_poisson_server_prepare.argtypes = [ctypes.c_int, I3]
    if L.p3dfft_hip_poisson_server_prepare(h, I3(*dims)) != 0:
        raise RuntimeError(last_error())


def poisson_submit(h, dims, rhs_ptr, sol_ptr):
    """Submit one solve; rhs/sol are device addresses of nx*ny*nz reals in
    memory order (0,1,2).  Returns the request id."""
    L = lib()
    L.p3dfft_hip_poisson_submit.argtypes = [ctypes.c_int, I3, VP, VP]
    L.p3dfft_hip_poisson_submit.restype = ctypes.c_longlong
    rid = L.p3dfft_hip_poisson_submit(h, I3(*dims), rhs_ptr, sol_ptr)
    if rid < 0:
        raise RuntimeError(last_error())
    return rid


def poisson_poll(h, cap=1024, wait_ms=0):
    """[(id, t_submit, t_admit, t_done, ok)] of requests completed since the
    last poll (seconds since the service started)."""
    L = lib()
    D = ctypes.c_double * cap
    ids, ts, ta, td, ok = (ctypes.c_longlong * cap)(), D(), D(), D(), (ctypes.c_int * cap)()
    L.p3dfft_hip_poisson_poll.argtypes = [ctypes.c_int, ctypes.c_longlong * cap, D, D, D, ctypes.c_int * cap,
                                          ctypes.c_int, ctypes.c_int]
    n = L.p3dfft_hip_poisson_poll(h, ids, ts, ta, td, ok, cap, int(wait_ms))
    if n < 0:
        raise RuntimeError(last_error())
    return [(ids[i], ts[i], ta[i], td[i], bool(ok[i])) for i in range(n)]


def poisson_server_stats(h):
    L = lib()
    L.p3dfft_hip_poisson_server_stats.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_longlong),
                                                  ctypes.POINTER(ctypes.c_longlong)]
    a, b = ctypes.c_longlong(), ctypes.c_longlong()
    rc = L.p3dfft_hip_poisson_server_stats(h, ctypes.byref(a), ctypes.byref(b))
    return {"launches": a.value, "tiles": b.value, "failed": rc == 1}


def poisson_server_state(h):
    """{pending, holding, groups_busy, lanes} of a service (diagnostics)."""
    L = lib()
    IP = ctypes.POINTER(ctypes.c_int)
    L.p3dfft_hip_poisson_server_state.argtypes = [ctypes.c_int, IP, IP, IP, IP]
    v = [ctypes.c_int() for _ in range(4)]
    if L.p3dfft_hip_poisson_server_state(h, *[ctypes.byref(x) for x in v]) != 0:
        check()
        raise RuntimeError("p3dfft_hip_poisson_server_state failed")
    return dict(zip(("pending", "holding", "groups_busy", "lanes"), (x.value for x in v)))


def poisson_server_destroy(h):
    L = lib()
    L.p3dfft_hip_poisson_server_destroy.argtypes = [ctypes.c_int]
    return L.p3dfft_hip_poisson_server_destroy(h)


def poisson_solve(dims, rhs_ptr, sol_ptr, prec=4, stream=None, sync=True):
    """One solve on the per-request path."""
    L = lib()
    L.p3dfft_hip_poisson_solve.argtypes = [ctypes.c_int, I3, VP, VP, VP, ctypes.c_int]
    if L.p3dfft_hip_poisson_solve(int(prec), I3(*dims), rhs_ptr, sol_ptr, stream, int(bool(sync))) != 0:
        raise RuntimeError(last_error())
