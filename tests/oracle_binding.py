"""ctypes binding of the CPU oracle (oracle/liboracle.so).

The oracle is the checker: tests, smoke() and bench.py's cpu_baseline leg
load it; the product library never does.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

_lib = None

INT3 = ctypes.c_int * 3
DP = ctypes.POINTER(ctypes.c_double)


class StageT(ctypes.Structure):
    _fields_ = [("trans_dim", ctypes.c_int), ("type_id", ctypes.c_int),
                ("mo_in", ctypes.c_int * 3), ("mo_out", ctypes.c_int * 3),
                ("dims_in", ctypes.c_int * 3), ("dims_out", ctypes.c_int * 3)]


class TypeT(ctypes.Structure):
    _fields_ = [("cls", ctypes.c_int), ("dt1", ctypes.c_int), ("dt2", ctypes.c_int),
                ("prec", ctypes.c_int), ("isign", ctypes.c_int), ("r2r_kind", ctypes.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-C", ORACLE_DIR, "-s"])
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _dp(a):
    return a.ctypes.data_as(DP)


def _i3(v):
    return INT3(*[int(x) for x in v])


def fft(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().or_fft(len(x), _dp(x), _dp(out), int(sign))
    return out


def dft_naive(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().or_dft_naive(len(x), _dp(x), _dp(out), int(sign))
    return out


def r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty(len(x) // 2 + 1, dtype=np.complex128)
    lib().or_r2c(len(x), _dp(x), _dp(out))
    return out


def c2r(h, n):
    h = np.ascontiguousarray(h, dtype=np.complex128)
    out = np.empty(n, dtype=np.float64)
    lib().or_c2r(int(n), _dp(h), _dp(out))
    return out


REDFT00, REDFT10, REDFT01, REDFT11, RODFT00, RODFT10, RODFT01, RODFT11 = range(8)
DCT_KIND = {1: REDFT00, 2: REDFT10, 3: REDFT01, 4: REDFT11}
DST_KIND = {1: RODFT00, 2: RODFT10, 3: RODFT01, 4: RODFT11}


def r2r(kind, x, naive=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    f = lib().or_r2r_naive if naive else lib().or_r2r
    f(int(kind), len(x), _dp(x), _dp(out))
    return out


def type_info(tid):
    t = TypeT()
    rc = lib().or_type_info(int(tid), ctypes.byref(t))
    if rc:
        raise ValueError(tid)
    return t


def find_order(types, mo1, mo2):
    L = INT3()
    lib().or_find_order(_i3(types), _i3(mo1), _i3(mo2), L)
    return list(L)


def swap0(mo, L):
    out = INT3()
    lib().or_swap0(out, _i3(mo), int(L))
    return list(out)


def excl(a, b):
    return lib().or_excl(int(a), int(b))


def plan3d(types, gdims1, mo1, mo2):
    st = (StageT * 4)()
    n = lib().or_plan3d(_i3(types), _i3(gdims1), _i3(mo1), _i3(mo2), st)
    return [dict(trans_dim=s.trans_dim, type_id=s.type_id, mo_in=list(s.mo_in),
                 mo_out=list(s.mo_out), dims_in=list(s.dims_in),
                 dims_out=list(s.dims_out)) for s in st[:n]]


def storage_shape(dims, mo):
    """numpy shape (slowest first) of an array with logical dims and MemOrder mo."""
    sd = [0, 0, 0]
    for i in range(3):
        sd[mo[i]] = dims[i]
    return (sd[2], sd[1], sd[0])


def to_storage(logical, mo):
    """logical [x][y][z] array -> contiguous storage array for MemOrder mo."""
    # storage axis s (0 fastest) holds logical dim imo[s]; numpy axis 2-s.
    imo = [0, 0, 0]
    for i in range(3):
        imo[mo[i]] = i
    return np.ascontiguousarray(np.transpose(logical, (imo[2], imo[1], imo[0])))


def from_storage(stored, mo):
    """inverse of to_storage."""
    imo = [0, 0, 0]
    for i in range(3):
        imo[mo[i]] = i
    perm = (imo[2], imo[1], imo[0])
    inv = [0, 0, 0]
    for a, p in enumerate(perm):
        inv[p] = a
    return np.transpose(stored, inv)


def exec_1d(type_id, d, dims_in, mo_in, dims_out, mo_out, data, deriv=False):
    t = type_info(type_id)
    odt = np.complex128 if t.dt2 == 2 else np.float64
    idt = np.complex128 if t.dt1 == 2 else np.float64
    data = np.ascontiguousarray(data, dtype=idt)
    out = np.empty(storage_shape(dims_out, mo_out), dtype=odt)
    rc = lib().or_exec_1d(int(type_id), int(d), _i3(dims_in), _i3(mo_in), _i3(dims_out),
                          _i3(mo_out), _dp(data), _dp(out), int(bool(deriv)))
    if rc:
        raise RuntimeError(f"or_exec_1d rc={rc}")
    return out


def exec3d(types, gdims1, mo1, mo2, data, idir=-1):
    """Storage-layout input (contiguous, MemOrder mo1) -> storage-layout output."""
    st = plan3d(types, gdims1, mo1, mo2)
    last = st[-1]
    t = type_info(last["type_id"])
    odt = np.complex128 if t.dt2 == 2 else np.float64
    t0 = type_info(st[0]["type_id"])
    idt = np.complex128 if t0.dt1 == 2 else np.float64
    data = np.ascontiguousarray(data, dtype=idt)
    out = np.empty(storage_shape(last["dims_out"], mo2), dtype=odt)
    rc = lib().or_exec3d(_i3(types), _i3(gdims1), _i3(mo1), _i3(mo2), _dp(data), _dp(out),
                         int(idir))
    if rc:
        raise RuntimeError(f"or_exec3d rc={rc}")
    return out


def compute_deriv(spec, gdims, ldims, glob_start, mem_order, dim_conj_sym, idir):
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.empty_like(spec)
    lib().or_compute_deriv(_dp(spec), _dp(out), _i3(gdims), _i3(ldims), _i3(glob_start),
                           _i3(mem_order), int(dim_conj_sym), int(idir))
    return out
