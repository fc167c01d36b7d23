"""C ABI of libp3dfft3_hip.so without a GPU: the library loads, exports every
function and global its public headers declare, publishes the reference's
type IDs, and keeps the reference's handle and grid conventions."""
import ctypes
import os
import re

import pytest

import p3dfft3 as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(p3dfft_\w+)\s*\(", txt)))


def declared_globals(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(P3DFFT_\w+)\b", txt)) - {"P3DFFT_CWRAP_H", "P3DFFT_H", "P3DFFT_HIP_H"})


def test_library_exports_every_declared_symbol():
    L = P.lib()
    fns = declared_functions("Cwrap.h") + declared_functions("Fwrap.h") + declared_functions("p3dfft_hip.h")
    assert len(fns) >= 40
    assert len(declared_functions("Fwrap.h")) == 13
    for f in fns + ["find_grid"]:
        assert hasattr(L, f), f
    gl = declared_globals("Cwrap.h")
    assert len(gl) == 44
    for g in gl:
        ctypes.c_int.in_dll(L, g)


def test_type_ids_registration_order():
    P.setup()
    # build/init.C:121-731 order
    assert P.type_id("EMPTY_TYPE_SINGLE") == 0
    assert P.type_id("EMPTY_TYPE_DOUBLE_COMPLEX") == 3
    assert P.type_id("R2CFFT_S") == 4 and P.type_id("R2CFFT_D") == 5
    assert P.type_id("C2RFFT_S") == 6 and P.type_id("C2RFFT_D") == 7
    assert P.type_id("CFFT_FORWARD_S") == 8 and P.type_id("CFFT_BACKWARD_D") == 11
    assert P.type_id("DCT1_REAL_S") == 12 and P.type_id("DCT1_COMPLEX_D") == 15
    assert P.type_id("DST1_REAL_S") == 16 and P.type_id("DCT2_REAL_S") == 20
    assert P.type_id("DST2_COMPLEX_D") == 27 and P.type_id("DCT3_REAL_D") == 29
    assert P.type_id("DST3_REAL_S") == 32 and P.type_id("DCT4_REAL_S") == 36
    assert P.type_id("DST4_COMPLEX_D") == 43


def test_data_grid_fields_on_1x1():
    P.setup()
    pg = P.init_proc_grid((1, 1, 1))
    g = P.init_data_grid((8, 6, 5), -1, pg, (0, 1, 2), (2, 1, 0))
    c = g.contents
    assert list(c.Gdims) == [8, 6, 5] and list(c.Ldims) == [8, 6, 5]
    assert list(c.GlobStart) == [0, 0, 0] and list(c.MemOrder) == [2, 1, 0]
    assert c.nd == 1 and c.numtasks == 1 and c.taskid == 0 and c.pgrid == pg
    assert list(c.ProcDims) == [1, 1, 1]
    P.free_data_grid(g)
    P.free_proc_grid(pg)


def test_inv_mo():
    mo = (ctypes.c_int * 3)(1, 2, 0)
    imo = (ctypes.c_int * 3)()
    P.lib().p3dfft_inv_mo(mo, imo)
    assert list(imo) == [2, 0, 1]


def test_free_proc_grid_shifts_handles():
    # build/wrap.C:339-344 erases the entry, so later handles shift down
    P.setup()
    a = P.init_proc_grid((1, 1, 1))
    b = P.init_proc_grid((1, 1, 1))
    assert b == a + 1
    P.free_proc_grid(a)
    c = P.init_proc_grid((1, 1, 1))
    assert c == b  # the table shrank by one, so the new handle reuses b's index
    P.free_proc_grid(c)
    P.free_proc_grid(a)


def test_multi_rank_grid_is_rejected():
    P.setup()
    P.last_error()
    pg = P.init_proc_grid((1, 2, 2))
    rc, msg = P.last_error()
    assert rc != 0 and "1x1" in msg
    P.free_proc_grid(pg)


def test_invalid_handles_record_errors():
    P.setup()
    P.last_error()
    P.lib().p3dfft_exec_3Dtrans_double(12345, None, None, 0)
    rc, msg = P.last_error()
    assert rc != 0 and "invalid plan handle" in msg
    assert P.lib().p3dfft_init_3Dtype((ctypes.c_int * 3)(5, 9, 99)) >= 0
    rc, msg = P.last_error()
    assert rc != 0 and "undefined 1D type ID 99" in msg


I3 = ctypes.c_int * 3


def _grid_f(gdims, conj, pg, mo):
    h, ld, gs = ctypes.c_int(), I3(), I3()
    P.lib().p3dfft_init_data_grid_f(ctypes.byref(h), ld, gs, I3(*gdims), ctypes.byref(ctypes.c_int(conj)),
                                    ctypes.byref(ctypes.c_int(pg)), I3(0, 1, 2), I3(*mo))
    return h.value, list(ld), list(gs)


def test_fortran_side_grid_handles():
    """p3dfft_init_data_grid_f (build/wrap.C:725-762): int grid handles, Ldims
    and GlobStart written back, and an equal grid (global dims, processor
    grid, Dmap, MemOrder; dim_conj_sym is not compared) reuses its handle
    (find_grid, build/wrap.C:279-294)."""
    P.setup()
    L = P.lib()
    pg = L.p3dfft_init_proc_grid_f(I3(1, 1, 1), ctypes.byref(ctypes.c_int(P.MPI_COMM_WORLD)))
    assert pg >= 0
    a, ld, gs = _grid_f([12, 10, 8], -1, pg, [0, 1, 2])
    assert a >= 0 and ld == [12, 10, 8] and gs == [0, 0, 0]
    b, _, _ = _grid_f([12, 10, 8], -1, pg, [0, 1, 2])
    assert b == a
    c, _, _ = _grid_f([12, 10, 8], 0, pg, [0, 1, 2])
    assert c == a  # the reference's key leaves dim_conj_sym out
    d, _, _ = _grid_f([12, 10, 8], -1, pg, [2, 1, 0])
    assert d != a
    assert L.find_grid(I3(12, 10, 8), pg, I3(0, 1, 2), I3(2, 1, 0)) == d
    assert L.find_grid(I3(12, 10, 9), pg, I3(0, 1, 2), I3(2, 1, 0)) == -1
    t = ctypes.c_int(-5)
    L.p3dfft_init_3Dtype_f(ctypes.byref(t), I3(P.type_id("R2CFFT_D"), P.type_id("CFFT_FORWARD_D"),
                                             P.type_id("CFFT_FORWARD_D")))
    assert t.value >= 0


def test_fortran_side_invalid_handles():
    P.setup()
    L = P.lib()
    P.last_error()
    plan = ctypes.c_int(7)
    L.p3dfft_plan_3Dtrans_f(ctypes.byref(plan), ctypes.byref(ctypes.c_int(9999)),
                            ctypes.byref(ctypes.c_int(9999)), ctypes.byref(ctypes.c_int(0)))
    rc, msg = P.last_error()
    assert plan.value == -1 and rc != 0 and "invalid data grid handle" in msg
    L.p3dfft_compute_deriv_double_f(None, None, ctypes.byref(ctypes.c_int(-1)), ctypes.byref(ctypes.c_int(1)))
    rc, msg = P.last_error()
    assert rc != 0 and "invalid data grid handle" in msg
