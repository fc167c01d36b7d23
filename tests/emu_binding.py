"""ctypes binding of the host build of the pass kernel (tests/emu/libemu.so).

The library is compiled from the product's own planner and kernel sources
(p3dfft.3_amd/csrc/host/pass_plan.cpp, kernels/fft_core.h) and runs every
workgroup sequentially on the CPU.  Test infrastructure only.
"""
import ctypes
import os
import subprocess

import numpy as np

import oracle_binding as orc

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
LIB = os.path.join(EMU_DIR, "libemu.so")
_lib = None
I3 = ctypes.c_int * 3


def lib():
    global _lib
    if _lib is None:
        subprocess.check_call(["make", "-C", EMU_DIR, "-s"])
        _lib = ctypes.CDLL(LIB)
        _lib.emu_error.restype = ctypes.c_char_p
    return _lib


def set_dct_split(on):
    """Plan complex DCT-I passes of 129/257/513/1025 points in the split form
    (the default) or as the FFT of the even extension."""
    lib().emu_set_dct_split(1 if on else 0)


def set_pad_threshold(nbytes):
    """Intermediate arrays of at least nbytes get padded rows (the planner's
    default is 32 MiB; 0 pads every eligible intermediate)."""
    lib().emu_set_pad_threshold(ctypes.c_longlong(int(nbytes)))


def ops_for(types):
    cls, sign, kind, idt, odt = [], [], [], [], []
    for t in types:
        ti = orc.type_info(t)
        cls.append(ti.cls)
        sign.append(-1 if ti.isign < 0 else 1)
        kind.append(ti.r2r_kind)
        idt.append(ti.dt1)
        odt.append(ti.dt2)
    return cls, sign, kind, idt, odt


def exec3d(types, in_dims, mo1, out_dims, mo2, data, prec, idir=-1, order=None, batch_threads=0):
    """Run the transform with the product's planner and kernel code on the CPU.
    data: storage-order array; returns the storage-order output."""
    cls, sign, kind, idt, odt = ops_for(types)
    if order is None:
        order = orc.find_order(types, mo1, mo2)
    t0 = orc.type_info(types[order[0]])
    tl = orc.type_info(types[order[2]])
    in_dt, out_dt = t0.dt1, tl.dt2
    rdt = np.float64 if prec == 8 else np.float32
    cdt = np.complex128 if prec == 8 else np.complex64
    data = np.ascontiguousarray(data, dtype=cdt if in_dt == 2 else rdt)
    out = np.zeros(orc.storage_shape(out_dims, mo2), dtype=cdt if out_dt == 2 else rdt)
    desc = ctypes.create_string_buffer(512)
    rc = lib().emu_exec(prec, I3(*cls), I3(*sign), I3(*kind), I3(*idt), I3(*odt), I3(*order),
                        I3(*in_dims), I3(*mo1), in_dt, I3(*out_dims), I3(*mo2), out_dt,
                        data.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                        int(idir), desc, 512, int(batch_threads))
    if rc != 0:
        raise RuntimeError(lib().emu_error().decode())
    return out, desc.value.decode()


def exec1d(type_id, d, in_dims, mo1, out_dims, mo2, data, prec, deriv=False):
    """One transform along logical axis d (the transplan path)."""
    ti = orc.type_info(type_id)
    empty = [0, 0, 0]
    cls, sign, kind = list(empty), [-1, -1, -1], list(empty)
    idt, odt = [ti.dt1] * 3, [ti.dt1] * 3
    cls[d], sign[d], kind[d], idt[d], odt[d] = ti.cls, (-1 if ti.isign < 0 else 1), ti.r2r_kind, ti.dt1, ti.dt2
    order = [d, (d + 1) % 3, (d + 2) % 3]
    rdt = np.float64 if prec == 8 else np.float32
    cdt = np.complex128 if prec == 8 else np.complex64
    data = np.ascontiguousarray(data, dtype=cdt if ti.dt1 == 2 else rdt)
    out = np.zeros(orc.storage_shape(out_dims, mo2), dtype=cdt if ti.dt2 == 2 else rdt)
    desc = ctypes.create_string_buffer(512)
    rc = lib().emu_exec(prec, I3(*cls), I3(*sign), I3(*kind), I3(*idt), I3(*odt), I3(*order),
                        I3(*in_dims), I3(*mo1), ti.dt1, I3(*out_dims), I3(*mo2), ti.dt2,
                        data.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                        d if deriv else -1, desc, 512, 0)
    if rc != 0:
        raise RuntimeError(lib().emu_error().decode())
    return out


def tables_ready_counts(types, in_dims, mo1, out_dims, mo2, prec):
    """(passes, not ready with null table addresses, not ready with all set)."""
    cls, sign, kind, idt, odt = ops_for(types)
    order = orc.find_order(types, mo1, mo2)
    in_dt = orc.type_info(types[order[0]]).dt1
    out_dt = orc.type_info(types[order[2]]).dt2
    b, a = ctypes.c_int(), ctypes.c_int()
    n = lib().emu_tables_ready(prec, I3(*cls), I3(*sign), I3(*kind), I3(*idt), I3(*odt), I3(*order),
                               I3(*in_dims), I3(*mo1), in_dt, I3(*out_dims), I3(*mo2), out_dt,
                               ctypes.byref(b), ctypes.byref(a))
    if n < 0:
        raise RuntimeError(lib().emu_error().decode())
    return n, b.value, a.value


def poisson(dims, rhs, prec, mo_real=(0, 1, 2), mo_spec=(1, 2, 0), batch_threads=0):
    """Poisson solve through the product's Poisson plan on the CPU: returns
    (solution in mo_real storage, spectrum after the -1/|k|^2 multiply in
    mo_spec storage, plan description)."""
    rdt = np.float64 if prec == 8 else np.float32
    cdt = np.complex128 if prec == 8 else np.complex64
    rhs = np.ascontiguousarray(rhs, dtype=rdt)
    sol = np.zeros(orc.storage_shape(dims, mo_real), dtype=rdt)
    sdims = [dims[0] // 2 + 1, dims[1], dims[2]]
    spec = np.zeros(orc.storage_shape(sdims, mo_spec), dtype=cdt)
    desc = ctypes.create_string_buffer(1024)
    rc = lib().emu_poisson(prec, I3(*dims), I3(*mo_real), I3(*mo_spec), rhs.ctypes.data_as(ctypes.c_void_p),
                           sol.ctypes.data_as(ctypes.c_void_p), spec.ctypes.data_as(ctypes.c_void_p), desc, 1024,
                           int(batch_threads))
    if rc != 0:
        raise RuntimeError(lib().emu_error().decode())
    return sol, spec, desc.value.decode()


def exec3d_retuned(types, in_dims, mo1, out_dims, mo2, data, prec, B, pick=0):
    """Batch-mode plan with every pass re-tiled by the autotuner's
    retune_pass (B lines, radix candidate `pick`); returns (output, number
    of re-tiled passes, description)."""
    cls, sign, kind, idt, odt = ops_for(types)
    order = orc.find_order(types, mo1, mo2)
    in_dt = orc.type_info(types[order[0]]).dt1
    out_dt = orc.type_info(types[order[2]]).dt2
    rdt = np.float64 if prec == 8 else np.float32
    cdt = np.complex128 if prec == 8 else np.complex64
    data = np.ascontiguousarray(data, dtype=cdt if in_dt == 2 else rdt)
    out = np.zeros(orc.storage_shape(out_dims, mo2), dtype=cdt if out_dt == 2 else rdt)
    desc = ctypes.create_string_buffer(512)
    n = lib().emu_exec_retuned(prec, I3(*cls), I3(*sign), I3(*kind), I3(*idt), I3(*odt), I3(*order),
                               I3(*in_dims), I3(*mo1), in_dt, I3(*out_dims), I3(*mo2), out_dt,
                               data.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                               int(B), int(pick), desc, 512)
    if n < 0:
        raise RuntimeError(lib().emu_error().decode())
    return out, n, desc.value.decode()
