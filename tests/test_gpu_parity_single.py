"""FP32 variants of the drop-in calls and the Fortran-side entry points,
through the C ABI on the GPU, against the FP64 oracle on the same (FP32
rounded) inputs.

Tolerances: FP32 paths <= 1e-5 relative L2 against the FP64 oracle; the
Fortran-side calls run the same plans as the C calls, so their FP64 results
are held to the FP64 drop-in bar (1e-12) and to the C calls' results.
"""
import ctypes
import sys

import numpy as np
import pytest

import cases as C
import oracle_binding as orc
import p3dfft3 as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

I3 = ctypes.c_int * 3
KINDS = ["DCT1", "DST1", "DCT2", "DST2", "DCT3", "DST3", "DCT4", "DST4"]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def lib_setup():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.setup()
    yield
    torch.cuda.synchronize()


def to_dev(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    torch.cuda.synchronize()
    return t


def f32(x):
    return x.astype(np.complex64 if np.iscomplexobj(x) else np.float32)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("cplx", [False, True])
def test_r2r_1d_plans_single(kind, cplx):
    """Every DCT*/DST*_S ID (real and complex) through p3dfft_plan_1Dtrans /
    p3dfft_exec_1Dtrans_single at N = 17, 64, 513 and three memory-order
    pairs (DCT4 IDs compute DCT-I, as in the reference)."""
    pg = P.init_proc_grid((1, 1, 1))
    for n, mo1, mo2 in ((17, [0, 1, 2], [0, 1, 2]), (64, [2, 0, 1], [0, 2, 1]), (513, [0, 1, 2], [1, 0, 2])):
        log("case: r2r single", kind, cplx, n, mo1, mo2)
        dims = [n, 6, 4]
        tid = C.r2r_id(kind, cplx, 4)
        g1 = P.init_data_grid(dims, -1, pg, (0, 1, 2), mo1)
        g2 = P.init_data_grid(dims, -1, pg, (0, 1, 2), mo2)
        h = P.plan_1Dtrans(g1, g2, tid, 0)
        rng = np.random.default_rng(1000 + n)
        shp = orc.storage_shape(dims, mo1)
        x = rng.uniform(-1, 1, shp) if not cplx else rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        x = f32(x)
        din = to_dev(x)
        dout = torch.zeros(orc.storage_shape(dims, mo2), dtype=torch.complex64 if cplx else torch.float32,
                           device="cuda")
        torch.cuda.synchronize()
        P.exec_1Dtrans(h, din.data_ptr(), dout.data_ptr(), 0, 4)
        ref = orc.exec_1d(C.r2r_id(kind, cplx, 8), 0, dims, mo1, dims, mo2,
                          x.astype(np.complex128 if cplx else np.float64))
        err = C.rel_l2(dout.cpu().numpy(), ref)
        assert err < 1e-5, (kind, cplx, n, err)


def test_compute_deriv_standalone_single():
    """p3dfft_compute_deriv_single (build/deriv.C:85-186), including the
    reference's storage-axis choice for 3-cycle memory orders: one complex
    multiply per element, so the FP32 result is within a few FP32 roundings
    of the FP64 oracle's."""
    pg = P.init_proc_grid((1, 1, 1))
    dims = [8, 6, 4]
    rng = np.random.default_rng(3)
    for mo in ([0, 1, 2], [2, 1, 0], [1, 2, 0]):
        g = P.init_data_grid(dims, -1, pg, (0, 1, 2), mo)
        for idir in range(3):
            shp = orc.storage_shape(dims, mo)
            z = f32(rng.standard_normal(shp) + 1j * rng.standard_normal(shp))
            din = to_dev(z)
            dout = torch.zeros_like(din)
            torch.cuda.synchronize()
            P.compute_deriv(din.data_ptr(), dout.data_ptr(), g, idir, 4)
            ref = orc.compute_deriv(z.astype(np.complex128), dims, dims, [0, 0, 0], mo, -1, idir)
            err = C.rel_l2(dout.cpu().numpy(), ref)
            assert err < 1e-6, (mo, idir, err)


@pytest.mark.parametrize("idir", [0, 1, 2])
def test_exec_3Dderiv_single_vs_oracle(idir):
    """p3dfft_exec_3Dderiv_single on a noisy field (sample/C++/test_deriv.C
    layout, (0,1,2) -> (1,2,0)), every direction, against the FP64 oracle's
    exec_deriv of the same FP32-rounded input."""
    n = [32, 24, 16]
    g = [np.sin(np.arange(m) * 2 * np.pi / m) for m in n]
    a = g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    a = a + np.random.default_rng(7).normal(0.0, 1e-2, a.shape)
    mo1, mo2 = [0, 1, 2], [1, 2, 0]
    x = f32(orc.to_storage(a, mo1))
    dims2 = [17, 24, 16]
    pg = P.init_proc_grid((1, 1, 1))
    g1 = P.init_data_grid(n, -1, pg, (0, 1, 2), mo1)
    g2 = P.init_data_grid(dims2, 0, pg, (0, 1, 2), mo2)
    h = P.plan_3Dtrans(g1, g2, P.init_3Dtype((C.R2C_S, C.CF_S, C.CF_S)))
    din = to_dev(x)
    dout = torch.zeros(orc.storage_shape(dims2, mo2), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    P.exec_3Dderiv(h, din.data_ptr(), dout.data_ptr(), idir, 0, 4)
    ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], n, mo1, mo2, x.astype(np.float64), idir=idir)
    err = C.rel_l2(dout.cpu().numpy(), ref)
    log("case: exec_3Dderiv_single idir", idir, "rel L2", err)
    assert err < 1e-5


# ------------------------------------------------------------ Fortran-side entry points

def _grid_f(gdims, conj, pg, mo):
    h, ld, gs = ctypes.c_int(), I3(), I3()
    P.lib().p3dfft_init_data_grid_f(ctypes.byref(h), ld, gs, I3(*gdims), ctypes.byref(ctypes.c_int(conj)),
                                    ctypes.byref(ctypes.c_int(pg)), I3(0, 1, 2), I3(*mo))
    assert h.value >= 0
    return h


def test_fortran_side_3d_and_deriv():
    """p3dfft_plan_3Dtrans_f / exec_3Dtrans_*_f / exec_3Dderiv_*_f
    (build/wrap.C:636-716, 764-781): int grid and plan handles passed by
    address, 1-based idir; results equal the C calls' and the oracle's."""
    L = P.lib()
    n, dims2 = [16, 12, 10], [9, 12, 10]
    mo1, mo2 = [2, 1, 0], [1, 2, 0]
    pg = L.p3dfft_init_proc_grid_f(I3(1, 1, 1), ctypes.byref(ctypes.c_int(P.MPI_COMM_WORLD)))
    g1, g2 = _grid_f(n, -1, pg, mo1), _grid_f(dims2, 0, pg, mo2)
    t = ctypes.c_int()
    L.p3dfft_init_3Dtype_f(ctypes.byref(t), I3(C.R2C_D, C.CF_D, C.CF_D))
    plan = ctypes.c_int(-1)
    L.p3dfft_plan_3Dtrans_f(ctypes.byref(plan), ctypes.byref(g1), ctypes.byref(g2), ctypes.byref(t))
    P.check()
    assert plan.value >= 0
    x = np.random.default_rng(11).uniform(-1, 1, orc.storage_shape(n, mo1))
    din = to_dev(x)
    ow = ctypes.byref(ctypes.c_int(0))
    for idir_f in (0, 1, 2, 3):  # 0: plain exec; 1..3: d/dx, d/dy, d/dz
        dout = torch.zeros(orc.storage_shape(dims2, mo2), dtype=torch.complex128, device="cuda")
        torch.cuda.synchronize()
        if idir_f == 0:
            L.p3dfft_exec_3Dtrans_double_f(ctypes.byref(plan), ctypes.c_void_p(din.data_ptr()),
                                           ctypes.c_void_p(dout.data_ptr()), ow)
        else:
            L.p3dfft_exec_3Dderiv_double_f(ctypes.byref(plan), ctypes.c_void_p(din.data_ptr()),
                                           ctypes.c_void_p(dout.data_ptr()), ctypes.byref(ctypes.c_int(idir_f)), ow)
        P.check()
        ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], n, mo1, mo2, x, idir=idir_f - 1)
        err = C.rel_l2(dout.cpu().numpy(), ref)
        log("case: fortran-side exec, idir (1-based)", idir_f, "rel L2", err)
        assert err < 1e-12
    # single precision twin
    ts = ctypes.c_int()
    L.p3dfft_init_3Dtype_f(ctypes.byref(ts), I3(C.R2C_S, C.CF_S, C.CF_S))
    plan_s = ctypes.c_int(-1)
    L.p3dfft_plan_3Dtrans_f(ctypes.byref(plan_s), ctypes.byref(g1), ctypes.byref(g2), ctypes.byref(ts))
    xs = f32(x)
    dins = to_dev(xs)
    douts = torch.zeros(orc.storage_shape(dims2, mo2), dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    L.p3dfft_exec_3Dderiv_single_f(ctypes.byref(plan_s), ctypes.c_void_p(dins.data_ptr()),
                                   ctypes.c_void_p(douts.data_ptr()), ctypes.byref(ctypes.c_int(2)), ow)
    P.check()
    ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], n, mo1, mo2, xs.astype(np.float64), idir=1)
    assert C.rel_l2(douts.cpu().numpy(), ref) < 1e-5


def test_fortran_side_1d_and_compute_deriv():
    """p3dfft_plan_1Dtrans_f / exec_1Dtrans_*_f and
    p3dfft_compute_deriv_*_f (build/wrap.C:584-634, 783-791)."""
    L = P.lib()
    dims = [33, 8, 6]
    pg = L.p3dfft_init_proc_grid_f(I3(1, 1, 1), ctypes.byref(ctypes.c_int(P.MPI_COMM_WORLD)))
    g1, g2 = _grid_f(dims, -1, pg, [0, 1, 2]), _grid_f(dims, -1, pg, [1, 0, 2])
    ow = ctypes.byref(ctypes.c_int(0))
    for prec in (8, 4):
        tid = C.r2r_id("DCT1", False, prec)
        plan = ctypes.c_int(-1)
        L.p3dfft_plan_1Dtrans_f(ctypes.byref(plan), ctypes.byref(g1), ctypes.byref(g2), ctypes.byref(ctypes.c_int(tid)),
                                ctypes.byref(ctypes.c_int(0)))
        P.check()
        assert plan.value >= 0
        x = np.random.default_rng(5).uniform(-1, 1, orc.storage_shape(dims, [0, 1, 2]))
        if prec == 4:
            x = f32(x)
        din = to_dev(x)
        dout = torch.zeros(orc.storage_shape(dims, [1, 0, 2]), dtype=din.dtype, device="cuda")
        torch.cuda.synchronize()
        f = L.p3dfft_exec_1Dtrans_double_f if prec == 8 else L.p3dfft_exec_1Dtrans_single_f
        f(ctypes.byref(plan), ctypes.c_void_p(din.data_ptr()), ctypes.c_void_p(dout.data_ptr()), ow)
        P.check()
        ref = orc.exec_1d(C.r2r_id("DCT1", False, 8), 0, dims, [0, 1, 2], dims, [1, 0, 2], x.astype(np.float64))
        assert C.rel_l2(dout.cpu().numpy(), ref) < (1e-12 if prec == 8 else 1e-5), prec
    # compute_deriv on a stored spectrum, grid by handle, idir 1-based
    sd = [8, 6, 4]
    for mo in ([0, 1, 2], [1, 2, 0]):
        gh = _grid_f(sd, -1, pg, mo)
        shp = orc.storage_shape(sd, mo)
        z = np.random.default_rng(9).standard_normal(shp) + 1j * np.random.default_rng(10).standard_normal(shp)
        for idir_f in (1, 2, 3):
            for prec in (8, 4):
                zz = z if prec == 8 else f32(z)
                din = to_dev(zz)
                dout = torch.zeros_like(din)
                torch.cuda.synchronize()
                f = L.p3dfft_compute_deriv_double_f if prec == 8 else L.p3dfft_compute_deriv_single_f
                f(ctypes.c_void_p(din.data_ptr()), ctypes.c_void_p(dout.data_ptr()), ctypes.byref(gh),
                  ctypes.byref(ctypes.c_int(idir_f)))
                P.check()
                ref = orc.compute_deriv(zz.astype(np.complex128), sd, sd, [0, 0, 0], mo, -1, idir_f - 1)
                assert C.rel_l2(dout.cpu().numpy(), ref) < (1e-14 if prec == 8 else 1e-6), (mo, idir_f, prec)
