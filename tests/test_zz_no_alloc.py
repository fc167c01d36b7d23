"""The 3D exec that allocates nothing (p3dfft_hip_exec_3Dtrans_noalloc)
against the allocating exec of the same plan (p3dfft_exec_3Dtrans_*_async):
the same pass kernels on the same inputs, so the results are bit-identical.
Two calls in a row, the second on a smaller grid (the fixed scratch array
keeps the first call's contents), an FP64 plan with the spectral derivative,
an in-place C2C plan, and the refusals (a plan whose intermediate arrays
exceed the fixed array, host arrays)."""
import sys

import numpy as np
import pytest

import oracle_binding as orc
import p3dfft3 as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def lib_setup():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.setup()
    yield
    torch.cuda.synchronize()


def r2c_plan(n, prec, mo1=(2, 1, 0), mo2=(1, 2, 0)):
    s = "D" if prec == 8 else "S"
    pg = P.init_proc_grid((1, 1, 1))
    g1 = P.init_data_grid(n, -1, pg, (0, 1, 2), mo1)
    g2 = P.init_data_grid([n[0] // 2 + 1, n[1], n[2]], 0, pg, (0, 1, 2), mo2)
    t = P.init_3Dtype((P.type_id(f"R2CFFT_{s}"), P.type_id(f"CFFT_FORWARD_{s}"), P.type_id(f"CFFT_FORWARD_{s}")))
    return P.plan_3Dtrans(g1, g2, t), orc.storage_shape(n, mo1), orc.storage_shape([n[0] // 2 + 1, n[1], n[2]], mo2)


def run_both(plan, x, out_shape, cdt, prec, idir=-1):
    """(allocating exec result, no-alloc exec result) of one plan on x."""
    st = P.stream()
    a = torch.full(out_shape, float("nan"), dtype=cdt, device="cuda")
    b = torch.full(out_shape, float("nan"), dtype=cdt, device="cuda")
    torch.cuda.synchronize()
    P.exec_async(plan, x.data_ptr(), a.data_ptr(), idir, 0, st, prec=prec)
    P.exec_noalloc(plan, x.data_ptr(), b.data_ptr(), idir, 0, st)
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_matches_exec_two_calls_second_smaller():
    gen = torch.Generator(device="cuda").manual_seed(11)
    for n in ([64, 48, 40], [32, 24, 20]):
        log("case: no-alloc R2C FP32", n)
        plan, shp_in, shp_out = r2c_plan(n, 4)
        x = torch.rand(shp_in, generator=gen, device="cuda", dtype=torch.float32) * 2 - 1
        a, b = run_both(plan, x, shp_out, torch.complex64, 4)
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


def test_noalloc_fp64_with_derivative():
    n = [48, 40, 36]
    log("case: no-alloc R2C FP64 d/dy", n)
    plan, shp_in, shp_out = r2c_plan(n, 8)
    gen = torch.Generator(device="cuda").manual_seed(12)
    x = torch.rand(shp_in, generator=gen, device="cuda", dtype=torch.float64)
    a, b = run_both(plan, x, shp_out, torch.complex128, 8, idir=1)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)


def test_noalloc_in_place_c2c():
    n = [40, 32, 24]
    log("case: no-alloc in-place C2C FP32", n)
    pg = P.init_proc_grid((1, 1, 1))
    g = P.init_data_grid(n, -1, pg, (0, 1, 2), (0, 1, 2))
    cf = P.type_id("CFFT_FORWARD_S")
    plan = P.plan_3Dtrans(g, g, P.init_3Dtype((cf, cf, cf)))
    gen = torch.Generator(device="cuda").manual_seed(13)
    shp = orc.storage_shape(n, (0, 1, 2))
    x = torch.complex(torch.randn(shp, generator=gen, device="cuda"), torch.randn(shp, generator=gen, device="cuda"))
    ref = torch.empty_like(x)
    y = x.clone()
    torch.cuda.synchronize()
    P.exec_async(plan, x.data_ptr(), ref.data_ptr(), -1, 0, P.stream(), prec=4)
    P.exec_noalloc(plan, y.data_ptr(), y.data_ptr(), -1, 1, P.stream())
    torch.cuda.synchronize()
    assert np.array_equal(ref.cpu().numpy(), y.cpu().numpy())


def test_noalloc_refuses_large_plan_and_host_arrays():
    # the 512^3 R2C plan keeps a whole half-complex volume (538 MB) between
    # its passes: more than the fixed array holds
    n = [512, 512, 512]
    log("case: no-alloc refusals", n)
    plan, shp_in, shp_out = r2c_plan(n, 4)
    log("  plan:", P.describe(plan))
    x = torch.zeros(shp_in, dtype=torch.float32, device="cuda")
    y = torch.empty(shp_out, dtype=torch.complex64, device="cuda")
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="limit"):
        P.exec_noalloc(plan, x.data_ptr(), y.data_ptr())
    del x, y
    torch.cuda.empty_cache()
    small, shp_in, shp_out = r2c_plan([16, 12, 10], 4)
    h_in = np.zeros(shp_in, dtype=np.float32)
    d_out = torch.empty(shp_out, dtype=torch.complex64, device="cuda")
    with pytest.raises(RuntimeError, match="device arrays"):
        P.exec_noalloc(small, h_in.ctypes.data, d_out.data_ptr())
