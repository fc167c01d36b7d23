"""HIP library (through its C ABI) against the oracle on the same seeded
inputs, plus the reference samples' known answers and size-independent
properties at the full 512^3 size.

Tolerances: FP64 paths <= 1e-12 relative L2 against the FP64 oracle (the
drop-in bar); FP32 paths <= 1e-5 relative L2 against the FP64 oracle.
"""
import sys

import numpy as np
import pytest

import cases as C
import oracle_binding as orc
import p3dfft3 as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {8: 1e-12, 4: 1e-5}


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
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a).cuda()
    torch.cuda.synchronize()
    return t


def empty_dev(shape, complex_, prec):
    if complex_:
        dt = torch.complex128 if prec == 8 else torch.complex64
    else:
        dt = torch.float64 if prec == 8 else torch.float32
    return torch.zeros(shape, dtype=dt, device="cuda")


class Plan3D:
    """A 3D plan over the C ABI, as the reference's C samples build one."""

    def __init__(self, types, dims1, mo1, dims2, mo2, conj1=-1, conj2=-1):
        self.pg = P.init_proc_grid((1, 1, 1))
        self.g1 = P.init_data_grid(dims1, conj1, self.pg, (0, 1, 2), mo1)
        self.g2 = P.init_data_grid(dims2, conj2, self.pg, (0, 1, 2), mo2)
        self.t = P.init_3Dtype(types)
        self.h = P.plan_3Dtrans(self.g1, self.g2, self.t)
        self.prec = orc.type_info(types[0]).prec

    def run(self, din, dout, idir=-1, OW=0):
        torch.cuda.synchronize()
        if idir >= 0:
            P.exec_3Dderiv(self.h, din.data_ptr(), dout.data_ptr(), idir, OW, self.prec)
        else:
            P.exec_3Dtrans(self.h, din.data_ptr(), dout.data_ptr(), OW, self.prec)
        return dout


def gpu_3d(types, dims, mo1, mo2, x, idir=-1, c2r_len=None):
    prec = orc.type_info(types[0]).prec
    od = C.out_dims(types, dims, c2r_len)
    pl = Plan3D(types, dims, mo1, od, mo2)
    t_last = orc.type_info(types[orc.find_order(types, mo1, mo2)[2]])
    dout = empty_dev(orc.storage_shape(od, mo2), t_last.dt2 == 2, prec)
    pl.run(to_dev(x), dout, idir)
    return dout.cpu().numpy(), pl


def cast(x, prec):
    if prec == 8:
        return x
    return x.astype(np.complex64 if np.iscomplexobj(x) else np.float32)


def wide(x):
    return x.astype(np.complex128 if np.iscomplexobj(x) else np.float64)


# ------------------------------------------------------------ config 1

def test_config1_64cubed_fp64_forward_and_roundtrip():
    """64^3 FP64, uniform[-1,1) from seed 1, R2C x + C2C y,z and back,
    mem order (0,1,2) in, (1,2,0) out."""
    log("case: 64^3 fp64 r2c/c2r (0,1,2)->(1,2,0)")
    n = [64, 64, 64]
    mo1, mo2 = [0, 1, 2], [1, 2, 0]
    x = np.random.default_rng(1).uniform(-1.0, 1.0, orc.storage_shape(n, mo1))
    fwd = [C.R2C_D, C.CF_D, C.CF_D]
    spec, pl = gpu_3d(fwd, n, mo1, mo2, x)
    ref = orc.exec3d(fwd, n, mo1, mo2, x)
    err = C.rel_l2(spec, ref)
    log("  forward rel L2 vs oracle", err, P.describe(pl.h))
    assert err < TOL[8]
    bwd = [C.C2R_D, C.CB_D, C.CB_D]
    back, _ = gpu_3d(bwd, [33, 64, 64], mo2, mo1, spec, c2r_len=64)
    refb = orc.exec3d(bwd, [33, 64, 64], mo2, mo1, ref)
    assert C.rel_l2(back, refb) < TOL[8]
    assert C.rel_l2(back / 64 ** 3, x) < 1e-14


def test_sample_sine_known_answer():
    """sample/C++/test3D_r2c.C: +-0.125i at (1, +-1, +-1) after 1/N^3."""
    n = [32, 24, 20]
    xs = [np.sin(np.arange(m) * 2 * np.pi / m) for m in n]
    a = xs[0][:, None, None] * xs[1][None, :, None] * xs[2][None, None, :]
    mo1, mo2 = [0, 1, 2], [1, 2, 0]
    spec, _ = gpu_3d([C.R2C_D, C.CF_D, C.CF_D], n, mo1, mo2, orc.to_storage(a, mo1))
    S = orc.from_storage(spec, mo2) / np.prod(n)
    ref = np.zeros(S.shape, complex)
    ref[1, 1, 1], ref[1, -1, 1], ref[1, 1, -1], ref[1, -1, -1] = 0.125j, -0.125j, -0.125j, 0.125j
    assert np.abs(S - ref).max() < 1e-14 * n[0] * 0.25


@pytest.mark.parametrize("mo1", C.PERMS)
def test_memory_order_sweep_fp64(mo1):
    """The reference's 36-pair memory-order matrix (extra/makejob.py)."""
    for mo2 in C.PERMS:
        log("case: r2c 12x10x6 mo", mo1, "->", mo2)
        x = C.random_input([C.R2C_D, C.CF_D, C.CF_D], [12, 10, 6], mo1, 5)
        got, pl = gpu_3d([C.R2C_D, C.CF_D, C.CF_D], [12, 10, 6], mo1, mo2, x)
        ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], [12, 10, 6], mo1, mo2, x)
        assert C.rel_l2(got, ref) < TOL[8], (mo1, mo2, P.describe(pl.h))


@pytest.mark.parametrize("fast", [True, False])
@pytest.mark.parametrize("prec", [8, 4])
def test_pow2_both_kernels_memory_orders(fast, prec):
    """Power-of-two sizes through the power-of-two kernel and, forced, the
    generic kernel, over a spread of memory-order pairs."""
    P.set_fast_path(fast)
    try:
        fwd = [C.R2C_D, C.CF_D, C.CF_D] if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
        for mo1, mo2 in (([0, 1, 2], [1, 2, 0]), ([2, 1, 0], [1, 2, 0]), ([1, 2, 0], [0, 1, 2]),
                         ([2, 0, 1], [0, 2, 1])):
            log("case: pow2 64x32x128", mo1, mo2, "fast", fast, "prec", prec)
            x = cast(C.random_input(fwd, [64, 32, 128], mo1, 12), prec)
            got, pl = gpu_3d(fwd, [64, 32, 128], mo1, mo2, x)
            ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], [64, 32, 128], mo1, mo2, wide(x))
            assert C.rel_l2(got, ref) < TOL[prec], (mo1, mo2, P.describe(pl.h))
            assert ("pow2" in P.describe(pl.h)) == fast
            bwd = [C.C2R_D, C.CB_D, C.CB_D] if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
            back, pl2 = gpu_3d(bwd, [33, 32, 128], mo2, mo1, got, c2r_len=64)
            assert C.rel_l2(back / (64 * 32 * 128), x) < (1e-13 if prec == 8 else 1e-6), P.describe(pl2.h)
    finally:
        P.set_fast_path(True)


def test_uneven_sizes_58x139x199():
    """extra/makejob.py uneven grid; 139 and 199 are prime (Bluestein)."""
    log("case: c2c/r2c 58x139x199")
    n = [58, 139, 199]
    x = C.random_input([C.R2C_D, C.CF_D, C.CF_D], n, [0, 1, 2], 9)
    got, pl = gpu_3d([C.R2C_D, C.CF_D, C.CF_D], n, [0, 1, 2], [1, 2, 0], x)
    ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], n, [0, 1, 2], [1, 2, 0], x)
    assert C.rel_l2(got, ref) < TOL[8], P.describe(pl.h)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("dims", [[16, 8, 4], [48, 40, 33], [17, 11, 13], [96, 120, 32]])
def test_c2c_3d(prec, dims):
    log("case: c2c", dims, prec)
    t = [C.CF_D] * 3 if prec == 8 else [C.CF_S] * 3
    x = cast(C.random_input(t, dims, [0, 1, 2], 2), prec)
    got, pl = gpu_3d(t, dims, [0, 1, 2], [2, 1, 0], x)
    ref = orc.exec3d([C.CF_D] * 3, dims, [0, 1, 2], [2, 1, 0], wide(x))
    assert C.rel_l2(got, ref) < TOL[prec], P.describe(pl.h)


@pytest.mark.parametrize("idir", [0, 1, 2])
def test_derivative_vs_oracle_and_analytic(idir):
    """sample/C++/test_deriv.C: d/d(idir) of sin*sin*sin -> cos in idir."""
    n = [32, 24, 16]
    g = [np.sin(np.arange(m) * 2 * np.pi / m) for m in n]
    a = g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    mo1, mo2 = [0, 1, 2], [1, 2, 0]
    fwd = [C.R2C_D, C.CF_D, C.CF_D]
    x = orc.to_storage(a, mo1)
    spec, _ = gpu_3d(fwd, n, mo1, mo2, x, idir=idir)
    ref = orc.exec3d(fwd, n, mo1, mo2, x, idir=idir)
    assert C.rel_l2(spec, ref) < TOL[8]
    back, _ = gpu_3d([C.C2R_D, C.CB_D, C.CB_D], [17, 24, 16], mo2, mo1, spec, c2r_len=32)
    got = orc.from_storage(back, mo1) / np.prod(n)
    h = list(g)
    h[idir] = np.cos(np.arange(n[idir]) * 2 * np.pi / n[idir])
    ana = h[0][:, None, None] * h[1][None, :, None] * h[2][None, None, :]
    assert np.abs(got - ana).max() < 1e-13


def test_mixed_r2c_c2c_dct1_reduced():
    """768x640x513 mixed-type configuration at reduced size 48x40x33."""
    types = [C.R2C_D, C.CF_D, C.r2r_id("DCT1", True, 8)]
    for prec in (8, 4):
        t = types if prec == 8 else [C.R2C_S, C.CF_S, C.r2r_id("DCT1", True, 4)]
        x = cast(C.random_input(types, [48, 40, 33], [0, 1, 2], 3), prec)
        got, pl = gpu_3d(t, [48, 40, 33], [0, 1, 2], [1, 2, 0], x)
        ref = orc.exec3d(types, [48, 40, 33], [0, 1, 2], [1, 2, 0], wide(x))
        assert C.rel_l2(got, ref) < TOL[prec], (prec, P.describe(pl.h))


KINDS = ["DCT1", "DST1", "DCT2", "DST2", "DCT3", "DST3", "DCT4", "DST4"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("cplx", [False, True])
def test_r2r_1d_plans(kind, cplx):
    """p3dfft_plan_1Dtrans / exec_1Dtrans for every DCT/DST ID (DCT4 IDs
    compute DCT-I, as in the reference)."""
    pg = P.init_proc_grid((1, 1, 1))
    for n, mo1, mo2 in ((17, [0, 1, 2], [0, 1, 2]), (64, [2, 0, 1], [0, 2, 1]), (513, [0, 1, 2], [1, 0, 2])):
        log("case: r2r", kind, cplx, n, mo1, mo2)
        dims = [n, 6, 4]
        tid = C.r2r_id(kind, cplx, 8)
        g1 = P.init_data_grid(dims, -1, pg, (0, 1, 2), mo1)
        g2 = P.init_data_grid(dims, -1, pg, (0, 1, 2), mo2)
        h = P.plan_1Dtrans(g1, g2, tid, 0)
        rng = np.random.default_rng(n)
        shp = orc.storage_shape(dims, mo1)
        x = rng.uniform(-1, 1, shp) if not cplx else rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        din = to_dev(x)
        dout = empty_dev(orc.storage_shape(dims, mo2), cplx, 8)
        torch.cuda.synchronize()
        P.exec_1Dtrans(h, din.data_ptr(), dout.data_ptr(), 0, 8)
        ref = orc.exec_1d(tid, 0, dims, mo1, dims, mo2, x)
        assert C.rel_l2(dout.cpu().numpy(), ref) < TOL[8], (kind, cplx, n)


def test_compute_deriv_standalone():
    """p3dfft_compute_deriv_double, including the reference's storage-axis
    choice for 3-cycle memory orders."""
    pg = P.init_proc_grid((1, 1, 1))
    dims = [8, 6, 4]
    rng = np.random.default_rng(2)
    for mo in ([0, 1, 2], [2, 1, 0], [1, 2, 0]):
        g = P.init_data_grid(dims, -1, pg, (0, 1, 2), mo)
        for idir in range(3):
            z = rng.standard_normal(orc.storage_shape(dims, mo)) + 1j * rng.standard_normal(
                orc.storage_shape(dims, mo))
            din = to_dev(z)
            dout = torch.zeros_like(din)
            torch.cuda.synchronize()
            P.compute_deriv(din.data_ptr(), dout.data_ptr(), g, idir, 8)
            ref = orc.compute_deriv(z, dims, dims, [0, 0, 0], mo, -1, idir)
            assert np.abs(dout.cpu().numpy() - ref).max() < 1e-13, (mo, idir)


def test_host_pointer_path():
    """Unchanged samples pass host arrays: the library stages them."""
    n = [16, 12, 10]
    x = C.random_input([C.R2C_D, C.CF_D, C.CF_D], n, [0, 1, 2], 4)
    pl = Plan3D([C.R2C_D, C.CF_D, C.CF_D], n, [0, 1, 2], [9, 12, 10], [1, 2, 0])
    out = np.zeros(orc.storage_shape([9, 12, 10], [1, 2, 0]), dtype=np.complex128)
    P.exec_3Dtrans(pl.h, x.ctypes.data, out.ctypes.data, 0, 8)
    ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], n, [0, 1, 2], [1, 2, 0], x)
    assert C.rel_l2(out, ref) < TOL[8]


def test_inplace_c2c_with_overwrite():
    """test3D_c2c_inplace.C: in == out with OW=1."""
    n = [16, 8, 12]
    t = [C.CF_D] * 3
    for mo1, mo2 in (([0, 1, 2], [0, 1, 2]), ([0, 1, 2], [2, 1, 0])):
        x = C.random_input(t, n, mo1, 6)
        pl = Plan3D(t, n, mo1, n, mo2)
        buf = to_dev(x)
        pl.run(buf, buf, OW=1)
        ref = orc.exec3d(t, n, mo1, mo2, x)
        assert C.rel_l2(buf.cpu().numpy().reshape(ref.shape), ref) < TOL[8], (mo1, mo2)


def test_input_preserved_without_overwrite():
    n = [32, 16, 8]
    x = C.random_input([C.C2R_D, C.CB_D, C.CB_D], [17, 16, 8], [1, 2, 0], 8)
    din = to_dev(x)
    keep = din.clone()
    gpu = Plan3D([C.C2R_D, C.CB_D, C.CB_D], [17, 16, 8], [1, 2, 0], n, [0, 1, 2])
    dout = empty_dev(orc.storage_shape(n, [0, 1, 2]), False, 8)
    gpu.run(din, dout, OW=0)
    assert torch.equal(din, keep)


# ------------------------------------------------------------ config 2 (full size)

def c2_field(n, noise, seed):
    x = np.arange(n) * 2 * np.pi / n
    f = np.sin(x)[:, None, None] * np.sin(2 * x)[None, :, None] * np.sin(3 * x)[None, None, :]
    if noise:
        f = f + np.random.default_rng(seed).normal(0.0, noise, f.shape)
    return f


@pytest.mark.parametrize("prec", [8, 4])
def test_config2_512cubed_roundtrip_and_derivative(prec):
    """512^3, input mem order (2,1,0), output (1,2,0): R2C + C2R round trip
    (noisy field) and d/dx against the analytic cos(x)sin(2y)sin(3z)."""
    n = 512
    log("case: 512^3 prec", prec)
    mo1, mo2 = [2, 1, 0], [1, 2, 0]
    fwd = [C.R2C_D, C.CF_D, C.CF_D] if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
    bwd = [C.C2R_D, C.CB_D, C.CB_D] if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
    f = cast(c2_field(n, 1e-3, 2), prec)
    x = orc.to_storage(f, mo1)
    del f
    dims2 = [n // 2 + 1, n, n]
    pf = Plan3D(fwd, [n] * 3, mo1, dims2, mo2)
    pb = Plan3D(bwd, dims2, mo2, [n] * 3, mo1)
    din = to_dev(x)
    spec = empty_dev(orc.storage_shape(dims2, mo2), True, prec)
    back = empty_dev(orc.storage_shape([n] * 3, mo1), False, prec)
    pf.run(din, spec)
    pb.run(spec, back)
    err = (torch.linalg.vector_norm((back / n ** 3 - din).double()) /
           torch.linalg.vector_norm(din.double())).item()
    log("  round trip rel L2", err)
    assert err < (1e-13 if prec == 8 else 1e-6)
    # Parseval on the half spectrum (kx = 0 and kx = n/2 planes counted once)
    s = spec.cpu().numpy()
    S = orc.from_storage(s, mo2)
    w = np.full(S.shape[0], 2.0)
    w[0] = w[-1] = 1.0
    e_spec = float((w[:, None, None] * np.abs(S) ** 2).sum()) / n ** 3
    e_real = float((din.double() ** 2).sum().item())
    assert abs(e_spec - e_real) / e_real < (1e-12 if prec == 8 else 1e-5)
    del s, S
    # derivative on the noise-free field
    clean = orc.to_storage(cast(c2_field(n, 0.0, 0), prec), mo1)
    din = to_dev(clean)
    del clean
    pf.run(din, spec, idir=0)
    pb.run(spec, back)
    xs = np.arange(n) * 2 * np.pi / n
    ana = np.cos(xs)[:, None, None] * np.sin(2 * xs)[None, :, None] * np.sin(3 * xs)[None, None, :]
    got = orc.from_storage(back.cpu().numpy(), mo1) / n ** 3
    err = np.abs(got - ana).max()
    rel = np.linalg.norm(got - ana) / np.linalg.norm(ana)
    log("  d/dx max abs error", err, "rel L2", rel)
    # FP32: i*k (|k| up to n/2 = 256) amplifies every rounding error in the
    # spectrum, starting with the FP32 rounding of the input (~6e-8 relative,
    # spread over all modes, k_rms ~ 150): the derivative field cannot be
    # closer than ~1e-5 to the exact one.  The bar is its relative L2 error
    # at 5e-5 (FP64: max abs 1e-12).
    if prec == 8:
        assert err < 1e-12
    else:
        assert rel < 5e-5


PRIMES = [7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47]


def _path_of_first_pass(pl):
    return P.describe(pl.h).split(" | ")[0]


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("p", PRIMES)
def test_prime_radix_kernels(p, prec):
    """Every in-register odd prime radix of the single-buffer path (one
    kernel per prime: 7..47 in FP32, 7..29 in FP64; FP64 31 on the generic
    path, FP64 37..47 by Bluestein) on a C2C of length 4p, against the
    oracle."""
    dims = [4 * p, 6, 5]
    types = [C.CF_D] * 3
    x = C.random_input(types, dims, [0, 1, 2], 100 + p)
    t = types if prec == 8 else [C.CF_S] * 3
    xin = x if prec == 8 else x.astype(np.complex64)
    got, pl = gpu_3d(t, dims, [0, 1, 2], [1, 2, 0], xin)
    ref = orc.exec3d(types, dims, [0, 1, 2], [1, 2, 0], x)
    first = _path_of_first_pass(pl)
    log("case: prime", p, "prec", prec, first)
    if prec == 4 or p <= 29:
        assert f"mid p{p}" in first, first
    elif p == 31:
        assert "mid" not in first and "bluestein" not in first, first
    else:
        assert "bluestein" in first, first
    assert C.rel_l2(got, ref) < TOL[prec]


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("n,mo2", [(2 * 17, [0, 1, 2]), (2 * 29, [1, 2, 0]), (4 * 43, [2, 1, 0]), (2 * 53, [1, 2, 0]),
                                   (250, [2, 0, 1]), (24, [0, 2, 1])])
def test_paired_real_lines(n, mo2, prec):
    """R2C and C2R on the single-buffer path transform two real lines per
    complex FFT: an odd number of lines (7, the last unpaired), line- and
    tile-order sides, prime radices and Bluestein (2 * 53); the C2R input is
    arbitrary complex data (its DC and Nyquist imaginary parts are ignored,
    as by a real inverse transform)."""
    dims = [n, 7, 3]
    fwd = [C.R2C_D, C.CF_D, C.CF_D]
    bwd = [C.C2R_D, C.CB_D, C.CB_D]
    tf = fwd if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
    tb = bwd if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
    x = C.random_input(fwd, dims, [0, 1, 2], 31)
    got, pl = gpu_3d(tf, dims, [0, 1, 2], mo2, x if prec == 8 else x.astype(np.float32))
    first = _path_of_first_pass(pl)
    log("case: paired", n, mo2, prec, first)
    assert "pair" in first, first
    assert C.rel_l2(got, orc.exec3d(fwd, dims, [0, 1, 2], mo2, x)) < TOL[prec]
    od = C.out_dims(fwd, dims)
    z = C.random_input(bwd, od, mo2, 32)
    got, pl = gpu_3d(tb, od, mo2, [0, 1, 2], z if prec == 8 else z.astype(np.complex64), c2r_len=n)
    assert C.rel_l2(got, orc.exec3d(bwd, od, mo2, [0, 1, 2], z)) < TOL[prec]


@pytest.mark.parametrize("prec", [4, 8])
def test_bluestein_above_47(prec):
    """Primes above 47 (53, 59, 61) take Bluestein on the single-buffer path."""
    dims = [53, 59, 61]
    types = [C.CF_D] * 3
    x = C.random_input(types, dims, [0, 1, 2], 7)
    t = types if prec == 8 else [C.CF_S] * 3
    got, pl = gpu_3d(t, dims, [0, 1, 2], [0, 1, 2], x if prec == 8 else x.astype(np.complex64))
    desc = P.describe(pl.h)
    log("case: bluestein 53 59 61", prec, desc)
    assert desc.count("bluestein") == 3, desc
    assert C.rel_l2(got, orc.exec3d(types, dims, [0, 1, 2], [0, 1, 2], x)) < TOL[prec]


# ------------------------------------------------- compile-time lengths

FIXED_ODD = [48, 96, 192, 384, 768, 80, 160, 320, 640, 120, 60, 72, 144, 112, 224, 240, 36]


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("M", FIXED_ODD)
def test_fixed_length_odd_kernels_vs_oracle(M, prec):
    """Compile-time lengths 3, 5 and 15 times a power of two on the GPU:
    C2C along a stride-1 axis (line order) and along a strided axis (tile
    order with a partial tile), and R2C of length 2M (line order, and tile
    order with paired two-line stores), against the oracle."""
    cf = C.CF_D if prec == 8 else C.CF_S
    for dims, mo1, mo2 in (([M, 3, 2], [0, 1, 2], [0, 1, 2]), ([5, M, 3], [0, 1, 2], [2, 0, 1])):
        log("case: fixed c2c", M, prec, dims, mo1, mo2)
        rng = np.random.default_rng(M)
        x = rng.standard_normal(orc.storage_shape(dims, mo1)) + 1j * rng.standard_normal(orc.storage_shape(dims, mo1))
        got, pl = gpu_3d([cf] * 3, dims, mo1, mo2, cast(x, prec))
        assert f"N={M}, pow2" in P.describe(pl.h), P.describe(pl.h)
        ref = orc.exec3d([C.CF_D] * 3, dims, mo1, mo2, wide(cast(x, prec)))
        assert C.rel_l2(got, ref) < TOL[prec], P.describe(pl.h)
    r2c = C.R2C_D if prec == 8 else C.R2C_S
    for dims, mo in (([2 * M, 3, 4], [0, 1, 2]), ([2 * M, 2, 16], [2, 1, 0])):
        log("case: fixed r2c", 2 * M, prec, dims, mo)
        x = np.random.default_rng(2 * M).uniform(-1, 1, orc.storage_shape(dims, mo))
        got, pl = gpu_3d([r2c, cf, cf], dims, mo, mo, cast(x, prec))
        assert f"r2c[axis 0, N={2 * M}, pow2" in P.describe(pl.h), P.describe(pl.h)
        ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], dims, mo, mo, wide(cast(x, prec)))
        assert C.rel_l2(got, ref) < TOL[prec], P.describe(pl.h)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("n", [129, 257, 513, 1025])
def test_dct1_complex_split_vs_oracle(n, prec):
    """Complex DCT-I of n points in the split form on the GPU (a (n-1)-point
    FFT and an (n-1)/2-point inverse FFT per line): tile order with paired
    lines (stride-1 tile axis, whole tiles), tile order with a partial tile,
    and line order, against the oracle."""
    tid = C.r2r_id("DCT1", True, prec)
    empty = C.EMPTY_DC if prec == 8 else C.EMPTY_SC
    for dims, mo1, mo2 in (([16, 3, n], [0, 1, 2], [0, 1, 2]), ([7, 3, n], [0, 1, 2], [2, 0, 1]),
                           ([n, 5, 3], [0, 1, 2], [0, 1, 2])):
        log("case: dct1 split", n, prec, dims, mo1, mo2)
        ax = dims.index(n)
        types = [empty if a != ax else tid for a in range(3)]
        rng = np.random.default_rng(n + prec)
        x = rng.standard_normal(orc.storage_shape(dims, mo1)) + 1j * rng.standard_normal(orc.storage_shape(dims, mo1))
        got, pl = gpu_3d(types, dims, mo1, mo2, cast(x, prec))
        assert "pow2 split" in P.describe(pl.h), P.describe(pl.h)
        ref = orc.exec_1d(C.r2r_id("DCT1", True, 8), ax, dims, mo1, dims, mo2, wide(cast(x, prec)))
        assert C.rel_l2(got, ref) < TOL[prec], P.describe(pl.h)
