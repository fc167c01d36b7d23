"""The product's planner and kernel code, built for the host and run one
workgroup at a time, against the oracle.  Catches indexing, radix, layout
and extension-rule mistakes on the CPU; the GPU tests then check the same
plans on the device."""
import numpy as np
import pytest

import cases as C
import emu_binding as emu
import oracle_binding as orc

TOL = {8: 1e-12, 4: 1e-5}


def run_both_3d(types, dims, mo1, mo2, prec, idir=-1, seed=0, c2r_len=None):
    od = C.out_dims(types, dims, c2r_len)
    x = C.random_input(types, dims, mo1, seed)
    if prec == 4:
        x = x.astype(np.complex64 if np.iscomplexobj(x) else np.float32)
    ref = orc.exec3d(types, dims, mo1, mo2, x.astype(np.complex128 if np.iscomplexobj(x) else np.float64),
                     idir=idir)
    got, desc = emu.exec3d(types, dims, mo1, od, mo2, x, prec, idir=idir)
    return got, ref, desc


@pytest.mark.parametrize("mo1", C.PERMS)
@pytest.mark.parametrize("mo2", C.PERMS)
def test_r2c_3d_all_memory_orders(mo1, mo2):
    got, ref, desc = run_both_3d([C.R2C_D, C.CF_D, C.CF_D], [12, 10, 6], mo1, mo2, 8)
    assert C.rel_l2(got, ref) < TOL[8], desc


@pytest.mark.parametrize("mo1", C.PERMS)
def test_c2r_3d_memory_orders(mo1):
    # backward from a spectrum; C2R length 12 on axis 0
    types = [C.C2R_D, C.CB_D, C.CB_D]
    dims = [7, 10, 6]
    for mo2 in ([0, 1, 2], [2, 1, 0], [1, 2, 0]):
        rng = np.random.default_rng(3)
        x = rng.standard_normal(orc.storage_shape(dims, mo1)) + 1j * rng.standard_normal(
            orc.storage_shape(dims, mo1))
        ref = orc.exec3d(types, dims, mo1, mo2, x)
        got, desc = emu.exec3d(types, dims, mo1, [12, 10, 6], mo2, x, 8)
        assert C.rel_l2(got, ref) < TOL[8], (mo1, mo2, desc)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("dims", [[16, 8, 4], [9, 7, 5], [32, 6, 3], [17, 11, 13], [64, 2, 1]])
def test_c2c_3d_sizes(prec, dims):
    t = [C.CF_D] * 3 if prec == 8 else [C.CF_S] * 3
    got, ref, desc = run_both_3d(t, dims, [0, 1, 2], [2, 1, 0], prec)
    assert C.rel_l2(got, ref) < TOL[prec], desc


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 15, 16, 17, 25, 31, 32, 37, 48, 60, 64,
                               96, 97, 120, 128, 139, 199, 256, 343, 385, 512])
def test_c2c_1d_lengths(n):
    dims = [n, 3, 2]
    mo = [0, 1, 2]
    rng = np.random.default_rng(n)
    x = rng.standard_normal(orc.storage_shape(dims, mo)) + 1j * rng.standard_normal(orc.storage_shape(dims, mo))
    for tid in (C.CF_D, C.CB_D):
        ref = orc.exec_1d(tid, 0, dims, mo, dims, mo, x)
        got = emu.exec1d(tid, 0, dims, mo, dims, mo, x, 8)
        assert C.rel_l2(got, ref) < TOL[8], (n, tid)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 9, 16, 17, 31, 64, 100, 139, 512])
def test_r2c_c2r_1d(n):
    dims = [n, 4, 3]
    h = n // 2 + 1
    rng = np.random.default_rng(n)
    for mo1, mo2 in (([0, 1, 2], [0, 1, 2]), ([1, 0, 2], [0, 2, 1]), ([0, 1, 2], [2, 1, 0])):
        x = rng.uniform(-1, 1, orc.storage_shape(dims, mo1))
        ref = orc.exec_1d(C.R2C_D, 0, dims, mo1, [h, 4, 3], mo2, x)
        got = emu.exec1d(C.R2C_D, 0, dims, mo1, [h, 4, 3], mo2, x, 8)
        assert C.rel_l2(got, ref) < TOL[8]
        z = rng.standard_normal(orc.storage_shape([h, 4, 3], mo2)) + 1j * rng.standard_normal(
            orc.storage_shape([h, 4, 3], mo2))
        ref = orc.exec_1d(C.C2R_D, 0, [h, 4, 3], mo2, dims, mo1, z)
        got = emu.exec1d(C.C2R_D, 0, [h, 4, 3], mo2, dims, mo1, z, 8)
        assert C.rel_l2(got, ref) < TOL[8]


KINDS = ["DCT1", "DST1", "DCT2", "DST2", "DCT3", "DST3", "DCT4", "DST4"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [2, 3, 5, 8, 16, 17, 33, 65, 100, 129, 513])
@pytest.mark.parametrize("cplx", [False, True])
def test_r2r_1d(kind, n, cplx):
    if kind in ("DCT1", "DCT4") and n < 2:
        pytest.skip("DCT-I needs n >= 2")
    tid = C.r2r_id(kind, cplx, 8)
    dims = [n, 3, 2]
    rng = np.random.default_rng(n + len(kind))
    for mo1, mo2 in (([0, 1, 2], [0, 1, 2]), ([2, 0, 1], [0, 2, 1])):
        shp = orc.storage_shape(dims, mo1)
        x = rng.uniform(-1, 1, shp) if not cplx else rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        ref = orc.exec_1d(tid, 0, dims, mo1, dims, mo2, x)
        got = emu.exec1d(tid, 0, dims, mo1, dims, mo2, x, 8)
        assert C.rel_l2(got, ref) < TOL[8], (kind, n, cplx, mo1, mo2)


@pytest.mark.parametrize("idir", [0, 1, 2])
@pytest.mark.parametrize("mo2", [[1, 2, 0], [0, 1, 2], [2, 1, 0]])
def test_derivative_3d(idir, mo2):
    got, ref, desc = run_both_3d([C.R2C_D, C.CF_D, C.CF_D], [12, 8, 6], [0, 1, 2], mo2, 8, idir=idir)
    assert C.rel_l2(got, ref) < TOL[8], desc


def test_derivative_odd_c2c():
    # odd C2C length: mid = (n-1)/2 is zeroed, as compute_deriv_loc does
    got, ref, desc = run_both_3d([C.CF_D] * 3, [9, 5, 7], [0, 1, 2], [0, 1, 2], 8, idir=2)
    assert C.rel_l2(got, ref) < TOL[8], desc


def test_mixed_r2c_c2c_dct1_complex():
    # the mixed-type configuration at reduced size: R2C x, C2C y, DCT-I z
    types = [C.R2C_D, C.CF_D, C.r2r_id("DCT1", True, 8)]
    for mo1, mo2 in (([0, 1, 2], [1, 2, 0]), ([2, 1, 0], [2, 1, 0])):
        got, ref, desc = run_both_3d(types, [24, 20, 17], mo1, mo2, 8)
        assert C.rel_l2(got, ref) < TOL[8], desc


def test_empty_axis_2d_transform():
    # test2D+empty.c: R2C x, EMPTY y, C2C z
    types = [C.R2C_D, C.EMPTY_DC, C.CF_D]
    got, ref, desc = run_both_3d(types, [10, 6, 8], [0, 1, 2], [1, 2, 0], 8)
    assert C.rel_l2(got, ref) < TOL[8], desc


def test_all_empty_reorder():
    types = [C.EMPTY_D] * 3
    dims = [5, 4, 3]
    x = np.random.default_rng(1).uniform(-1, 1, orc.storage_shape(dims, [0, 1, 2]))
    got, desc = emu.exec3d(types, dims, [0, 1, 2], dims, [2, 0, 1], x, 8, order=[0, 1, 2])
    ref = orc.to_storage(orc.from_storage(x, [0, 1, 2]), [2, 0, 1])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("prec", [4, 8])
def test_r2c_c2r_roundtrip_fp32_fp64(prec):
    tf = [C.R2C_D, C.CF_D, C.CF_D] if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
    tb = [C.C2R_D, C.CB_D, C.CB_D] if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
    dims = [32, 16, 8]
    mo1, mo2 = [2, 1, 0], [1, 2, 0]
    x = np.random.default_rng(7).uniform(-1, 1, orc.storage_shape(dims, mo1))
    spec, _ = emu.exec3d(tf, dims, mo1, [17, 16, 8], mo2, x, prec)
    back, _ = emu.exec3d(tb, [17, 16, 8], mo2, dims, mo1, spec, prec)
    assert C.rel_l2(back / np.prod(dims), x) < (1e-13 if prec == 8 else 1e-6)


@pytest.mark.parametrize("dims", [[32, 48, 64], [96, 120, 128], [128, 32, 96], [48, 48, 120]])
def test_batch_mode_plans(dims):
    """Passes planned for the batch (step) kernel: 256-thread workgroups,
    power-of-two tiles of up to 256 lines, for the batched-transform sizes."""
    x = (np.random.default_rng(sum(dims)).standard_normal(orc.storage_shape(dims, [0, 1, 2])) +
         1j * np.random.default_rng(1).standard_normal(orc.storage_shape(dims, [0, 1, 2])))
    for prec in (8, 4):
        t = [C.CF_D] * 3 if prec == 8 else [C.CF_S] * 3
        xi = x if prec == 8 else x.astype(np.complex64)
        got, desc = emu.exec3d(t, dims, [0, 1, 2], dims, [0, 1, 2], xi, prec, batch_threads=256)
        ref = orc.exec3d([C.CF_D] * 3, dims, [0, 1, 2], [0, 1, 2], xi.astype(np.complex128))
        assert C.rel_l2(got, ref) < TOL[prec], desc


@pytest.mark.parametrize("types,dims", [
    ([C.CF_S] * 3, [48, 96, 120]),          # generic path: twiddles
    ([C.R2C_D, C.CF_D, C.CF_D], [64, 32, 16]),  # power-of-two path: twiddles + split factors
    ([C.CF_D] * 3, [139, 8, 4]),            # Bluestein: chirp tables
    ([C.r2r_id("DCT2", False, 8)] * 3, [12, 10, 9]),  # post-twiddles
])
def test_launch_guard_requires_device_tables(types, dims):
    """Every transform pass has tables, so with the planner's null table
    addresses none may be launched; with addresses for all, all may."""
    od = C.out_dims(types, dims)
    n, before, after = emu.tables_ready_counts(types, dims, [0, 1, 2], od, [0, 1, 2], orc.type_info(types[0]).prec)
    assert n == 3 and before == 3 and after == 0


@pytest.mark.parametrize("B,pick", [(2, 0), (4, 1), (8, 2), (16, 3)])
@pytest.mark.parametrize("dims", [[48, 60, 36], [56, 40, 24]])
def test_autotuner_retiled_passes_keep_parity(dims, B, pick):
    """Every re-tiling the autotuner may choose (tile size, radix order) for
    single-buffer (48, 60, 36, 40, 24) and generic (56 = 8 * 7) passes
    gives the oracle's result."""
    t = [C.CF_D] * 3
    x = C.random_input(t, dims, [0, 1, 2], 11)
    got, n, desc = emu.exec3d_retuned(t, dims, [0, 1, 2], dims, [0, 1, 2], x, 8, B, pick)
    ref = orc.exec3d(t, dims, [0, 1, 2], [0, 1, 2], x)
    assert n >= 1, desc
    assert C.rel_l2(got, ref) < 1e-12, desc


def _run_case(types, dims, prec, seed, batch_threads=0):
    t = [x - 1 if prec == 4 else x for x in types]
    x = C.random_input(types, dims, [0, 1, 2], seed)
    od = C.out_dims(types, dims)
    xin = x.astype(np.complex64 if np.iscomplexobj(x) else np.float32) if prec == 4 else x
    got, desc = emu.exec3d(t, dims, [0, 1, 2], od, [0, 1, 2], xin, prec, batch_threads=batch_threads)
    ref = orc.exec3d(types, dims, [0, 1, 2], [0, 1, 2], x)
    return C.rel_l2(got, ref), desc


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("types,dims", [
    ([C.R2C_D, C.CF_D, C.CF_D], [232, 14, 26]),     # 8 * 29, 2 * 7, 2 * 13
    ([C.CF_D] * 3, [31, 17, 19]),                   # single primes
    ([C.R2C_D, C.CF_D, C.CF_D], [176, 22, 26]),     # radices 11 and 13
    ([C.r2r_id("DCT2", True, 8)] * 3, [23, 19, 29]),
    ([C.CF_D] * 3, [49, 121, 21]),                  # 7 * 7, 11 * 11, 3 * 7
])
def test_odd_prime_radices_single_buffer(types, dims, prec):
    """A length with one distinct odd prime factor 7..29 (FP64) or 7..47
    (FP32) runs on the single-buffer path with that prime as an in-register
    radix (its own kernel instantiation); FP64 31 runs on the generic path."""
    err, desc = _run_case(types, dims, prec, 5)
    assert "bluestein" not in desc, desc
    if prec == 4:
        assert "mid p" in desc, desc
    assert err < (1e-12 if prec == 8 else 1e-5), desc


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("types,dims", [
    ([C.R2C_D, C.CF_D, C.CF_D], [232, 14, 26]),
    ([C.CF_D] * 3, [31, 17, 19]),
    ([C.CF_D] * 3, [77, 91, 8]),                    # 7 * 11, 7 * 13: two distinct primes
])
def test_odd_prime_radices_generic(types, dims, prec):
    """Odd prime radices 7..31 as direct DFTs parallel over outputs on the
    two-buffer generic path (batch plans, and lengths with two distinct odd
    primes)."""
    bt = 0 if dims[0] == 77 else 256
    err, desc = _run_case(types, dims, prec, 5, batch_threads=bt)
    assert "bluestein" not in desc and "mid p" not in desc, desc
    assert err < (1e-12 if prec == 8 else 1e-5), desc


@pytest.mark.parametrize("types,dims", [
    ([C.R2C_D, C.CF_D, C.CF_D], [344, 6, 4]),       # 8 * 43
    ([C.CF_D] * 3, [37, 41, 47]),
    ([C.r2r_id("DST1", False, 8)] * 3, [36, 7, 5]),  # W = 74 = 2 * 37
])
def test_large_prime_radices_fp32(types, dims):
    """FP32 primes 37..47 run as in-register radices, not Bluestein."""
    err, desc = _run_case(types, dims, 4, 6)
    assert "bluestein" not in desc and "mid p" in desc, desc
    assert err < 1e-5, desc


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("types,dims,bt", [
    ([C.R2C_D, C.CF_D, C.CF_D], [344, 6, 4], 0),    # FP64: 43 above the in-register limit
    ([C.CF_D] * 3, [53, 59, 61], 0),                # primes above 47
    ([C.CF_D] * 3, [37, 41, 47], 256),              # batch plans
    ([C.r2r_id("DST1", False, 8)] * 3, [36, 7, 5], 256),
])
def test_bluestein_single_buffer(types, dims, bt, prec):
    """Other large primes take Bluestein on the single-buffer path."""
    if prec == 4 and bt == 0 and dims[0] == 344:
        pytest.skip("FP32 takes 43 as a radix")
    err, desc = _run_case(types, dims, prec, 6, batch_threads=bt)
    assert "bluestein" in desc, desc
    assert err < (1e-12 if prec == 8 else 1e-5), desc


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("dims", [[24, 7, 5], [30, 3, 1], [106, 5, 3], [344, 3, 2], [250, 9, 4], [6, 1, 1]])
@pytest.mark.parametrize("mo2", [[0, 1, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]])
def test_r2c_c2r_paired_lines(dims, mo2, prec):
    """R2C and C2R on the single-buffer path transform two real lines as one
    complex line: odd line counts (the last line unpaired), tile-order and
    line-order sides, Bluestein (106 = 2 * 53) and prime radices (344 = 8 * 43,
    250 = 2 * 5^3); the C2R input is arbitrary complex data, whose DC and
    Nyquist imaginary parts a real inverse transform ignores."""
    fwd = [C.R2C_D, C.CF_D, C.CF_D]
    bwd = [C.C2R_D, C.CB_D, C.CB_D]
    tf = fwd if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
    tb = bwd if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
    tol = 1e-12 if prec == 8 else 1e-5
    x = C.random_input(fwd, dims, [0, 1, 2], 21)
    od = C.out_dims(fwd, dims)
    xin = x if prec == 8 else x.astype(np.float32)
    got, desc = emu.exec3d(tf, dims, [0, 1, 2], od, mo2, xin, prec)
    ref = orc.exec3d(fwd, dims, [0, 1, 2], mo2, x)
    assert "pair" in desc.split(" | ")[0], desc
    assert C.rel_l2(got, ref) < tol, desc
    z = C.random_input(bwd, od, mo2, 22)
    zin = z if prec == 8 else z.astype(np.complex64)
    got, desc = emu.exec3d(tb, od, mo2, dims, [0, 1, 2], zin, prec)
    ref = orc.exec3d(bwd, od, mo2, [0, 1, 2], z)
    assert C.rel_l2(got, ref) < tol, desc


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("dims,mo1,mo2", [([64, 32, 64], [2, 1, 0], [1, 2, 0]), ([128, 16, 32], [2, 1, 0], [2, 1, 0]),
                                          ([32, 8, 48], [1, 2, 0], [1, 2, 0])])
def test_r2c_tile_load_real_pairs(dims, mo1, mo2, prec):
    """Power-of-two R2C with a tile-order load along a stride-1 axis reads two
    adjacent lines' reals per load, and C2R with a tile-order store writes
    two per store, when every tile is whole and the strides are even (the
    first two cases); the third has a partial last tile and moves one real
    at a time."""
    fwd = [C.R2C_D, C.CF_D, C.CF_D]
    t = fwd if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
    x = C.random_input(fwd, dims, mo1, 3)
    od = C.out_dims(fwd, dims)
    got, desc = emu.exec3d(t, dims, mo1, od, mo2, x if prec == 8 else x.astype(np.float32), prec)
    assert "pow2" in desc.split(" | ")[0], desc
    assert C.rel_l2(got, orc.exec3d(fwd, dims, mo1, mo2, x)) < (1e-12 if prec == 8 else 1e-5), desc
    # and back: the C2R pass's tile-order store writes two lines' reals per store
    bwd = [C.C2R_D, C.CB_D, C.CB_D]
    tb = bwd if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
    z = C.random_input(bwd, od, mo2, 4)
    got, desc = emu.exec3d(tb, od, mo2, dims, mo1, z if prec == 8 else z.astype(np.complex64), prec)
    assert "pow2" in desc.split(" | ")[-1], desc
    assert C.rel_l2(got, orc.exec3d(bwd, od, mo2, mo1, z)) < (1e-12 if prec == 8 else 1e-5), desc


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("mo1,mo2", [([2, 1, 0], [1, 2, 0]), ([0, 1, 2], [0, 1, 2]), ([1, 0, 2], [2, 0, 1])])
def test_padded_intermediate_arrays(prec, mo1, mo2):
    """Intermediate arrays with padded rows (the planner pads large ones;
    here every eligible one): rows of 32 FP32 or 16 FP64 complex values are
    a multiple of 256 bytes and get 128 bytes of padding.  Forward R2C and
    C2C, backward C2R, against the oracle."""
    emu.set_pad_threshold(0)
    try:
        s = "D" if prec == 8 else "S"
        n = 32 if prec == 4 else 16
        dims = [2 * (n - 1), n, n]  # R2C x: nh = n
        x = np.random.default_rng(prec).uniform(-1, 1, orc.storage_shape(dims, mo1))
        tf = [C.R2C_D, C.CF_D, C.CF_D] if prec == 8 else [C.R2C_S, C.CF_S, C.CF_S]
        tb = [C.C2R_D, C.CB_D, C.CB_D] if prec == 8 else [C.C2R_S, C.CB_S, C.CB_S]
        sd = [n, n, n]
        ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], dims, mo1, mo2, x)
        spec, desc = emu.exec3d(tf, dims, mo1, sd, mo2, x, prec)
        assert "pitch" in desc, desc
        assert C.rel_l2(spec, ref) < TOL[prec], desc
        back, desc_b = emu.exec3d(tb, sd, mo2, dims, mo1, spec, prec)
        assert C.rel_l2(back / np.prod(dims), x) < (1e-13 if prec == 8 else 1e-6), desc_b
        xc = (np.random.default_rng(5).standard_normal(orc.storage_shape(sd, mo1)) +
              1j * np.random.default_rng(6).standard_normal(orc.storage_shape(sd, mo1)))
        cf = [C.CF_D] * 3 if prec == 8 else [C.CF_S] * 3
        got, desc_c = emu.exec3d(cf, sd, mo1, sd, mo2, xc, prec)
        ref_c = orc.exec3d([C.CF_D] * 3, sd, mo1, mo2, xc)
        assert C.rel_l2(got, ref_c) < TOL[prec], desc_c
    finally:
        emu.set_pad_threshold(32 << 20)



@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("M", [48, 96, 192, 384, 768, 80, 160, 320, 640, 120, 60, 72, 144, 112, 224, 240, 36])
def test_fixed_length_odd_radix(M, prec):
    """Compile-time lengths 3 and 5 times a power of two (the odd radix pass
    has ceil(16 / R) butterfly slots per thread): C2C along a stride-1 axis
    (line order) and a strided axis (tile order, a partial last tile), R2C
    of length 2M and its C2R, against the oracle."""
    s = "D" if prec == 8 else "S"
    cf = C.CF_D if prec == 8 else C.CF_S
    cases = [([M, 3, 2], [0, 1, 2], [0, 1, 2]), ([5, M, 3], [0, 1, 2], [2, 0, 1])]
    for dims, mo1, mo2 in cases:
        x = (np.random.default_rng(M).standard_normal(orc.storage_shape(dims, mo1)) +
             1j * np.random.default_rng(M + 1).standard_normal(orc.storage_shape(dims, mo1)))
        got, desc = emu.exec3d([cf, cf, cf], dims, mo1, dims, mo2, x, prec)
        assert f"N={M}, pow2" in desc, desc
        ref = orc.exec3d([C.CF_D] * 3, dims, mo1, mo2, x)
        assert C.rel_l2(got, ref) < TOL[prec], desc
    r2c = C.R2C_D if prec == 8 else C.R2C_S
    c2r = C.C2R_D if prec == 8 else C.C2R_S
    for dims, mo1, mo2 in (([2 * M, 3, 2], [0, 1, 2], [0, 1, 2]), ([2 * M, 6, 3], [2, 1, 0], [2, 1, 0])):
        x = np.random.default_rng(2 * M).uniform(-1, 1, orc.storage_shape(dims, mo1))
        od = [M + 1, dims[1], dims[2]]
        spec, desc = emu.exec3d([r2c, cf, cf], dims, mo1, od, mo2, x, prec)
        assert f"r2c[axis 0, N={2 * M}, pow2" in desc, desc
        ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], dims, mo1, mo2, x)
        assert C.rel_l2(spec, ref) < TOL[prec], desc
        cb = C.CB_D if prec == 8 else C.CB_S
        back, desc_b = emu.exec3d([c2r, cb, cb], od, mo2, dims, mo1, spec, prec)
        assert C.rel_l2(back / np.prod(dims), x) < (1e-13 if prec == 8 else 1e-6), desc_b


@pytest.mark.parametrize("n", [9, 17, 49, 129, 513])
@pytest.mark.parametrize("prec", [4, 8])
def test_fixed_length_dct1_complex(n, prec):
    """Complex DCT-I of n points on the compile-time-length path (the
    2 (n - 1)-point FFT of the even extension): line order (d stride-1),
    tile order with paired lines (e stride-1, whole tiles) and tile order
    with a partial last tile, against the oracle."""
    tid = C.r2r_id("DCT1", True, prec)
    cases = [([n, 6, 3], [0, 1, 2], [0, 1, 2]),   # line order both sides
             ([16, 5, n], [0, 1, 2], [0, 1, 2]),  # tile order, e stride-1, whole tiles
             ([7, 3, n], [0, 1, 2], [2, 0, 1]),   # tile load (partial tile), line store
             ([n, 5, 4], [1, 2, 0], [0, 2, 1])]   # tile load, line store
    rng = np.random.default_rng(n)
    for dims, mo1, mo2 in cases:
        shp = orc.storage_shape(dims, mo1)
        x = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        ax = dims.index(n)
        empty = C.EMPTY_DC if prec == 8 else C.EMPTY_SC
        got, desc = emu.exec3d([empty if a != ax else tid for a in range(3)], dims, mo1, dims, mo2, x, prec)
        if n - 1 >= 8:
            assert f"r2r[axis {ax}, N={n}, pow2" in desc, desc
        ref = orc.exec_1d(C.r2r_id("DCT1", True, 8), ax, dims, mo1, dims, mo2, x)
        assert C.rel_l2(got, ref) < TOL[prec], desc


@pytest.mark.parametrize("n", [129, 257, 513, 1025])
@pytest.mark.parametrize("prec", [4, 8])
def test_dct1_complex_split(n, prec):
    """Complex DCT-I of n = P + 1 points in the split form (a P-point FFT of
    the even extension of x_n + x_{P-n} and a P/2-point inverse FFT for the
    odd outputs): line order, tile order with whole and partial tiles,
    against the oracle and against the unsplit form."""
    tid = C.r2r_id("DCT1", True, prec)
    empty = C.EMPTY_DC if prec == 8 else C.EMPTY_SC
    cases = [([n, 5, 3], [0, 1, 2], [0, 1, 2]),   # line order both sides
             ([16, 3, n], [0, 1, 2], [0, 1, 2]),  # tile order, whole tiles
             ([7, 3, n], [0, 1, 2], [2, 0, 1]),   # tile load (partial tile), line store
             ([n, 5, 4], [1, 2, 0], [0, 2, 1])]   # tile load, line store
    rng = np.random.default_rng(n + prec)
    for dims, mo1, mo2 in cases:
        shp = orc.storage_shape(dims, mo1)
        x = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        ax = dims.index(n)
        types = [empty if a != ax else tid for a in range(3)]
        got, desc = emu.exec3d(types, dims, mo1, dims, mo2, x, prec)
        assert f"r2r[axis {ax}, N={n}, pow2 split" in desc, desc
        ref = orc.exec_1d(C.r2r_id("DCT1", True, 8), ax, dims, mo1, dims, mo2, x)
        assert C.rel_l2(got, ref) < TOL[prec], desc
        emu.set_dct_split(False)
        try:
            plain, desc_p = emu.exec3d(types, dims, mo1, dims, mo2, x, prec)
        finally:
            emu.set_dct_split(True)
        assert "split" not in desc_p, desc_p
        assert C.rel_l2(got, plain) < TOL[prec], (desc, desc_p)


@pytest.mark.parametrize("prec", [4, 8])
@pytest.mark.parametrize("dims", [[64, 4, 16], [512, 2, 64], [96, 3, 32], [768, 2, 8]])
def test_r2c_tile_order_paired_store(dims, prec):
    """R2C along the slowest axis with a stride-1 tile axis and whole tiles
    (the paired tile-order store: outputs k and M - k of two adjacent lines
    per entry), against the oracle."""
    mo = [2, 1, 0]
    x = np.random.default_rng(dims[0]).uniform(-1, 1, orc.storage_shape(dims, mo))
    r2c = C.R2C_D if prec == 8 else C.R2C_S
    cf = C.CF_D if prec == 8 else C.CF_S
    od = [dims[0] // 2 + 1, dims[1], dims[2]]
    got, desc = emu.exec3d([r2c, cf, cf], dims, mo, od, mo, x, prec)
    assert f"r2c[axis 0, N={dims[0]}, pow2" in desc, desc
    ref = orc.exec3d([C.R2C_D, C.CF_D, C.CF_D], dims, mo, mo, x)
    assert C.rel_l2(got, ref) < TOL[prec], desc


def test_planner_avoids_ragged_tile_axes():
    """Intermediate memory orders of equal primary cost: the planner takes
    the one with the fewest passes whose tile axis extent is not a multiple
    of 8.  R2C x / C2C y / DCT-I z with an odd spectrum x extent (25) runs
    its DCT-I and C2C passes along the y axis (40); a grid whose axes are
    all multiples of 8 keeps the input's order."""
    types = [C.R2C_S, C.CF_S, C.r2r_id("DCT1", True, 4)]
    dims = [48, 40, 33]
    x = np.random.default_rng(7).uniform(-1, 1, orc.storage_shape(dims, [0, 1, 2]))
    got, desc = emu.exec3d(types, dims, [0, 1, 2], [25, 40, 33], [0, 1, 2], x, 4)
    assert "r2c[axis 0, N=48, mid pair, mo (012)->(102)" in desc, desc
    assert "mo (102)->(102)" in desc and "mo (102)->(012)" in desc, desc
    ref = orc.exec3d([C.R2C_D, C.CF_D, C.r2r_id("DCT1", True, 8)], dims, [0, 1, 2], [0, 1, 2], x)
    assert C.rel_l2(got, ref) < TOL[4], desc
    back, desc_b = emu.exec3d([C.C2R_S, C.CB_S, C.r2r_id("DCT1", True, 4)], [25, 40, 33], [0, 1, 2], dims,
                              [0, 1, 2], got, 4)
    assert "mo (012)->(102)" in desc_b and "mo (102)->(012)" in desc_b, desc_b
    dims = [64, 64, 64]
    x = np.random.default_rng(8).uniform(-1, 1, orc.storage_shape(dims, [2, 1, 0]))
    got, desc = emu.exec3d([C.R2C_S, C.CF_S, C.CF_S], dims, [2, 1, 0], [33, 64, 64], [1, 2, 0], x, 4)
    assert "mo (210)->(210)" in desc and "mo (210)->(120)" in desc, desc
