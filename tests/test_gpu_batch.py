"""Batched 3D transforms (include/p3dfft_hip.h, p3dfft_hip_batch_*) against
the FP64 oracle, item by item.

The batch workload: 256 C2C forward transforms, each axis drawn from
{32, 48, 64, 96, 120, 128} (seed 4), complex64 N(0,1) input.  Each item is
checked against the FP64 oracle on the same (widened) input, FP32 bar
1e-5 relative L2; FP64 batches use the 1e-12 bar.  Further cases: graph
replay against eager execution, the autotuner's choices, in-place items,
and that an item's result does not depend on what else is in the batch.
"""
import os
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import cases as C
import oracle_binding as orc
import p3dfft3 as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SIZES = [32, 48, 64, 96, 120, 128]
TOL = {8: 1e-12, 4: 1e-5}


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def lib_setup(tmp_path_factory):
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    os.environ["P3DFFT_HIP_TUNE_CACHE"] = str(tmp_path_factory.mktemp("tune") / "tune.txt")
    P.setup()
    yield
    torch.cuda.synchronize()


def batch_dims(n, seed, sizes=SIZES):
    return np.random.default_rng(seed).choice(sizes, size=(n, 3)).tolist()


def batch_inputs(dims, seed, prec, real=False):
    rng = np.random.default_rng(seed)
    xs = []
    for d in dims:
        shape = orc.storage_shape(d, [0, 1, 2])
        if real:
            x = rng.standard_normal(shape)
            xs.append(x.astype(np.float32 if prec == 4 else np.float64))
        else:
            x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
            xs.append(x.astype(np.complex64 if prec == 4 else np.complex128))
    return xs


def oracle_all(types_d, dims, mo1, mo2, xs):
    """FP64 oracle on every item; the oracle's C calls release the GIL."""
    wide = [x.astype(np.complex128 if np.iscomplexobj(x) else np.float64) for x in xs]
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda a: orc.exec3d(types_d, a[0], mo1, mo2, a[1]), zip(dims, wide)))


def run_batch(types, dims, odims, mo1, mo2, xs, out_complex, prec, tune=False, inplace=False, graph=False,
              reps=1):
    h = P.batch_plan(dims, odims, types, mo1, mo2, tune)
    try:
        if graph:
            P.batch_set_graph(h, True)
        din = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in xs]
        if inplace:
            dout = din
        else:
            cdt = torch.complex128 if prec == 8 else torch.complex64
            rdt = torch.float64 if prec == 8 else torch.float32
            dout = [torch.empty(orc.storage_shape(o, mo2), dtype=cdt if out_complex else rdt, device="cuda")
                    for o in odims]
        torch.cuda.synchronize()
        for _ in range(reps):
            P.batch_exec(h, [t.data_ptr() for t in din], [t.data_ptr() for t in dout], None, True)
        return [t.cpu().numpy() for t in dout], P.batch_describe(h)
    finally:
        P.batch_destroy(h)


def test_batch_256_c2c_fp32_vs_oracle():
    """The full 256-item mixed-size C2C FP32 batch, per-item FP64 oracle error."""
    dims = batch_dims(256, 4)
    log("case: batch 256 c2c fp32 seed 4, points", sum(int(np.prod(d)) for d in dims))
    xs = batch_inputs(dims, 4, 4)
    got, desc = run_batch([C.CF_S] * 3, dims, dims, [0, 1, 2], [0, 1, 2], xs, True, 4)
    log("  ", desc)
    ref = oracle_all([C.CF_D] * 3, dims, [0, 1, 2], [0, 1, 2], xs)
    errs = [C.rel_l2(g, r) for g, r in zip(got, ref)]
    worst = int(np.argmax(errs))
    log("  max rel L2", errs[worst], "item", worst, dims[worst], "median", float(np.median(errs)))
    assert max(errs) < TOL[4]


@pytest.mark.parametrize("mo2", [[0, 1, 2], [2, 1, 0], [1, 2, 0]])
def test_batch_c2c_fp64_memory_orders(mo2):
    dims = batch_dims(12, 40, [8, 12, 17, 32, 48, 64])
    log("case: batch 12 c2c fp64 ->", mo2)
    xs = batch_inputs(dims, 41, 8)
    got, desc = run_batch([C.CF_D] * 3, dims, dims, [0, 1, 2], mo2, xs, True, 8)
    ref = oracle_all([C.CF_D] * 3, dims, [0, 1, 2], mo2, xs)
    errs = [C.rel_l2(g, r) for g, r in zip(got, ref)]
    log("  max rel L2", max(errs), desc)
    assert max(errs) < TOL[8]


@pytest.mark.parametrize("prec", [4, 8])
def test_batch_r2c_vs_oracle(prec):
    """Real-to-complex batches (the Poisson workload's forward transform)."""
    dims = [[64, 64, 64], [96, 72, 80], [128, 64, 96], [40, 40, 48], [200, 136, 64]]
    log("case: batch r2c prec", prec)
    xs = batch_inputs(dims, 5, prec, real=True)
    t = [C.R2C_S, C.CF_S, C.CF_S] if prec == 4 else [C.R2C_D, C.CF_D, C.CF_D]
    odims = [C.out_dims(t, d) for d in dims]
    got, desc = run_batch(t, dims, odims, [0, 1, 2], [0, 1, 2], xs, True, prec)
    ref = oracle_all([C.R2C_D, C.CF_D, C.CF_D], dims, [0, 1, 2], [0, 1, 2], xs)
    errs = [C.rel_l2(g, r) for g, r in zip(got, ref)]
    log("  max rel L2", max(errs), desc)
    assert max(errs) < TOL[prec]
    # and back: C2R batch of the oracle spectra, scaled, returns the input
    tb = [C.C2R_S, C.CB_S, C.CB_S] if prec == 4 else [C.C2R_D, C.CB_D, C.CB_D]
    specs = [r.astype(np.complex64 if prec == 4 else np.complex128) for r in ref]
    back, _ = run_batch(tb, odims, dims, [0, 1, 2], [0, 1, 2], specs, False, prec)
    errs = [C.rel_l2(b / np.prod(d), x) for b, d, x in zip(back, dims, xs)]
    log("  round trip max rel L2", max(errs))
    assert max(errs) < TOL[prec]


def test_batch_graph_replays_match_eager():
    """Graph capture on the first exec, replays after: every replay, with
    eager work and synchronization between, equals the eager batch."""
    dims = batch_dims(16, 6)
    log("case: batch graph 16 c2c fp32")
    types = [C.CF_S] * 3
    h_eager = P.batch_plan(dims, dims, types)
    h_graph = P.batch_plan(dims, dims, types)
    P.batch_set_graph(h_graph, True)
    try:
        for rep in range(4):
            xs = batch_inputs(dims, 100 + rep, 4)
            din = [torch.from_numpy(x).cuda() for x in xs]
            e_out = [torch.empty_like(t) for t in din]
            g_out = [torch.empty_like(t) for t in din]
            P.batch_exec(h_eager, [t.data_ptr() for t in din], [t.data_ptr() for t in e_out], None, True)
            junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
            P.batch_exec(h_graph, [t.data_ptr() for t in din], [t.data_ptr() for t in g_out], None, True)
            torch.cuda.synchronize()
            assert float(junk) == float(junk)
            for a, b in zip(e_out, g_out):
                assert torch.equal(a, b), f"replay {rep} differs from eager"
    finally:
        P.batch_destroy(h_eager)
        P.batch_destroy(h_graph)


def test_batch_tuned_vs_oracle():
    """Autotuned tiles and radix orders (non-power-of-two sizes: single-buffer
    passes, and a generic-path pass for the radix 7 in 56) keep parity, and a
    second plan reads the same choices from the cache."""
    dims = [[48, 96, 120], [120, 48, 96], [96, 120, 48], [60, 36, 100], [56, 48, 40]]
    log("case: batch tuned c2c fp32")
    xs = batch_inputs(dims, 7, 4)
    got, desc = run_batch([C.CF_S] * 3, dims, dims, [0, 1, 2], [0, 1, 2], xs, True, 4, tune=True)
    log("  ", desc)
    ref = oracle_all([C.CF_D] * 3, dims, [0, 1, 2], [0, 1, 2], xs)
    errs = [C.rel_l2(g, r) for g, r in zip(got, ref)]
    assert max(errs) < TOL[4]
    assert os.path.exists(os.environ["P3DFFT_HIP_TUNE_CACHE"])
    got2, _ = run_batch([C.CF_S] * 3, dims, dims, [0, 1, 2], [0, 1, 2], xs, True, 4, tune=True)
    for a, b in zip(got, got2):
        assert np.array_equal(a, b)


def test_batch_inplace():
    dims = batch_dims(8, 8)
    log("case: batch in place c2c fp64")
    xs = batch_inputs(dims, 9, 8)
    got, _ = run_batch([C.CF_D] * 3, dims, dims, [0, 1, 2], [0, 1, 2], xs, True, 8, inplace=True)
    ref = oracle_all([C.CF_D] * 3, dims, [0, 1, 2], [0, 1, 2], xs)
    assert max(C.rel_l2(g, r) for g, r in zip(got, ref)) < TOL[8]


def test_batch_item_alone_equals_item_in_mix():
    """An item's bits do not depend on its batch neighbours."""
    dims = batch_dims(24, 10)
    log("case: batch alone vs mix")
    xs = batch_inputs(dims, 11, 4)
    mix, _ = run_batch([C.CF_S] * 3, dims, dims, [0, 1, 2], [0, 1, 2], xs, True, 4)
    for i in (0, 7, 23):
        alone, _ = run_batch([C.CF_S] * 3, [dims[i]], [dims[i]], [0, 1, 2], [0, 1, 2], [xs[i]], True, 4)
        assert np.array_equal(alone[0], mix[i]), f"item {i} {dims[i]}"


@pytest.mark.parametrize("prec", [4, 8])
def test_batch_mixed_type_dct1_vs_oracle(prec):
    """Batches of the mixed-type transform (R2C x, C2C y, DCT1_COMPLEX z) of
    different sizes: z lengths with the split DCT-I (129, 257, 513) and
    without it (65, 97) share grouped launches with C2C passes of the same
    kernel length; each item against the FP64 oracle."""
    dims = [[48, 40, 129], [64, 32, 257], [32, 24, 513], [40, 48, 65], [96, 16, 97], [24, 40, 257]]
    log("case: batch mixed-type dct1 prec", prec)
    xs = batch_inputs(dims, 9, prec, real=True)
    d1 = C.r2r_id("DCT1", True, prec)
    t = [C.R2C_S, C.CF_S, d1] if prec == 4 else [C.R2C_D, C.CF_D, d1]
    odims = [C.out_dims(t, d) for d in dims]
    got, desc = run_batch(t, dims, odims, [0, 1, 2], [0, 1, 2], xs, True, prec)
    log("  ", desc)
    ref = oracle_all([C.R2C_D, C.CF_D, C.r2r_id("DCT1", True, 8)], dims, [0, 1, 2], [0, 1, 2], xs)
    errs = [C.rel_l2(g, r) for g, r in zip(got, ref)]
    log("  max rel L2", max(errs))
    assert max(errs) < TOL[prec]
