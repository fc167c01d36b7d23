"""Periodic Poisson solves on the GPU (include/p3dfft_hip.h: the Poisson
service and the per-request path) against the FP64 oracle solve
(cases.poisson_ref: the oracle's 3D R2C/C2R transforms with the -1/|k|^2
multiply done in numpy FP64).

Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5 against the FP64 oracle;
a request's result is bit-identical alone and in a mix of other requests.
"""
import sys
import time

import numpy as np
import pytest

import cases as C
import oracle_binding as orc
import p3dfft3 as P

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = {8: 1e-12, 4: 1e-5}
# grids of 2^18 points or more run dedicated pass launches in the service,
# smaller ones ([40, 56, 48], [48, 40, 36], [56, 32, 40]) share step-kernel launches
SIZES = [[64, 64, 64], [72, 48, 96], [40, 56, 48], [74, 64, 64], [128, 64, 96], [88, 88, 48], [48, 40, 36],
         [56, 32, 40]]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def lib_setup():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.setup()
    yield
    torch.cuda.synchronize()


def rhs_of(dims, seed, prec):
    x = orc.to_storage(C.gaussians_rhs(dims, seed), [0, 1, 2])
    t = torch.from_numpy(np.ascontiguousarray(x)).to("cuda", torch.float64 if prec == 8 else torch.float32)
    torch.cuda.synchronize()
    return x, t


_ref_cache = {}


def oracle(dims, seed):
    key = (tuple(dims), seed)
    if key not in _ref_cache:
        x = orc.to_storage(C.gaussians_rhs(dims, seed), [0, 1, 2])
        _ref_cache[key] = C.poisson_ref(dims, x)[0]
    return _ref_cache[key]


def run_server(reqs, prec, pool_bytes=4 << 30, max_inflight=64):
    """reqs: [(dims, seed)] submitted back to back; returns (solutions, records)."""
    h = P.poisson_server_create(prec, pool_bytes, max_inflight)
    try:
        data = [rhs_of(d, s, prec) for d, s in reqs]
        sols = [torch.empty_like(t) for _, t in data]
        torch.cuda.synchronize()
        ids = [P.poisson_submit(h, d, t.data_ptr(), o.data_ptr()) for (d, _), (_, t), o in zip(reqs, data, sols)]
        recs = {}
        deadline = time.time() + 60
        while len(recs) < len(ids):
            assert time.time() < deadline, "requests did not complete"
            for r in P.poisson_poll(h, wait_ms=1000):
                recs[r[0]] = r
        assert all(recs[i][4] for i in ids), "a request failed"
        st = P.poisson_server_stats(h)
        return [o.cpu().numpy() for o in sols], [recs[i] for i in ids], st
    finally:
        P.poisson_server_destroy(h)


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("dims", SIZES[:4])
def test_per_request_solve_vs_oracle(dims, prec):
    log("case: per-request poisson", dims, prec)
    _, t = rhs_of(dims, 1, prec)
    out = torch.empty_like(t)
    P.poisson_solve(dims, t.data_ptr(), out.data_ptr(), prec)
    err = C.rel_l2(out.cpu().numpy(), oracle(dims, 1))
    log("  rel L2", err)
    assert err < TOL[prec]


@pytest.mark.parametrize("prec", [4, 8])
def test_service_mixed_sizes_vs_oracle(prec):
    reqs = [(SIZES[i % len(SIZES)], 1) for i in range(12)]
    log("case: service", prec, len(reqs), "requests")
    sols, recs, st = run_server(reqs, prec)
    log("  stats", st)
    errs = [C.rel_l2(s, oracle(d, seed)) for s, (d, seed) in zip(sols, reqs)]
    log("  max rel L2", max(errs))
    assert max(errs) < TOL[prec]
    # six passes per request, whether a request ran alone or in a group
    assert 0 < st["launches"] <= 12 * 6


def test_service_groups_small_requests():
    """Small grids submitted back to back share step-kernel launches (six
    per group instead of six per request), with every result the same as
    the request alone."""
    d = [48, 40, 36]
    alone, _, _ = run_server([(d, 5)], 4)
    reqs = [(d, 5)] * 32
    sols, _, st = run_server(reqs, 4)
    log("  stats", st)
    assert st["launches"] < 32 * 6
    for s in sols:
        assert np.array_equal(s, alone[0])


def test_service_alone_equals_mix():
    d = [72, 48, 96]
    log("case: alone vs mix")
    alone, _, _ = run_server([(d, 3)], 4)
    mix_reqs = [(SIZES[i], 7 + i) for i in range(len(SIZES))]
    mix_reqs.insert(3, (d, 3))
    mix, recs, _ = run_server(mix_reqs, 4)
    assert np.array_equal(alone[0], mix[3])
    # and a request that arrives while others are in flight
    h = P.poisson_server_create(4, 4 << 30, 64)
    try:
        big = [rhs_of([128, 128, 128], 11 + i, 4)[1] for i in range(4)]
        outs = [torch.empty_like(t) for t in big]
        _, t = rhs_of(d, 3, 4)
        o = torch.empty_like(t)
        torch.cuda.synchronize()
        for b, ob in zip(big, outs):
            P.poisson_submit(h, [128, 128, 128], b.data_ptr(), ob.data_ptr())
        rid = P.poisson_submit(h, d, t.data_ptr(), o.data_ptr())
        got = {}
        deadline = time.time() + 60
        while len(got) < 5:
            assert time.time() < deadline, "requests did not complete"
            for r in P.poisson_poll(h, wait_ms=1000):
                got[r[0]] = r
        assert np.array_equal(o.cpu().numpy(), alone[0])
    finally:
        P.poisson_server_destroy(h)


def test_service_memory_admission():
    """A pool that holds one 128x64x96 request's working memory: requests run
    one at a time, all complete, results unchanged."""
    d = [128, 64, 96]
    work = 2 * (((d[0] // 2 + 1) * d[1] * d[2] * 8 + 255) // 256 * 256)
    reqs = [(d, 20 + i) for i in range(4)]
    log("case: admission, pool", work)
    sols, recs, _ = run_server(reqs, 4, pool_bytes=work + 4096)
    spans = sorted((r[2], r[3]) for r in recs)
    for (a0, d0), (a1, d1) in zip(spans, spans[1:]):
        assert a1 >= d0, "two requests held working memory at once"
    full, _, _ = run_server(reqs, 4)
    for a, b in zip(sols, full):
        assert np.array_equal(a, b)


def test_service_rejects_oversized_request():
    h = P.poisson_server_create(4, 1 << 20, 4)
    try:
        _, t = rhs_of([64, 64, 64], 1, 4)
        o = torch.empty_like(t)
        with pytest.raises(RuntimeError):
            P.poisson_submit(h, [64, 64, 64], t.data_ptr(), o.data_ptr())
    finally:
        P.poisson_server_destroy(h)


def test_service_large_grid_dedicated_launches():
    """A grid of at least 2^18 points runs its passes as their own launches,
    mixed with small requests in the same steps: parity with the oracle, and
    the same bits alone and in the mix."""
    big = [256, 256, 128]
    log("case: service large grid", big)
    alone, _, _ = run_server([(big, 30)], 4)
    reqs = [(SIZES[6], 31), (big, 30), (SIZES[1], 32), (SIZES[2], 33), (SIZES[7], 34)]
    mix, _, st = run_server(reqs, 4)
    log("  steps", st)
    assert np.array_equal(alone[0], mix[1])
    err = C.rel_l2(mix[1], oracle(big, 30))
    log("  rel L2 vs oracle", err)
    assert err < TOL[4]
    for s, (d, seed) in zip(mix, reqs):
        assert C.rel_l2(s, oracle(d, seed)) < TOL[4]


def test_large_edges_vs_oracle():
    """Edges at the top of the serving workload's range, with the largest
    in-register prime radices (344 = 8 * 43, 376 = 8 * 47, 368 = 16 * 23):
    the per-request path and the service against the oracle's solve, and the
    two paths give the same bits (same plans, same kernels)."""
    d = [344, 376, 368]
    log("case: poisson large edges", d)
    _, t = rhs_of(d, 40, 4)
    out = torch.empty_like(t)
    torch.cuda.synchronize()
    P.poisson_solve(d, t.data_ptr(), out.data_ptr(), 4)
    ref = oracle(d, 40)
    e1 = C.rel_l2(out.cpu().numpy(), ref)
    sols, _, _ = run_server([(d, 40), (SIZES[0], 41)], 4)
    e2 = C.rel_l2(sols[0], ref)
    log("  rel L2 vs oracle: per-request", e1, "service", e2)
    assert e1 < TOL[4] and e2 < TOL[4]
    assert np.array_equal(sols[0], out.cpu().numpy())
