"""Parity at the full sizes of the 512^3 and 768 x 640 x 513 workloads,
against the FP64 oracle on the same inputs.

512^3 (input memory order (2,1,0), spectrum (1,2,0)): the whole forward
spectrum of the noisy field sin(x)sin(2y)sin(3z) + N(0, 1e-3), rounded once to
FP32 so that the FP64 and FP32 library paths and the oracle all transform the
same numbers: FP64 <= 1e-12 relative L2 (the drop-in bar), FP32 <= 1e-5
relative L2.  The spectral d/dx of the same field in FP32 (exec_deriv, idir 0)
against the oracle's exec_deriv: multiplying mode k by i k turns a spectrum
error e_k into k e_k, so a spectrum within relative L2 eps of the oracle's
gives a derivative within eps * k_rms * |F| / |D| of the oracle's derivative
(k_rms: root mean square of the derivative's wavenumber over the stored
modes); the bar is that bound at eps = 1e-5.

768 x 640 x 513: each 1D stage of the mixed plan (R2C x of 768, C2C y of 640,
DCT1_COMPLEX z of 513) runs alone at the full size, with the memory orders the
3D plan uses for it, and 64 whole lines of its output are compared with the
oracle's transform of the same lines (FP32 <= 1e-5, FP64 <= 1e-12 relative
L2 over the lines).

The oracle runs in worker threads (its C code releases the GIL) while the GPU
runs.
"""
import re
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import cases as C
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


def rel_l2_chunked(a, b, chunk=1 << 24):
    """|a - b| / |b| over flat arrays, in chunks (keeps host memory bounded)."""
    a, b = a.reshape(-1), b.reshape(-1)
    num = den = 0.0
    for i in range(0, b.size, chunk):
        bb = b[i:i + chunk].astype(np.complex128)
        num += float(np.sum(np.abs(a[i:i + chunk].astype(np.complex128) - bb) ** 2))
        den += float(np.sum(np.abs(bb) ** 2))
    return (num / den) ** 0.5


def plan3d(types, dims1, mo1, dims2, mo2, conj2):
    pg = P.init_proc_grid((1, 1, 1))
    g1 = P.init_data_grid(dims1, -1, pg, (0, 1, 2), mo1)
    g2 = P.init_data_grid(dims2, conj2, pg, (0, 1, 2), mo2)
    return P.plan_3Dtrans(g1, g2, P.init_3Dtype(types))


def test_config2_512cubed_spectrum_vs_oracle():
    n = 512
    mo1, mo2 = [2, 1, 0], [1, 2, 0]
    dims2 = [n // 2 + 1, n, n]
    xs = np.arange(n) * 2 * np.pi / n
    f = np.sin(xs)[:, None, None] * np.sin(2 * xs)[None, :, None] * np.sin(3 * xs)[None, None, :]
    f = (f + np.random.default_rng(2).normal(0.0, 1e-3, f.shape)).astype(np.float32)
    x32 = orc.to_storage(f, mo1)
    del f
    x64 = x32.astype(np.float64)
    fwd_d = [C.R2C_D, C.CF_D, C.CF_D]
    log("case: 512^3 forward spectrum vs oracle (oracle in two threads)")
    with ThreadPoolExecutor(max_workers=2) as ex:
        fut = ex.submit(orc.exec3d, fwd_d, [n] * 3, mo1, mo2, x64)
        fut_d = ex.submit(orc.exec3d, fwd_d, [n] * 3, mo1, mo2, x64, 0)
        shp2 = orc.storage_shape(dims2, mo2)
        res = {}
        for prec, types, xin, cdt in ((8, fwd_d, x64, torch.complex128), (4, [C.R2C_S, C.CF_S, C.CF_S], x32,
                                                                         torch.complex64)):
            h = plan3d(types, [n] * 3, mo1, dims2, mo2, 0)
            din = torch.from_numpy(xin).cuda()
            out = torch.empty(shp2, dtype=cdt, device="cuda")
            torch.cuda.synchronize()
            P.exec_3Dtrans(h, din.data_ptr(), out.data_ptr(), 0, prec)
            res[prec] = out.cpu().numpy()
            if prec == 4:
                P.exec_3Dderiv(h, din.data_ptr(), out.data_ptr(), 0, 0, 4)
                res["d4"] = out.cpu().numpy()
            log("  GPU prec", prec, "done:", P.describe(h))
            del din, out
            torch.cuda.empty_cache()
        ref = fut.result()
        refd = fut_d.result()
    e64 = rel_l2_chunked(res[8], ref)
    e32 = rel_l2_chunked(res[4], ref)
    log(f"  forward spectrum rel L2 vs oracle: FP64 {e64:.3e}, FP32 {e32:.3e}")
    assert e64 < 1e-12
    assert e32 < 1e-5
    # derivative: wavenumbers of the x axis (R2C: g = 2 (nh - 1), the
    # Nyquist mode k = g/2 set to 0), build/exec.C:228-287
    kx = np.arange(n // 2 + 1, dtype=np.float64)
    kx[n // 2] = 0.0
    k_rms = float(np.sqrt(np.mean(kx ** 2)))
    nF = float(np.sqrt(sum(np.sum(np.abs(ref.reshape(-1)[i:i + (1 << 24)]) ** 2)
                           for i in range(0, ref.size, 1 << 24))))
    nD = float(np.sqrt(sum(np.sum(np.abs(refd.reshape(-1)[i:i + (1 << 24)]) ** 2)
                           for i in range(0, refd.size, 1 << 24))))
    ed = rel_l2_chunked(res["d4"], refd)
    bar = 1e-5 * k_rms * nF / nD
    log(f"  FP32 d/dx rel L2 vs oracle {ed:.3e}; bar 1e-5 * k_rms {k_rms:.1f} * |F|/|D| {nF / nD:.3f} = {bar:.3e}")
    assert ed < bar


def _passes(desc):
    """(kind, axis, N, mo_in, mo_out) of each pass of a plan description."""
    out = []
    for m in re.finditer(r"(\w+)\[axis (\d), N=(\d+)[^\]]*?mo \((\d)(\d)(\d)\)->\((\d)(\d)(\d)\)", desc):
        g = m.groups()
        out.append((g[0], int(g[1]), int(g[2]), [int(v) for v in g[3:6]], [int(v) for v in g[6:9]]))
    return out


def _line(t_storage, mo, axis, coord):
    """The whole line along logical `axis` through logical coords `coord`
    (entry `axis` ignored) of an array in storage layout."""
    idx = [0, 0, 0]
    for a in range(3):
        idx[2 - mo[a]] = slice(None) if a == axis else coord[a]
    return t_storage[tuple(idx)]


@pytest.mark.parametrize("prec", [4, 8])
def test_config3_full_size_lines_vs_oracle(prec):
    dims = [768, 640, 513]
    s = "D" if prec == 8 else "S"
    types = [P.type_id(f"R2CFFT_{s}"), P.type_id(f"CFFT_FORWARD_{s}"), P.type_id(f"DCT1_COMPLEX_{s}")]
    types_d = [C.R2C_D, C.CF_D, C.r2r_id("DCT1", True, 8)]
    nh = dims[0] // 2 + 1
    h3 = plan3d(types, dims, [0, 1, 2], [nh, dims[1], dims[2]], [0, 1, 2], 0)
    desc = P.describe(h3)
    passes = _passes(desc)
    log("case: 768x640x513 stages, prec", prec, ":", desc)
    assert sorted(p[1] for p in passes) == [0, 1, 2], desc
    pg = P.init_proc_grid((1, 1, 1))
    rng = np.random.default_rng(30 + prec)
    gen = torch.Generator(device="cuda").manual_seed(prec)
    rdt, cdt = (torch.float64, torch.complex128) if prec == 8 else (torch.float32, torch.complex64)
    for kind, axis, N, moi, moo in passes:
        tid, tid_d = types[axis], types_d[axis]
        din_dims = [dims[0] if axis == 0 else nh, dims[1], dims[2]]
        dout_dims = [nh, dims[1], dims[2]]
        g1 = P.init_data_grid(din_dims, -1, pg, (0, 1, 2), moi)
        g2 = P.init_data_grid(dout_dims, -1, pg, (0, 1, 2), moo)
        h = P.plan_1Dtrans(g1, g2, tid, axis)
        shp_in = orc.storage_shape(din_dims, moi)
        if axis == 0:
            x = torch.rand(shp_in, generator=gen, device="cuda", dtype=rdt) * 2 - 1
        else:
            x = torch.complex(torch.randn(shp_in, generator=gen, device="cuda", dtype=rdt),
                              torch.randn(shp_in, generator=gen, device="cuda", dtype=rdt))
        y = torch.empty(orc.storage_shape(dout_dims, moo), dtype=cdt, device="cuda")
        torch.cuda.synchronize()
        P.exec_1Dtrans(h, x.data_ptr(), y.data_ptr(), 0, prec)
        got, want = [], []
        for _ in range(64):
            coord = [int(rng.integers(0, din_dims[a])) for a in range(3)]
            xl = _line(x, moi, axis, coord).cpu().numpy().astype(np.float64 if axis == 0 else np.complex128)
            yl = _line(y, moo, axis, coord).cpu().numpy()
            n_in, n_out = len(xl), len(yl)
            ref = orc.exec_1d(tid_d, 0, [n_in, 1, 1], [0, 1, 2], [n_out, 1, 1], [0, 1, 2], xl)
            got.append(yl.astype(np.complex128))
            want.append(ref.reshape(-1))
        err = C.rel_l2(np.concatenate(got), np.concatenate(want))
        log(f"  {kind} axis {axis} N={N} mo {moi}->{moo}: 64 lines rel L2 {err:.3e}")
        assert err < (1e-12 if prec == 8 else 1e-5), (kind, axis, err)
        del x, y
        torch.cuda.empty_cache()
