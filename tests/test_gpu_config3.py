"""The mixed-type workload at its full size: 768 x 640 x 513 real grid, R2C in
x, C2C in y, DCT-I (DCT1_COMPLEX) in z, FP64 and FP32, forward and back.

At this size the oracle's full 3D transform takes minutes, so parity is shown
by (a) individual spectrum entries computed in FP64 on the CPU directly from
the definitions (FFTW's R2C / forward DFT / REDFT00 sums, separably), (b) the
round trip with scale 1/(768 * 640 * 2 * 512), and (c) the FP32 spectrum
against the FP64 one (the error of the FP32 path against FP64, stated in the
log).  The same kernels are checked against the oracle on every entry at
reduced sizes in test_gpu_parity.py (test_mixed_r2c_c2c_dct1_reduced).
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

DIMS = [768, 640, 513]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def lib_setup():
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    P.setup()
    yield
    torch.cuda.synchronize()


def spectrum_entry(f_host, kx, ky, kz):
    """FP64 spectrum entry from the definitions: FFTW R2C (x), forward DFT
    (y), REDFT00 (z): sum_x sum_y sum_z f e^{-2 pi i kx x/nx} e^{-2 pi i ky y/ny}
    w_kz(z) with w_k(0) = 1, w_k(N-1) = (-1)^k, else 2 cos(pi j k/(N-1))."""
    nz, ny, nx = f_host.shape
    j = np.arange(nz)
    wz = 2 * np.cos(np.pi * j * kz / (nz - 1))
    wz[0], wz[-1] = 1.0, (-1.0) ** kz
    a = np.tensordot(wz, f_host, axes=(0, 0))  # [y][x]
    a = np.exp(-2j * np.pi * ky * np.arange(ny) / ny) @ a  # [x]
    return np.exp(-2j * np.pi * kx * np.arange(nx) / nx) @ a


def plans(prec):
    s = "D" if prec == 8 else "S"
    pg = P.init_proc_grid((1, 1, 1))
    nx, ny, nz = DIMS
    g_real = P.init_data_grid(DIMS, -1, pg, (0, 1, 2), (0, 1, 2))
    g_spec = P.init_data_grid([nx // 2 + 1, ny, nz], 0, pg, (0, 1, 2), (0, 1, 2))
    tf = P.init_3Dtype((P.type_id(f"R2CFFT_{s}"), P.type_id(f"CFFT_FORWARD_{s}"), P.type_id(f"DCT1_COMPLEX_{s}")))
    tb = P.init_3Dtype((P.type_id(f"C2RFFT_{s}"), P.type_id(f"CFFT_BACKWARD_{s}"), P.type_id(f"DCT1_COMPLEX_{s}")))
    return P.plan_3Dtrans(g_real, g_spec, tf), P.plan_3Dtrans(g_spec, g_real, tb)


def run(prec, f64):
    pf, pb = plans(prec)
    x = f64 if prec == 8 else f64.float()
    nx, ny, nz = DIMS
    spec = torch.empty((nz, ny, nx // 2 + 1), dtype=torch.complex128 if prec == 8 else torch.complex64,
                       device="cuda")
    back = torch.empty_like(x)
    st = P.stream()
    torch.cuda.synchronize()
    P.exec_async(pf, x.data_ptr(), spec.data_ptr(), -1, 0, st, prec=prec)
    P.exec_async(pb, spec.data_ptr(), back.data_ptr(), -1, 0, st, prec=prec)
    torch.cuda.synchronize()
    # timed repeat, synchronized on both sides
    t0 = time.perf_counter()
    P.exec_async(pf, x.data_ptr(), spec.data_ptr(), -1, 0, st, prec=prec)
    P.exec_async(pb, spec.data_ptr(), back.data_ptr(), -1, 0, st, prec=prec)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    log(f"  prec {prec}: round trip {t * 1e3:.2f} ms; plans {P.describe(pf)} || {P.describe(pb)}")
    return spec, back


def test_config3_full_size():
    log("case: 768x640x513 R2C/C2C/DCT-I, FP64 and FP32")
    f = C.chebyshev_field(torch, DIMS, 3)
    scale = 1.0 / (768 * 640 * 2 * 512)
    spec64, back64 = run(8, f)
    rt64 = (torch.linalg.vector_norm(back64 * scale - f) / torch.linalg.vector_norm(f)).item()
    log("  FP64 round trip rel L2", rt64)
    assert rt64 < 1e-12
    # spectrum entries from the definitions (CPU, FP64)
    fh = f.cpu().numpy()
    sh = spec64.cpu().numpy()  # [z][y][x]
    worst = 0.0
    ref_norm = float(np.abs(sh).max())
    for kx, ky, kz in [(0, 0, 0), (1, 0, 0), (3, 5, 7), (64, 1, 2), (383, 639, 512), (384, 320, 256), (17, 600, 511)]:
        e = spectrum_entry(fh, kx, ky, kz)
        worst = max(worst, abs(sh[kz, ky, kx] - e) / ref_norm)
    log("  FP64 spectrum entries vs definition, max |diff| / max |F|", worst)
    assert worst < 1e-12
    del back64
    spec32, back32 = run(4, f)
    rt32 = (torch.linalg.vector_norm(back32.double() * scale - f) / torch.linalg.vector_norm(f)).item()
    e32 = (torch.linalg.vector_norm(spec32.to(torch.complex128) - spec64) / torch.linalg.vector_norm(spec64)).item()
    log("  FP32 spectrum vs FP64 rel L2", e32, "; FP32 round trip rel L2", rt32)
    assert e32 < 1e-5 and rt32 < 1e-5
