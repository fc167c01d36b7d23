"""Shared case construction for the parity tests (CPU host-build and GPU)."""
import numpy as np

import oracle_binding as orc

PERMS = [[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]]

# type IDs (registration order of the reference's setup())
EMPTY_S, EMPTY_D, EMPTY_SC, EMPTY_DC = 0, 1, 2, 3
R2C_S, R2C_D, C2R_S, C2R_D = 4, 5, 6, 7
CF_S, CF_D, CB_S, CB_D = 8, 9, 10, 11


def r2r_id(kind_name, complex_, prec):
    """kind_name in DCT1..DCT4, DST1..DST4 -> type ID."""
    base = {"DCT1": 12, "DST1": 16, "DCT2": 20, "DST2": 24, "DCT3": 28, "DST3": 32,
            "DCT4": 36, "DST4": 40}[kind_name]
    return base + (2 if complex_ else 0) + (1 if prec == 8 else 0)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def out_dims(types, dims, c2r_len=None):
    """Logical output dims of a 3D transform: R2C axis -> n/2+1; C2R axis ->
    c2r_len (the real length) or 2(n-1)."""
    o = list(dims)
    for a, t in enumerate(types):
        ti = orc.type_info(t)
        if ti.dt1 < ti.dt2:
            o[a] = dims[a] // 2 + 1
        elif ti.dt1 > ti.dt2:
            o[a] = c2r_len if c2r_len is not None else 2 * (dims[a] - 1)
    return o


def random_input(types, dims, mo, seed, hermitian_axis=None):
    """Storage-order input for a 3D transform with the given types."""
    rng = np.random.default_rng(seed)
    order = orc.find_order(types, mo, mo)
    t0 = orc.type_info(types[order[0]])
    shape = orc.storage_shape(dims, mo)
    if t0.dt1 == 1:
        return rng.uniform(-1.0, 1.0, shape)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def poisson_ref(dims, rhs, mo_real=(0, 1, 2), mo_spec=(1, 2, 0)):
    """FP64 oracle Poisson solve on [0, 2 pi)^3: the oracle's R2C x / C2C y, z
    transform, multiply by -1/|k|^2 (0 at k = 0) and 1/(nx ny nz), and the
    oracle's inverse.  Returns (solution in mo_real storage, spectrum after
    the multiply in mo_spec storage)."""
    nx, ny, nz = dims
    spec = orc.exec3d([R2C_D, CF_D, CF_D], list(dims), list(mo_real), list(mo_spec), np.asarray(rhs, np.float64))
    lg = orc.from_storage(spec, list(mo_spec))
    kx = np.arange(nx // 2 + 1, dtype=np.float64)
    ky = np.fft.fftfreq(ny, 1.0 / ny)
    kz = np.fft.fftfreq(nz, 1.0 / nz)
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    fac = np.zeros_like(k2)
    fac[k2 > 0] = -1.0 / k2[k2 > 0]
    spec2 = orc.to_storage(lg * fac / (nx * ny * nz), list(mo_spec))
    sol = orc.exec3d([C2R_D, CB_D, CB_D], [nx // 2 + 1, ny, nz], list(mo_spec), list(mo_real), spec2)
    return sol, spec2


def gaussians_rhs(dims, seed, count=8):
    """Sum of `count` random periodic Gaussians on [0, 2 pi)^3 minus its mean
    (logical [x][y][z] array)."""
    rng = np.random.default_rng(seed)
    axes = [np.arange(n) * (2 * np.pi / n) for n in dims]
    f = np.zeros(dims)
    for _ in range(count):
        c = rng.uniform(0, 2 * np.pi, 3)
        w = rng.uniform(0.3, 0.8)
        a = rng.uniform(-1, 1)
        g = [np.exp(-(np.angle(np.exp(1j * (ax - ci))) ** 2) / (2 * w * w)) for ax, ci in zip(axes, c)]
        f += a * g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    return f - f.mean()


def chebyshev_field(torch, dims, seed):
    """Real field on the periodic x, y grid and the Chebyshev-Gauss-Lobatto z
    grid: a sum of 4 separable terms, each axis a series with coefficients
    decaying as 1/(k+1)^2 (uniform phases in x and y, random signs in z);
    storage order (0,1,2) ([z][y][x]), FP64, on the GPU."""
    nx, ny, nz = dims
    g = torch.Generator(device="cuda").manual_seed(seed)
    dev, f64 = "cuda", torch.float64

    def periodic(n, kmax):
        k = torch.arange(kmax, device=dev, dtype=f64)
        ph = torch.rand(kmax, generator=g, device=dev, dtype=f64) * 2 * torch.pi
        x = torch.arange(n, device=dev, dtype=f64) * (2 * torch.pi / n)
        return (torch.cos(k[:, None] * x[None, :] + ph[:, None]) / (k[:, None] + 1) ** 2).sum(0)

    def cheb(n):
        k = torch.arange(n, device=dev, dtype=f64)
        sg = torch.where(torch.rand(n, generator=g, device=dev) < 0.5, -1.0, 1.0).to(f64)
        j = torch.arange(n, device=dev, dtype=f64)
        return (torch.cos(torch.pi * k[:, None] * j[None, :] / (n - 1)) * (sg / (k + 1) ** 2)[:, None]).sum(0)

    f = torch.zeros((nz, ny, nx), device=dev, dtype=f64)
    for _ in range(4):
        f += cheb(nz)[:, None, None] * periodic(ny, 64)[None, :, None] * periodic(nx, 64)[None, None, :]
    return f
