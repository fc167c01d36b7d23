"""The Poisson plan (forward chain with the fused -1/|k|^2 multiply, backward
chain) through the host build of the kernels, against the FP64 oracle and a
closed form.  CPU only."""
import numpy as np
import pytest

import cases as C
import emu_binding as emu
import oracle_binding as orc


def analytic(dims):
    """rhs = sin(2x) sin(3y) cos(z) + cos(x) -> u = -sin(2x)sin(3y)cos(z)/14 - cos(x)."""
    ax = [np.arange(n) * (2 * np.pi / n) for n in dims]
    s = np.sin(2 * ax[0])[:, None, None] * np.sin(3 * ax[1])[None, :, None] * np.cos(ax[2])[None, None, :]
    c = np.cos(ax[0])[:, None, None] * np.ones((1, dims[1], dims[2]))
    return s + c, -s / 14.0 - c


@pytest.mark.parametrize("dims", [[16, 12, 8], [24, 40, 16]])
def test_oracle_poisson_closed_form(dims):
    rhs, u = analytic(dims)
    sol, _ = C.poisson_ref(dims, orc.to_storage(rhs, [0, 1, 2]))
    assert np.abs(orc.from_storage(sol, [0, 1, 2]) - u).max() < 1e-13


@pytest.mark.parametrize("batch_threads", [0, 256])
@pytest.mark.parametrize("dims", [[16, 12, 8], [32, 32, 32], [24, 40, 16], [64, 32, 48]])
@pytest.mark.parametrize("prec", [8, 4])
def test_emu_poisson_vs_oracle(dims, prec, batch_threads):
    rhs = orc.to_storage(C.gaussians_rhs(dims, 5), [0, 1, 2])
    ref, ref_spec = C.poisson_ref(dims, rhs)
    sol, spec, desc = emu.poisson(dims, rhs, prec, batch_threads=batch_threads)
    tol = 1e-12 if prec == 8 else 1e-5
    assert C.rel_l2(spec, ref_spec) < tol, desc
    assert C.rel_l2(sol, ref) < tol, desc


def test_emu_poisson_closed_form():
    dims = [32, 24, 16]
    rhs, u = analytic(dims)
    sol, _, _ = emu.poisson(dims, orc.to_storage(rhs, [0, 1, 2]), 8)
    assert np.abs(orc.from_storage(sol, [0, 1, 2]) - u).max() < 1e-13
