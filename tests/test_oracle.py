"""Pins the CPU oracle before anything is checked against it.

Each check is chosen so that a plausible slip in the restatement (a sign, an
index, a dropped term, a wrong kind or ID) fails: FFTW-generated vectors,
numpy pocketfft fixtures, long-double definitions, the reference samples'
closed-form answers, and hand traces of the reference planner.
"""
import os

import numpy as np
import pytest

import oracle_binding as orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


# ---------------------------------------------------------------- 1D DFT

def test_fft_vs_numpy_fixtures():
    d = np.load(os.path.join(GOLD, "dft_vectors.npz"), allow_pickle=False)
    for n in d["sizes"]:
        n = int(n)
        z = d[f"c_in_{n}"]
        assert rel_l2(orc.fft(z, -1), d[f"c_fwd_{n}"]) < 1e-14, n
        assert rel_l2(orc.fft(z, +1), d[f"c_bwd_{n}"]) < 1e-14, n
        assert rel_l2(orc.r2c(d[f"r_in_{n}"]), d[f"r2c_{n}"]) < 1e-14, n
        assert rel_l2(orc.c2r(d[f"h_in_{n}"], n), d[f"c2r_{n}"]) < 1e-14, n


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 12, 17, 30, 49, 64])
def test_fft_vs_long_double_definition(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for s in (-1, 1):
        assert rel_l2(orc.fft(z, s), orc.dft_naive(z, s)) < 1e-15


def test_dft_naive_closed_forms():
    # delta -> all ones; single mode -> single spike at the right sign
    n = 16
    e = np.zeros(n, complex)
    e[0] = 1
    assert np.allclose(orc.dft_naive(e, -1), np.ones(n), atol=1e-18)
    x = np.exp(2j * np.pi * 3 * np.arange(n) / n)
    X = orc.dft_naive(x, -1)
    ref = np.zeros(n, complex)
    ref[3] = n
    assert np.abs(X - ref).max() < 1e-13


def test_c2r_ignores_dc_and_nyquist_imag():
    n = 8
    h = np.arange(n // 2 + 1) + 0j
    h2 = h.copy()
    h2[0] += 5j
    h2[-1] += 3j
    assert np.abs(orc.c2r(h, n) - orc.c2r(h2, n)).max() < 1e-13


# ---------------------------------------------------------------- r2r

FFTW_NAME_TO_KIND = {("dct", 1): orc.REDFT00, ("dct", 2): orc.REDFT10,
                     ("dct", 3): orc.REDFT01, ("dct", 4): orc.REDFT11,
                     ("dst", 1): orc.RODFT00, ("dst", 2): orc.RODFT10,
                     ("dst", 3): orc.RODFT01, ("dst", 4): orc.RODFT11}


@pytest.mark.parametrize("naive", [True, False])
def test_r2r_vs_fftw_generated_vectors(naive):
    d = np.load(os.path.join(GOLD, "fftw_r2r_double.npz"), allow_pickle=False)
    checked = 0
    for (name, t), kind in FFTW_NAME_TO_KIND.items():
        for n in d["sizes"]:
            n = int(n)
            if name == "dct" and t == 1 and n < 2:
                continue
            key = f"{name}_{t}_{n}"
            ref = d[key]
            x = np.arange(n, dtype=np.float64)  # FFTW reference input a[i] = i
            y = orc.r2r(kind, x, naive=naive)
            assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), key
            checked += 1
    assert checked >= 110


def test_r2r_vs_scipy_fixtures():
    d = np.load(os.path.join(GOLD, "r2r_vectors.npz"), allow_pickle=False)
    for n in d["sizes"]:
        n = int(n)
        x = d[f"x_{n}"]
        for t in (1, 2, 3, 4):
            if f"dct{t}_{n}" in d.files:
                assert rel_l2(orc.r2r(orc.DCT_KIND[t], x), d[f"dct{t}_{n}"]) < 1e-14
            assert rel_l2(orc.r2r(orc.DST_KIND[t], x), d[f"dst{t}_{n}"]) < 1e-14


@pytest.mark.parametrize("n", [2, 3, 6, 11, 40, 97])
def test_r2r_fast_vs_naive(n):
    rng = np.random.default_rng(100 + n)
    x = rng.uniform(-1, 1, n)
    for kind in range(8):
        assert rel_l2(orc.r2r(kind, x), orc.r2r(kind, x, naive=True)) < 1e-14


def test_r2r_inverse_pairs():
    # unnormalized FFTW inverse pairs: DCT-I 2(N-1), DST-I 2(N+1), II/III 2N, IV 2N
    rng = np.random.default_rng(5)
    n = 24
    x = rng.uniform(-1, 1, n)
    pairs = [(orc.REDFT00, orc.REDFT00, 2 * (n - 1)), (orc.RODFT00, orc.RODFT00, 2 * (n + 1)),
             (orc.REDFT10, orc.REDFT01, 2 * n), (orc.RODFT10, orc.RODFT01, 2 * n),
             (orc.RODFT11, orc.RODFT11, 2 * n), (orc.REDFT11, orc.REDFT11, 2 * n)]
    for a, b, s in pairs:
        assert rel_l2(orc.r2r(b, orc.r2r(a, x)) / s, x) < 1e-14


# ---------------------------------------------------------------- registry

def test_type_registry_order_and_dct4_quirk():
    assert orc.lib().or_num_types() == 44
    # build/init.C:121-731 registration order
    assert orc.type_info(5).cls == 1 and orc.type_info(5).prec == 8  # R2CFFT_D
    assert orc.type_info(6).cls == 2 and orc.type_info(6).prec == 4  # C2RFFT_S
    assert orc.type_info(9).isign == -1 and orc.type_info(11).isign == 1
    assert orc.type_info(13).r2r_kind == orc.REDFT00 and orc.type_info(13).dt1 == 1
    assert orc.type_info(15).dt1 == 2                                # DCT1_COMPLEX_D
    assert orc.type_info(17).r2r_kind == orc.RODFT00
    assert orc.type_info(21).r2r_kind == orc.REDFT10
    assert orc.type_info(25).r2r_kind == orc.RODFT10
    assert orc.type_info(29).r2r_kind == orc.REDFT01
    assert orc.type_info(33).r2r_kind == orc.RODFT01
    for i in (36, 37, 38, 39):       # DCT4_* registered with plan_dct1_* (quirk)
        assert orc.type_info(i).r2r_kind == orc.REDFT00
    for i in (40, 41, 42, 43):
        assert orc.type_info(i).r2r_kind == orc.RODFT11
    with pytest.raises(ValueError):
        orc.type_info(44)


# ---------------------------------------------------------------- planner

def test_swap0_excl():
    assert orc.swap0([0, 1, 2], 0) == [0, 1, 2]
    assert orc.swap0([0, 1, 2], 1) == [1, 0, 2]
    assert orc.swap0([1, 2, 0], 1) == [2, 0, 1]
    assert orc.swap0([2, 0, 1], 2) == [2, 1, 0]
    assert orc.swap0([2, 1, 0], 0) == [0, 2, 1]
    assert orc.excl(0, 2) == 1 and orc.excl(1, 0) == 2 and orc.excl(1, 2) == 0


R2C_D, C2R_D, CF_D, CB_D = 5, 7, 9, 11


def chain(types, mo1, mo2, g=(8, 8, 8)):
    return [(s["trans_dim"], s["mo_in"], s["mo_out"]) for s in orc.plan3d(types, g, mo1, mo2)]


def test_plan_chain_samples_forward_backward():
    # hand trace of build/templ.C for the samples' (0,1,2) <-> (1,2,0)
    assert chain([R2C_D, CF_D, CF_D], [0, 1, 2], [1, 2, 0]) == [
        (0, [0, 1, 2], [0, 1, 2]), (1, [0, 1, 2], [1, 0, 2]), (2, [1, 0, 2], [1, 2, 0])]
    assert chain([C2R_D, CB_D, CB_D], [1, 2, 0], [0, 1, 2]) == [
        (1, [1, 2, 0], [2, 0, 1]), (2, [2, 0, 1], [2, 1, 0]), (0, [2, 1, 0], [0, 1, 2])]


def test_plan_chain_reversed_input_order():
    assert chain([R2C_D, CF_D, CF_D], [2, 1, 0], [1, 2, 0]) == [
        (0, [2, 1, 0], [0, 2, 1]), (1, [0, 2, 1], [1, 0, 2]), (2, [1, 0, 2], [1, 2, 0])]
    # backward ends with an EMPTY reorder (final_seq)
    st = orc.plan3d([C2R_D, CB_D, CB_D], [5, 8, 8], [1, 2, 0], [2, 1, 0])
    assert [(s["trans_dim"], s["mo_in"], s["mo_out"]) for s in st] == [
        (1, [1, 2, 0], [2, 0, 1]), (2, [2, 0, 1], [2, 1, 0]), (0, [2, 1, 0], [0, 2, 1]),
        (0, [0, 2, 1], [2, 1, 0])]
    assert st[-1]["type_id"] == 1            # EMPTY_TYPE_DOUBLE
    assert st[2]["dims_out"] == [8, 8, 8]    # (5-1)*2


def test_plan_dims_r2c():
    st = orc.plan3d([R2C_D, CF_D, CF_D], [16, 6, 5], [0, 1, 2], [1, 2, 0])
    assert st[0]["dims_out"] == [9, 6, 5]
    assert st[-1]["dims_out"] == [9, 6, 5]


def test_plan_all_36_memory_order_pairs_end_in_grid2_order():
    perms = [[0, 1, 2], [0, 2, 1], [1, 0, 2], [1, 2, 0], [2, 0, 1], [2, 1, 0]]
    for mo1 in perms:
        for mo2 in perms:
            for types in ([R2C_D, CF_D, CF_D], [CF_D, CF_D, CF_D], [C2R_D, CB_D, CB_D]):
                st = orc.plan3d(types, [8, 6, 4], mo1, mo2)
                assert st[0]["mo_in"] == mo1
                assert st[-1]["mo_out"] == mo2
                dims = sorted(s["trans_dim"] for s in st[:3])
                assert dims == [0, 1, 2]
                for s in st[:3]:
                    # every transform stage leaves or takes its axis stride-1
                    assert s["mo_in"][s["trans_dim"]] == 0 or s["mo_out"][s["trans_dim"]] == 0 \
                        or s["mo_in"] == s["mo_out"]


# ---------------------------------------------------------------- 3D exec

def test_exec3d_vs_numpy_fixtures_all_layouts():
    d = np.load(os.path.join(GOLD, "fft3d_vectors.npz"), allow_pickle=False)
    perms = [[0, 1, 2], [2, 1, 0], [1, 2, 0], [0, 2, 1]]
    for i in range(4):
        s = [int(v) for v in d[f"shape_{i}"]]
        a = d[f"real_{i}"]
        ref = d[f"rcc_{i}"]
        for mo1 in perms:
            for mo2 in perms:
                out = orc.exec3d([R2C_D, CF_D, CF_D], s, mo1, mo2, orc.to_storage(a, mo1))
                got = orc.from_storage(out, mo2)
                assert rel_l2(got, ref) < 1e-14, (s, mo1, mo2)
        z = d[f"cplx_{i}"]
        out = orc.exec3d([CF_D, CF_D, CF_D], s, [0, 1, 2], [2, 1, 0], orc.to_storage(z, [0, 1, 2]))
        assert rel_l2(orc.from_storage(out, [2, 1, 0]), d[f"ccc_{i}"]) < 1e-14


def sine_wave(n):
    x = np.arange(n[0]) * 2 * np.pi / n[0]
    y = np.arange(n[1]) * 2 * np.pi / n[1]
    z = np.arange(n[2]) * 2 * np.pi / n[2]
    return np.sin(x)[:, None, None] * np.sin(y)[None, :, None] * np.sin(z)[None, None, :]


def test_sample_known_answer_r2c_roundtrip():
    # sample/C++/test3D_r2c.C:281-327: normalized spectrum is +-0.125i at
    # (1, +-1, +-1) with only k=1 on the conj-sym axis; round trip restores input.
    n = [16, 12, 10]
    a = sine_wave(n)
    mo1, mo2 = [0, 1, 2], [1, 2, 0]
    spec = orc.exec3d([R2C_D, CF_D, CF_D], n, mo1, mo2, orc.to_storage(a, mo1))
    S = orc.from_storage(spec, mo2) / np.prod(n)
    ref = np.zeros(S.shape, complex)
    ref[1, 1, 1] = 0.125j  # sin*sin*sin -> (-i/2)^3 = i/8 at (1,1,1)
    ref[1, -1, 1] = -0.125j
    ref[1, 1, -1] = -0.125j
    ref[1, -1, -1] = 0.125j
    assert np.abs(S - ref).max() < 1e-14 * n[0] * 0.25
    back = orc.exec3d([C2R_D, CB_D, CB_D], [n[0] // 2 + 1, n[1], n[2]], mo2, mo1,
                      spec)
    assert np.abs(orc.from_storage(back, mo1) / np.prod(n) - a).max() < 1e-14 * np.prod(n) * 0.25


@pytest.mark.parametrize("idir", [0, 1, 2])
def test_sample_known_answer_derivative(idir):
    # sample/C++/test_deriv.C:385-420: d/d(idir) of the sine product
    n = [16, 12, 10]
    a = sine_wave(n)
    mo1, mo2 = [0, 1, 2], [1, 2, 0]
    spec = orc.exec3d([R2C_D, CF_D, CF_D], n, mo1, mo2, orc.to_storage(a, mo1), idir=idir)
    back = orc.exec3d([C2R_D, CB_D, CB_D], [n[0] // 2 + 1, n[1], n[2]], mo2, mo1, spec)
    got = orc.from_storage(back, mo1) / np.prod(n)
    g = [np.sin(np.arange(m) * 2 * np.pi / m) for m in n]
    g[idir] = np.cos(np.arange(n[idir]) * 2 * np.pi / n[idir])
    ref = g[0][:, None, None] * g[1][None, :, None] * g[2][None, None, :]
    assert np.abs(got - ref).max() < 1e-13


def test_derivative_nyquist_and_upper_half():
    # compute_deriv_loc: k < mid -> i k, k == mid -> 0, k > mid -> i (k - g)
    n = [6, 4, 3]
    rng = np.random.default_rng(3)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    mo = [0, 1, 2]
    out = orc.exec_1d(CF_D, 1, n, mo, n, mo, orc.to_storage(z, mo), deriv=True)
    plain = orc.exec_1d(CF_D, 1, n, mo, n, mo, orc.to_storage(z, mo), deriv=False)
    kk = np.array([0, 1, 0, -1])  # g=4, mid=2
    ref = orc.from_storage(plain, mo) * (1j * kk)[None, :, None]
    assert np.abs(orc.from_storage(out, mo) - ref).max() < 1e-13


def test_compute_deriv_standalone_and_axis_quirk():
    # identity / involution memory orders: derivative along the requested axis
    g = [8, 6, 4]
    rng = np.random.default_rng(9)
    z = rng.standard_normal(g) + 1j * rng.standard_normal(g)
    for mo in ([0, 1, 2], [2, 1, 0], [1, 0, 2]):
        for idir in range(3):
            out = orc.from_storage(orc.compute_deriv(orc.to_storage(z, mo), g, g, [0, 0, 0],
                                                     mo, -1, idir), mo)
            n = g[idir]
            k = np.arange(n)
            mult = np.where(k < n // 2, k, np.where(k == n // 2, 0, k - n)).astype(float)
            shp = [1, 1, 1]
            shp[idir] = n
            assert np.abs(out - z * (1j * mult).reshape(shp)).max() < 1e-13, (mo, idir)
    # 3-cycle memory order (1,2,0): build/deriv.C:91-94 picks storage rank
    # i with MemOrder[i]==idir; for idir=0 that is rank 2 (logical y), with g
    # taken from Gdims[0].
    mo = [1, 2, 0]
    out = orc.from_storage(orc.compute_deriv(orc.to_storage(z, mo), g, g, [0, 0, 0], mo, -1, 0), mo)
    ky = np.arange(g[1])
    gx = g[0]
    mult = np.where(ky < gx // 2, ky, np.where(ky == gx // 2, 0, ky - gx)).astype(float)
    assert np.abs(out - z * (1j * mult)[None, :, None]).max() < 1e-13


def test_exec_1d_r2r_layouts():
    # 1D cosine KAT (sample/C++/test1D_cos.C): cos(pi x/(N-1)) -> 0.5 at k=1
    # after scaling by 0.5/(N-1), for a transposing layout change.
    n = [9, 5, 4]
    x = np.cos(np.pi * np.arange(n[0]) / (n[0] - 1))
    a = np.broadcast_to(x[:, None, None], n).copy()
    mo1, mo2 = [0, 1, 2], [2, 0, 1]
    out = orc.exec_1d(13, 0, n, mo1, n, mo2, orc.to_storage(a, mo1))
    got = orc.from_storage(out, mo2) * 0.5 / (n[0] - 1)
    ref = np.zeros(n)
    ref[1] = 0.5
    assert np.abs(got - ref).max() < 1e-14 * n[0] * 0.25
