"""Writes the golden fixtures under tests/golden/ (run once, where numpy and
scipy are installed; the tests only read the .npz files).

Fixtures:
  fftw_r2r_double.npz  FFTW-generated DCT/DST I-IV vectors (input a[i] = i,
                       FFTW REDFT/RODFT kinds, 14 sizes) as shipped inside
                       scipy's fftpack test-suite, re-exported as plain arrays.
  dft_vectors.npz      seeded complex/real inputs and numpy (pocketfft, FP64)
                       c2c forward/backward, r2c and c2r outputs.
  r2r_vectors.npz      seeded inputs and scipy.fft dct/dst types 1-4
                       (unnormalized = FFTW REDFT/RODFT conventions).
  fft3d_vectors.npz    small 3D cases: real field, R2C along x then C2C along
                       y and z (numpy rfft/fft), in logical [x][y][z] order.

No file of the reference is read or executed here.
"""
import os

import numpy as np
import scipy
import scipy.fft

HERE = os.path.dirname(os.path.abspath(__file__))


def fftw_r2r():
    src = os.path.join(os.path.dirname(scipy.__file__), "fftpack", "tests",
                       "fftw_double_ref.npz")
    d = np.load(src, allow_pickle=False)
    out = {"sizes": np.asarray(d["sizes"], dtype=np.int64)}
    for k in d.files:
        if k.startswith("dct_") or k.startswith("dst_"):
            out[k] = np.asarray(d[k], dtype=np.float64)
    np.savez(os.path.join(HERE, "fftw_r2r_double.npz"), **out)


DFT_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 17, 25, 31, 32, 48, 60,
             64, 96, 97, 120, 121, 128, 139, 169, 199, 243, 256, 343, 344, 385,
             512, 513, 640, 768, 1000, 1024, 2048]


def dft_vectors():
    rng = np.random.default_rng(20240601)
    out = {"sizes": np.array(DFT_SIZES, dtype=np.int64)}
    for n in DFT_SIZES:
        z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        r = rng.uniform(-1.0, 1.0, n)
        h = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
        h[0] = h[0].real
        if n % 2 == 0:
            h[-1] = h[-1].real
        out[f"c_in_{n}"] = z
        out[f"c_fwd_{n}"] = np.fft.fft(z)
        out[f"c_bwd_{n}"] = np.fft.ifft(z) * n
        out[f"r_in_{n}"] = r
        out[f"r2c_{n}"] = np.fft.rfft(r)
        out[f"h_in_{n}"] = h
        out[f"c2r_{n}"] = np.fft.irfft(h, n) * n
    np.savez(os.path.join(HERE, "dft_vectors.npz"), **out)


R2R_SIZES = [2, 3, 4, 5, 7, 8, 16, 17, 33, 64, 100, 129, 257, 385, 513]


def r2r_vectors():
    rng = np.random.default_rng(7)
    out = {"sizes": np.array(R2R_SIZES, dtype=np.int64)}
    for n in R2R_SIZES:
        x = rng.uniform(-1.0, 1.0, n)
        out[f"x_{n}"] = x
        for t in (1, 2, 3, 4):
            if not (t == 1 and n < 2):
                out[f"dct{t}_{n}"] = scipy.fft.dct(x, type=t)
            out[f"dst{t}_{n}"] = scipy.fft.dst(x, type=t)
    np.savez(os.path.join(HERE, "r2r_vectors.npz"), **out)


SHAPES_3D = [(8, 6, 5), (12, 10, 9), (16, 16, 16), (9, 7, 4)]


def fft3d_vectors():
    rng = np.random.default_rng(11)
    out = {}
    for i, s in enumerate(SHAPES_3D):
        a = rng.uniform(-1.0, 1.0, s)
        spec = np.fft.fft(np.fft.fft(np.fft.rfft(a, axis=0), axis=1), axis=2)
        out[f"shape_{i}"] = np.array(s, dtype=np.int64)
        out[f"real_{i}"] = a
        out[f"rcc_{i}"] = spec
        z = rng.standard_normal(s) + 1j * rng.standard_normal(s)
        out[f"cplx_{i}"] = z
        out[f"ccc_{i}"] = np.fft.fftn(z)
    np.savez(os.path.join(HERE, "fft3d_vectors.npz"), **out)


if __name__ == "__main__":
    fftw_r2r()
    dft_vectors()
    r2r_vectors()
    fft3d_vectors()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
