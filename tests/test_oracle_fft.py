"""Pins the oracle's own FFT (which replaces FFTW) against numpy.fft and a direct DFT."""
import numpy as np
import pytest

import oracle as O


def direct_dft(x, sign):
    n = x.shape[0]
    j = np.arange(n)
    Wm = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return Wm @ x


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16, 32, 64, 128, 256])
def test_fft1d_vs_direct(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    for sign in (-1, 1):
        ref = direct_dft(x, sign)
        assert np.abs(O.fft1(x, sign) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("shape", [(4, 4), (8, 16), (64, 64), (128, 128)])
def test_fft2d_vs_numpy(shape):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fwd = O.fft2(x, -1)
    bwd = O.fft2(x, +1)
    assert np.abs(fwd - np.fft.fft2(x)).max() < 1e-12 * np.abs(fwd).max()
    assert np.abs(bwd - np.fft.ifft2(x) * x.size).max() < 1e-12 * np.abs(bwd).max()


def test_pad_extract_roundtrip():
    rng = np.random.default_rng(3)
    N = 8
    f = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    p = O.zero_pad(f)
    assert p.shape == (2 * N, 2 * N)
    assert np.count_nonzero(p) == N * N
    assert np.array_equal(O.extract(p), f)
    # frequency (l, k) of the N-grid lands on the padded frequency with the same value
    kg = lambda k, n: k if k < n // 2 else k - n
    for l in range(N):
        for k in range(N):
            assert p[kg(l, N) % (2 * N), kg(k, N) % (2 * N)] == f[l, k]


def test_r2c_fill_matches_full_fft():
    rng = np.random.default_rng(5)
    w = rng.standard_normal((32, 32))
    assert np.abs(O.r2c_fill(w) - np.fft.fft2(w) / 32 ** 2).max() < 1e-15
    x = np.fft.fft2(w) / 32 ** 2
    assert np.abs(O.c2r_real(x) - w).max() < 1e-13
