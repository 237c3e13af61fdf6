"""Pins the diagnostics oracle (oracle/diag_ref.py): eigenvalues against numpy's eigvalsh of the
2x2 matrices built as visualization/visualizeHomBoltzmann.m:26-33 does, entropy closed forms,
moments of a shifted Gaussian."""
import numpy as np

from oracle import diag_ref as D
from conftest import rand_state


def test_eigenvalues_match_eigvalsh():
    rng = np.random.default_rng(0)
    W = rand_state(rng, 2, 8, 6.0)
    out = D.cell_diagnostics(W, 6.0)
    for c in range(2):
        M = np.zeros((8, 8, 2, 2), complex)
        M[..., 0, 0] = W[c, 0] + W[c, 3]
        M[..., 1, 1] = W[c, 0] - W[c, 3]
        M[..., 1, 0] = W[c, 1] + 1j * W[c, 2]
        M[..., 0, 1] = W[c, 1] - 1j * W[c, 2]
        ev = np.linalg.eigvalsh(M)
        assert abs(out[c, 8] - ev[..., 0].min()) < 1e-14
        assert abs(out[c, 9] - ev[..., 1].max()) < 1e-14
        S = (D.binary_entropy(ev[..., 0]) + D.binary_entropy(ev[..., 1])).sum() * (12.0 / 8) ** 2
        assert abs(out[c, 7] - S) < 1e-12 * abs(S)


def test_entropy_closed_forms():
    N, L = 16, 6.0
    W = np.zeros((2, 4, N, N))
    W[0, 0] = 0.5                       # eigenvalues 1/2, 1/2: s = ln 2 each
    W[1, 0] = 0.5
    W[1, 3] = 0.5                       # eigenvalues 1 and 0: s = 0
    out = D.cell_diagnostics(W, L)
    assert abs(out[0, 7] - 2 * np.log(2) * (2 * L) ** 2) < 1e-12
    assert out[1, 7] == 0.0
    assert out[1, 8] == 0.0 and out[1, 9] == 1.0


def test_moments_of_shifted_gaussian():
    N, L = 64, 6.0
    v = np.array([k if k < N // 2 else k - N for k in range(N)]) * (2 * L / N)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    u, s2 = (0.7, -0.4), 0.49          # variance small enough that the box truncates below 1e-12
    g = np.exp(-((vx - u[0]) ** 2 + (vy - u[1]) ** 2) / (2 * s2)) / (2 * np.pi * s2)
    W = np.zeros((1, 4, N, N))
    W[0, 0] = g
    out = D.cell_diagnostics(W, L)[0]
    assert abs(out[0] - 1) < 1e-10
    assert abs(out[4] - u[0]) < 1e-10 and abs(out[5] - u[1]) < 1e-10
    assert abs(out[6] - (s2 + 0.5 * (u[0] ** 2 + u[1] ** 2))) < 1e-10    # <|v|^2/2> = T + |u|^2/2
