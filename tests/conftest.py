import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "inhom-quantum-boltzmann-fourier-c_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

GOLDEN = os.path.join(ROOT, "tests", "golden", "ref")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP) device")
    config.addinivalue_line("markers", "slow: CPU test that takes more than ~10 s")


def read_dat(name, dtype=np.float64, shape=None):
    a = np.fromfile(os.path.join(GOLDEN, name), dtype=dtype)
    return a.reshape(shape) if shape is not None else a


def l1rel(a, b):
    """L1-relative error sum|a-b| / sum|b| over complex entries (test/Cc_test.c:177-189)."""
    return float(np.abs(np.asarray(a) - np.asarray(b)).sum() / np.abs(np.asarray(b)).sum())


def rand_state(rng, ncell, N, L=6.0):
    """Seeded physical Wigner states: w0 a positive Gaussian bump, (w1,w2,w3) = 0.8 w0 n, |n| = 1."""
    v = np.array([k if k < N // 2 else k - N for k in range(N)]) * 2 * L / N
    vx, vy = np.meshgrid(v, v, indexing="ij")
    out = np.empty((ncell, 4, N, N))
    for c in range(ncell):
        cx_, cy_ = rng.uniform(-1.0, 1.0, 2)
        s = rng.uniform(0.8, 1.2)
        w0 = rng.uniform(0.2, 0.5) * np.exp(-((vx - cx_) ** 2 + (vy - cy_) ** 2) / (2 * s * s))
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        out[c, 0] = w0
        for i in range(3):
            out[c, i + 1] = 0.8 * w0 * n[i]
    return out


def maxwell_vx(N, L):
    """x velocities of the Maxwell grid, [N*N] with x on the slow index (src/simulation.c:155-161)."""
    sc = 2 * L / N
    row = np.array([sc * (l + 1) if l < N // 2 else sc * (l - N) for l in range(N)])
    return np.repeat(row, N)


def maxwell_transport(W, L, h, dt, lam, WL, WR):
    """Half transport with Maxwell boundaries assembled from the oracle's limiter and trace fluxes
    (src/simulation.c:882-1051).  The Maxwell grid drops the zero x-velocity row and repeats the -L
    row as +L; the boundary states are normalised by their incoming trace flux and the ghost cells
    take the outgoing trace fluxes of the end cells.  `dt` is the full step (the limiter runs dt/2)."""
    import oracle as O
    nv, _, N, _ = W.shape
    A = maxwell_vx(N, L)
    half = N * N // 2

    def to_m(c):
        c = c.reshape(-1)
        return np.concatenate([c[N:N + half], c[half:]])

    def from_m(m, c):
        out = c.reshape(-1).copy()
        out[N:half] = m[:half - N]
        out[half:] = m[half:]
        return out.reshape(N, N)

    WLm = np.stack([to_m(WL[j]) for j in range(4)])
    WRm = np.stack([to_m(WR[j]) for j in range(4)])
    WLm = WLm * (1.0 / O.rightward_flux(A, WLm[0]))
    WRm = WRm * (1.0 / O.leftward_flux(A, WRm[0]))
    fL = O.leftward_flux(A, to_m(W[0, 0]))
    fR = O.rightward_flux(A, to_m(W[-1, 0]))
    out = np.empty_like(W)
    for j in range(4):
        Wx = np.stack([to_m(W[i, j]) for i in range(nv)])
        Wx1 = O.slope_limiter_maxwell(h, 0.5 * dt, A, lam, fL, fR, WLm[j], WRm[j], Wx).reshape(nv, -1)
        for i in range(nv):
            out[i, j] = from_m(Wx1[i], W[i, j])
    return out
