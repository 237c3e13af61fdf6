"""Seeded synthetic initial states for the benchmark configurations.

All generators are deterministic functions of (seed, sizes), so every machine and the
oracle see identical inputs.  States are physical Wigner states
[ncell, 4, N, N] (wignerV_t per cell): w0 is the trace part, (w1, w2, w3) the Pauli
vector, and every state has |(w1,w2,w3)| = 0.8 w0, so the eigenvalues w0 (1 +- 0.8)
stay in [0, 1] when max w0 <= 0.5.

Velocity grid: v = [0..N/2-1, -N/2..-1] * 2L/N per direction, v_x on the slow index l
(src/simulation.c:151-160).
"""
from __future__ import annotations

import numpy as np

L_BOX = 6.0       # velocity box [-6, 6)^2
R_BALL = 3.75     # collision-ball radius, R = L * 7.5/12 (the reference's ratio, src/main_hom.c:42-43)
J_ANG = 32        # angular quadrature points (reference default, src/main_hom.c:40)
M_GL = 16         # Gauss-Legendre order


def velocity_grid(N: int, L: float = L_BOX):
    v = np.array([k if k < N // 2 else k - N for k in range(N)], dtype=np.float64) * (2 * L / N)
    return np.meshgrid(v, v, indexing="ij")   # vx[l, k] = v[l], vy[l, k] = v[k]


def homogeneous(N: int = 32, seed: int = 0, L: float = L_BOX) -> np.ndarray:
    """Homogeneous state: sum of two isotropic Gaussians with centres in [-1.5, 1.5]^2, sigma in [0.6, 1.2],
    w0 amplitude in [0.2, 0.5], Pauli direction uniform on the sphere.  Returns [1, 4, N, N]."""
    rng = np.random.default_rng(seed)
    vx, vy = velocity_grid(N, L)
    G = np.zeros((N, N))
    for _ in range(2):
        c = rng.uniform(-1.5, 1.5, 2)
        s = rng.uniform(0.6, 1.2)
        G += np.exp(-((vx - c[0]) ** 2 + (vy - c[1]) ** 2) / (2 * s * s))
    w0 = rng.uniform(0.2, 0.5) * G / G.max()
    n = rng.standard_normal(3)
    n /= np.linalg.norm(n)
    W = np.empty((1, 4, N, N))
    W[0, 0] = w0
    for i in range(3):
        W[0, i + 1] = 0.8 * w0 * n[i]
    return W


def periodic(N: int, Nx: int, seed: int = 0, kmax: int = 4, L: float = L_BOX, t: float = 0.0,
             cells=None) -> np.ndarray:
    """Periodic inhomogeneous state on cells x_i = (i + 1/2)/Nx of [0, 1).  w0 is a Maxwellian in v
    whose density and temperature are 1 + 0.3 sin(2 pi x) + sum_{k<=kmax} 0.05 sin(2 pi k x + phase_k)
    (seeded phases), scaled so that max w0 = 0.5 at t = 0; the Pauli direction rotates about a
    seeded axis by an integer number of turns over the period.  With t > 0 the state is the free
    streaming of that initial state, W(x, v, t) = W(x - v_x t, v, 0) (transport without
    collisions), which is how the surrogate's training states leave local equilibrium.
    Returns [Nx, 4, N, N], or only the listed cells."""
    rng = np.random.default_rng(seed)
    ph_rho = rng.uniform(0, 2 * np.pi, kmax)
    ph_T = rng.uniform(0, 2 * np.pi, kmax)
    turns = int(rng.integers(1, 4))
    tilt = rng.uniform(0.3, 1.2)
    vx, vy = velocity_grid(N, L)
    v2 = vx ** 2 + vy ** 2

    def profiles(x):
        rho = 1 + 0.3 * np.sin(2 * np.pi * x) + sum(0.05 * np.sin(2 * np.pi * (k + 1) * x + ph_rho[k]) for k in range(kmax))
        T = 1 + 0.3 * np.sin(2 * np.pi * x + 1.0) + sum(0.05 * np.sin(2 * np.pi * (k + 1) * x + ph_T[k]) for k in range(kmax))
        phi = 2 * np.pi * turns * x
        n = np.stack([np.sin(tilt) * np.cos(phi), np.sin(tilt) * np.sin(phi), np.full(np.shape(x), np.cos(tilt))])
        return rho / (2 * np.pi * T) * np.exp(-v2 / (2 * T)), n

    x = (np.arange(Nx) + 0.5) / Nx
    w0_init, _ = profiles(x[:, None, None])
    scale = 0.5 / w0_init.max()
    if cells is not None:
        x = x[np.asarray(cells)]
    xs = x[:, None, None] - vx[None] * t          # foot of the characteristic, per cell and velocity
    w0, n = profiles(xs)
    w0 = w0 * scale
    W = np.empty((len(x), 4, N, N))
    W[:, 0] = w0
    for i in range(3):
        W[:, i + 1] = 0.8 * w0 * n[i]
    return W


def cfl_dt(N: int, Nx: int, cfl: float = 0.5, L: float = L_BOX) -> float:
    """dt = CFL * h / max|v_x|, h = 1/Nx; max|v_x| = L at index N/2 (src/simulation.c:153)."""
    return cfl * (1.0 / Nx) / L


def moments(W: np.ndarray, L: float = L_BOX) -> np.ndarray:
    """[mass, spin_x, spin_y, spin_z, momentum_x, momentum_y, energy] per state, summed over the
    last two axes with dv = (2L/N)^2 (visualization/wignerMoments.m:15-37 without normalisation)."""
    N = W.shape[-1]
    vx, vy = velocity_grid(N, L)
    dv = (2 * L / N) ** 2
    w = W.reshape(-1, 4, N, N)
    out = np.stack([w[:, 0].sum((-1, -2)), w[:, 1].sum((-1, -2)), w[:, 2].sum((-1, -2)), w[:, 3].sum((-1, -2)),
                    (vx * w[:, 0]).sum((-1, -2)), (vy * w[:, 0]).sum((-1, -2)),
                    (0.5 * (vx ** 2 + vy ** 2) * w[:, 0]).sum((-1, -2))], axis=1) * dv
    return out.reshape(W.shape[:-3] + (7,))
