"""Per-cell diagnostics of physical Wigner states -- TEST INFRASTRUCTURE ONLY (the oracle of
qbf_cell_diagnostics).

Restates visualization/wignerMoments.m:15-37 (moments without the trace normalisation),
visualization/entropy.m (binary entropy, zero outside (0, 1)) and visualization/visualizeHomBoltzmann.m:26-45
and 98-105 (eigenvalues of the 2x2 matrix w0 I + w.sigma, entropy summed over both and over v, times dv).
Velocity grid v = [0..N/2-1, -N/2..-1] * 2L/N with v_x on the slow index (src/simulation.c:151-160).
"""
import numpy as np


def binary_entropy(p):
    p = np.asarray(p, dtype=np.float64)
    out = np.zeros_like(p)
    m = (p > 0) & (p < 1)
    q = p[m]
    out[m] = -(q * np.log(q) + (1 - q) * np.log1p(-q))
    return out


def cell_diagnostics(W, L):
    """W: [ncell, 4, N, N] -> [ncell, 10]: mass, spin x/y/z, momentum x/y, energy, entropy,
    min eigenvalue, max eigenvalue."""
    W = np.asarray(W, dtype=np.float64)
    N = W.shape[-1]
    v = np.array([k if k < N // 2 else k - N for k in range(N)], dtype=np.float64) * (2 * L / N)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    dv = (2 * L / N) ** 2
    w0 = W[:, 0]
    r = np.sqrt(W[:, 1] ** 2 + W[:, 2] ** 2 + W[:, 3] ** 2)
    lp, lm = w0 + r, w0 - r
    out = np.stack([W[:, 0].sum((1, 2)), W[:, 1].sum((1, 2)), W[:, 2].sum((1, 2)), W[:, 3].sum((1, 2)),
                    (vx * w0).sum((1, 2)), (vy * w0).sum((1, 2)), (0.5 * (vx ** 2 + vy ** 2) * w0).sum((1, 2)),
                    (binary_entropy(lp) + binary_entropy(lm)).sum((1, 2))], axis=1) * dv
    return np.concatenate([out, lm.min((1, 2))[:, None], lp.max((1, 2))[:, None]], axis=1)
