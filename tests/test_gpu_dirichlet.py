"""Dirichlet boundary step on the GPU (qbf_transport_dirichlet, qbf_inhom_step_dirichlet) against
the oracle: the transport bit for bit (same expression order, no contraction), the full step to
1e-12 (FP64 collision of the interior cells)."""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel, rand_state

pytestmark = pytest.mark.gpu


def _vgrid(N, L):
    return np.array([k if k < N // 2 else k - N for k in range(N)]) * (2 * L / N)


@pytest.mark.parametrize("nx", [2, 3, 9])
def test_transport_dirichlet_bit_exact(nx):
    import torch
    import qbf
    N, L, h, dt = 16, 6.0, 0.08, 0.004
    W = rand_state(np.random.default_rng(nx), nx, N, L)
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    got = e.transport_dirichlet(torch.from_numpy(W).cuda(), L, h, dt).cpu().numpy()
    v = _vgrid(N, L)
    want = np.empty_like(W)
    for c in range(4):
        for l in range(N):
            for k in range(N):
                want[:, c, l, k] = O.slope_limiter_dirichlet(h, dt, v[l], W[:, c, l, k])
    assert np.array_equal(got, want)


def test_inhom_step_dirichlet_vs_oracle():
    import torch
    import qbf
    N, L, nx, h, dt = 16, 6.0, 5, 0.08, 0.002
    q = O.Quad(N, 6, L, 3.75, 4)
    rng = np.random.default_rng(11)
    W = rand_state(rng, nx, N, L)
    B = rng.standard_normal(3 * nx) * 0.3
    want = O.inhom_step_dirichlet(q, W, h, dt, B)
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    got = e.inhom_step_dirichlet(torch.from_numpy(W).cuda(), L, h, dt, torch.from_numpy(B).cuda()).cpu().numpy()
    err = l1rel(got, want)
    print("inhom step Dirichlet vs oracle", err)
    assert err < 1e-12
