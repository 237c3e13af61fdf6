"""Maxwell boundary step on the GPU (qbf_transport_maxwell, qbf_inhom_step_maxwell) against the
oracle: the transport bit for bit (trace fluxes summed in index order, same expression order, no
contraction), the full step to 1e-12 (FP64 collision of every cell)."""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel, maxwell_transport, rand_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,lam,N", [(1, 0.5, 16), (2, 0.0, 16), (5, 1.0, 16), (7, 0.35, 32), (3, 0.6, 64)])
def test_transport_maxwell_bit_exact(nx, lam, N):
    import torch
    import qbf
    L, h, dt = 6.0, 0.08, 0.004
    rng = np.random.default_rng(nx + 10 * N)
    W = rand_state(rng, nx, N, L)
    WL, WR = rand_state(rng, 2, N, L)
    print(f"Maxwell transport nx={nx} lambda={lam} N={N}", flush=True)
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    got = e.transport_maxwell(cu(W), L, h, 0.5 * dt, lam, cu(WL), cu(WR)).cpu().numpy()
    want = maxwell_transport(W, L, h, dt, lam, WL, WR)
    assert np.array_equal(got, want)


def test_inhom_step_maxwell_vs_oracle():
    import torch
    import qbf
    N, L, nx, h, dt, lam = 16, 6.0, 5, 0.08, 0.002, 0.4
    q = O.Quad(N, 6, L, 3.75, 4)
    rng = np.random.default_rng(12)
    W = rand_state(rng, nx, N, L)
    WL, WR = rand_state(rng, 2, N, L)
    B = rng.standard_normal(3 * nx) * 0.3
    want = O.inhom_step_maxwell(q, W, h, dt, lam, WL, WR, B)
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    got = e.inhom_step_maxwell(cu(W), L, h, dt, lam, cu(WL), cu(WR), cu(B)).cpu().numpy()
    err = l1rel(got, want)
    print("inhom step Maxwell vs oracle", err)
    assert err < 1e-12
