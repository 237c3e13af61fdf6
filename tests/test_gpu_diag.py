"""qbf_cell_diagnostics against the oracle (oracle/diag_ref.py) on random and workload states."""
import numpy as np
import pytest

from conftest import rand_state
from oracle import diag_ref as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [16, 32, 64])
def test_cell_diagnostics_vs_oracle(N):
    import torch
    import qbf
    from qbf import workloads as wl
    rng = np.random.default_rng(N)
    W = np.concatenate([rand_state(rng, 3, N, 6.0), wl.periodic(N, 8, seed=1, t=0.02)])
    e = qbf.Engine(N, qbf.FP64, max_cells=W.shape[0])
    got = e.cell_diagnostics(torch.from_numpy(W).cuda(), 6.0).cpu().numpy()
    want = D.cell_diagnostics(W, 6.0)
    scale = np.abs(want).max(axis=0) + 1e-300
    err = (np.abs(got - want) / scale).max()
    print("max relative error", err)
    assert err < 1e-12
