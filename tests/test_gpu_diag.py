"""qbf_cell_diagnostics against the oracle (oracle/diag_ref.py) on random and workload states."""
import numpy as np
import pytest

from conftest import rand_state
from oracle import diag_ref as D

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [16, 32, 64, 128])
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


@pytest.mark.parametrize("N", [32, 64, 128])
def test_fermi_dirac_fit_vs_oracle(N):
    """qbf_fermi_dirac against asymptoticFD.m as restated in the oracle (least squares on the three
    moment equations; the kernel solves the equivalent one-dimensional equation in beta)."""
    import torch
    import qbf
    from qbf import workloads as wl
    L = wl.L_BOX
    rng = np.random.default_rng(N + 1)
    W = np.concatenate([rand_state(rng, 3, N, L), wl.periodic(N, 4, seed=1, t=0.02)])
    W = np.concatenate([W, np.zeros((1, 4, N, N))])           # no Fermi-Dirac state: status 1
    e = qbf.Engine(N, qbf.FP64, max_cells=W.shape[0])
    prm, wfd = e.fermi_dirac(torch.from_numpy(W).cuda(), L)
    prm, wfd = prm.cpu().numpy(), wfd.cpu().numpy()
    for c in range(W.shape[0] - 1):
        WFD, beta, mu, vel, lam, res = D.asymptotic_fd(W[c], L)
        print(f"cell {c}: beta {prm[c, 0]:.15g} vs {beta:.15g}, mu {prm[c, 1:3]} vs {mu}")
        assert prm[c, 11] == 0 and res < 1e-9
        assert abs(prm[c, 0] - beta) < 1e-9 * beta
        assert np.abs(prm[c, 1:3] - mu).max() < 1e-9 * (1 + np.abs(mu).max())
        assert np.abs(prm[c, 3:5] - vel).max() < 1e-12 * (1 + np.abs(vel).max())
        assert np.abs(prm[c, 8:10] - lam).max() < 1e-12 * np.abs(lam).max()
        assert np.abs(wfd[c] - WFD).max() < 1e-9
    assert prm[-1, 11] == 1 and not np.any(wfd[-1])


@pytest.mark.parametrize("N", [32, 128])
def test_fd_distance_and_cell_average_vs_oracle(N):
    """qbf_fd_distance (per-cell and global equilibrium) and qbf_cell_average against the oracle,
    on the oracle's own Fermi-Dirac states."""
    import torch
    import qbf
    L, n = 6.0, 4
    rng = np.random.default_rng(9)
    W = rand_state(rng, n, N, L)
    e = qbf.Engine(N, qbf.FP64, max_cells=n)
    Wd = torch.from_numpy(W).cuda()
    avg = e.cell_average(Wd).cpu().numpy()[0]
    assert np.array_equal(avg, D.cell_average(W))
    loc = np.stack([D.asymptotic_fd(W[c], L)[0] for c in range(n)])
    glob = D.asymptotic_fd(avg, L)[0]
    got_loc = e.fd_distance(Wd, L, torch.from_numpy(loc).cuda()).cpu().numpy()
    got_glob = e.fd_distance(Wd, L, torch.from_numpy(glob).cuda()).cpu().numpy()
    for c in range(n):
        for got, ref in ((got_loc[c], D.fd_distance(W[c], loc[c], L)), (got_glob[c], D.fd_distance(W[c], glob, L))):
            print(f"cell {c}: distance {got[0]:.12g} vs {ref[0]:.12g}, relative entropy {got[1]:.12g} vs {ref[1]:.12g}")
            assert abs(got[0] - ref[0]) < 1e-12 * ref[0]
            assert abs(got[1] - ref[1]) < 1e-11 * abs(ref[1])
