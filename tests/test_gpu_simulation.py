"""qbf_simulate_periodic, qbf_simulate_dirichlet and qbf_simulate_maxwell (device-resident loops with
strided frame output, src/simulation.c:1288-1341) against repeated single steps, and the per-cell
diagnostics along a trajectory (mass conservation of the periodic problem, the eigenvalue range)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stride", [1, 2])
def test_simulate_periodic_matches_single_steps(stride):
    import torch
    import qbf
    from qbf import workloads as wl
    N, nx, steps = 32, 16, 5
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    e.set_quadrature(*qbf.quadrature(N, 8, wl.L_BOX, wl.R_BALL, 4))
    # the Fourier collision conserves mass only to quadrature accuracy (about 1e-3 over these
    # steps at J=8, M=4); with the conservation correction the periodic step keeps it to rounding
    e.set_moment_correction(qbf.CONS_ALL)
    h, dt = 1.0 / nx, wl.cfl_dt(N, nx)
    W0 = torch.from_numpy(wl.periodic(N, nx, seed=2)).cuda()
    frames = e.simulate_periodic(W0, wl.L_BOX, h, dt, steps, stride)
    assert frames.shape[0] == (steps - 1) // stride + 1
    W = W0.clone()
    ref = [W0.cpu().numpy()]
    for it in range(1, steps):
        W = e.inhom_step_periodic(W, wl.L_BOX, h, dt)
        ref.append(W.cpu().numpy())
    for f in range(frames.shape[0]):
        assert np.array_equal(frames[f], ref[f * stride]), f
    d0 = e.cell_diagnostics(W0, wl.L_BOX).cpu().numpy()
    d1 = e.cell_diagnostics(torch.from_numpy(ref[-1]).cuda(), wl.L_BOX).cpu().numpy()
    assert abs(d1[:, 0].sum() - d0[:, 0].sum()) < 1e-12 * abs(d0[:, 0].sum())   # periodic transport + corrected collision keep mass
    assert d1[:, 8].min() > -1e-3 and d1[:, 9].max() < 1.0                        # eigenvalues stay in [0, 1]


@pytest.mark.parametrize("kind,stride", [("dirichlet", 1), ("dirichlet", 2), ("maxwell", 1), ("maxwell", 3)])
def test_simulate_dirichlet_maxwell_match_single_steps(kind, stride):
    import torch
    import qbf
    from qbf import workloads as wl
    from conftest import rand_state
    N, nx, steps, lam = 32, 8, 5, 0.4
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    e.set_quadrature(*qbf.quadrature(N, 8, wl.L_BOX, wl.R_BALL, 4))
    h, dt = 1.0 / nx, wl.cfl_dt(N, nx)
    W0 = torch.from_numpy(wl.periodic(N, nx, seed=4)).cuda()
    rng = np.random.default_rng(7)
    WL, WR = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in rand_state(rng, 2, N, wl.L_BOX))
    if kind == "dirichlet":
        frames = e.simulate_dirichlet(W0, wl.L_BOX, h, dt, steps, stride)
        step = lambda W: e.inhom_step_dirichlet(W, wl.L_BOX, h, dt)   # noqa: E731
    else:
        frames = e.simulate_maxwell(W0, wl.L_BOX, h, dt, lam, WL, WR, steps, stride)
        step = lambda W: e.inhom_step_maxwell(W, wl.L_BOX, h, dt, lam, WL, WR)   # noqa: E731
    assert frames.shape[0] == (steps - 1) // stride + 1
    W = W0.clone()
    ref = [W0.cpu().numpy()]
    for it in range(1, steps):
        W = step(W)
        ref.append(W.cpu().numpy())
    for f in range(frames.shape[0]):
        assert np.array_equal(frames[f], ref[f * stride]), f
    assert not np.array_equal(ref[-1], ref[0])
    e.close()
