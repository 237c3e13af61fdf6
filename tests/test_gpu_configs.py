"""The configurations at their real sizes.  The two FP64 plain ones (velocity grid 32 x 32 on [-6, 6)^2,
Gauss-Legendre order 16 for I3):

* homogeneous, two seeded Gaussians, midpoint collision step dt = 0.01, 200 steps
  (src/main_hom.c, src/simulation.c:1219-1281): Cc and Cd at t = 0 and the first two steps against
  the oracle (1e-12), then the 200 steps on the GPU: mass (w0) kept to 1e-8 (the plain method
  conserves it only to the quadrature's accuracy), eigenvalues of W(v) in [0, 1] to 1e-5, entropy
  increasing in every step (H-theorem);
* periodic, Nx = 128 cells on [0, 1), CFL 0.5, 100 steps (src/main_inhom.c,
  src/simulation.c:1288-1301): the first Strang step against the oracle on all 128 cells (1e-12;
  quadrature J = 8, M = 4, which keeps the oracle at seconds), then 100 steps with J = 32, M = 16 on
  the GPU: the device-resident loop equals repeated single steps, total mass drifts by less than
  1e-8, eigenvalues stay in [0, 1] to 1e-5;

and the scale configuration's shape (N = 128, Nx = 2048), see its test.
"""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel
from oracle import diag_ref as D

pytestmark = pytest.mark.gpu

N, J, M = 32, 32, 16


def _eigen_range(W):
    r = np.sqrt((W[:, 1:] ** 2).sum(axis=1))
    return float((W[:, 0] - r).min()), float((W[:, 0] + r).max())


def test_homogeneous_config_200_steps():
    import torch
    import qbf
    from qbf import workloads as wl
    L, R, dt = wl.L_BOX, wl.R_BALL, 0.01
    q = O.Quad(N, J, L, R, M)
    e = qbf.Engine(N, qbf.FP64, max_cells=1)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    W0 = wl.homogeneous(N, seed=0)
    WF = e.phys_to_fourier(torch.from_numpy(W0).cuda())
    WFh = WF.cpu().numpy()[0]
    cc = e.collision(WF, qbf.CC).cpu().numpy()[0]
    cd = e.collision(WF, qbf.CD).cpu().numpy()[0]
    cc_ref, cd_ref = O.CcInt(q, WFh), O.CdInt(q, WFh)
    # the generator polarises every velocity along one direction, so the commutator part Cc vanishes
    # analytically (both results are rounding noise): its error is measured against |Cc| + |Cd|
    scale = np.abs(cc_ref).sum() + np.abs(cd_ref).sum()
    ecc, ecd = np.abs(cc - cc_ref).sum() / scale, l1rel(cd, cd_ref)
    print(f"|Cc| / |Cd| = {np.abs(cc_ref).sum() / np.abs(cd_ref).sum():.2e}; Cc, Cd at t = 0 vs oracle {ecc:.3e} {ecd:.3e}",
          flush=True)
    assert ecc < 1e-12 and ecd < 1e-12
    ref = O.simulation_homogeneous(q, W0[0], dt, 3)
    W = torch.from_numpy(W0).cuda()
    frames = [W0[0]]
    for it in range(200):
        WF = e.collision_step(e.phys_to_fourier(W), dt)
        W = e.fourier_to_phys(WF)
        frames.append(W.cpu().numpy()[0])
    for it in (1, 2):
        err = l1rel(frames[it], ref[it])
        print(f"step {it} vs oracle {err:.3e}", flush=True)
        assert err < 1e-12
    d = D.cell_diagnostics(np.stack(frames), L)
    mass = d[:, 0]
    print(f"mass drift {abs(mass[-1] - mass[0]) / mass[0]:.3e}, eigenvalues {d[:, 8].min():.3e} .. {d[:, 9].max():.6f}, "
          f"entropy {d[0, 7]:.6f} -> {d[-1, 7]:.6f}")
    assert abs(mass[-1] - mass[0]) < 1e-8 * mass[0]                 # measured 5.3e-10
    assert d[:, 8].min() > -1e-5 and d[:, 9].max() < 1 + 1e-5       # measured -5.4e-7 .. 0.69
    print(f"smallest entropy increase of one step {np.diff(d[:, 7]).min():.3e}")
    assert np.all(np.diff(d[:, 7]) > 0)                              # every step (measured >= 0.066)


def test_periodic_config_128_cells_100_steps():
    import torch
    import qbf
    from qbf import workloads as wl
    L, R, Nx = wl.L_BOX, wl.R_BALL, 128
    h, dt = 1.0 / Nx, wl.cfl_dt(N, Nx)
    W0 = wl.periodic(N, Nx, seed=1)
    # first step on all cells against the oracle
    q = O.Quad(N, 8, L, R, 4)
    e = qbf.Engine(N, qbf.FP64, max_cells=Nx)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    got = e.inhom_step_periodic(torch.from_numpy(W0).cuda(), L, h, dt).cpu().numpy()
    err = l1rel(got, O.inhom_step_periodic(q, W0, h, dt))
    print(f"step 1 on {Nx} cells vs oracle {err:.3e}", flush=True)
    assert err < 1e-12
    # 100 steps with the configuration's quadrature, device-resident
    e.set_quadrature(*qbf.quadrature(N, J, L, R, M))
    frames = e.simulate_periodic(torch.from_numpy(W0).cuda(), L, h, dt, 100, stride=33)
    assert frames.shape[0] == 4
    W = torch.from_numpy(W0).cuda()
    for it in range(1, 34):
        W = e.inhom_step_periodic(W, L, h, dt)
    assert np.array_equal(W.cpu().numpy(), frames[1])
    m0, m1 = W0[:, 0].sum(), frames[-1][:, 0].sum()
    lo, hi = _eigen_range(frames[-1])
    print(f"mass drift after 99 steps {abs(m1 - m0) / m0:.3e}, eigenvalues {lo:.3e} .. {hi:.6f}")
    assert abs(m1 - m0) < 1e-8 * m0                                   # measured 3.8e-10
    assert lo > -1e-5 and hi < 1 + 1e-5                               # measured -6.6e-7 .. 0.87


def test_scale_config_n128_2048_cells():
    """The scale configuration's shape: N = 128 (padded 256 x 256), Nx = 2048 cells, generator modes
    k <= 8, FP32 convolutions with the FP64 conservation correction, 4 device-resident steps.  The
    quadrature is J = 4, M = 2 here (seconds); the full J = 32, M = 16 step is timed by bench.py
    (--N 128 --Nx 2048).  Checks: the device-resident loop equals repeated single steps, total mass
    and energy are conserved (transport in flux form, collision corrected), eigenvalues stay in
    [0, 1], and the first FP32 step agrees with the FP64 step (both conservation-corrected) on all 2048
    cells."""
    import torch
    import qbf
    from qbf import workloads as wl
    N128, Nx = 128, 2048
    L, h, dt = wl.L_BOX, 1.0 / Nx, wl.cfl_dt(N128, Nx)
    tabs = qbf.quadrature(N128, 4, L, wl.R_BALL, 2)
    W0 = wl.periodic(N128, Nx, seed=5, kmax=8)
    Wd = torch.from_numpy(W0).cuda()
    print("N=128, Nx=2048: FP32 engine, 4 device-resident steps", flush=True)
    e = qbf.Engine(N128, qbf.FP32, max_cells=Nx)
    e.set_quadrature(*tabs)
    e.set_moment_correction(qbf.CONS_ALL)
    frames = e.simulate_periodic(Wd, L, h, dt, 4)
    assert frames.shape[0] == 4
    w1 = e.inhom_step_periodic(Wd, L, h, dt).cpu().numpy()
    assert np.array_equal(w1, frames[1])
    e.close()
    m = np.stack([wl.moments(f).sum(axis=0) for f in frames])      # [step][mass, spin3, momentum2, energy]
    dm = np.abs(m[:, 0] - m[0, 0]).max() / m[0, 0]
    de = np.abs(m[:, 6] - m[0, 6]).max() / m[0, 6]
    lo, hi = _eigen_range(frames[-1])
    print(f"mass drift {dm:.2e}, energy drift {de:.2e}, eigenvalues {lo:.2e} .. {hi:.6f}", flush=True)
    assert dm < 1e-12 and de < 1e-12
    assert lo > -1e-5 and hi < 1 + 1e-5
    del frames
    print("N=128, Nx=2048: FP64 step (same conservation correction)", flush=True)
    e = qbf.Engine(N128, qbf.FP64, max_cells=Nx)
    e.set_quadrature(*tabs)
    e.set_moment_correction(qbf.CONS_ALL)
    w64 = e.inhom_step_periodic(Wd, L, h, dt).cpu().numpy()
    e.close()
    err = l1rel(w1, w64)
    print(f"FP32 vs FP64 after one step: {err:.2e} (increment {np.abs(w1 - w64).sum() / np.abs(w64 - W0).sum():.2e})")
    assert err < 1e-7
