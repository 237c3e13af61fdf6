"""N = 128 (padded 256 x 256 grids, the kernels of qbf_big.hip) against the CPU oracle, through the
C ABI of libqbf.so.  Tolerances as for the smaller grids: FP64 1e-12 L1-relative, FP32 2e-5.  The
quadrature orders are small (J = 4, M = 2) so that the oracle finishes in seconds; the kernels run
the same code for every order.  (Odd J puts a pair of I1 angles at cos(theta1 - theta2) ~ 1e-17, whose
principal-value weights of ~1e17 turn Cc into cancellation noise in any implementation.)
"""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel, rand_state

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qbf  # noqa: E402

N, L, R = 128, 6.0, 3.75
TOL64, TOL32 = 1e-12, 2e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def e64():
    e = qbf.Engine(N, qbf.FP64, max_cells=3)
    yield e
    e.close()


def test_n128_state_transforms(e64):
    rng = np.random.default_rng(5)
    W = rng.standard_normal((5, N, N))
    print("N=128 state transforms, 5 grids", flush=True)
    WF = host(e64.phys_to_fourier(dev(W)))
    assert np.abs(WF - np.fft.fft2(W) / N ** 2).max() < 1e-15
    back = host(e64.fourier_to_phys(dev(WF)))
    assert np.abs(back - W).max() < 1e-13


def test_n128_integrals_vs_oracle(e64):
    rng = np.random.default_rng(6)
    f, g, h = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N)) for _ in range(3))
    args = tuple(dev(a) for a in (f, g, h))
    print("N=128 I1..I4 vs oracle", flush=True)
    p1, p2 = O.fourier_I1(N, 4, L, R)
    e64.set_i1_table(qbf.TAB_I1, p1, p2)
    err = l1rel(host(e64.I1(*args, qbf.TAB_I1)), O.I1integral(f, g, h, p1, p2))
    print(f"I1 {err:.2e}", flush=True)
    assert err < TOL64
    p1, p2 = O.fourier_I4(N, 4, L, R)
    e64.set_i1_table(qbf.TAB_I4, p1, p2)
    err = l1rel(host(e64.I1(*args, qbf.TAB_I4)), O.I1integral(f, g, h, p1, p2))
    print(f"I4 {err:.2e}", flush=True)
    assert err < TOL64
    p1, p2 = O.fourier_I2(N, 4, L, R)
    e64.set_i2_table(p1, p2)
    err = l1rel(host(e64.I2(*args)), O.I2integral(f, g, h, p1, p2))
    print(f"I2 {err:.2e}", flush=True)
    assert err < TOL64
    p1, p2 = O.fourier_I3(N, 4, L, R, 2)
    e64.set_i3_table(p1, p2)
    err = l1rel(host(e64.I3(*args)), O.I3integral(f, g, h, p1, p2))
    print(f"I3 {err:.2e}", flush=True)
    assert err < TOL64


@pytest.mark.parametrize("prec,tol", [(qbf.FP64, TOL64), (qbf.FP32, TOL32)])
def test_n128_collision_step_vs_oracle(prec, tol):
    rng = np.random.default_rng(128 + prec)
    q = O.Quad(N, 4, L, R, 2)
    Wp = rand_state(rng, 2, N)
    WF = np.stack([np.stack([O.r2c_fill(Wp[c, i]) for i in range(4)]) for c in range(2)])
    B = np.array([[0.3, -0.2, 0.5], [0.0, 0.1, -0.4]])
    print(f"N=128 collision step, precision {prec}", flush=True)
    e = qbf.Engine(N, prec, max_cells=2)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    out = host(e.collision_step(dev(WF), 0.01, dev(B)))
    e.close()
    ref = np.stack([O.collision_step(q, WF[c], 0.01, B[c]) for c in range(2)])
    err, derr = l1rel(out, ref), l1rel(out - WF, ref - WF)
    print(f"step {err:.2e}, increment {derr:.2e}")
    assert err < tol
    assert derr < 1e3 * tol


def test_n128_inhom_step_periodic_vs_oracle(e64):
    from qbf import workloads as wl
    q = O.Quad(N, 4, L, R, 2)
    e64.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    W0 = wl.periodic(N, 3, seed=8, kmax=8)
    h, dt = 1.0 / 3, wl.cfl_dt(N, 3)
    print("N=128 periodic Strang step on 3 cells", flush=True)
    got = host(e64.inhom_step_periodic(dev(W0), L, h, dt))
    err = l1rel(got, O.inhom_step_periodic(q, W0, h, dt))
    print(f"step {err:.2e}")
    assert err < TOL64


def _nyquist_free(WF):
    WF = WF.clone()
    WF[..., N // 2, :] = 0
    WF[..., :, N // 2] = 0
    return WF


def test_n128_paired_path_exact_on_nyquist_free_states_fp64():
    """k_i3acc_big with PAIR (and the paired I1 / I4 chains): on spectra of real states with empty
    Nyquist lines every grid K_q acts on is real, and the paired evaluation equals the complex one up
    to rounding, and the oracle."""
    from qbf import workloads as wl
    q = O.Quad(N, 4, L, R, 2)
    e = qbf.Engine(N, qbf.FP64, max_cells=2)
    try:
        e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
        WF = _nyquist_free(e.phys_to_fourier(dev(wl.periodic(N, 2, seed=1, kmax=8))))
        print("N=128 paired path, FP64, 2 cells", flush=True)
        e.set_real_pairs(0)
        c0 = e.collision(WF)
        e.set_real_pairs(2)
        c1 = e.collision(WF)
        used, _ = e.real_pairs_status()
        d = float((c1 - c0).abs().sum() / c0.abs().sum())
        print(f"paired vs complex {d:.3e}", flush=True)
        assert used and d < 1e-13
        W0 = WF[0].cpu().numpy()
        want = O.CcInt(q, W0) + O.CdInt(q, W0)
        err = l1rel(host(c1)[0], want)
        print(f"paired vs oracle {err:.3e}")
        assert err < TOL64
    finally:
        e.close()


def test_n128_paired_path_fp32_bounded_by_measured_ratio():
    """FP32 at N = 128 on workload states (modes k <= 8): the paired path deviates from the complex one
    by at most the dropped imaginary parts, bounded by the measured ratio, and the automatic mode
    (the FP32 default, tolerance 2e-5) takes it when the ratio is below the tolerance."""
    from qbf import workloads as wl
    nx = 4
    tabs = qbf.quadrature(N, 4, L, R, 2)
    e32 = qbf.Engine(N, qbf.FP32, max_cells=nx)
    e64 = qbf.Engine(N, qbf.FP64, max_cells=nx)
    try:
        e32.set_quadrature(*tabs)
        e64.set_quadrature(*tabs)
        WF = e64.phys_to_fourier(dev(wl.periodic(N, nx, seed=3, kmax=8)))
        print("N=128 paired path, FP32, 4 cells", flush=True)
        ref = e64.collision(WF)
        e32.set_real_pairs(0)
        cc = e32.collision(WF)
        e32.set_real_pairs(1, 2e-5)
        cp = e32.collision(WF)
        used, ratio = e32.real_pairs_status()
        l1 = lambda a, b: float((a - b).abs().sum() / b.abs().sum())  # noqa: E731
        ec, ep, d = l1(cc, ref), l1(cp, ref), l1(cp, cc)
        print(f"ratio {ratio:.3e} used {used}: complex vs FP64 {ec:.3e}, paired vs FP64 {ep:.3e}, "
              f"paired vs complex {d:.3e}")
        assert used == (ratio <= 2e-5)
        assert d < 1e-5 + 10 * ratio
        assert ep < 2 * ec + 10 * ratio
    finally:
        e32.close()
        e64.close()


def test_n128_boundary_steps_vs_oracle(e64):
    """Dirichlet and Maxwell Strang steps at N = 128 (transport kernels take N at run time; the
    collision of the interior / all cells runs the N = 128 kernels) against the oracle."""
    import torch
    q = O.Quad(N, 4, L, R, 2)
    e64.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    rng = np.random.default_rng(21)
    nx, h, dt, lam = 3, 0.08, 0.002, 0.4
    W = rand_state(rng, nx, N, L)
    WL, WR = rand_state(rng, 2, N, L)
    B = rng.standard_normal(3 * nx) * 0.3
    print("N=128 Dirichlet and Maxwell steps on 3 cells", flush=True)
    got = host(e64.inhom_step_dirichlet(dev(W), L, h, dt, dev(B)))
    err = l1rel(got, O.inhom_step_dirichlet(q, W, h, dt, B))
    print(f"Dirichlet {err:.2e}", flush=True)
    assert err < TOL64
    got = host(e64.inhom_step_maxwell(dev(W), L, h, dt, lam, dev(WL), dev(WR), dev(B)))
    err = l1rel(got, O.inhom_step_maxwell(q, W, h, dt, lam, WL, WR, B))
    print(f"Maxwell {err:.2e}")
    assert err < TOL64
