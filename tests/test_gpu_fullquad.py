"""Parity at the configurations' own quadrature (J = 32 angles, Gauss-Legendre order M = 16) against
the oracle, the reference's own formulation (src/collision.c:74-254, src/simulation.c:105-136,
307-555):

* N = 64 (the mixed-precision and surrogate configurations, seed 3 of the periodic workload, the
  bench's cells 0 and 200): the FP64 collision operator Cc + Cd to 1e-12 and the FP64 midpoint
  collision step, the FP32 operator (the paired path, which the bench's exact steps take) to the
  FP32 path's own accuracy;
* N = 32, Nx = 128 (the periodic FP64 configuration, CFL 0.5): the first Strang step on two cells,
  with the oracle's collisions evaluated on the five cells each of them depends on;
* the Maxwell half transport against the oracle's, bit for bit.

The oracle needs about 14 s per N = 64 evaluation and core; its evaluations run in threads."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from conftest import l1rel, rand_state

pytestmark = pytest.mark.gpu

J, M = 32, 16


def _fourier(W):
    return np.stack([O.r2c_fill(W[c]) for c in range(4)])


def test_n64_operator_and_step_at_full_quadrature():
    import torch
    import qbf
    from qbf import workloads as wl
    N, cells = 64, [0, 200]
    L, R = wl.L_BOX, wl.R_BALL
    dt = wl.cfl_dt(N, 512)
    print(f"N={N} J={J} M={M}: oracle evaluations on cells {cells} (threads)", flush=True)
    q = O.Quad(N, J, L, R, M)
    Wp = wl.periodic(N, 512, seed=3, cells=cells)
    WF = np.stack([_fourier(Wp[i]) for i in range(len(cells))])
    with ThreadPoolExecutor(max_workers=5) as ex:
        fcc = [ex.submit(O.CcInt, q, WF[i]) for i in range(2)]
        fcd = [ex.submit(O.CdInt, q, WF[i]) for i in range(2)]
        fst = ex.submit(O.collision_step, q, WF[0], dt)
        tabs = qbf.quadrature(N, J, L, R, M)
        dev = torch.from_numpy(WF).cuda()
        e64 = qbf.Engine(N, qbf.FP64, max_cells=2)
        e64.set_quadrature(*tabs)
        C64 = e64.collision(dev).cpu().numpy()
        S64 = e64.collision_step(dev[:1].contiguous(), dt).cpu().numpy()[0]
        e32 = qbf.Engine(N, qbf.FP32, max_cells=2)
        e32.set_quadrature(*tabs)
        C32 = e32.collision(dev).cpu().numpy()
        used, ratio = e32.real_pairs_status()
        ref = np.stack([fcc[i].result() + fcd[i].result() for i in range(2)])
        step_ref = fst.result()
    e64.close()
    e32.close()
    err64 = l1rel(C64, ref)
    err32 = l1rel(C32, ref)
    inc = l1rel(S64 - WF[0], step_ref - WF[0])
    st = l1rel(S64, step_ref)
    print(f"FP64 operator vs oracle {err64:.3e}; FP64 step: state {st:.3e}, increment {inc:.3e}; "
          f"FP32 operator (paired path {used}, imaginary/real {ratio:.2e}) {err32:.3e}", flush=True)
    assert err64 < 1e-12
    assert st < 1e-12
    # the increment dt C is 1e-4 of the state, so its relative error carries the state's rounding / 1e-4
    assert inc < 1e-9
    assert used
    # FP32 convolutions: rounding 6e-8 per operation through 28 FFT stages and the cancelling sums of
    # CdInt (measured 2.1e-5 on this workload)
    assert err32 < 4e-5


def test_periodic_config_first_step_at_full_quadrature():
    import torch
    import qbf
    from qbf import workloads as wl
    N, Nx, cells = 32, 128, [10, 77]
    L, R = wl.L_BOX, wl.R_BALL
    h, dt = 1.0 / Nx, wl.cfl_dt(N, Nx)
    print(f"periodic N={N} Nx={Nx} J={J} M={M}: oracle on cells {cells} and their neighbours", flush=True)
    q = O.Quad(N, J, L, R, M)
    W0 = wl.periodic(N, Nx, seed=0)
    with ThreadPoolExecutor(max_workers=1) as ex:
        fut = ex.submit(O.inhom_step_periodic_cells, q, W0, h, dt, cells)
        e = qbf.Engine(N, qbf.FP64, max_cells=Nx)
        e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
        got = e.inhom_step_periodic(torch.from_numpy(W0).cuda(), L, h, dt).cpu().numpy()
        want = fut.result()
    e.close()
    err = l1rel(got[cells], want[cells])
    inc = l1rel(got[cells] - W0[cells], want[cells] - W0[cells])
    print(f"first Strang step on cells {cells}: state {err:.3e}, increment {inc:.3e}", flush=True)
    assert err < 1e-12
    assert inc < 1e-9


@pytest.mark.parametrize("N,nx,lam", [(16, 6, 0.35), (32, 4, 1.0), (64, 3, 0.0)])
def test_transport_maxwell_vs_oracle_assembly(N, nx, lam):
    """The GPU Maxwell half transport against the oracle's own (the one its Maxwell step applies,
    pinned by the reflection symmetry), bit for bit."""
    import torch
    import qbf
    L, h, dt = 6.0, 0.08, 0.004
    rng = np.random.default_rng(N + nx)
    W = rand_state(rng, nx, N, L)
    WL, WR = rand_state(rng, 2, N, L)
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()   # noqa: E731
    got = e.transport_maxwell(cu(W), L, h, 0.5 * dt, lam, cu(WL), cu(WR)).cpu().numpy()
    e.close()
    assert np.array_equal(got, O.transport_maxwell(W, L, h, dt, lam, WL, WR))
