"""HIP-graph capture of the periodic Strang step (qbf_graph_capture_inhom_step_periodic).  Every
replay, with the program's own eager work and synchronisations between the replays, must equal the
eager step on the same input bit for bit (the same kernels in the same order; the engine's kernels
have no atomics).  The step path records no memset nodes: the accumulators are cleared by the
kernels themselves."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qbf  # noqa: E402
from qbf import workloads as wl  # noqa: E402

CASES = [(32, qbf.FP64, 0), (64, qbf.FP32, 2), (64, qbf.FP32, 0), (128, qbf.FP32, 0)]


@pytest.mark.parametrize("N,prec,pairs", CASES)
def test_graph_replays_equal_eager_steps(N, prec, pairs):
    Nx = 6
    L, h, dt = wl.L_BOX, 1.0 / Nx, wl.cfl_dt(N, Nx)
    print(f"graph capture N={N} precision {prec} pairs {pairs}", flush=True)
    e = qbf.Engine(N, prec, max_cells=Nx)
    e.set_quadrature(*qbf.quadrature(N, 4, L, wl.R_BALL, 2))
    e.set_moment_correction(qbf.CONS_ALL)
    if N != 128:
        e.set_real_pairs(pairs, 0.0)
    rng = np.random.default_rng(N + prec)
    B = torch.from_numpy(0.3 * rng.standard_normal((Nx, 3))).cuda()
    states = [torch.from_numpy(wl.periodic(N, Nx, seed=s)).cuda() for s in range(4)]
    Win = states[0].clone()
    g, out = e.capture_inhom_step_periodic(Win, L, h, dt, Bext=B)
    torch.cuda.synchronize()
    ref = e.inhom_step_periodic(states[0], L, h, dt, Bext=B)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)                  # the eager step run by the capture call
    for k in (1, 2, 3, 1):
        Win.copy_(states[k])                      # eager work between the replays
        out.zero_()
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        ref = e.inhom_step_periodic(states[k], L, h, dt, Bext=B)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), k
    g.close()
    e.close()


def test_capture_needs_a_fixed_pair_choice():
    e = qbf.Engine(16, qbf.FP32, max_cells=2)     # paired path chosen by measurement (the default)
    try:
        q = qbf.quadrature(16, 4, wl.L_BOX, wl.R_BALL, 2)
        e.set_quadrature(*q)
        W = torch.from_numpy(wl.periodic(16, 2, seed=1)).cuda()
        with pytest.raises(qbf.QBFError):
            e.capture_inhom_step_periodic(W, wl.L_BOX, 0.5, 1e-4)
    finally:
        e.close()
