"""HIP-graph capture of the periodic Strang step (qbf_graph_capture_inhom_step_periodic) and of the
surrogate Strang step (qbf_graph_capture_surrogate_inhom_step_periodic).  Every
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

CASES = [(32, qbf.FP64, 0), (64, qbf.FP32, 2), (64, qbf.FP32, 0), (64, qbf.BF16X3, 2), (128, qbf.FP32, 0),
         (128, qbf.FP32, 2)]


@pytest.mark.parametrize("N,prec,pairs", CASES)
def test_graph_replays_equal_eager_steps(N, prec, pairs):
    Nx = 6
    L, h, dt = wl.L_BOX, 1.0 / Nx, wl.cfl_dt(N, Nx)
    print(f"graph capture N={N} precision {prec} pairs {pairs}", flush=True)
    e = qbf.Engine(N, prec, max_cells=Nx)
    e.set_quadrature(*qbf.quadrature(N, 4, L, wl.R_BALL, 2))
    e.set_moment_correction(qbf.CONS_ALL)
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


def test_replays_are_ordered_with_the_current_stream():
    """No synchronisation between eager work and replays: the input written on torch's stream just
    before launch() is what the replay reads, and a read queued after launch() sees its result."""
    N, Nx = 32, 5
    L, h, dt = wl.L_BOX, 1.0 / Nx, wl.cfl_dt(N, Nx)
    e = qbf.Engine(N, qbf.FP64, max_cells=Nx)
    e.set_quadrature(*qbf.quadrature(N, 4, L, wl.R_BALL, 2))
    states = [torch.from_numpy(wl.periodic(N, Nx, seed=s)).cuda() for s in range(3)]
    refs = [e.inhom_step_periodic(s, L, h, dt) for s in states]
    Win = states[0].clone()
    g, out = e.capture_inhom_step_periodic(Win, L, h, dt)
    torch.cuda.synchronize()
    got = []
    for k in (1, 2, 0):
        Win.copy_(states[k])
        g.launch()
        got.append(out.clone())
    torch.cuda.synchronize()
    for k, x in zip((1, 2, 0), got):
        assert torch.equal(x, refs[k]), k
    g.close()
    e.close()


def test_a_live_graph_pins_the_engine_buffers():
    """While a captured graph is alive the engine refuses to grow (reallocate) the buffers the graph
    points to; after the graph is destroyed it grows them again.  An engine destroyed before its graph
    leaves the graph dead: launch() raises instead of touching freed memory."""
    N = 32
    L, dt = wl.L_BOX, 1e-4
    e = qbf.Engine(N, qbf.FP64, max_cells=8)
    e.set_quadrature(*qbf.quadrature(N, 4, L, wl.R_BALL, 2))
    W4 = torch.from_numpy(wl.periodic(N, 4, seed=1)).cuda()
    W8 = torch.from_numpy(wl.periodic(N, 8, seed=1)).cuda()
    g, out = e.capture_inhom_step_periodic(W4, L, 0.25, dt)
    with pytest.raises(qbf.QBFError, match="captured graph"):
        e.inhom_step_periodic(W8, L, 0.125, dt)
    again = e.inhom_step_periodic(W4, L, 0.25, dt)      # no growth: allowed
    g.launch()
    torch.cuda.synchronize()
    assert torch.equal(out, again)
    g.close()
    e.inhom_step_periodic(W8, L, 0.125, dt)           # grows now
    g2, _ = e.capture_inhom_step_periodic(W4, L, 0.25, dt)
    e.close()
    with pytest.raises(qbf.QBFError):
        g2.launch()
    g2.close()


@pytest.mark.parametrize("N", [64, 128])
def test_surrogate_step_graph_replays_equal_eager_steps(N):
    """The surrogate Strang step (qbf_graph_capture_surrogate_inhom_step_periodic): replays equal eager
    steps bit for bit, and an offset field rewritten in place between replays is what a replay uses."""
    from qbf.cnn import CNN
    Nx = 6
    L, h, dt = wl.L_BOX, 1.0 / Nx, wl.cfl_dt(N, Nx)
    print(f"surrogate graph N={N}", flush=True)
    e = qbf.Engine(N, qbf.FP32, max_cells=Nx)
    e.set_quadrature(*qbf.quadrature(N, 4, L, wl.R_BALL, 2))
    net = CNN(max_batch=Nx, seed=3, N=N)
    rng = np.random.default_rng(N)
    D = torch.from_numpy((1e-3 * rng.standard_normal((Nx, 4, N, N))).astype(np.float32)).cuda()
    net.set_offset(D)
    states = [torch.from_numpy(wl.periodic(N, Nx, seed=s)).cuda() for s in range(3)]
    isc, osc = 2.0, 0.05
    Win = states[0].clone()
    g, out = net.capture_inhom_step_periodic(e, Win, L, h, dt, isc, osc, mask=qbf.CONS_ALL)
    torch.cuda.synchronize()
    ref = net.inhom_step_periodic(e, states[0], L, h, dt, isc, osc, mask=qbf.CONS_ALL)
    torch.cuda.synchronize()
    assert torch.equal(out, ref)
    for k, scale in ((1, 1.0), (2, -2.0), (1, 0.5)):
        Win.copy_(states[k])
        D.mul_(scale)                               # the offset field rewritten in place
        out.zero_()
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        ref = net.inhom_step_periodic(e, states[k], L, h, dt, isc, osc, mask=qbf.CONS_ALL)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), k
        assert not torch.equal(out, states[k])
    g.close()
    net.close()
    e.close()


@pytest.mark.parametrize("N", [64, 128])
def test_policy_in_graph_mode_equals_the_eager_policy(N):
    """SurrogateStepper(graphs=True): the exact and the surrogate Strang steps are captured the second
    time they run between the same buffers and replayed afterwards, with the residual checks and the
    anchoring eager in between.  Trajectory, decisions and residuals equal the eager policy's bit for
    bit; the anchor after an exact-step replay reads that replay through last_collision, and the
    surrogate replays read the offset field rewritten in place."""
    from qbf import surrogate as S
    from qbf.cnn import CNN
    Nx = 6
    L, h, dt = wl.L_BOX, 1.0 / Nx, wl.cfl_dt(N, Nx)
    print(f"policy graph mode N={N}", flush=True)
    e = qbf.Engine(N, qbf.FP32, max_cells=Nx)
    e.set_quadrature(*qbf.quadrature(N, 4, L, wl.R_BALL, 2))
    e.set_real_pairs(2, 0.0)
    net = CNN(max_batch=Nx, seed=5, N=N)
    W0 = torch.from_numpy(wl.periodic(N, Nx, seed=2)).cuda()
    isc, osc = 2.0, 0.05

    def run(graphs):
        st = S.SurrogateStepper(e, net, isc, osc, L, h, dt, exact_every=3, sample_frac=0.34, tol=1e9, seed=1,
                                graphs=graphs)
        bufs = [W0.clone(), torch.empty_like(W0)]
        traj = []
        for i in range(14):
            st.step(bufs[i % 2], out=bufs[1 - i % 2])
            traj.append(bufs[1 - i % 2].clone())
        torch.cuda.synchronize()
        info = (dict(st.counts), list(st.residuals), st.replays, len(st._graphs))
        st.close()
        return traj, info

    ref, (c0, r0, _, _) = run(False)
    got, (c1, r1, replays, captured) = run(True)
    # exact steps 0, 3, 6, 9, 12 and surrogate steps between, alternating buffers: each kind meets both
    # buffer orders, is captured on its second run in an order and replayed on the third
    assert captured == 4 and replays == 6, (captured, replays)
    assert c0 == c1 == {"exact": 5, "surrogate": 9, "fallback": 0}, (c0, c1)
    assert r0 == r1
    for k, (a, b) in enumerate(zip(ref, got)):
        assert torch.equal(a, b), k
    net.close()
    e.close()
