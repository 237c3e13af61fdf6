"""Host logic of the surrogate stepping policy (qbf.surrogate.SurrogateStepper) on the CPU, with stand-in
engine and network objects that record the calls the policy makes (no device code runs):

* the cadence of exact steps, the residual check and its fallback on a non-finite or large residual;
* graph mode: each step kind runs eagerly the first time it meets a pair of buffers, is captured the
  second time and replayed afterwards; an engine in the automatic paired-path mode is fixed to the last
  eager exact step's choice before the first capture; the anchor offset is rewritten in place; close()
  destroys the graphs."""
import math

import pytest

torch = pytest.importorskip("torch")
from qbf import surrogate as S  # noqa: E402


class FakeGraph:
    def __init__(self, log, name):
        self.log, self.name, self.closed = log, name, False

    def launch(self):
        self.log.append(("replay", self.name))

    def close(self):
        self.closed = True


class FakeEngine:
    def __init__(self, pairs_used=True):
        self.log, self.pair_mode, self.pairs_used, self.mask = [], 1, pairs_used, None
        self.graphs = []

    def set_moment_correction(self, mask):
        self.mask = mask

    def inhom_step_periodic(self, W, L, h, dt, out=None):
        self.log.append(("exact", W.data_ptr(), out.data_ptr()))
        out.copy_(W + 1)
        return out

    def capture_inhom_step_periodic(self, W, L, h, dt, out=None):
        assert self.pair_mode in (0, 2), "captured with the automatic paired-path choice"
        self.log.append(("capture exact", W.data_ptr(), out.data_ptr()))
        out.copy_(W + 1)
        g = FakeGraph(self.log, "exact")
        self.graphs.append(g)
        return g, out

    def real_pairs_status(self):
        return self.pairs_used, 1e-8

    def set_real_pairs(self, mode, tol=0.0):
        self.log.append(("pairs", mode))
        self.pair_mode = mode

    def last_collision(self, ncell, with_state=True):
        return torch.zeros(ncell, 4, 2, 2, dtype=torch.float64), torch.ones(ncell, 4, 2, 2, dtype=torch.float64)


class FakeNet:
    def __init__(self, engine, residual=(1e-9, 1.0, 1.0)):
        self.e, self.res, self._offset, self.offsets_set = engine, residual, None, 0
        self.graphs = []

    def set_offset(self, D=None):
        self._offset = D
        self.offsets_set += 1

    def update_offset(self, D):
        if self._offset is None or self._offset.shape != D.shape:
            self.set_offset(D)
        else:
            self._offset.copy_(D)

    def forward(self, X, iscale, oscale):
        return torch.zeros_like(X)

    def residual(self, engine, W, cells, iscale, oscale):
        self.e.log.append(("check", len(cells)))
        return self.res

    def inhom_step_periodic(self, engine, W, L, h, dt, iscale, oscale, mask, out=None):
        self.e.log.append(("surrogate", W.data_ptr(), out.data_ptr()))
        out.copy_(W + 2)
        return out

    def capture_inhom_step_periodic(self, engine, W, L, h, dt, iscale, oscale, mask, out=None):
        self.e.log.append(("capture surrogate", W.data_ptr(), out.data_ptr()))
        out.copy_(W + 2)
        g = FakeGraph(self.e.log, "surrogate")
        self.graphs.append(g)
        return g, out


def run(st, n, W0):
    bufs = [W0.clone(), torch.empty_like(W0)]
    for i in range(n):
        st.step(bufs[i % 2], out=bufs[1 - i % 2])
    return bufs


def kinds(log):
    return [x[0] if x[0] != "replay" else "replay " + x[1] for x in log if x[0] not in ("check", "pairs")]


def test_eager_policy_cadence_and_fallbacks():
    e = FakeEngine()
    net = FakeNet(e)
    st = S.SurrogateStepper(e, net, 1.0, 1.0, 6.0, 0.1, 1e-3, exact_every=3, sample_frac=0.5)
    W0 = torch.zeros(4, 4, 2, 2, dtype=torch.float64)
    run(st, 7, W0)
    assert st.counts == {"exact": 3, "surrogate": 4, "fallback": 0}
    assert kinds(e.log) == ["exact", "surrogate", "surrogate"] * 2 + ["exact"]
    assert sum(1 for x in e.log if x[0] == "check") == 4 and all(x[1] == 2 for x in e.log if x[0] == "check")
    assert net.offsets_set == 1 + 3                 # cleared, then a new field after every exact step
    for bad in ((math.nan, 1.0, 1.0), (math.inf, 1.0, 1.0), (1.0, 1.0, 0.0), (10.0, 1.0, 1.0)):
        e2 = FakeEngine()
        st2 = S.SurrogateStepper(e2, FakeNet(e2, bad), 1.0, 1.0, 6.0, 0.1, 1e-3, exact_every=3, sample_frac=0.5)
        run(st2, 3, W0)
        assert st2.counts == {"exact": 1, "surrogate": 0, "fallback": 2}, bad


@pytest.mark.parametrize("used", [True, False])
def test_graph_mode_captures_on_the_second_run_and_replays_after(used):
    e = FakeEngine(pairs_used=used)
    net = FakeNet(e)
    st = S.SurrogateStepper(e, net, 1.0, 1.0, 6.0, 0.1, 1e-3, exact_every=3, sample_frac=0.5, graphs=True)
    W0 = torch.zeros(4, 4, 2, 2, dtype=torch.float64)
    bufs = run(st, 14, W0)
    # exact steps 0, 3, 6, 9, 12; buffer orders alternate, so each kind meets both orders
    assert kinds(e.log) == [
        "exact", "surrogate", "surrogate",                          # steps 0-2: first runs (eager)
        "exact", "capture surrogate", "capture surrogate",          # 3: B->A first; 4, 5: second runs
        "capture exact", "replay surrogate", "replay surrogate",    # 6: A->B second run
        "capture exact", "replay surrogate", "replay surrogate",    # 9: B->A second run
        "replay exact", "replay surrogate"]
    assert st.replays == 6 and len(st._graphs) == 4
    assert ("pairs", 2 if used else 0) in e.log and e.pair_mode == (2 if used else 0)
    # the steps' arithmetic ran in order whether eager, captured or (here, by the stand-ins) not replayed
    assert st.counts == {"exact": 5, "surrogate": 9, "fallback": 0}
    # the anchor offset: set once, then rewritten in place (the captured surrogate steps read it)
    assert net.offsets_set == 1 + 1                 # cleared, then set once
    st.close()
    assert all(g.closed for g in e.graphs + net.graphs) and not st._graphs
    del bufs


def test_graph_mode_keeps_a_fixed_pair_choice():
    e = FakeEngine()
    e.set_real_pairs(0)
    e.log.clear()
    st = S.SurrogateStepper(e, FakeNet(e), 1.0, 1.0, 6.0, 0.1, 1e-3, exact_every=2, sample_frac=0.5, graphs=True)
    run(st, 8, torch.zeros(2, 4, 2, 2, dtype=torch.float64))
    assert not any(x[0] == "pairs" for x in e.log) and e.pair_mode == 0


def test_fillers_never_capture():
    e = FakeEngine()
    st = S.SurrogateStepper(e, None, 1.0, 1.0, 6.0, 0.1, 1e-3, exact_every=3, filler="frozen", graphs=True)
    assert not st.graphs


def test_graph_mode_needs_the_output_buffer():
    e = FakeEngine()
    st = S.SurrogateStepper(e, FakeNet(e), 1.0, 1.0, 6.0, 0.1, 1e-3, exact_every=2, graphs=True)
    with pytest.raises(S.QBFError):
        st.step(torch.zeros(2, 4, 2, 2, dtype=torch.float64))
