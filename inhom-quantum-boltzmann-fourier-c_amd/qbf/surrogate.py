"""CNN surrogate of the collision operator (N = 64 or 128): training data, training on the GPU,
and the stepping policy that keeps it honest.

Training data are physical states of the periodic workload family (qbf.workloads.periodic)
sampled along exact-solver trajectories (trajectory_set), or at several free-streaming times
(training_set), with the exact collision operator C = Cc + Cd of the FP32 engine as target.  The network learns C(W) ~ oscale * net(iscale * W) with Adam on its own
GPU kernels (qbf.cnn).

Stepping policy (SurrogateStepper.step), per Strang step of the periodic problem:
  * every `exact_every`-th step is an exact step (FP32 convolution engine, FP64 update);
  * otherwise a fraction `sample_frac` of the cells (at least one) is checked first, against
    the exact collision operator C on those cells: the relative error that the surrogate adds to
    the state until the next exact step,
        r = exact_every * dt * sum|C_net - C| / sum|W|   (L1 over the sampled cells),
    and above `tol` the step falls back to the exact step.  (A relative error of the operator
    itself, sum|C_net - C| / sum|C|, is meaningless near local equilibrium, where C -> 0.)
  * a surrogate step is the midpoint collision step with the network, followed by the FP64
    conservation correction (mass, spin, momentum, energy), between two exact half transports;
  * anchoring (the default): after every exact step the operator is anchored to the exact one,
        C(W) ~ C(W0) + net(W) - net(W0),
    with W0 the state the exact step's collision started from and C(W0) taken from that step
    (Engine.last_collision: no extra evaluation); the network then supplies only the change of
    the operator since the exact step, and its systematic error cancels.

Graph mode (`graphs=True`): the exact Strang step and the surrogate Strang step are captured as
HIP graphs (Engine.capture_inhom_step_periodic, CNN.capture_inhom_step_periodic) the second time
they run between the same input and output buffers, and replayed from then on; the residual
checks and the anchoring stay eager between the replays, and the offset field is rewritten in
place.  A captured step needs a fixed paired-path choice: an engine in the automatic mode is set,
before the first capture, to the choice its last eager exact step made (Engine.set_real_pairs mode
0 or 2).  The engine's buffers cannot grow while the stepper's graphs live (close() destroys them).

Network-free fillers (`filler`), the baselines that show what the network gains: between the
exact steps the collision is either dropped ("none": transport only) or frozen at the last exact
operator ("frozen": C(W) ~ C(W0), with the same conservation correction); neither runs a residual
check.
"""
from __future__ import annotations

import math
import sys
import time

import numpy as np

from . import CONS_ALL, FP32, Engine, QBFError, quadrature
from . import workloads as wl
from .cnn import CNN

N_SURR = 64


def exact_collision_phys(engine, W):
    """C = Cc + Cd of physical states W (float64 cuda [n, 4, N, N]) in physical space."""
    WF = engine.phys_to_fourier(W)
    CF = engine.collision(WF)
    return engine.fourier_to_phys(CF)


def make_engine(N=N_SURR, max_cells=512, J=wl.J_ANG, M=wl.M_GL, precision=FP32):
    e = Engine(N, precision, max_cells=max_cells)
    e.set_quadrature(*quadrature(N, J, wl.L_BOX, wl.R_BALL, M))
    return e


DILATION = (1, 2, 4, 8, 4, 1)      # receptive field 43 x 43 pixels (an option of the kernels)


def make_net(max_batch, seed=0, N=N_SURR):
    """The surrogate network: six undilated 3x3 layers without per-cell normalisation.  (Measured
    at N=64, Nx=512 over 40 bare surrogate steps: the dilated and the normalised variants fit the
    operator as well but drift 2-30x more in energy.)"""
    return CNN(max_batch=max_batch, seed=seed, N=N)


def training_set(engine, seeds, times, cells_per_state=64, Nx=512, chunk=None, rng_seed=0, kmax=4):
    """(X, Y): float32 cuda tensors [n, 4, N, N] of states (the engine's N) and their exact collision
    operator."""
    import torch
    rng = np.random.default_rng(rng_seed)
    chunk = chunk or engine.max_cells
    X, Y = [], []
    for s in seeds:
        for t in times:
            cells = np.sort(rng.choice(Nx, cells_per_state, replace=False))
            W = wl.periodic(engine.N, Nx, seed=int(s), t=float(t), cells=cells, kmax=kmax)
            X.append(torch.from_numpy(W).cuda())
    Xd = torch.cat(X)
    for i in range(0, Xd.shape[0], chunk):
        Y.append(exact_collision_phys(engine, Xd[i:i + chunk].contiguous()))
    Yd = torch.cat(Y)
    return Xd.float().contiguous(), Yd.float().contiguous()


def trajectory_set(engine, seeds, t_end=0.04, Nx=32, per_seed=None, rng_seed=0, kmax=4):
    """(X, Y): float32 cuda tensors [n, 4, N, N] of states sampled along exact-solver trajectories
    and their exact collision operator.  For each seed the periodic workload on Nx cells (h = 1/Nx,
    CFL 0.5) is stepped with the exact engine (its precision and conservation setting) until t_end;
    every step's collision yields pairs for all its cells: the state after the first half transport
    and the operator there (Engine.last_collision, taken from the step itself: no extra evaluation).
    per_seed: pairs sampled per seed (all if None)."""
    import torch
    rng = np.random.default_rng(rng_seed)
    h, dt = 1.0 / Nx, wl.cfl_dt(engine.N, Nx)
    nsteps = max(1, int(round(t_end / dt)))
    X, Y = [], []
    for s in seeds:
        W = torch.from_numpy(wl.periodic(engine.N, Nx, seed=int(s), kmax=kmax)).cuda()
        xs, ys = [], []
        buf = torch.empty_like(W)
        for _ in range(nsteps):
            engine.inhom_step_periodic(W, wl.L_BOX, h, dt, out=buf)
            Win, C = engine.last_collision(Nx)
            xs.append(Win.float())
            ys.append(C.float())
            W, buf = buf, W
        xs, ys = torch.cat(xs), torch.cat(ys)
        if per_seed is not None and per_seed < xs.shape[0]:
            idx = torch.from_numpy(np.sort(rng.choice(xs.shape[0], per_seed, replace=False))).cuda()
            xs, ys = xs.index_select(0, idx), ys.index_select(0, idx)
        X.append(xs)
        Y.append(ys)
        if len(X) % 8 == 0:
            print(f"  training trajectories: {len(X)}/{len(seeds)} seeds", file=sys.stderr, flush=True)
    return torch.cat(X).contiguous(), torch.cat(Y).contiguous()


def scales(X, Y, norm=False):
    """iscale maps the states to O(1); oscale the collision operator (per unit a^2 with the
    per-cell normalisation, a = max |W0| of the cell)."""
    if not norm:
        return 1.0 / float(X.abs().max()), float(Y.abs().max())
    a = X[:, 0].abs().amax(dim=(1, 2))
    return 1.0 / float(a.max()), float((Y.abs().amax(dim=(1, 2, 3)) / (a * a)).max())


def rel_l1(net, X, Y, iscale, oscale, batch=256):
    """sum |net(X) - Y| / sum |Y|"""
    num = den = 0.0
    batch = min(batch, net.max_batch)
    for i in range(0, X.shape[0], batch):
        P = net.forward(X[i:i + batch].contiguous(), iscale, oscale)
        num += float((P - Y[i:i + batch]).abs().sum())
        den += float(Y[i:i + batch].abs().sum())
    return num / den


def train(net, X, Y, iscale, oscale, steps=3000, batch=32, lr=2e-3, lr_min=2e-5, seed=0, log_every=0):
    """Adam with a cosine learning-rate schedule on the mean squared error of the normalised
    target Y / oscale.  Returns the loss history."""
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    # without per-cell normalisation the target is normalised here; with it, by the kernels
    Tn = Y.contiguous() if net.norm else (Y / oscale).contiguous()
    osc = oscale if net.norm else 1.0
    n = X.shape[0]
    hist = []
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).cuda()
        xb = X.index_select(0, idx).contiguous()
        tb = Tn.index_select(0, idx).contiguous()
        loss, _ = net.gradients(xb, tb, iscale, osc)
        a = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / max(1, steps - 1)))
        net.adam(a)
        hist.append(loss)
        if log_every and it % log_every == 0:
            print(f"  train step {it}: loss {loss:.3e} lr {a:.2e}", file=sys.stderr, flush=True)
        if not math.isfinite(loss):
            raise QBFError(f"training diverged at step {it}")
    return hist


class SurrogateStepper:
    """Strang steps of the periodic problem with the surrogate collision step and its safeguards."""

    FILLERS = ("network", "frozen", "none")

    def __init__(self, engine, net, iscale, oscale, L, h, dt, exact_every=10, sample_frac=0.02, tol=1e-3,
                 mask=CONS_ALL, check=True, seed=0, anchor=True, filler="network", graphs=False):
        if filler not in self.FILLERS:
            raise QBFError(f"filler must be one of {self.FILLERS}")
        if filler != "network" and not exact_every:
            raise QBFError("a network-free filler needs exact steps (exact_every >= 1)")
        self.e, self.net = engine, net
        self.iscale, self.oscale = iscale, oscale
        self.L, self.h, self.dt = L, h, dt
        self.exact_every, self.sample_frac, self.tol = exact_every, sample_frac, tol
        self.filler = filler
        self.mask, self.check, self.anchor = mask, check and filler == "network", anchor and filler == "network"
        self.C0 = None               # the last exact operator (frozen filler)
        if net is not None:
            net.set_offset(None)
        self.rng = np.random.default_rng(seed)
        self.nstep = 0
        self.counts = {"exact": 0, "surrogate": 0, "fallback": 0}
        self.residuals = []          # r as defined above
        self.op_residuals = []       # sum|C_net - C| / sum|C|
        self.graphs = graphs and filler == "network"
        self._graphs = {}            # (kind, input address, output address, cells) -> Graph
        self._seen = set()           # the same keys, run eagerly once
        self.replays = 0
        self._exact_pairs = False    # the paired-path choice of the last eager exact step

    def close(self):
        """Destroy the captured graphs (the engine's buffers can grow again)."""
        for g in self._graphs.values():
            g.close()
        self._graphs.clear()
        self._seen.clear()

    def _run(self, kind, W, out, eager, capture):
        """One step of the given kind from W into out: eagerly, or (graph mode) captured the second time
        these buffers meet and replayed afterwards."""
        if not self.graphs:
            return eager(out)
        if out is None:
            raise QBFError("graph mode steps between fixed buffers: pass the output buffer (out=...)")
        key = (kind, W.data_ptr(), out.data_ptr(), W.shape[0])
        g = self._graphs.get(key)
        if g is not None:
            g.launch()
            self.replays += 1
        elif key in self._seen:
            if kind == "exact" and self.e.pair_mode == 1:
                self.e.set_real_pairs(2 if self._exact_pairs else 0)
            self._graphs[key], _ = capture(out)       # runs the step once, then captures it
        else:
            eager(out)
            self._seen.add(key)
            if kind == "exact":
                self._exact_pairs = self.e.real_pairs_status()[0]
        return out

    def _exact(self, W, out):
        self.e.set_moment_correction(self.mask)
        out = self._run("exact", W, out, lambda o: self.e.inhom_step_periodic(W, self.L, self.h, self.dt, out=o),
                        lambda o: self.e.capture_inhom_step_periodic(W, self.L, self.h, self.dt, out=o))
        if self.filler == "frozen":
            self.C0 = self.e.last_collision(W.shape[0], with_state=False)[1]
        if self.anchor:
            W0, C0 = self.e.last_collision(W.shape[0])
            N0 = self.net.forward(W0.float().contiguous(), self.iscale, self.oscale)
            D = (C0 - N0.double()).float().contiguous()
            if self.graphs:
                self.net.update_offset(D)     # in place: the captured surrogate steps read it
            else:
                self.net.set_offset(D)
        return out

    def step(self, W, out=None):
        ncell = W.shape[0]
        kind = "surrogate"
        if self.exact_every and self.nstep % self.exact_every == 0:
            kind = "exact"
        elif self.check and self.sample_frac > 0:
            ns = min(64, max(1, int(round(self.sample_frac * ncell))))
            cells = np.sort(self.rng.choice(ncell, ns, replace=False))
            d, c, w = self.net.residual(self.e, W, cells, self.iscale, self.oscale)
            # an empty or non-finite sample (w == 0, NaN or Inf anywhere in the sums) cannot vouch for
            # the network: r is then infinite or NaN, and only a finite r <= tol skips the exact step
            r = max(1, self.exact_every) * self.dt * d / w if w > 0 else math.inf
            self.residuals.append(r)
            self.op_residuals.append(d / c if c > 0 else math.inf)
            if not (math.isfinite(r) and r <= self.tol):
                kind = "fallback"
        if kind == "surrogate" and self.filler == "network":
            args = (self.e, W, self.L, self.h, self.dt, self.iscale, self.oscale, self.mask)
            out = self._run("surrogate", W, out, lambda o: self.net.inhom_step_periodic(*args, out=o),
                            lambda o: self.net.capture_inhom_step_periodic(*args, out=o))
        elif kind == "surrogate":
            B = self.e.transport_periodic(W, self.L, self.h, 0.5 * self.dt)
            if self.filler == "frozen":
                # midpoint step with a constant operator: W + dt C0, then the conservation correction
                Bn = B + self.dt * self.C0
                if self.mask:
                    self.e.moment_correct(B, Bn, self.L, self.mask)
                B = Bn
            out = self.e.transport_periodic(B, self.L, self.h, 0.5 * self.dt, out=out)
        else:
            out = self._exact(W, out)
        self.counts[kind] += 1
        self.nstep += 1
        return out


def timed(fn, *a, **k):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = fn(*a, **k)
    torch.cuda.synchronize()
    return r, time.perf_counter() - t0
