"""BF16 hi/lo split path (engine precision BF16X3, N = 64): the line DFTs of the I3 K-operator run
as DFT-16 GEMMs on the matrix cores with split operands x = hi + lo (bf16 each), products
hi*hi + hi*lo + lo*hi accumulated in FP32 (k_i3acc_mfma); everything else as the FP32 path.

Parity against the oracle (src/integrals.c:332-379 for I3, src/collision.c:74-254 for Cc + Cd) on
small quadratures, against the oracle-pinned FP64 engine at the configurations' J = 32, M = 16,
and on both the paired and the complex I3 path.

Tolerances.  A split operand keeps 16 significant bits (hi: 8, lo: the next 8), so an operand is
represented to 2^-17 relative and a product misses its lo*lo part, 2^-16 relative: about 1.5e-5
per product against FP32's 6e-8.  Each K-operator application runs four DFT-16 stages on the
matrix cores (an FFT pair along each axis).  For a single I3 integral the errors of the terms are
independent and average; in Cc + Cd the I2, I3 and I4 parts cancel to a result about ten times
smaller than the parts (the FP32 path's 2.1e-5 against its 1.7e-6 per K application shows the
same factor), so the operator error is about 4 x 1.5e-5 x 10 = 6e-4 at most; the tests allow 2e-3
for the operator and 3e-4 for a single integral.
"""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel

pytestmark = pytest.mark.gpu


def _fourier(W):
    return np.stack([O.r2c_fill(W[c]) for c in range(4)])


def test_bf16x3_i3_integral_vs_oracle():
    """A single I3 integral (the complex K-operator kernel, nout = 1) on smooth grids."""
    import torch
    import qbf
    from qbf import workloads as wl
    N, J, M = 64, 4, 3
    L, R = wl.L_BOX, wl.R_BALL
    q = O.Quad(N, J, L, R, M)
    Wp = wl.periodic(N, 3, seed=5)
    f, g, h = (O.r2c_fill(Wp[i, 0]) for i in range(3))
    want = O.I3integral(f, g, h, q.I3[0], q.I3[1])
    res = {}
    for name, prec in (("bf16x3", qbf.BF16X3), ("fp32", qbf.FP32)):
        e = qbf.Engine(N, prec, max_cells=1)
        e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
        cu = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()   # noqa: E731
        res[name] = e.I3(cu(f), cu(g), cu(h)).cpu().numpy()
        e.close()
    eb, ef = l1rel(res["bf16x3"], want), l1rel(res["fp32"], want)
    print(f"I3 (N={N}, J={J}, M={M}) vs oracle: BF16 split {eb:.3e}, FP32 {ef:.3e}", flush=True)
    assert ef < 2e-6
    assert eb < 3e-4
    # not the FP32 kernel under another name: the split path's error is of its own size
    assert eb > 4 * ef


@pytest.mark.parametrize("pairs", [0, 2])
def test_bf16x3_operator_vs_oracle(pairs):
    """Cc + Cd on two cells of the N = 64 workload, complex (0) and paired (2) I3 path."""
    import torch
    import qbf
    from qbf import workloads as wl
    N, J, M = 64, 4, 3
    q = O.Quad(N, J, wl.L_BOX, wl.R_BALL, M)
    Wp = wl.periodic(N, 512, seed=3, cells=[0, 200])
    WF = np.stack([_fourier(Wp[i]) for i in range(2)])
    want = np.stack([O.CcInt(q, WF[i]) + O.CdInt(q, WF[i]) for i in range(2)])
    e = qbf.Engine(N, qbf.BF16X3, max_cells=2)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    e.set_real_pairs(pairs, 1.0)
    got = e.collision(torch.from_numpy(WF).cuda()).cpu().numpy()
    used, ratio = e.real_pairs_status()
    e.close()
    err = l1rel(got, want)
    print(f"Cc + Cd (N={N}, J={J}, M={M}, paired {used}) vs oracle: {err:.3e}", flush=True)
    assert used == (pairs == 2)
    assert err < 2e-3


def test_bf16x3_full_quadrature_vs_fp64():
    """J = 32, M = 16 (the mixed-precision configuration's quadrature) on 16 cells of the benchmark
    workload: the split path against the FP64 engine (itself pinned to the oracle at 1e-12 on these
    cells' quadrature) and against the FP32 path; one collision step with the FP64 conservation
    correction keeps the moments."""
    import torch
    import qbf
    from qbf import workloads as wl
    N, J, M, nc = 64, 32, 16, 16
    cells = list(range(0, 512, 512 // nc))
    tabs = qbf.quadrature(N, J, wl.L_BOX, wl.R_BALL, M)
    Wp = torch.from_numpy(wl.periodic(N, 512, seed=3, cells=cells)).cuda()
    out = {}
    for name, prec in (("fp64", qbf.FP64), ("fp32", qbf.FP32), ("bf16x3", qbf.BF16X3)):
        e = qbf.Engine(N, prec, max_cells=nc)
        e.set_quadrature(*tabs)
        WF = e.phys_to_fourier(Wp)
        out[name] = e.collision(WF)
        if name == "bf16x3":
            used, ratio = e.real_pairs_status()
            e.set_real_pairs(0)
            out["bf16x3_complex"] = e.collision(WF)
            e.set_real_pairs(1, 2e-5)
            dt = wl.cfl_dt(N, 512)
            e.set_moment_correction(qbf.CONS_ALL)
            W1 = e.inhom_step_periodic(Wp, wl.L_BOX, 1.0 / 512, dt)
            m0, m1 = wl.moments(Wp.cpu().numpy()), wl.moments(W1.cpu().numpy())
        e.close()
    ref = out["fp64"]
    eb = float((out["bf16x3"] - ref).abs().sum() / ref.abs().sum())
    ec = float((out["bf16x3_complex"] - ref).abs().sum() / ref.abs().sum())
    ef = float((out["fp32"] - ref).abs().sum() / ref.abs().sum())
    # the step conserves what the correction fixes: the total over all cells of mass and energy is
    # moved only by the transport (periodic, exact sums), so it stays to rounding
    dm = float(np.abs(m1[:, [0, 6]].sum(0) - m0[:, [0, 6]].sum(0)).max() / np.abs(m0[:, [0, 6]].sum(0)).max())
    print(f"J={J} M={M}: BF16 split paired ({used}, ratio {ratio:.2e}) {eb:.3e}, complex {ec:.3e}, "
          f"FP32 {ef:.3e} against FP64; step moments drift {dm:.2e}", flush=True)
    assert ef < 4e-5
    assert eb < 2e-3 and ec < 2e-3
    assert dm < 1e-12
