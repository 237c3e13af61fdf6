"""Paired I3 path (qbf_set_real_pairs).  On spectra of real states with empty Nyquist lines every
grid K_q acts on is real and the pairing is exact up to rounding (FP64: against the complex path and
the oracle).  Otherwise it drops imaginary parts bounded by the measured ratio (FP32 workload states),
and the automatic mode declines states whose ratio exceeds the tolerance."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _l1rel(a, b):
    return float((a - b).abs().sum() / b.abs().sum())


def _nyquist_free(WF, N):
    WF = WF.clone()
    WF[..., N // 2, :] = 0
    WF[..., :, N // 2] = 0
    return WF


def test_pairs_exact_on_nyquist_free_states_fp64():
    import torch
    import qbf
    from qbf import workloads as wl
    N, nx = 32, 4
    q = O.Quad(N, 8, wl.L_BOX, wl.R_BALL, 4)
    e = qbf.Engine(N, qbf.FP64, max_cells=nx)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    WF = _nyquist_free(e.phys_to_fourier(torch.from_numpy(wl.periodic(N, nx, seed=1)).cuda()), N)
    e.set_real_pairs(0)
    c0 = e.collision(WF)
    e.set_real_pairs(1, 1e-13)
    c1 = e.collision(WF)
    used, ratio = e.real_pairs_status()
    print(f"ratio {ratio:.3e}, used {used}, paired vs complex {_l1rel(c1, c0):.3e}")
    assert used and ratio < 1e-13
    assert _l1rel(c1, c0) < 1e-13
    got = c1.cpu().numpy()
    for c in range(2):
        want = O.CcInt(q, WF[c].cpu().numpy()) + O.CdInt(q, WF[c].cpu().numpy())
        err = np.abs(got[c] - want).sum() / np.abs(want).sum()
        print(f"cell {c}: paired vs oracle {err:.3e}")
        assert err < 1e-12


def test_pairs_fp32_error_bounded_by_measured_ratio():
    import torch
    import qbf
    from qbf import workloads as wl
    N, nx = 64, 16
    tabs = qbf.quadrature(N, 8, wl.L_BOX, wl.R_BALL, 4)
    e32 = qbf.Engine(N, qbf.FP32, max_cells=nx)
    e64 = qbf.Engine(N, qbf.FP64, max_cells=nx)
    e32.set_quadrature(*tabs)
    e64.set_quadrature(*tabs)
    WF = e64.phys_to_fourier(torch.from_numpy(wl.periodic(N, nx, seed=3, t=0.02)).cuda())
    ref = e64.collision(WF)
    e32.set_real_pairs(0)
    cc = e32.collision(WF)
    e32.set_real_pairs(1, 1.0)
    cp = e32.collision(WF)
    used, ratio = e32.real_pairs_status()
    ec, ep, d = _l1rel(cc, ref), _l1rel(cp, ref), _l1rel(cp, cc)
    print(f"ratio {ratio:.3e}: complex vs FP64 {ec:.3e}, paired vs FP64 {ep:.3e}, paired vs complex {d:.3e}")
    assert used
    assert d < 1e-5 + 10 * ratio
    assert ep < 2 * ec + 10 * ratio


def test_pairs_auto_mode_declines_rough_states():
    import torch
    import qbf
    from qbf import workloads as wl
    N, nx = 32, 4
    e = qbf.Engine(N, qbf.FP32, max_cells=nx)
    e.set_quadrature(*qbf.quadrature(N, 8, wl.L_BOX, wl.R_BALL, 4))
    W = wl.periodic(N, nx, seed=2)
    W = W + 1e-3 * np.random.default_rng(0).standard_normal(W.shape)      # grid-scale roughness
    WF = e.phys_to_fourier(torch.from_numpy(W).cuda())
    e.set_real_pairs(1, 1e-6)
    c1 = e.collision(WF)
    used, ratio = e.real_pairs_status()
    print(f"rough state: ratio {ratio:.3e}, used {used}")
    assert not used and ratio > 1e-6
    e.set_real_pairs(0)
    assert torch.equal(c1, e.collision(WF))          # the declined call ran the complex path
    e.set_real_pairs(1, 1e-6)
    e.collision(_nyquist_free(WF, N))
    used, ratio = e.real_pairs_status()
    print(f"Nyquist-free state: ratio {ratio:.3e}, used {used}")
    assert used
