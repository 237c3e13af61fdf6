"""FP64 conservation correction of the mixed-precision path (qbf_moment_correct)."""
import numpy as np
import pytest

from conftest import rand_state

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
import qbf  # noqa: E402
from qbf import workloads as wl  # noqa: E402


def basis(N, L):
    vx, vy = wl.velocity_grid(N, L)
    return np.stack([np.ones_like(vx), vx, vy, 0.5 * (vx ** 2 + vy ** 2)])


@pytest.fixture(scope="module")
def eng():
    return qbf.Engine(32, qbf.FP64, max_cells=4)


def test_selected_moments_vanish_and_change_is_minimal(eng):
    N, L = 32, 6.0
    rng = np.random.default_rng(5)
    Wold = rand_state(rng, 3, N, L)
    Wnew = Wold + 1e-3 * rng.standard_normal(Wold.shape) * Wold
    psi = basis(N, L)
    for mask in (qbf.CONS_MASS, qbf.CONS_MASS | qbf.CONS_ENERGY, qbf.CONS_ALL):
        out = eng.moment_correct(torch.from_numpy(Wold).cuda(), torch.from_numpy(Wnew.copy()).cuda(), L, mask)
        out = out.cpu().numpy()
        d = out - Wold
        sel = [0] + ([1, 2] if mask & 2 else []) + ([3] if mask & 4 else [])
        for c in range(3):
            scale = np.abs(psi[:, None] * d[c, 0][None]).sum()
            for j in sel:
                assert abs((psi[j] * d[c, 0]).sum()) < 1e-13 * scale
            for s in range(1, 4):
                assert abs(d[c, s].sum()) < 1e-13 * np.abs(d[c, s]).sum()
            # the change of the trace part lies in span{|w0| psi_j, j selected}; least squares in numpy
            ch = (out - Wnew)[c, 0].ravel()
            A = np.stack([np.abs(Wold[c, 0]).ravel() * psi[j].ravel() for j in sel], axis=1)
            coef, *_ = np.linalg.lstsq(A, ch, rcond=None)
            assert np.abs(A @ coef - ch).max() < 1e-10 * max(1e-300, np.abs(ch).max())


def test_conserving_increment_is_unchanged_and_mask0_is_identity(eng):
    N, L = 32, 6.0
    rng = np.random.default_rng(6)
    Wold = rand_state(rng, 2, N, L)
    # an increment with zero mass/momentum/energy/spin moments: project a random one in numpy
    d = 1e-3 * rng.standard_normal(Wold.shape) * Wold
    psi = basis(N, L).reshape(4, -1)
    for c in range(2):
        x = d[c, 0].ravel()
        d[c, 0] = (x - psi.T @ np.linalg.lstsq(psi.T, x, rcond=None)[0]).reshape(N, N)
        for s in range(1, 4):
            d[c, s] -= d[c, s].mean()
    Wnew = Wold + d
    out = eng.moment_correct(torch.from_numpy(Wold).cuda(), torch.from_numpy(Wnew.copy()).cuda(), L, qbf.CONS_ALL)
    assert np.abs(out.cpu().numpy() - Wnew).max() < 1e-14
    out0 = eng.moment_correct(torch.from_numpy(Wold).cuda(), torch.from_numpy(Wnew.copy()).cuda(), L, 0)
    assert np.array_equal(out0.cpu().numpy(), Wnew)
