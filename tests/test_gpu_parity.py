"""Parity of the HIP path (through the C ABI of libqbf.so) with the reference's fixtures and the CPU oracle.

Tolerance: the FP64 path must match the oracle to 1e-12 L1-relative; the GPU
evaluates an exactly reassociated formulation, so only the rounding differs.  FP32 kernels are
checked at 2e-5 L1-relative.
"""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel, rand_state, read_dat

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import qbf  # noqa: E402

N32, L, R = 32, 12.0, 7.5
TOL64 = 1e-12
TOL32 = 2e-5


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.fixture(scope="module")
def eng32():
    return qbf.Engine(32, qbf.FP64, max_cells=16)


@pytest.fixture(scope="module")
def fgh():
    c = np.complex128
    return tuple(read_dat(n, c, (N32, N32)) for n in ("f.dat", "g.dat", "h.dat"))


@pytest.fixture(scope="module")
def WF32():
    W = np.stack([read_dat(f"W{i}.dat", shape=(N32, N32)) for i in range(4)])
    return np.stack([O.r2c_fill(W[i]) for i in range(4)])


def test_integrals_vs_fixtures_and_oracle(eng32, fgh):
    f, g, h = (dev(a) for a in fgh)
    p1, p2 = O.fourier_I1(N32, 32, L, R)
    eng32.set_i1_table(qbf.TAB_I1, p1, p2)
    out = host(eng32.I1(f, g, h, qbf.TAB_I1))
    assert l1rel(out, read_dat("I1_ref.dat", np.complex128, (N32, N32))) < TOL64
    assert l1rel(out, O.I1integral(*fgh, p1, p2)) < TOL64
    p1, p2 = O.fourier_I4(N32, 32, L, R)
    eng32.set_i1_table(qbf.TAB_I4, p1, p2)
    out = host(eng32.I1(f, g, h, qbf.TAB_I4))
    assert l1rel(out, read_dat("I4_ref.dat", np.complex128, (N32, N32))) < TOL64
    p1, p2 = O.fourier_I2(N32, 32, L, R)
    eng32.set_i2_table(p1, p2)
    out = host(eng32.I2(f, g, h))
    assert l1rel(out, read_dat("I2_ref2.dat", np.complex128, (N32, N32))) < TOL64
    wt = read_dat("weight.dat")
    q1 = np.stack([read_dat(f"psiR1_{j + 1}.dat", shape=(N32, N32)) for j in range(7)])
    q2 = np.stack([read_dat(f"psiR2_{j + 1}.dat", shape=(N32, N32)) * wt[j] for j in range(7)])
    eng32.set_i2_table(q1, q2)
    out = host(eng32.I2(f, g, h))
    assert l1rel(out, read_dat("I2_ref.dat", np.complex128, (N32, N32))) < TOL64
    p1, p2 = O.fourier_I3(N32, 32, L, R, 7)
    eng32.set_i3_table(p1, p2)
    out = host(eng32.I3(f, g, h))
    assert l1rel(out, read_dat("I3_ref.dat", np.complex128, (N32, N32))) < TOL64
    assert l1rel(out, O.I3integral(*fgh, p1, p2)) < TOL64


def test_Cc_Cd_vs_fixtures(eng32, WF32):
    q = O.Quad(N32, 17, L, R, 15, J1=48)
    eng32.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    W = dev(WF32[None])
    Cc = host(eng32.collision(W, qbf.CC))[0]
    ref = np.stack([read_dat(f"Cc{i}_ref.dat", np.complex128, (N32, N32)) for i in range(4)])
    assert l1rel(Cc, ref) < TOL64
    assert np.all(Cc[0] == 0)
    Cd = host(eng32.collision(W, qbf.CD))[0]
    ref = np.stack([read_dat(f"Cd{i}_ref.dat", np.complex128, (N32, N32)) for i in range(4)])
    assert l1rel(Cd, ref) < TOL64


def test_collision_step_vs_fixture(eng32, WF32):
    q = O.Quad(N32, 32, L, R, 32)
    eng32.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    out = host(eng32.collision_step(dev(WF32[None]), 1e-3))[0]
    ref = np.stack([read_dat(f"WFnext{i}_ref.dat", np.complex128, (N32, N32)) for i in range(4)])
    assert l1rel(out, ref) < TOL64


def test_inhom_step_vs_fixture(eng32):
    q = O.Quad(N32, 16, L, R, 32)
    eng32.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    W = read_dat("Wx.dat", shape=(8, 4, N32, N32))
    out = host(eng32.inhom_step_periodic(dev(W), L, 0.08, 0.002))
    assert l1rel(out, read_dat("Wx_next_ref.dat", shape=(8, 4, N32, N32))) < TOL64


def test_trajectory_homogeneous_100_steps(eng32):
    """All 100 steps of data/Wevolv_hom.dat (written by the reference C code, src/main_hom.c:40-50)."""
    q = O.Quad(N32, 32, L, R, 32)
    eng32.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    ev = read_dat("Wevolv_hom.dat", shape=(101, 4, N32, N32))
    W = dev(ev[0][None])
    errs = []
    for it in range(1, 101):
        WF = eng32.phys_to_fourier(W)
        WF = eng32.collision_step(WF, 1e-3)
        W = eng32.fourier_to_phys(WF)
        errs.append(l1rel(host(W)[0], ev[it]))
    print("max frame error", max(errs), "last", errs[-1])
    assert max(errs) < 1e-11


def test_trajectory_periodic_4_steps(eng32):
    """data/Wevolv_inhom_periodic.dat: 10 cells, bin/params_periodic_example.txt (J=M=32, h=0.1, dt=0.008)."""
    q = O.Quad(N32, 32, L, R, 32)
    eng32.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    ev = read_dat("Wevolv_inhom_periodic.dat", shape=(5, 10, 4, N32, N32))
    W = dev(ev[0])
    for it in range(1, 5):
        W = eng32.inhom_step_periodic(W, L, 0.1, 0.008)
        assert l1rel(host(W), ev[it]) < 1e-11


def test_transport_bit_exact(eng32):
    rng = np.random.default_rng(1)
    W = rng.standard_normal((13, 4, N32, N32))
    out = host(eng32.transport_periodic(dev(W), L, 0.05, 0.003))
    v = np.array([k if k < N32 // 2 else k - N32 for k in range(N32)]) * 2 * L / N32
    ref = np.empty_like(W)
    for j in range(4):
        for l in range(N32):
            for k in range(N32):
                ref[:, j, l, k] = O.slope_limiter_periodic(0.05, 0.003, v[l], W[:, j, l, k])
    assert np.array_equal(out, ref)


def test_phys_fourier_roundtrip(eng32):
    rng = np.random.default_rng(2)
    W = rng.standard_normal((3, 4, N32, N32))
    WF = host(eng32.phys_to_fourier(dev(W)))
    assert np.abs(WF - np.fft.fft2(W) / N32 ** 2).max() < 1e-15
    back = host(eng32.fourier_to_phys(dev(WF)))
    assert np.abs(back - W).max() < 1e-13


@pytest.mark.parametrize("N,prec,tol", [(16, qbf.FP64, TOL64), (16, qbf.FP32, TOL32), (32, qbf.FP32, TOL32),
                                        (64, qbf.FP32, TOL32), (64, qbf.FP64, TOL64)])
def test_random_collision_step_vs_oracle(N, prec, tol):
    rng = np.random.default_rng(100 + N + prec)
    q = O.Quad(N, 4, 6.0, 3.75, 3)
    Wp = rand_state(rng, 2, N)
    WF = np.stack([np.stack([O.r2c_fill(Wp[c, i]) for i in range(4)]) for c in range(2)])
    B = np.array([[0.3, -0.2, 0.5], [0.0, 0.1, -0.4]])
    ref = np.stack([O.collision_step(q, WF[c], 0.01, B[c]) for c in range(2)])
    e = qbf.Engine(N, prec, max_cells=2)
    e.set_quadrature(q.I1, q.I2, q.I3, q.I4)
    out = host(e.collision_step(dev(WF), 0.01, dev(B)))
    d_ref = ref - WF
    d_out = out - WF
    assert l1rel(out, ref) < tol
    assert l1rel(d_out, d_ref) < 1e3 * tol   # the increment itself
    e.close()
