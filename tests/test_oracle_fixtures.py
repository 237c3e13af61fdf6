"""Pins the CPU oracle against the reference's own fixtures (test/data/*.dat, data/*.dat).

The reference's C tests only print an error norm (no assertion); here each one
becomes a thresholded test.  Residuals measured for this oracle are ~1e-15
(L1-relative) on every fixture; the thresholds leave two orders of margin.
Parameters are those of the reference test programs cited per test.
"""
import numpy as np
import pytest

import oracle as O
from conftest import l1rel, read_dat

N, L, R = 32, 12.0, 7.5
TOL = 1e-13


@pytest.fixture(scope="module")
def fgh():
    c = np.complex128
    return tuple(read_dat(n, c, (N, N)) for n in ("f.dat", "g.dat", "h.dat"))


@pytest.fixture(scope="module")
def WF():
    W = np.stack([read_dat(f"W{i}.dat", shape=(N, N)) for i in range(4)])
    return np.stack([O.r2c_fill(W[i]) for i in range(4)])


def test_lgwt():  # test/lgwt_test.c:49-72
    x, w = O.legendre_gauss(15, 0.5, 2.0)
    assert np.abs(x - read_dat("x_ref.dat")).sum() + np.abs(w - read_dat("w_ref.dat")).sum() < 1e-14
    # nodes are numpy's Gauss-Legendre nodes in descending order
    xn, wn = np.polynomial.legendre.leggauss(15)
    assert np.abs(x - ((0.5 * (1 - xn) + 2.0 * (1 + xn)) / 2)[::-1]).max() < 1e-14
    assert np.abs(w - wn[::-1] * 0.75).max() < 1e-14


def test_I1(fgh):  # test/I1_test.c:45-76, J=32
    p1, p2 = O.fourier_I1(N, 32, L, R)
    assert l1rel(O.I1integral(*fgh, p1, p2), read_dat("I1_ref.dat", np.complex128, (N, N))) < TOL


def test_I2_custom_weights(fgh):  # test/I2_test.c:47-98, J=7, hand-supplied psi grids
    wt = read_dat("weight.dat")
    p1 = np.stack([read_dat(f"psiR1_{j + 1}.dat", shape=(N, N)) for j in range(7)])
    p2 = np.stack([read_dat(f"psiR2_{j + 1}.dat", shape=(N, N)) * wt[j] for j in range(7)])
    assert l1rel(O.I2integral(*fgh, p1, p2), read_dat("I2_ref.dat", np.complex128, (N, N))) < TOL


def test_I2_fourier(fgh):  # test/I2_test2.c, J=32
    p1, p2 = O.fourier_I2(N, 32, L, R)
    assert l1rel(O.I2integral(*fgh, p1, p2), read_dat("I2_ref2.dat", np.complex128, (N, N))) < TOL


def test_I3(fgh):  # test/I3_test.c, J=32, M=7
    p1, p2 = O.fourier_I3(N, 32, L, R, 7)
    assert l1rel(O.I3integral(*fgh, p1, p2), read_dat("I3_ref.dat", np.complex128, (N, N))) < TOL


def test_I4(fgh):  # test/I4_test.c, J=32, I4 = I1integral with FourierI4 tables
    p1, p2 = O.fourier_I4(N, 32, L, R)
    assert l1rel(O.I1integral(*fgh, p1, p2), read_dat("I4_ref.dat", np.complex128, (N, N))) < TOL


def test_Cc_Cd(WF):  # test/Cc_test.c (J=48), test/Cd_test.c (J=17, M=15)
    q = O.Quad(N, 17, L, R, 15, J1=48)
    Cc = O.CcInt(q, WF)
    ref = np.stack([read_dat(f"Cc{i}_ref.dat", np.complex128, (N, N)) for i in range(4)])
    assert l1rel(Cc, ref) < TOL
    assert np.all(Cc[0] == 0)  # trace part of the commutator vanishes identically
    Cd = O.CdInt(q, WF)
    ref = np.stack([read_dat(f"Cd{i}_ref.dat", np.complex128, (N, N)) for i in range(4)])
    assert l1rel(Cd, ref) < TOL


@pytest.mark.slow
def test_collision_step(WF):  # test/coll_step_test.c, J=M=32, dt=1e-3, Bext=0
    q = O.Quad(N, 32, L, R, 32)
    out = O.collision_step(q, WF, 1e-3)
    ref = np.stack([read_dat(f"WFnext{i}_ref.dat", np.complex128, (N, N)) for i in range(4)])
    assert l1rel(out, ref) < TOL


@pytest.mark.slow
def test_inhom_step_periodic():  # test/sim_step_test.c: 8 cells, J=16, M=32, h=0.08, dt=0.002
    q = O.Quad(N, 16, L, R, 32)
    W = read_dat("Wx.dat", shape=(8, 4, N, N))
    out = O.inhom_step_periodic(q, W, 0.08, 0.002)
    assert l1rel(out, read_dat("Wx_next_ref.dat", shape=(8, 4, N, N))) < TOL


@pytest.mark.slow
def test_trajectory_homogeneous_frame1():
    """data/Wevolv_hom.dat was written by the reference C code itself (src/main_hom.c:40-50)."""
    q = O.Quad(N, 32, L, R, 32)
    ev = read_dat("Wevolv_hom.dat", shape=(101, 4, N, N))
    out = O.simulation_homogeneous(q, ev[0], 1e-3, 2)
    assert np.array_equal(out[0], ev[0])
    assert l1rel(out[1], ev[1]) < TOL


@pytest.mark.slow
def test_trajectory_periodic_frame1():
    """data/Wevolv_inhom_periodic.dat (bin/params_periodic_example.txt: J=M=32, h=0.1, dt=0.008)."""
    q = O.Quad(N, 32, L, R, 32)
    ev = read_dat("Wevolv_inhom_periodic.dat", shape=(5, 10, 4, N, N))
    out = O.inhom_step_periodic(q, ev[0], 0.1, 0.008)
    assert l1rel(out, ev[1]) < TOL


def test_slope_limiters():  # test/slope_lim_test.c and slope_lim_test2.c
    Un = read_dat("Un.dat")
    assert np.array_equal(O.slope_limiter_periodic(0.08, 0.02, -1.2, Un), read_dat("Un1_periodic_ref.dat"))
    assert np.array_equal(O.slope_limiter_dirichlet(0.08, 0.02, -1.2, Un), read_dat("Un1_Dirichlet_ref.dat"))
    out = O.slope_limiter_maxwell(0.01, 0.005, [1.0, 0.5, -0.5, -1.0], 0.4, 0.7, 0.45,
                                  [0.2, -np.pi / 4, 0, 0], [0, 0, 1 / np.e, -0.8], read_dat("UnM.dat"))
    assert np.array_equal(out, read_dat("Un1_Maxwell_ref.dat"))


def test_oracle_inhom_step_dirichlet_structure():
    """With two cells there is no interior cell: the Dirichlet step is exactly two Dirichlet half
    transports (src/simulation.c:715 skips the end cells); with three, the collision of the middle
    cell changes the result."""
    import oracle as O
    from conftest import rand_state
    N, L, h, dt = 8, 6.0, 0.08, 0.002
    q = O.Quad(N, 4, L, 3.75, 2)
    v = np.array([k if k < N // 2 else k - N for k in range(N)]) * (2 * L / N)

    def transport(W):
        out = np.empty_like(W)
        for c in range(4):
            for l in range(N):
                for k in range(N):
                    out[:, c, l, k] = O.slope_limiter_dirichlet(h, 0.5 * dt, v[l], W[:, c, l, k])
        return out

    rng = np.random.default_rng(3)
    W2 = rand_state(rng, 2, N, L)
    assert np.array_equal(O.inhom_step_dirichlet(q, W2, h, dt), transport(transport(W2)))
    W3 = rand_state(rng, 3, N, L)
    got = O.inhom_step_dirichlet(q, W3, h, dt)
    ref = transport(transport(W3))
    assert np.abs(got - ref).max() > 0


def test_oracle_trace_fluxes_closed_form():
    """For U = 1 the rightward flux is N * sum_{l<N/2} (l+1) * 2L/N and the leftward one
    N * sum_{l>=N/2} (N-l) * 2L/N; both sums run in index order, one rounding per product and
    per addition (src/finite_volume.c:186-214)."""
    from conftest import maxwell_vx
    N, L = 8, 6.0
    A = maxwell_vx(N, L)
    sc = 2 * L / N
    assert O.rightward_flux(A, np.ones(N * N)) == pytest.approx(N * sc * sum(l + 1 for l in range(N // 2)), rel=1e-15)
    assert O.leftward_flux(A, np.ones(N * N)) == pytest.approx(N * sc * sum(N - l for l in range(N // 2, N)), rel=1e-15)
    rng = np.random.default_rng(4)
    U = rng.uniform(0, 1, N * N)
    acc_r = acc_l = 0.0
    for i in range(N * N // 2):
        acc_r += float(A[i] * U[i])
    for i in range(N * N // 2, N * N):
        acc_l += float(abs(A[i]) * U[i])
    assert O.rightward_flux(A, U) == acc_r and O.leftward_flux(A, U) == acc_l


def test_oracle_inhom_step_maxwell_composition():
    """The Maxwell step against its assembly from the limiter (pinned by Un1_Maxwell_ref.dat), the
    trace fluxes and the single-cell collision (pinned by the collision fixtures): transport,
    collision of every cell, transport (src/simulation.c:837-1165)."""
    from conftest import maxwell_transport, rand_state
    N, L, h, dt, lam, nv = 8, 6.0, 0.08, 0.004, 0.35, 4
    q = O.Quad(N, 4, L, 3.75, 2)
    rng = np.random.default_rng(6)
    W = rand_state(rng, nv, N, L)
    WL, WR = rand_state(rng, 2, N, L)
    Wt = maxwell_transport(W, L, h, dt, lam, WL, WR)
    Wc = np.stack([O.simulation_homogeneous(q, Wt[i], dt, 2)[1] for i in range(nv)])
    got = O.inhom_step_maxwell(q, W, h, dt, lam, WL, WR)
    assert np.array_equal(got, maxwell_transport(Wc, L, h, dt, lam, WL, WR))
    # the zero x-velocity row is not transported: it changes by the collision only
    assert np.array_equal(got[:, :, 0, :], Wc[:, :, 0, :])


def _reflect(W):
    """x -> 1 - x (cell order reversed) and v_x -> -v_x (row l -> (N - l) mod N) of states [.., 4, N, N]."""
    N = W.shape[-1]
    return W[..., (-np.arange(N)) % N, :][::-1].copy() if W.ndim == 4 else W[..., (-np.arange(N)) % N, :].copy()


def test_oracle_maxwell_transport_commutes_with_reflection():
    """Mirror symmetry of the Maxwell half transport, which pins its velocity reshuffle by
    mathematics: reflecting the state (x -> 1 - x, v_x -> -v_x) and exchanging the two boundary
    states (each reflected in v_x) gives the reflected result.  The Maxwell grid is symmetric
    (x velocities +-dv ... +-L, src/simulation.c:155-161), the limiter is odd under A -> -A with the
    cell order reversed, and the trace fluxes exchange sides.  The -L row (l = N/2) is excluded: the
    Maxwell grid carries it twice, as -L and +L, and keeps the -L result, whose mirror image is the
    discarded +L one.  A wrong reshuffle (a dropped or shifted row, a sign of the duplicated row)
    breaks the symmetry by O(1)."""
    N, L, h, dt, lam, nv = 8, 6.0, 0.08, 0.01, 0.35, 7
    rng = np.random.default_rng(8)
    W = 0.3 * rng.random((nv, 4, N, N)) + 0.1
    WL, WR = 0.3 * rng.random((4, N, N)) + 0.1, 0.3 * rng.random((4, N, N)) + 0.1
    a = O.transport_maxwell(W, L, h, dt, lam, WL, WR)
    b = O.transport_maxwell(_reflect(W), L, h, dt, lam, _reflect(WR), _reflect(WL))
    rows = [l for l in range(N) if l != N // 2]
    err = np.abs(_reflect(a)[:, :, rows] - b[:, :, rows]).max() / np.abs(a).max()
    print("reflection defect", err)
    assert err < 1e-14
    assert np.abs(a - W).max() > 1e-3 * np.abs(W).max()     # the transport moved something
    # the step of the oracle applies this transport on both sides of the collision
    q = O.Quad(N, 4, L, 3.75, 2)
    Wc = np.stack([O.simulation_homogeneous(q, a[i], dt, 2)[1] for i in range(nv)])
    assert np.array_equal(O.inhom_step_maxwell(q, W, h, dt, lam, WL, WR),
                          O.transport_maxwell(Wc, L, h, dt, lam, WL, WR))


def test_oracle_periodic_step_on_a_subset_of_cells():
    """inhom_step_periodic_cells (collisions only within two cells of the listed ones) equals the full
    periodic step on the listed cells, bit for bit (the limiter's stencil reaches two cells)."""
    from conftest import rand_state
    N, L, nv = 8, 6.0, 11
    q = O.Quad(N, 4, L, 3.75, 2)
    rng = np.random.default_rng(9)
    W = rand_state(rng, nv, N, L)
    full = O.inhom_step_periodic(q, W, 1.0 / nv, 0.004)
    cells = [0, 5, 10]
    part = O.inhom_step_periodic_cells(q, W, 1.0 / nv, 0.004, cells)
    assert np.array_equal(full[cells], part[cells])
