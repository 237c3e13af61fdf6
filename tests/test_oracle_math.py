"""Pins the oracle where no fixture reaches: brute-force direct convolutions at tiny N
(no FFT involved), closed forms and invariants."""
import numpy as np
import pytest

import oracle as O


def cconv(a, b):
    """Cyclic 2-D convolution by direct summation: out[n] = sum_m a[m] b[n-m mod P]."""
    P = a.shape[0]
    out = np.zeros(a.shape, np.complex128)
    for m0 in range(P):
        for m1 in range(P):
            if a[m0, m1] != 0:
                out += a[m0, m1] * np.roll(np.roll(b, m0, axis=0), m1, axis=1)
    return out


def rand_c(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(11)
    N = 4
    return N, rng, rand_c(rng, N, N), rand_c(rng, N, N), rand_c(rng, N, N)


def test_I1_bruteforce(small):
    # I1 = sum_q Extract( pad(h) (*) ( psiR1_q . ( pad(f) (*) pad(psiR2_q g) ) ) )   (src/integrals.c:90-133)
    N, rng, f, g, h = small
    p1, p2 = O.fourier_I1(N, 3, 5.0, 2.0)
    ref = sum(O.extract(cconv(O.zero_pad(h), p1[q] * cconv(O.zero_pad(f), O.zero_pad(p2[q] * g))))
              for q in range(p1.shape[0]))
    assert np.abs(O.I1integral(f, g, h, p1, p2) - ref).max() < 1e-12 * np.abs(ref).max()


def test_I2_bruteforce(small):
    # I2 = sum_q Extract( pad(psiR1_q f) (*) pad(psiR2_q g) (*) pad(h) )   (src/integrals.c:211-251)
    N, rng, f, g, h = small
    p1, p2 = rng.standard_normal((5, N, N)), rng.standard_normal((5, N, N))
    ref = sum(O.extract(cconv(cconv(O.zero_pad(p1[q] * f), O.zero_pad(p2[q] * g)), O.zero_pad(h)))
              for q in range(5))
    assert np.abs(O.I2integral(f, g, h, p1, p2) - ref).max() < 1e-12 * np.abs(ref).max()


def test_I3_bruteforce(small):
    # I3 = sum_q Extract( pad(psiR1_q f) (*) ( psiR2_q . ( pad(g) (*) pad(psiR1_q h) ) ) )   (src/integrals.c:332-379)
    N, rng, f, g, h = small
    p1, p2 = O.fourier_I3(N, 3, 5.0, 2.0, 2)
    ref = sum(O.extract(cconv(O.zero_pad(p1[q] * f), p2[q] * cconv(O.zero_pad(g), O.zero_pad(p1[q] * h))))
              for q in range(p1.shape[0]))
    assert np.abs(O.I3integral(f, g, h, p1, p2) - ref).max() < 1e-12 * np.abs(ref).max()


def test_quadrature_closed_forms():
    # I3 weights are a phase (unit modulus) times a sinc that equals 2R w_m pi/J at frequency 0;
    # the GL weights on [-R, R] sum to 2R and the nodes are symmetric.
    N, J, M, L, R = 8, 4, 5, 6.0, 3.75
    p1, p2 = O.fourier_I3(N, J, L, R, M)
    assert np.allclose(np.abs(p1), 1.0, atol=1e-15)
    x, w = O.legendre_gauss(M, -R, R)
    assert abs(w.sum() - 2 * R) < 1e-13 and np.allclose(x, -x[::-1], atol=1e-14)
    for j in range(J):
        for m in range(M):
            assert abs(p2[m + j * M][0, 0] - 2 * R * w[m] * np.pi / J) < 1e-14
    # I1: psiR1 is duplicated over j, psiR2 over i, weights -(2pi/J)^2 / cos(theta1_i - theta2_j)
    q1, q2 = O.fourier_I1(N, J, L, R)
    for i in range(J):
        for j in range(J):
            assert np.array_equal(q1[i + J * j], q1[i])
            th1, th2 = i * np.pi / J, j * np.pi / J + 0.5 * np.pi / J
            wij = -(2 * np.pi / J) ** 2 / np.cos(th1 - th2)
            base = q2[J * j] / (-(2 * np.pi / J) ** 2 / np.cos(0 - th2))
            assert np.allclose(q2[i + J * j], base * wij, rtol=1e-13, atol=1e-15)


def test_magnetic_rotation_cross_product():
    rng = np.random.default_rng(2)
    N = 8
    W = rand_c(rng, 4, N, N)
    B = rng.standard_normal(3)
    CB = O.magnetic_rotation(W, B)
    w = W[1:]
    ref = 2 * np.stack([B[1] * w[2] - B[2] * w[1], B[2] * w[0] - B[0] * w[2], B[0] * w[1] - B[1] * w[0]])
    assert np.all(CB[0] == 0) and np.abs(CB[1:] - ref).max() < 1e-14


def test_collision_step_dt0_is_identity():
    rng = np.random.default_rng(4)
    N = 8
    q = O.Quad(N, 3, 6.0, 3.75, 2)
    W = rand_c(rng, 4, N, N) * 0.1
    assert np.array_equal(O.collision_step(q, W, 0.0), W)


def test_periodic_limiter_invariants():
    rng = np.random.default_rng(9)
    U = rng.standard_normal(64)
    for A in (-1.3, -0.2, 0.0, 0.7, 2.0):
        out = O.slope_limiter_periodic(0.1, 0.02, A, U)
        assert abs(out.sum() - U.sum()) < 1e-12
        c = np.full(64, 0.37)
        assert np.allclose(O.slope_limiter_periodic(0.1, 0.02, A, c), c, rtol=0, atol=1e-15)
    # pure shift when nu = 1: upwind scheme moves the profile by one cell
    out = O.slope_limiter_periodic(0.1, 0.1, 1.0, U)
    assert np.allclose(out, np.roll(U, 1), atol=1e-13)


def test_moments_closed_form():
    # a Gaussian centred at u with variance T: mass, momentum and energy in closed form
    N, L = 64, 8.0
    v = np.array([k if k < N // 2 else k - N for k in range(N)]) * 2 * L / N
    vx, vy = np.meshgrid(v, v, indexing="ij")  # vx depends on the slow index l
    u = (0.3, -0.5)
    T = 0.8
    w0 = np.exp(-((vx - u[0]) ** 2 + (vy - u[1]) ** 2) / (2 * T)) / (2 * np.pi * T)
    W = np.stack([w0, 0.1 * w0, 0 * w0, -0.2 * w0])
    m = O.moments(W, L)
    assert abs(m[0] - 1) < 1e-10 and abs(m[1] - 0.1) < 1e-10 and abs(m[3] + 0.2) < 1e-10
    assert abs(m[4] - u[0]) < 1e-10 and abs(m[5] - u[1]) < 1e-10
    assert abs(m[6] - (T + 0.5 * (u[0] ** 2 + u[1] ** 2))) < 1e-9
