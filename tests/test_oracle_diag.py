"""Pins the diagnostics oracle (oracle/diag_ref.py): eigenvalues against numpy's eigvalsh of the
2x2 matrices built as visualization/visualizeHomBoltzmann.m:26-33 does, entropy closed forms,
moments of a shifted Gaussian."""
import numpy as np
import pytest

from oracle import diag_ref as D
from conftest import rand_state


def test_eigenvalues_match_eigvalsh():
    rng = np.random.default_rng(0)
    W = rand_state(rng, 2, 8, 6.0)
    out = D.cell_diagnostics(W, 6.0)
    for c in range(2):
        M = np.zeros((8, 8, 2, 2), complex)
        M[..., 0, 0] = W[c, 0] + W[c, 3]
        M[..., 1, 1] = W[c, 0] - W[c, 3]
        M[..., 1, 0] = W[c, 1] + 1j * W[c, 2]
        M[..., 0, 1] = W[c, 1] - 1j * W[c, 2]
        ev = np.linalg.eigvalsh(M)
        assert abs(out[c, 8] - ev[..., 0].min()) < 1e-14
        assert abs(out[c, 9] - ev[..., 1].max()) < 1e-14
        S = (D.binary_entropy(ev[..., 0]) + D.binary_entropy(ev[..., 1])).sum() * (12.0 / 8) ** 2
        assert abs(out[c, 7] - S) < 1e-12 * abs(S)


def test_entropy_closed_forms():
    N, L = 16, 6.0
    W = np.zeros((2, 4, N, N))
    W[0, 0] = 0.5                       # eigenvalues 1/2, 1/2: s = ln 2 each
    W[1, 0] = 0.5
    W[1, 3] = 0.5                       # eigenvalues 1 and 0: s = 0
    out = D.cell_diagnostics(W, L)
    assert abs(out[0, 7] - 2 * np.log(2) * (2 * L) ** 2) < 1e-12
    assert out[1, 7] == 0.0
    assert out[1, 8] == 0.0 and out[1, 9] == 1.0


def test_moments_of_shifted_gaussian():
    N, L = 64, 6.0
    v = np.array([k if k < N // 2 else k - N for k in range(N)]) * (2 * L / N)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    u, s2 = (0.7, -0.4), 0.49          # variance small enough that the box truncates below 1e-12
    g = np.exp(-((vx - u[0]) ** 2 + (vy - u[1]) ** 2) / (2 * s2)) / (2 * np.pi * s2)
    W = np.zeros((1, 4, N, N))
    W[0, 0] = g
    out = D.cell_diagnostics(W, L)[0]
    assert abs(out[0] - 1) < 1e-10
    assert abs(out[4] - u[0]) < 1e-10 and abs(out[5] - u[1]) < 1e-10
    assert abs(out[6] - (s2 + 0.5 * (u[0] ** 2 + u[1] ** 2))) < 1e-10    # <|v|^2/2> = T + |u|^2/2


# ------------------------------------------------------------------ Fermi-Dirac equilibrium diagnostics
def test_li2_special_values_and_identities():
    """Li2(-1) = -pi^2/12, the power series at small |x|, the inversion formula
    Li2(x) + Li2(1/x) = -pi^2/6 - ln^2(-x)/2 for x < 0, and Li2'(x) = -ln(1 - x)/x."""
    assert abs(D.li2(-1.0) + np.pi ** 2 / 12) < 1e-15
    x = -0.1
    assert abs(D.li2(x) - sum(x ** k / k ** 2 for k in range(1, 40))) < 1e-16
    for x in (-0.3, -2.0, -17.5, -1e4):
        assert abs(D.li2(x) + D.li2(1 / x) + np.pi ** 2 / 6 + 0.5 * np.log(-x) ** 2) < 1e-12 * (1 + np.log(-x) ** 2)
    for x in (-0.5, -3.0, -40.0):
        h = 1e-5 * abs(x)
        d = (D.li2(x + h) - D.li2(x - h)) / (2 * h)
        assert abs(d + np.log1p(-x) / x) < 1e-8


@pytest.mark.parametrize("beta,z1,z2", [(1.3, 0.4, 2.5), (0.2, 0.01, 0.05), (6.0, 30.0, 300.0), (1.0, 1.0, 1.0)])
def test_fd_params_round_trip(beta, z1, z2):
    """Densities and energy of a Fermi-Dirac pair (calculateFDparams.m:65-83) fitted back: the fit
    returns the beta and mu = ln(z) / beta it was made from."""
    z = np.array([z1, z2])
    lam = D.density_fd(z, beta)
    E = D.energy_fd(z, beta).sum()
    b, mu, res = D.fd_params(lam, E / lam.sum())
    assert abs(b - beta) < 1e-10 * beta
    assert np.abs(mu - np.log(z) / beta).max() < 1e-9 * (1 + np.abs(np.log(z) / beta).max())
    assert res < 1e-10


def test_fd_low_temperature_limit():
    """At large beta the energy is the degenerate-gas value sum lambda^2 / (4 pi) plus the Sommerfeld
    term (pi^2 / 6) g T^2 per species, with the 2D density of states g = 2 pi (exact up to
    exp(-beta mu) corrections)."""
    lam = np.array([0.7, 1.9])
    for beta in (300.0, 400.0, 1000.0):
        z = np.expm1(lam * beta / (2 * np.pi))
        E = D.energy_fd(z, beta).sum()
        assert abs(E - (lam ** 2).sum() / (4 * np.pi) - 2 * (np.pi ** 2 / 6) * 2 * np.pi / beta ** 2) < 1e-12 * E


def test_create_fd_pauli_structure():
    """U = identity: W_FD = diag(f1, f2) -> Pauli (f1 + f2, 0, 0, f1 - f2) / 2; a rotated basis turns
    the polarisation into (f2 - f1) / 2 times the eigenvector direction of the larger eigenvalue."""
    N, L, beta, mu, vel = 16, 6.0, 1.1, np.array([0.3, 1.2]), np.array([0.2, -0.1])
    W = D.create_fd(N, L, beta, mu, vel, np.eye(2))
    v = D._vgrid(N, L)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    e = 0.5 * ((vx - vel[0]) ** 2 + (vy - vel[1]) ** 2)
    f1, f2 = 1 / (np.exp(beta * (e - mu[0])) + 1), 1 / (np.exp(beta * (e - mu[1])) + 1)
    assert np.abs(W - np.stack([(f1 + f2) / 2, 0 * e, 0 * e, (f1 - f2) / 2])).max() < 1e-15
    n = np.array([0.48, -0.6, 0.64])
    rho = D.pauli_to_matrix(np.array([1.0, *n]))
    lam, U = np.linalg.eigh(rho)
    W = D.create_fd(N, L, beta, mu, vel, U)
    for i in range(3):
        assert np.abs(W[i + 1] - 0.5 * (f2 - f1) * n[i]).max() < 1e-14


def test_asymptotic_fd_keeps_the_moments():
    """asymptoticFD.m: the Fermi-Dirac state has the moments of W, up to the truncation of its tails
    by the velocity box [-L, L]^2 (about 3e-10 here at beta = 1.75; 3e-7 for the hotter beta = 0.9
    of seed 2)."""
    rng = np.random.default_rng(1)
    W = rand_state(rng, 1, 32, 6.0)[0]
    WFD, beta, mu, vel, lam, res = D.asymptotic_fd(W, 6.0)
    r0, v0, e0 = D.wigner_moments(W, 6.0)
    r1, v1, e1 = D.wigner_moments(WFD, 6.0)
    assert res < 1e-10 and beta > 0
    assert np.abs(r1 - r0).max() < 1e-8 * np.abs(r0).max()
    assert np.abs(v1 - v0).max() < 1e-8 and abs(e1 - e0) < 1e-8 * e0


def test_relentropy_reduces_to_classical_and_matrix_log():
    """Commuting (diagonal) matrices give the binary relative entropy of the eigenvalues; S(rho||rho)
    = 0; non-commuting ones match tr rho (log rho - log tau) + ... with scipy's logm."""
    from scipy.linalg import logm
    p, q = np.array([0.2, 0.7]), np.array([0.4, 0.55])
    cl = (p * np.log(p / q) + (1 - p) * np.log((1 - p) / (1 - q))).sum()
    assert abs(D.relentropy(np.diag(p), np.diag(q)) - cl) < 1e-15
    rho = D.pauli_to_matrix(np.array([0.45, 0.1, -0.2, 0.15]))
    tau = D.pauli_to_matrix(np.array([0.5, -0.1, 0.05, 0.3]))
    assert abs(D.relentropy(rho, rho)) < 1e-15
    I = np.eye(2)
    ref = np.real(np.trace(rho @ (logm(rho) - logm(tau))) + np.trace((I - rho) @ (logm(I - rho) - logm(I - tau))))
    assert abs(D.relentropy(rho, tau) - ref) < 1e-13
    assert D.relentropy(rho, tau) > 0


def test_fd_distance_uses_the_matrix_one_norm():
    """norm(A, 1) of MATLAB on the [v_y, v_x] grid: a difference spread over v_y at one v_x counts
    fully (one column), the same spread over v_x at one v_y counts once."""
    N, L = 8, 6.0
    dv = (2 * L / N) ** 2
    F = np.zeros((4, N, N))
    F[0] = 0.5
    W = F.copy()
    W[2, 3, :] += 0.01                        # one v_x (l = 3), all v_y
    assert abs(D.fd_distance(W, F, L)[0] - 0.01 * N * dv) < 1e-15
    W = F.copy()
    W[2, :, 5] += 0.01                        # one v_y (k = 5), all v_x
    assert abs(D.fd_distance(W, F, L)[0] - 0.01 * dv) < 1e-15


def test_cell_average():
    rng = np.random.default_rng(5)
    W = rng.standard_normal((5, 4, 4, 4))
    assert np.abs(D.cell_average(W) - W.mean(axis=0)).max() < 1e-15
