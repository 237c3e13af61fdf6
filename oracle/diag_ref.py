"""Per-cell diagnostics of physical Wigner states -- TEST INFRASTRUCTURE ONLY (the oracle of
qbf_cell_diagnostics).

Restates visualization/wignerMoments.m:15-37 (moments without the trace normalisation),
visualization/entropy.m (binary entropy, zero outside (0, 1)) and visualization/visualizeHomBoltzmann.m:26-45
and 98-105 (eigenvalues of the 2x2 matrix w0 I + w.sigma, entropy summed over both and over v, times dv).
Velocity grid v = [0..N/2-1, -N/2..-1] * 2L/N with v_x on the slow index (src/simulation.c:151-160).

Fermi-Dirac equilibrium diagnostics (oracle of qbf_fermi_dirac, qbf_fd_distance, qbf_cell_average):
  wigner_moments ....... visualization/wignerMoments.m:15-37 (rho = pauliToMatrix(w), vel, en)
  fd_params ............ visualization/calculateFDparams.m: the three moment equations in (z_up, z_dn, beta)
                         solved by nonlinear least squares with the analytic Jacobian of mdiff, from
                         (1, 1, 1) and again from beta = 0.5 when the residual exceeds 1e-4;
                         Li2 from scipy.special.spence (Li2(x) = spence(1 - x))
  create_fd ............ visualization/createFD.m (U diag(FD) U^H per velocity, matrixToPauli.m)
  asymptotic_fd ........ visualization/asymptoticFD.m (eig of rho in ascending order)
  relentropy ........... visualization/relentropy.m (eigendecompositions, clamping to [eps, 1 - eps])
  fd_distance .......... visualization/visualizeHomBoltzmann.m:120-147 (norm(., 1) of MATLAB on the
                         [k, l] grid, relative entropy over WFD_0 > 1e-6, both times dv)
  cell_average ......... visualization/visualizeInhomBoltzmann.m:80-89
"""
import numpy as np

SIGMA = np.array([[[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]]], dtype=np.complex128)


def binary_entropy(p):
    p = np.asarray(p, dtype=np.float64)
    out = np.zeros_like(p)
    m = (p > 0) & (p < 1)
    q = p[m]
    out[m] = -(q * np.log(q) + (1 - q) * np.log1p(-q))
    return out


def cell_diagnostics(W, L):
    """W: [ncell, 4, N, N] -> [ncell, 10]: mass, spin x/y/z, momentum x/y, energy, entropy,
    min eigenvalue, max eigenvalue."""
    W = np.asarray(W, dtype=np.float64)
    N = W.shape[-1]
    v = np.array([k if k < N // 2 else k - N for k in range(N)], dtype=np.float64) * (2 * L / N)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    dv = (2 * L / N) ** 2
    w0 = W[:, 0]
    r = np.sqrt(W[:, 1] ** 2 + W[:, 2] ** 2 + W[:, 3] ** 2)
    lp, lm = w0 + r, w0 - r
    out = np.stack([W[:, 0].sum((1, 2)), W[:, 1].sum((1, 2)), W[:, 2].sum((1, 2)), W[:, 3].sum((1, 2)),
                    (vx * w0).sum((1, 2)), (vy * w0).sum((1, 2)), (0.5 * (vx ** 2 + vy ** 2) * w0).sum((1, 2)),
                    (binary_entropy(lp) + binary_entropy(lm)).sum((1, 2))], axis=1) * dv
    return np.concatenate([out, lm.min((1, 2))[:, None], lp.max((1, 2))[:, None]], axis=1)


def _vgrid(N, L):
    return np.array([k if k < N // 2 else k - N for k in range(N)], dtype=np.float64) * (2 * L / N)


def pauli_to_matrix(w):
    """[..., 4] -> [..., 2, 2] (pauliToMatrix.m)"""
    w = np.asarray(w)
    return w[..., 0, None, None] * np.eye(2) + np.einsum("...i,ijk->...jk", w[..., 1:].astype(np.complex128), SIGMA)


def matrix_to_pauli(M):
    """[..., 2, 2] -> [..., 4] (matrixToPauli.m)"""
    tr = lambda A: np.trace(A, axis1=-2, axis2=-1)  # noqa: E731
    return 0.5 * np.real(np.stack([tr(M)] + [tr(M @ SIGMA[i]) for i in range(3)], axis=-1))


def wigner_moments(W, L):
    """W [4, N, N] -> (rho 2x2, vel (2,), en): wignerMoments.m"""
    N = W.shape[-1]
    v = _vgrid(N, L)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    dv = (2 * L / N) ** 2
    rho = pauli_to_matrix(np.array([W[j].sum() * dv for j in range(4)]))
    t = np.real(np.trace(rho))
    vel = 2 * np.array([(vx * W[0]).sum(), (vy * W[0]).sum()]) * dv / t
    en = 2 * (0.5 * ((vx - vel[0]) ** 2 + (vy - vel[1]) ** 2) * W[0]).sum() * dv / t
    return rho, vel, en


def li2(x):
    from scipy.special import spence
    return spence(1.0 - np.asarray(x, dtype=np.float64))


def density_fd(z, beta):
    return 2 * np.pi / beta * np.log1p(z)


def energy_fd(z, beta):
    return -2 * np.pi / beta ** 2 * li2(-z)


def fd_params(lam, en):
    """calculateFDparams.m: (beta, mu[2], residual) for eigenvalues lam (ascending) and the
    normalised energy en."""
    from scipy.optimize import least_squares
    lam = np.asarray(lam, dtype=np.float64)
    E = en * lam.sum()

    def mdiff(x):
        zu, zd, b = x
        return np.array([density_fd(zu, b) - lam[0], density_fd(zd, b) - lam[1],
                         energy_fd(zu, b) + energy_fd(zd, b) - E])

    def jac(x):
        zu, zd, b = x
        dru = [2 * np.pi / (b * (1 + zu)), -density_fd(zu, b) / b]
        drd = [2 * np.pi / (b * (1 + zd)), -density_fd(zd, b) / b]
        deu = [density_fd(zu, b) / (b * zu), -2 * energy_fd(zu, b) / b]
        ded = [density_fd(zd, b) / (b * zd), -2 * energy_fd(zd, b) / b]
        return np.array([[dru[0], 0, dru[1]], [0, drd[0], drd[1]], [deu[0], ded[0], deu[1] + ded[1]]])

    best = None
    for b0 in (1.0, 0.5):
        r = least_squares(mdiff, [1.0, 1.0, b0], jac=jac, bounds=([1e-300, 1e-300, 1e-300], np.inf),
                          xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=10000)
        res = np.sqrt(2 * r.cost)
        if best is None or res < best[1]:
            best = (r.x, res)
        if res <= 1e-4:
            break
    x, res = best
    return x[2], np.log(x[:2]) / x[2], res


def fd_occupation(beta, mu, e):
    with np.errstate(over="ignore"):
        return 1.0 / (np.exp(beta * (e - mu)) + 1.0)


def create_fd(N, L, beta, mu, vel, U):
    """createFD.m: [4, N, N] Pauli components of U diag(FD(mu1), FD(mu2)) U^H at v - vel."""
    v = _vgrid(N, L)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    e = 0.5 * ((vx - vel[0]) ** 2 + (vy - vel[1]) ** 2)
    D = np.zeros((N, N, 2, 2), dtype=np.complex128)
    D[..., 0, 0] = fd_occupation(beta, mu[0], e)
    D[..., 1, 1] = fd_occupation(beta, mu[1], e)
    M = U @ D @ U.conj().T
    return np.moveaxis(matrix_to_pauli(M), -1, 0)


def asymptotic_fd(W, L):
    """asymptoticFD.m: (WFD [4, N, N], beta, mu, vel, lambda, residual) for one state W [4, N, N]."""
    rho, vel, en = wigner_moments(W, L)
    lam, U = np.linalg.eigh(rho)
    beta, mu, res = fd_params(lam, en)
    return create_fd(W.shape[-1], L, beta, mu, vel, U), beta, mu, vel, lam, res


def relentropy(rho, tau):
    """relentropy.m for two 2x2 Hermitian matrices."""
    eps = np.finfo(np.float64).eps
    mu, U = np.linalg.eigh(rho)
    nu, V = np.linalg.eigh(tau)
    mu = np.where(mu <= 0, eps, np.where(mu >= 1, 1 - eps, mu))
    nu = np.where(nu <= 0, eps, np.where(nu >= 1, 1 - eps, nu))
    Wm = U.conj().T @ V
    r = np.diag(mu)
    log_t = Wm @ np.diag(np.log(nu)) @ Wm.conj().T
    log_1t = Wm @ np.diag(np.log(1 - nu)) @ Wm.conj().T
    return float(np.real(np.trace(r @ (np.diag(np.log(mu)) - log_t))
                         + np.trace((np.eye(2) - r) @ (np.diag(np.log(1 - mu)) - log_1t))))


def fd_distance(W, WFD, L):
    """(L1 distance, relative entropy) of W [4, N, N] to WFD [4, N, N] (visualizeHomBoltzmann.m:120-147);
    MATLAB's [k1, k2] grid is [v_y, v_x], i.e. the transpose of W[j]."""
    N = W.shape[-1]
    dv = (2 * L / N) ** 2
    dist = sum(np.linalg.norm((W[j] - WFD[j]).T, 1) for j in range(4)) * dv
    kl = 0.0
    for l in range(N):
        for k in range(N):
            if WFD[0, l, k] > 1e-6:
                kl += relentropy(pauli_to_matrix(W[:, l, k]), pauli_to_matrix(WFD[:, l, k])) * dv
    return dist, kl


def cell_average(W):
    acc = np.zeros(W.shape[1:])
    for c in range(W.shape[0]):
        acc = acc + W[c]
    return acc / W.shape[0]
