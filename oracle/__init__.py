"""CPU oracle for the collision-plus-transport step -- TEST INFRASTRUCTURE ONLY.

ctypes bindings to ``oracle/lib/liboracle.so`` (built from the C restatement of
the reference in this directory by ``oracle/Makefile``).  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module; the product never does.

Arrays: real grids are float64 ``[..., N, N]`` (index ``[l, k]``, k fast);
Fourier grids are complex128 ``[..., N, N]``; a Wigner state is ``[4, N, N]``
(component-major), as in include/wigner.h of the reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "liboracle.so")
_lib = None

_dp = C.POINTER(C.c_double)


def build(force: bool = False) -> str:
    """Compile the oracle (gcc, no GPU); force: recompile even where it is up to date."""
    subprocess.run(["make", "-s", *(["-B"] if force else []), "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u, d, i = C.c_uint, C.c_double, C.c_int
        sig = {
            "orc_fft1d": (None, [i, _dp, i]),
            "orc_fft2d": (None, [i, i, _dp, i]),
            "orc_zero_pad": (None, [i, _dp, _dp]),
            "orc_extract": (None, [i, _dp, _dp]),
            "orc_LegendreGaussQuad": (None, [u, d, d, _dp, _dp]),
            "orc_FourierI1": (None, [i, u, d, d, _dp, _dp]),
            "orc_FourierI2": (None, [i, u, d, d, _dp, _dp]),
            "orc_FourierI3": (None, [i, u, d, d, u, _dp, _dp]),
            "orc_FourierI4": (None, [i, u, d, d, _dp, _dp]),
            "orc_I1integral": (None, [i, _dp, _dp, _dp, u, _dp, _dp, _dp]),
            "orc_I2integral": (None, [i, _dp, _dp, _dp, u, _dp, _dp, _dp]),
            "orc_I3integral": (None, [i, _dp, _dp, _dp, u, _dp, _dp, _dp]),
            "orc_CcInt": (None, [C.c_void_p, _dp, _dp]),
            "orc_CdInt": (None, [C.c_void_p, _dp, _dp]),
            "orc_MagneticRotation": (None, [i, _dp, _dp, _dp]),
            "orc_CollisionStep": (None, [C.c_void_p, _dp, d, _dp, _dp]),
            "orc_r2c_fill": (None, [i, _dp, _dp]),
            "orc_c2r_real": (None, [i, _dp, _dp]),
            "orc_InhomStepPeriodic": (None, [C.c_void_p, u, d, _dp, d, d, _dp, _dp]),
            "orc_InhomStepPeriodicCells": (None, [C.c_void_p, u, d, _dp, d, d, _dp, u, C.POINTER(u), _dp]),
            "orc_InhomStepDirichlet": (None, [C.c_void_p, u, d, _dp, d, d, _dp, _dp]),
            "orc_TransportMaxwell": (None, [i, u, d, _dp, d, d, d, _dp, _dp, _dp]),
            "orc_InhomStepMaxwell": (None, [C.c_void_p, u, d, _dp, d, d, d, _dp, _dp, _dp, _dp]),
            "orc_SimulationHomogeneous": (None, [C.c_void_p, _dp, d, u, _dp, _dp]),
            "orc_SlopeLimiterStepPeriodic": (None, [d, d, d, u, _dp, _dp]),
            "orc_SlopeLimiterStepDirichlet": (None, [d, d, d, u, _dp, _dp]),
            "orc_CalculateRightwardFlux": (d, [u, _dp, _dp]),
            "orc_CalculateLeftwardFlux": (d, [u, _dp, _dp]),
            "orc_SlopeLimiterStepMaxwell": (None, [d, d, u, _dp, u, d, d, d, _dp, _dp, _dp, _dp]),
            "orc_moments": (None, [i, d, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


# ---------------------------------------------------------------- FFT / grid
def fft2(x: np.ndarray, sign: int) -> np.ndarray:
    y = _c(x, np.complex128).copy()
    lib().orc_fft2d(y.shape[0], y.shape[1], _p(y.view(np.float64)), sign)
    return y


def fft1(x: np.ndarray, sign: int) -> np.ndarray:
    y = _c(x, np.complex128).copy()
    lib().orc_fft1d(y.shape[0], _p(y.view(np.float64)), sign)
    return y


def zero_pad(f: np.ndarray) -> np.ndarray:
    N = f.shape[0]
    f = _c(f, np.complex128)
    out = np.empty((2 * N, 2 * N), np.complex128)
    lib().orc_zero_pad(N, _p(f.view(np.float64)), _p(out.view(np.float64)))
    return out


def extract(f: np.ndarray) -> np.ndarray:
    N = f.shape[0] // 2
    f = _c(f, np.complex128)
    out = np.empty((N, N), np.complex128)
    lib().orc_extract(N, _p(f.view(np.float64)), _p(out.view(np.float64)))
    return out


# ---------------------------------------------------------------- quadrature
def legendre_gauss(M: int, a: float, b: float):
    x = np.empty(M)
    w = np.empty(M)
    lib().orc_LegendreGaussQuad(M, a, b, _p(x), _p(w))
    return x, w


def fourier_I1(N, J, L, R):
    P = 2 * N
    p1 = np.empty((J * J, P, P))
    p2 = np.empty((J * J, N, N))
    lib().orc_FourierI1(N, J, L, R, _p(p1), _p(p2))
    return p1, p2


def fourier_I2(N, J, L, R):
    p1 = np.empty((J, N, N))
    p2 = np.empty((J, N, N))
    lib().orc_FourierI2(N, J, L, R, _p(p1), _p(p2))
    return p1, p2


def fourier_I3(N, J, L, R, M):
    P = 2 * N
    p1 = np.empty((J * M, N, N), np.complex128)
    p2 = np.empty((J * M, P, P))
    lib().orc_FourierI3(N, J, L, R, M, _p(p1.view(np.float64)), _p(p2))
    return p1, p2


def fourier_I4(N, J, L, R):
    P = 2 * N
    p1 = np.empty((J, P, P))
    p2 = np.empty((J, N, N))
    lib().orc_FourierI4(N, J, L, R, _p(p1), _p(p2))
    return p1, p2


class _QuadStruct(C.Structure):
    _fields_ = [("N", C.c_int),
                ("numI1", C.c_uint), ("I1_psiR1", _dp), ("I1_psiR2", _dp),
                ("numI2", C.c_uint), ("I2_psiR1", _dp), ("I2_psiR2", _dp),
                ("numI3", C.c_uint), ("I3_psiR1", _dp), ("I3_psiR2", _dp),
                ("numI4", C.c_uint), ("I4_psiR1", _dp), ("I4_psiR2", _dp)]


class Quad:
    """The four quadrature tables of the collision operator (FourierI1..I4)."""

    def __init__(self, N, J, L, R, M, J1=None):
        self.N, self.J, self.L, self.R, self.M = N, J, L, R, M
        J1 = J if J1 is None else J1
        self.I1 = fourier_I1(N, J1, L, R)
        self.I2 = fourier_I2(N, J, L, R)
        self.I3 = fourier_I3(N, J, L, R, M)
        self.I4 = fourier_I4(N, J, L, R)
        self._s = _QuadStruct(
            N,
            self.I1[0].shape[0], _p(self.I1[0]), _p(self.I1[1]),
            self.I2[0].shape[0], _p(self.I2[0]), _p(self.I2[1]),
            self.I3[0].shape[0], _p(self.I3[0].view(np.float64)), _p(self.I3[1]),
            self.I4[0].shape[0], _p(self.I4[0]), _p(self.I4[1]))

    @property
    def ptr(self):
        return C.byref(self._s)


# ---------------------------------------------------------------- integrals
def _integral(name, f, g, h, psi1, psi2):
    N = f.shape[0]
    f, g, h = (_c(a, np.complex128) for a in (f, g, h))
    psi1 = np.ascontiguousarray(psi1)
    psi2 = np.ascontiguousarray(psi2)
    out = np.empty((N, N), np.complex128)
    p1 = psi1.view(np.float64) if psi1.dtype == np.complex128 else psi1
    getattr(lib(), name)(N, _p(f.view(np.float64)), _p(g.view(np.float64)), _p(h.view(np.float64)),
                         psi1.shape[0], _p(p1), _p(psi2), _p(out.view(np.float64)))
    return out


def I1integral(f, g, h, psi1, psi2):
    return _integral("orc_I1integral", f, g, h, psi1, psi2)


def I2integral(f, g, h, psi1, psi2):
    return _integral("orc_I2integral", f, g, h, psi1, psi2)


def I3integral(f, g, h, psi1, psi2):
    return _integral("orc_I3integral", f, g, h, psi1, psi2)


# ---------------------------------------------------------------- operators
def CcInt(q: Quad, W: np.ndarray) -> np.ndarray:
    W = _c(W, np.complex128)
    out = np.empty_like(W)
    lib().orc_CcInt(q.ptr, _p(W.view(np.float64)), _p(out.view(np.float64)))
    return out


def CdInt(q: Quad, W: np.ndarray) -> np.ndarray:
    W = _c(W, np.complex128)
    out = np.empty_like(W)
    lib().orc_CdInt(q.ptr, _p(W.view(np.float64)), _p(out.view(np.float64)))
    return out


def magnetic_rotation(W, Bext):
    W = _c(W, np.complex128)
    out = np.empty_like(W)
    B = _c(Bext, np.float64)
    lib().orc_MagneticRotation(W.shape[-1], _p(W.view(np.float64)), _p(B), _p(out.view(np.float64)))
    return out


def collision_step(q: Quad, W, dt, Bext=(0.0, 0.0, 0.0)):
    W = _c(W, np.complex128)
    out = np.empty_like(W)
    B = _c(Bext, np.float64)
    lib().orc_CollisionStep(q.ptr, _p(W.view(np.float64)), dt, _p(B), _p(out.view(np.float64)))
    return out


def r2c_fill(w: np.ndarray) -> np.ndarray:
    w = _c(w, np.float64)
    out = np.empty(w.shape, np.complex128)
    lib().orc_r2c_fill(w.shape[0], _p(w), _p(out.view(np.float64)))
    return out


def c2r_real(x: np.ndarray) -> np.ndarray:
    x = _c(x, np.complex128)
    out = np.empty(x.shape, np.float64)
    lib().orc_c2r_real(x.shape[0], _p(x.view(np.float64)), _p(out))
    return out


def inhom_step_periodic(q: Quad, W, h, dt, Bext=None):
    W = _c(W, np.float64)
    nv = W.shape[0]
    B = np.zeros(3 * nv) if Bext is None else _c(Bext, np.float64)
    out = np.empty_like(W)
    lib().orc_InhomStepPeriodic(q.ptr, nv, q.L, _p(W), h, dt, _p(B), _p(out))
    return out


def inhom_step_periodic_cells(q: Quad, W, h, dt, cells, Bext=None):
    """inhom_step_periodic, exact on the listed cells only (collisions evaluated within two cells of
    them); the other cells of the result are undefined."""
    W = _c(W, np.float64)
    nv = W.shape[0]
    B = np.zeros(3 * nv) if Bext is None else _c(Bext, np.float64)
    idx = np.ascontiguousarray(cells, np.uint32)
    out = np.zeros_like(W)
    lib().orc_InhomStepPeriodicCells(q.ptr, nv, q.L, _p(W), h, dt, _p(B), idx.size,
                                     idx.ctypes.data_as(C.POINTER(C.c_uint)), _p(out))
    return out


def inhom_step_dirichlet(q: Quad, W, h, dt, Bext=None):
    W = _c(W, np.float64)
    nv = W.shape[0]
    B = np.zeros(3 * nv) if Bext is None else _c(Bext, np.float64)
    out = np.empty_like(W)
    lib().orc_InhomStepDirichlet(q.ptr, nv, q.L, _p(W), h, dt, _p(B), _p(out))
    return out


def inhom_step_maxwell(q: Quad, W, h, dt, lam, WL, WR, Bext=None):
    W = _c(W, np.float64)
    nv = W.shape[0]
    B = np.zeros(3 * nv) if Bext is None else _c(Bext, np.float64)
    out = np.empty_like(W)
    lib().orc_InhomStepMaxwell(q.ptr, nv, q.L, _p(W), h, dt, lam, _p(_c(WL, np.float64)), _p(_c(WR, np.float64)),
                               _p(B), _p(out))
    return out


def transport_maxwell(W, L, h, dt, lam, WL, WR):
    """One half transport (dt / 2) with Maxwell boundaries, as inhom_step_maxwell applies it."""
    W = _c(W, np.float64)
    out = np.empty_like(W)
    lib().orc_TransportMaxwell(W.shape[-1], W.shape[0], L, _p(W), h, dt, lam, _p(_c(WL, np.float64)),
                               _p(_c(WR, np.float64)), _p(out))
    return out


def rightward_flux(A, U):
    A = _c(A, np.float64)
    return lib().orc_CalculateRightwardFlux(A.size, _p(A), _p(_c(U, np.float64)))


def leftward_flux(A, U):
    A = _c(A, np.float64)
    return lib().orc_CalculateLeftwardFlux(A.size, _p(A), _p(_c(U, np.float64)))


def simulation_homogeneous(q: Quad, W0, dt, numsteps, Bext=(0.0, 0.0, 0.0)):
    W0 = _c(W0, np.float64)
    out = np.empty((numsteps,) + W0.shape)
    B = _c(Bext, np.float64)
    lib().orc_SimulationHomogeneous(q.ptr, _p(W0), dt, numsteps, _p(B), _p(out))
    return out


# ---------------------------------------------------------------- transport
def slope_limiter_periodic(h, dt, A, Un):
    Un = _c(Un, np.float64)
    out = np.empty_like(Un)
    lib().orc_SlopeLimiterStepPeriodic(h, dt, A, Un.size, _p(Un), _p(out))
    return out


def slope_limiter_dirichlet(h, dt, A, Un):
    Un = _c(Un, np.float64)
    out = np.empty_like(Un)
    lib().orc_SlopeLimiterStepDirichlet(h, dt, A, Un.size, _p(Un), _p(out))
    return out


def slope_limiter_maxwell(h, dt, A, lam, fluxL, fluxR, UmL, UmR, Un):
    A = _c(A, np.float64)
    m = A.size
    Un = _c(Un, np.float64)
    out = np.empty_like(Un)
    lib().orc_SlopeLimiterStepMaxwell(h, dt, m, _p(A), Un.size // m, lam, fluxL, fluxR,
                                      _p(_c(UmL, np.float64)), _p(_c(UmR, np.float64)), _p(Un), _p(out))
    return out


def moments(W, L):
    W = _c(W, np.float64)
    mom = np.empty(7)
    lib().orc_moments(W.shape[-1], L, _p(W), _p(mom))
    return mom
