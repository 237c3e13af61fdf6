"""Python host bindings of the MI355X collision-plus-transport engine (libqbf.so).

A thin ctypes layer over the C API of include/qbf.h.  Device memory is owned by
torch tensors on ``cuda:0``; the engine launches on torch's current stream.
There is no CPU fallback: if ``lib/libqbf.so`` is missing or no HIP device is
present, every entry point raises.

Tensor layouts (see include/qbf.h):
  Fourier state  complex128 [ncell, 4, N, N]     (wignerF_t per cell)
  physical state float64    [ncell, 4, N, N]     (wignerV_t per cell)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# QBF_LIB_PATH selects another build of the library (for example one with other kernel shapes)
LIB_PATH = os.environ.get("QBF_LIB_PATH") or os.path.join(PKG_DIR, "lib", "libqbf.so")

FP64, FP32, BF16X3 = 0, 1, 2
NOALLOC_STATE_BYTES = 32 << 20      # QBF_NOALLOC_STATE_BYTES of include/qbf.h
CC, CD, CFULL = 1, 2, 3
TAB_I1, TAB_I4 = 0, 1
CONS_MASS, CONS_MOMENTUM, CONS_ENERGY = 1, 2, 4
CONS_ALL = 7

_lib = None
_dp = C.POINTER(C.c_double)


class QBFError(RuntimeError):
    pass


def build(verbose: bool = False, force: bool = False) -> str:
    """Compile csrc/ with hipcc for gfx950 into lib/libqbf.so (works without a GPU); force: recompile
    every source even where the library is up to date."""
    out = subprocess.run(["make", "-s", *(["-B"] if force else []), "-C", PKG_DIR], capture_output=not verbose,
                         text=True)
    if out.returncode != 0:
        raise QBFError("build failed:\n" + (out.stdout or "") + (out.stderr or ""))
    return LIB_PATH


class _Tables(C.Structure):
    _fields_ = [("numI1", C.c_uint), ("I1_psiR1", _dp), ("I1_psiR2", _dp),
                ("numI2", C.c_uint), ("I2_psiR1", _dp), ("I2_psiR2", _dp),
                ("numI3", C.c_uint), ("I3_psiR1", _dp), ("I3_psiR2", _dp),
                ("numI4", C.c_uint), ("I4_psiR1", _dp), ("I4_psiR2", _dp)]


def lib():
    """Load libqbf.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QBFError(f"{LIB_PATH} not found: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        vp, i, d, u = C.c_void_p, C.c_int, C.c_double, C.c_uint
        sig = {
            "qbf_last_error": (C.c_char_p, []),
            "qbf_engine_create": (i, [i, i, i, C.POINTER(vp)]),
            "qbf_engine_destroy": (None, [vp]),
            "qbf_set_stream": (i, [vp, vp]),
            "qbf_sync": (i, [vp]),
            "qbf_set_quadrature": (i, [vp, C.POINTER(_Tables)]),
            "qbf_set_i1_table": (i, [vp, i, u, _dp, _dp]),
            "qbf_set_i2_table": (i, [vp, u, _dp, _dp]),
            "qbf_set_i3_table": (i, [vp, u, _dp, _dp]),
            "qbf_collision": (i, [vp, vp, i, i, vp, vp, vp, vp, d]),
            "qbf_collision_step": (i, [vp, vp, i, d, vp, vp]),
            "qbf_I1": (i, [vp, i, vp, vp, vp, vp]),
            "qbf_I2": (i, [vp, vp, vp, vp, vp]),
            "qbf_I3": (i, [vp, vp, vp, vp, vp]),
            "qbf_phys_to_fourier": (i, [vp, vp, i, vp]),
            "qbf_fourier_to_phys": (i, [vp, vp, i, vp]),
            "qbf_transport_periodic": (i, [vp, vp, i, d, d, d, vp]),
            "qbf_inhom_step_periodic": (i, [vp, vp, i, d, d, d, vp, vp]),
            "qbf_graph_capture_inhom_step_periodic": (i, [vp, vp, i, d, d, d, vp, vp, C.POINTER(vp)]),
            "qbf_graph_launch": (i, [vp]),
            "qbf_last_collision": (i, [vp, i, vp, vp]),
            "qbf_graph_destroy": (None, [vp]),
            "qbf_inhom_step_dirichlet": (i, [vp, vp, i, d, d, d, vp, vp]),
            "qbf_transport_dirichlet": (i, [vp, vp, i, d, d, d, vp]),
            "qbf_transport_maxwell": (i, [vp, vp, i, d, d, d, d, vp, vp, vp]),
            "qbf_inhom_step_maxwell": (i, [vp, vp, i, d, d, d, d, vp, vp, vp, vp]),
            "qbf_set_moment_correction": (i, [vp, i]),
            "qbf_moment_correct": (i, [vp, vp, vp, i, d, i]),
            "qbf_cell_diagnostics": (i, [vp, vp, i, d, vp]),
            "qbf_fermi_dirac": (i, [vp, vp, i, d, vp, vp]),
            "qbf_set_real_pairs": (i, [vp, i, d]),
            "qbf_real_pairs_status": (i, [vp, C.POINTER(i), C.POINTER(d)]),
            "qbf_fd_distance": (i, [vp, vp, i, d, vp, i, vp]),
            "qbf_cell_average": (i, [vp, vp, i, vp]),
            "qbf_simulate_periodic": (i, [vp, vp, i, d, d, d, u, u, vp, vp]),
            "qbf_simulate_periodic_noalloc": (i, [vp, vp, i, d, d, d, u, u, vp, vp]),
            "qbf_simulate_dirichlet": (i, [vp, vp, i, d, d, d, u, u, vp, vp]),
            "qbf_simulate_maxwell": (i, [vp, vp, i, d, d, d, d, vp, vp, u, u, vp, vp]),
            "qbf_profile": (i, [vp, i]),
            "qbf_profile_read": (i, [vp, C.POINTER(d), C.POINTER(i), C.POINTER(d)]),
            "qbf_legendre_gauss": (None, [u, d, d, _dp, _dp]),
            "qbf_fourier_I1": (None, [i, u, d, d, _dp, _dp]),
            "qbf_fourier_I2": (None, [i, u, d, d, _dp, _dp]),
            "qbf_fourier_I3": (None, [i, u, d, d, u, _dp, _dp]),
            "qbf_fourier_I4": (None, [i, u, d, d, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class Graph:
    """A captured step (Engine.capture_inhom_step_periodic); keeps its engine, buffers and stream
    alive.  A replay runs after the work queued on the engine's current stream and before the work
    queued there afterwards."""

    def __init__(self, h, stream, keep, engine):
        self._h, self.stream, self._keep, self._engine = h, stream, keep, engine

    def launch(self):
        if not getattr(self, "_h", None):
            raise QBFError("graph destroyed")
        self._engine._bind()
        _check(lib().qbf_graph_launch(self._h))

    def close(self):
        if getattr(self, "_h", None):
            lib().qbf_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _check(rc: int):
    if rc != 0:
        raise QBFError(lib().qbf_last_error().decode())


def _ptr(t, dtype=None, shape=None, what="tensor"):
    """Device address of a contiguous cuda tensor (None passes through), after checking its dtype
    (a torch dtype) and shape (a tuple; None entries are free) when given."""
    if t is None:
        return None
    if not (t.is_cuda and t.is_contiguous()):
        raise QBFError(f"{what}: expected a contiguous cuda tensor")
    if dtype is not None and t.dtype != dtype:
        raise QBFError(f"{what}: expected {dtype}, got {t.dtype}")
    if shape is not None and (t.dim() != len(shape) or any(a is not None and a != b for a, b in zip(shape, t.shape))):
        raise QBFError(f"{what}: expected shape {shape}, got {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


def _np(a, dtype=np.float64):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data_as(_dp)


def quadrature(N: int, J: int, L: float, R: float, M: int, J1: int | None = None):
    """The four quadrature tables (FourierI1..I4 of the reference, src/quadrature.c) as numpy
    arrays in the reference's layout, computed by the library's host code."""
    J1 = J if J1 is None else J1
    P = 2 * N
    L_ = lib()
    i1 = (np.empty((J1 * J1, P, P)), np.empty((J1 * J1, N, N)))
    L_.qbf_fourier_I1(N, J1, L, R, i1[0].ctypes.data_as(_dp), i1[1].ctypes.data_as(_dp))
    i2 = (np.empty((J, N, N)), np.empty((J, N, N)))
    L_.qbf_fourier_I2(N, J, L, R, i2[0].ctypes.data_as(_dp), i2[1].ctypes.data_as(_dp))
    i3 = (np.empty((J * M, N, N), np.complex128), np.empty((J * M, P, P)))
    L_.qbf_fourier_I3(N, J, L, R, M, i3[0].view(np.float64).ctypes.data_as(_dp), i3[1].ctypes.data_as(_dp))
    i4 = (np.empty((J, P, P)), np.empty((J, N, N)))
    L_.qbf_fourier_I4(N, J, L, R, i4[0].ctypes.data_as(_dp), i4[1].ctypes.data_as(_dp))
    return i1, i2, i3, i4


def legendre_gauss(M: int, a: float, b: float):
    x = np.empty(M)
    w = np.empty(M)
    lib().qbf_legendre_gauss(M, a, b, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    return x, w


class Engine:
    """One engine per (N, precision); holds quadrature tables and work buffers on the GPU."""

    def __init__(self, N: int, precision: int = FP64, max_cells: int = 1):
        import torch
        if not torch.cuda.is_available():
            raise QBFError("no HIP device available")
        self.N, self.precision, self.max_cells = N, precision, max_cells
        h = C.c_void_p()
        _check(lib().qbf_engine_create(N, precision, max_cells, C.byref(h)))
        self._h = h
        self._torch = torch
        self.pair_mode = 0 if precision == FP64 else 1     # the library's default (qbf_set_real_pairs)

    def close(self):
        if getattr(self, "_h", None):
            lib().qbf_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------------- profiling of the I3 kernel
    def profile(self, enable: bool = True):
        _check(lib().qbf_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        ms, n, k = C.c_double(), C.c_int(), C.c_double()
        _check(lib().qbf_profile_read(self._h, C.byref(ms), C.byref(n), C.byref(k)))
        return ms.value, n.value, k.value

    # --------------------------------------------------------------- conservation correction
    def set_moment_correction(self, mask: int):
        _check(lib().qbf_set_moment_correction(self._h, mask))

    def moment_correct(self, Wold, Wnew, L, mask=CONS_ALL):
        self._bind()
        n = Wold.shape[0]
        _check(lib().qbf_moment_correct(self._h, self._R(Wold, "Wold"), self._R(Wnew, "Wnew", n), n, L, mask))
        return Wnew

    def simulate_periodic(self, W0, L, h, dt, numsteps, stride=1, Bext=None, noalloc=False):
        """Frames 0, stride, 2 stride, ... (< numsteps) of the periodic problem from W0, as a
        numpy array [nframes, ncell, 4, N, N] (the state stays on the device between steps).
        noalloc: the allocation-free variant (qbf_simulate_periodic_noalloc: states of at most
        NOALLOC_STATE_BYTES, engine buffers already sized by a step on as many cells)."""
        self._bind()
        nf = (numsteps - 1) // stride + 1 if numsteps > 0 else 0
        out = np.empty((max(nf, 1),) + tuple(W0.shape))
        fn = lib().qbf_simulate_periodic_noalloc if noalloc else lib().qbf_simulate_periodic
        _check(fn(self._h, self._R(W0), W0.shape[0], L, h, dt, numsteps, stride, self._B(Bext, W0.shape[0]),
                  out.ctypes.data_as(C.c_void_p)))
        return out[:nf]

    def simulate_dirichlet(self, W0, L, h, dt, numsteps, stride=1, Bext=None):
        """As simulate_periodic with Dirichlet ends (qbf_inhom_step_dirichlet per step)."""
        self._bind()
        nf = (numsteps - 1) // stride + 1 if numsteps > 0 else 0
        out = np.empty((max(nf, 1),) + tuple(W0.shape))
        _check(lib().qbf_simulate_dirichlet(self._h, self._R(W0), W0.shape[0], L, h, dt, numsteps, stride,
                                            self._B(Bext, W0.shape[0]), out.ctypes.data_as(C.c_void_p)))
        return out[:nf]

    def simulate_maxwell(self, W0, L, h, dt, lam, WL, WR, numsteps, stride=1, Bext=None):
        """As simulate_periodic with Maxwell boundaries (qbf_inhom_step_maxwell per step)."""
        self._bind()
        nf = (numsteps - 1) // stride + 1 if numsteps > 0 else 0
        out = np.empty((max(nf, 1),) + tuple(W0.shape))
        _check(lib().qbf_simulate_maxwell(self._h, self._R(W0), W0.shape[0], L, h, dt, lam, self._R1(WL, "WL"),
                                          self._R1(WR, "WR"), numsteps, stride, self._B(Bext, W0.shape[0]),
                                          out.ctypes.data_as(C.c_void_p)))
        return out[:nf]

    def cell_diagnostics(self, W, L):
        """[ncell, 10] float64 cuda tensor: mass, spin x/y/z, momentum x/y, energy, entropy,
        min and max eigenvalue per cell of physical states W."""
        self._bind()
        out = self._torch.empty((W.shape[0], 10), dtype=self._torch.float64, device=W.device)
        _check(lib().qbf_cell_diagnostics(self._h, self._R(W), W.shape[0], L, _ptr(out)))
        return out

    def set_real_pairs(self, mode: int, tol: float = 0.0):
        """Paired I3 path: 0 off, 1 when the measured imaginary/real ratio <= tol, 2 always."""
        _check(lib().qbf_set_real_pairs(self._h, mode, tol))
        self.pair_mode = mode

    def real_pairs_status(self):
        """(used, ratio) of the last collision evaluation."""
        u, r = C.c_int(0), C.c_double(0)
        _check(lib().qbf_real_pairs_status(self._h, C.byref(u), C.byref(r)))
        return bool(u.value), r.value

    def fermi_dirac(self, W, L, with_states=True):
        """Fermi-Dirac equilibria with the conserved moments of each state of W [n, 4, N, N]:
        (params [n, 12]: beta, mu1, mu2, vel_x, vel_y, n_x, n_y, n_z, lambda1, lambda2, en, status;
        states [n, 4, N, N] or None)."""
        self._bind()
        t = self._torch
        prm = t.empty((W.shape[0], 12), dtype=t.float64, device=W.device)
        wfd = t.empty_like(W) if with_states else None
        _check(lib().qbf_fermi_dirac(self._h, self._R(W), W.shape[0], L, _ptr(prm), _ptr(wfd)))
        return prm, wfd

    def fd_distance(self, W, L, WFD):
        """[n, 2]: L1 distance (matrix 1-norm per component) and quantum relative entropy of each
        state of W to WFD: one state [4, N, N] (or [1, 4, N, N]) for all, or one per state."""
        self._bind()
        t = self._torch
        stride = 1 if (WFD.dim() == 4 and WFD.shape[0] == W.shape[0] and W.shape[0] > 1) else 0
        out = t.empty((W.shape[0], 2), dtype=t.float64, device=W.device)
        WFD = WFD.contiguous()
        if WFD.dtype != t.float64 or WFD.numel() != (W.shape[0] if stride else 1) * 4 * self.N * self.N:
            raise QBFError("WFD: expected float64 [4, N, N] or one state per cell")
        _check(lib().qbf_fd_distance(self._h, self._R(W), W.shape[0], L, _ptr(WFD), stride, _ptr(out)))
        return out

    def cell_average(self, W):
        """[1, 4, N, N]: the average of the states of W over the cells."""
        self._bind()
        out = self._torch.empty((1,) + tuple(W.shape[1:]), dtype=W.dtype, device=W.device)
        _check(lib().qbf_cell_average(self._h, self._R(W), W.shape[0], _ptr(out)))
        return out

    # --------------------------------------------------------------- stream
    def _bind(self):
        if not getattr(self, "_h", None):
            raise QBFError("engine closed")
        s = self._torch.cuda.current_stream().cuda_stream
        _check(lib().qbf_set_stream(self._h, C.c_void_p(s)))

    # checked device addresses: states [ncell, 4, N, N] (Fourier: complex128, physical: float64),
    # single grids [N, N] complex128, per-cell magnetic fields [ncell, 3] float64
    def _F(self, t, what="Fourier state", n=None):
        return _ptr(t, self._torch.complex128, (n, 4, self.N, self.N), what)

    def _R(self, t, what="physical state", n=None):
        return _ptr(t, self._torch.float64, (n, 4, self.N, self.N), what)

    def _G(self, t, what="grid"):
        if t is not None and t.numel() != self.N * self.N:
            raise QBFError(f"{what}: expected {self.N} x {self.N} elements, got {t.numel()}")
        return _ptr(t, self._torch.complex128, None, what)

    def _R1(self, t, what):
        if t is None or t.numel() != 4 * self.N * self.N:
            raise QBFError(f"{what}: expected one physical state [4, {self.N}, {self.N}]")
        return _ptr(t, self._torch.float64, None, what)

    def _B(self, t, n):
        if t is not None and t.numel() != 3 * n:
            raise QBFError(f"Bext: expected {3 * n} values ([ncell, 3]), got {t.numel()}")
        return _ptr(t, self._torch.float64, None, "Bext")

    # --------------------------------------------------------------- tables (numpy, reference layout)
    def set_quadrature(self, I1, I2, I3, I4):
        keep = []
        vals = []
        for (p1, p2), cplx in ((I1, False), (I2, False), (I3, True), (I4, False)):
            a1, q1 = _np(np.asarray(p1).view(np.float64) if cplx else p1)
            a2, q2 = _np(p2)
            keep += [a1, a2]
            vals += [np.asarray(p2).shape[0], q1, q2]
        t = _Tables(*vals)
        _check(lib().qbf_set_quadrature(self._h, C.byref(t)))

    def set_i1_table(self, slot, p1, p2):
        a1, q1 = _np(p1)
        a2, q2 = _np(p2)
        _check(lib().qbf_set_i1_table(self._h, slot, a2.shape[0], q1, q2))

    def set_i2_table(self, p1, p2):
        a1, q1 = _np(p1)
        a2, q2 = _np(p2)
        _check(lib().qbf_set_i2_table(self._h, a1.shape[0], q1, q2))

    def set_i3_table(self, p1, p2):
        a1, q1 = _np(np.asarray(p1, np.complex128).view(np.float64))
        a2, q2 = _np(p2)
        _check(lib().qbf_set_i3_table(self._h, a2.shape[0], q1, q2))

    # --------------------------------------------------------------- operators (torch cuda tensors)
    def collision(self, W, mode=CFULL, base=None, Wcb=None, Bext=None, cf=1.0, out=None):
        torch = self._torch
        self._bind()
        out = torch.empty_like(W) if out is None else out
        n = W.shape[0]
        _check(lib().qbf_collision(self._h, self._F(W), n, mode, self._F(out, "out", n), self._F(base, "base", n),
                                   self._F(Wcb, "Wcb", n), self._B(Bext, n), cf))
        return out

    def collision_step(self, W, dt, Bext=None, out=None):
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        n = W.shape[0]
        _check(lib().qbf_collision_step(self._h, self._F(W), n, dt, self._B(Bext, n), self._F(out, "out", n)))
        return out

    def I1(self, f, g, h, slot=TAB_I1):
        self._bind()
        out = self._torch.empty_like(f)
        _check(lib().qbf_I1(self._h, slot, self._G(f, "f"), self._G(g, "g"), self._G(h, "h"), _ptr(out)))
        return out

    def I2(self, f, g, h):
        self._bind()
        out = self._torch.empty_like(f)
        _check(lib().qbf_I2(self._h, self._G(f, "f"), self._G(g, "g"), self._G(h, "h"), _ptr(out)))
        return out

    def I3(self, f, g, h):
        self._bind()
        out = self._torch.empty_like(f)
        _check(lib().qbf_I3(self._h, self._G(f, "f"), self._G(g, "g"), self._G(h, "h"), _ptr(out)))
        return out

    def phys_to_fourier(self, Wp):
        self._bind()
        out = self._torch.empty(Wp.shape, dtype=self._torch.complex128, device=Wp.device)
        ng = Wp.numel() // (self.N * self.N)
        if ng * self.N * self.N != Wp.numel() or ng < 1:
            raise QBFError(f"phys_to_fourier: {Wp.numel()} elements are not whole {self.N} x {self.N} grids")
        _check(lib().qbf_phys_to_fourier(self._h, _ptr(Wp, self._torch.float64, None, "physical grids"), ng, _ptr(out)))
        return out

    def fourier_to_phys(self, WF):
        self._bind()
        out = self._torch.empty(WF.shape, dtype=self._torch.float64, device=WF.device)
        ng = WF.numel() // (self.N * self.N)
        if ng * self.N * self.N != WF.numel() or ng < 1:
            raise QBFError(f"fourier_to_phys: {WF.numel()} elements are not whole {self.N} x {self.N} grids")
        _check(lib().qbf_fourier_to_phys(self._h, _ptr(WF, self._torch.complex128, None, "Fourier grids"), ng, _ptr(out)))
        return out

    def transport_periodic(self, W, L, h, dt, out=None):
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        _check(lib().qbf_transport_periodic(self._h, self._R(W), W.shape[0], L, h, dt, self._R(out, "out", W.shape[0])))
        return out

    def transport_dirichlet(self, W, L, h, dt, out=None):
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        _check(lib().qbf_transport_dirichlet(self._h, self._R(W), W.shape[0], L, h, dt, self._R(out, "out", W.shape[0])))
        return out

    def inhom_step_dirichlet(self, W, L, h, dt, Bext=None, out=None):
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        n = W.shape[0]
        _check(lib().qbf_inhom_step_dirichlet(self._h, self._R(W), n, L, h, dt, self._B(Bext, n), self._R(out, "out", n)))
        return out

    def transport_maxwell(self, W, L, h, dt, lam, WL, WR, out=None):
        """Slope-limited transport with Maxwell reflection; WL, WR: boundary states [4, N, N] on the device."""
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        n = W.shape[0]
        _check(lib().qbf_transport_maxwell(self._h, self._R(W), n, L, h, dt, lam, self._R1(WL, "WL"), self._R1(WR, "WR"),
                                           self._R(out, "out", n)))
        return out

    def inhom_step_maxwell(self, W, L, h, dt, lam, WL, WR, Bext=None, out=None):
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        n = W.shape[0]
        _check(lib().qbf_inhom_step_maxwell(self._h, self._R(W), n, L, h, dt, lam, self._R1(WL, "WL"), self._R1(WR, "WR"),
                                            self._B(Bext, n), self._R(out, "out", n)))
        return out

    def inhom_step_periodic(self, W, L, h, dt, Bext=None, out=None):
        self._bind()
        out = self._torch.empty_like(W) if out is None else out
        n = W.shape[0]
        _check(lib().qbf_inhom_step_periodic(self._h, self._R(W), n, L, h, dt, self._B(Bext, n), self._R(out, "out", n)))
        return out

    def last_collision(self, ncell, with_state=True):
        """After inhom_step_periodic on ncell cells: (the state its collision step started from, or
        None; the collision operator there), float64 cuda [ncell, 4, N, N], from the step's FP64
        midpoint (no further evaluation)."""
        self._bind()
        shape = (ncell, 4, self.N, self.N)
        Win = self._torch.empty(shape, dtype=self._torch.float64, device="cuda") if with_state else None
        Cp = self._torch.empty(shape, dtype=self._torch.float64, device="cuda")
        _check(lib().qbf_last_collision(self._h, ncell, _ptr(Win), _ptr(Cp)))
        return Win, Cp

    def capture_inhom_step_periodic(self, W, L, h, dt, Bext=None, out=None):
        """HIP graph of one periodic Strang step from the buffer W into the buffer out (returned with
        the graph).  The step runs once eagerly first.  Capture and replays use a dedicated torch
        stream (the default stream cannot be captured); Graph.launch() is asynchronous on it."""
        torch = self._torch
        out = torch.empty_like(W) if out is None else out
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        g = C.c_void_p()
        with torch.cuda.stream(stream):
            self._bind()
            n = W.shape[0]
            _check(lib().qbf_graph_capture_inhom_step_periodic(self._h, self._R(W), n, L, h, dt, self._B(Bext, n),
                                                               self._R(out, "out", n), C.byref(g)))
        return Graph(g, stream, (W, out, Bext), self), out
