"""ctypes binding of the CNN surrogate kernels (qbf_cnn_* in include/qbf.h).

The network maps a physical Wigner state of one cell, fp32 planar [4, N, N] (N = 64 or 128), to a
[4, N, N] field (trained to the collision operator in physical space).  All tensors are torch
float32 cuda tensors of shape [B, 4, N, N]; the kernels run on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import QBFError, lib

NL, CH, CIO, CREAL = 6, 64, 32, 4
GRIDS = (64, 128)           # image sides the kernels are built for

_sig_done = False


def _lib():
    global _sig_done
    L = lib()
    if not _sig_done:
        vp, i, f, sz = C.c_void_p, C.c_int, C.c_float, C.c_size_t
        fp = C.POINTER(C.c_float)
        sig = {
            "qbf_cnn_last_error": (C.c_char_p, []),
            "qbf_cnn_create": (i, [i, i, C.POINTER(vp)]),
            "qbf_cnn_destroy": (None, [vp]),
            "qbf_cnn_set_stream": (i, [vp, vp]),
            "qbf_cnn_param_count": (sz, []),
            "qbf_cnn_configure": (i, [vp, C.POINTER(i), i]),
            "qbf_cnn_set_params": (i, [vp, fp]),
            "qbf_cnn_get_params": (i, [vp, fp, i]),
            "qbf_cnn_forward": (i, [vp, vp, i, f, vp, f]),
            "qbf_cnn_gradients": (i, [vp, vp, vp, i, f, f, vp, fp]),
            "qbf_cnn_adam": (i, [vp, f, f, f, f]),
            "qbf_surrogate_collision_step": (i, [vp, vp, vp, i, C.c_double, f, f, C.c_double, i, vp]),
            "qbf_surrogate_residual": (i, [vp, vp, vp, i, C.POINTER(i), i, f, f, C.POINTER(C.c_double)]),
            "qbf_surrogate_inhom_step_periodic": (i, [vp, vp, vp, i, C.c_double, C.c_double, C.c_double, f, f, i,
                                                      vp]),
            "qbf_surrogate_set_offset": (i, [vp, vp, i]),
            "qbf_graph_capture_surrogate_inhom_step_periodic": (i, [vp, vp, vp, i, C.c_double, C.c_double, C.c_double,
                                                                    f, f, i, vp, C.POINTER(vp)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _sig_done = True
    return L


def _check(rc):
    if rc != 0:
        raise QBFError(_lib().qbf_cnn_last_error().decode())


def _ptr(t):
    if not (t.is_cuda and t.is_contiguous() and t.dtype == __import__("torch").float32):
        raise QBFError(f"expected a contiguous float32 cuda tensor, got {t.dtype} on {t.device}")
    return C.c_void_p(t.data_ptr())


def param_count() -> int:
    return int(_lib().qbf_cnn_param_count())


def init_params(seed: int = 0, norm: bool = False) -> np.ndarray:
    """He-normal weights on the used channels (4 input channels, 5 with per-cell normalisation),
    zero biases and zero padded entries; the last layer is scaled by 0.1 so that an untrained
    network predicts small values."""
    rng = np.random.default_rng(seed)
    out = []
    for l in range(NL):
        ci = CIO if l == 0 else CH
        co = CIO if l == NL - 1 else CH
        ci_r = (CREAL + 1 if norm else CREAL) if l == 0 else CH
        co_r = CREAL if l == NL - 1 else CH
        W = np.zeros((9, co, ci), np.float32)
        W[:, :co_r, :ci_r] = rng.standard_normal((9, co_r, ci_r)) * np.sqrt(2.0 / (9 * ci_r))
        if l == NL - 1:
            W *= 0.1
        out += [W.ravel(), np.zeros(co, np.float32)]
    return np.concatenate(out)


class CNN:
    """Six-layer periodic 3x3 CNN (bf16 weights/activations, fp32 accumulation) on the GPU."""

    def __init__(self, max_batch: int, params: np.ndarray | None = None, seed: int = 0, dil=None,
                 norm: bool = False, N: int = 64):
        import torch
        if not torch.cuda.is_available():
            raise QBFError("no HIP device available")
        self._torch = torch
        h = C.c_void_p()
        _check(_lib().qbf_cnn_create(N, max_batch, C.byref(h)))
        self.N = N
        self._h = h
        self.max_batch = max_batch
        self.dil = tuple(dil) if dil is not None else (1,) * NL
        self.norm = bool(norm)
        d = (C.c_int * NL)(*self.dil)
        _check(_lib().qbf_cnn_configure(self._h, d, 1 if norm else 0))
        self.set_params(init_params(seed, norm) if params is None else params)

    def close(self):
        if getattr(self, "_h", None):
            _lib().qbf_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _bind(self):
        _check(_lib().qbf_cnn_set_stream(self._h, C.c_void_p(self._torch.cuda.current_stream().cuda_stream)))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != param_count():
            raise QBFError(f"expected {param_count()} parameters, got {p.size}")
        self._bind()
        _check(_lib().qbf_cnn_set_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def get_params(self, grads: bool = False) -> np.ndarray:
        self._bind()
        p = np.empty(param_count(), np.float32)
        _check(_lib().qbf_cnn_get_params(self._h, p.ctypes.data_as(C.POINTER(C.c_float)), 1 if grads else 0))
        return p

    def forward(self, x, iscale: float = 1.0, oscale: float = 1.0, out=None):
        """out = oscale * net(iscale * x); x float32 cuda [B, 4, N, N]."""
        self._bind()
        out = self._torch.empty_like(x) if out is None else out
        _check(_lib().qbf_cnn_forward(self._h, _ptr(x), x.shape[0], iscale, _ptr(out), oscale))
        return out

    def gradients(self, x, target, iscale: float = 1.0, oscale: float = 1.0, out=None):
        """Forward and backward of mean((out - target)^2); returns (loss, out).  The parameter
        gradients stay on the device (get_params(grads=True) reads them)."""
        self._bind()
        out = self._torch.empty_like(x) if out is None else out
        loss = C.c_float()
        _check(_lib().qbf_cnn_gradients(self._h, _ptr(x), _ptr(target), x.shape[0], iscale, oscale, _ptr(out),
                                        C.byref(loss)))
        return loss.value, out

    def adam(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self._bind()
        _check(_lib().qbf_cnn_adam(self._h, lr, b1, b2, eps))

    # ------------------------------------------------------------------ stepping with the surrogate
    def set_offset(self, D=None):
        """Offset field of the surrogate operator, C(W) ~ net(W) + D (float32 cuda [ncell, 4, N, N] on
        the cells of the states passed to the stepping calls), or None."""
        if D is not None and (D.dtype != self._torch.float32 or D.dim() != 4 or tuple(D.shape[1:]) != (4, self.N, self.N)):
            raise QBFError(f"offset field must be float32 [ncell, 4, {self.N}, {self.N}], got {D.dtype} {tuple(D.shape)}")
        self._offset = D
        _check(_lib().qbf_surrogate_set_offset(self._h, None if D is None else _ptr(D), 0 if D is None else D.shape[0]))

    def update_offset(self, D):
        """Set the offset field to the values of D: written in place into the current field when it
        has D's shape (graphs captured with that field then see the new values), else set_offset(D)
        (graphs captured before keep reading the old field)."""
        cur = getattr(self, "_offset", None)
        if cur is None or cur.shape != D.shape or cur.dtype != D.dtype:
            self.set_offset(D)
        else:
            cur.copy_(D)

    def collision_step(self, engine, W, dt, iscale, oscale, L, mask=0, out=None):
        """Midpoint collision step of physical states W (float64 cuda [ncell, 4, N, N]) with the
        network as collision operator, followed by the moment correction selected by mask."""
        engine._bind()
        out = self._torch.empty_like(W) if out is None else out
        _check(_lib().qbf_surrogate_collision_step(self._h, engine._h, _ptrd(W), W.shape[0], dt, iscale, oscale, L,
                                                   mask, _ptrd(out)))
        return out

    def residual(self, engine, W, cells, iscale, oscale):
        """(sum|C_net - C|, sum|C|, sum|W|) over the given cells of W (float64 cuda
        [ncell, 4, N, N]), with C the engine's exact collision operator."""
        engine._bind()
        idx = np.ascontiguousarray(cells, np.int32)
        r = (C.c_double * 3)()
        _check(_lib().qbf_surrogate_residual(self._h, engine._h, _ptrd(W), W.shape[0],
                                             idx.ctypes.data_as(C.POINTER(C.c_int)), idx.size, iscale, oscale, r))
        return r[0], r[1], r[2]

    def inhom_step_periodic(self, engine, W, L, h, dt, iscale, oscale, mask=0, out=None):
        """Strang step (half transport, surrogate collision step, half transport)."""
        engine._bind()
        out = self._torch.empty_like(W) if out is None else out
        _check(_lib().qbf_surrogate_inhom_step_periodic(self._h, engine._h, _ptrd(W), W.shape[0], L, h, dt, iscale,
                                                        oscale, mask, _ptrd(out)))
        return out

    def capture_inhom_step_periodic(self, engine, W, L, h, dt, iscale, oscale, mask=0, out=None):
        """HIP graph of one surrogate Strang step from the buffer W into the buffer out (returned with
        the graph), as Engine.capture_inhom_step_periodic: the step runs once eagerly first, capture
        and replays use a dedicated torch stream.  The graph keeps this network's weights and its
        offset field by address (an offset field must then be updated in place)."""
        from . import Graph
        torch = self._torch
        out = torch.empty_like(W) if out is None else out
        stream = torch.cuda.Stream()
        stream.wait_stream(torch.cuda.current_stream())
        g = C.c_void_p()
        with torch.cuda.stream(stream):
            engine._bind()
            _check(_lib().qbf_graph_capture_surrogate_inhom_step_periodic(
                self._h, engine._h, _ptrd(W), W.shape[0], L, h, dt, iscale, oscale, mask, _ptrd(out), C.byref(g)))
        return Graph(g, stream, (W, out, self._offset, self), engine), out


def _ptrd(t):
    if not (t.is_cuda and t.is_contiguous() and t.dtype == __import__("torch").float64):
        raise QBFError(f"expected a contiguous float64 cuda tensor, got {t.dtype} on {t.device}")
    return C.c_void_p(t.data_ptr())
