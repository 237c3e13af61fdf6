"""CPU oracle of the CNN surrogate (forward, loss, backward) -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the network run by csrc/qbf_cnn.hip: six periodic 3x3 convolutions
4 -> 64 -> 64 -> 64 -> 64 -> 64 -> 4 with the tanh-GELU between them, loss = mean squared error.
With ``bf16=True`` values are rounded to bfloat16 (round to nearest even) at the same points
as the GPU (layer inputs, weights, gradients of the pre-activations); arithmetic is float64.
With ``bf16=False`` everything is float64, which is what the finite-difference test uses to pin
the backward pass (tests/test_oracle_cnn.py).

Options (same as qbf_cnn_configure): per-layer dilation d_l (tap offsets (ky-1, kx-1) * d_l) and
per-cell normalisation: with it, a_b = max |x[b, 0]| (the trace channel) and the network sees
x / a_b on channels 0..3 and iscale * a_b as a constant channel 4; its output is multiplied by
oscale * a_b^2, and the loss compares the raw network output with target / (oscale * a_b^2).

Parameter layout (same as qbf_cnn_get_params): per layer, weights [9][cout_pad][cin_pad]
(tap t = 3*ky + kx, offset (ky-1, kx-1)) then bias [cout_pad]; cin_pad = 32 for layer 0,
cout_pad = 32 for layer 5, 64 otherwise.
"""
from __future__ import annotations

import numpy as np

NL, CH, CIO, CREAL = 6, 64, 32, 4


def cin(l):
    return CIO if l == 0 else CH


def cout(l):
    return CIO if l == NL - 1 else CH


def param_count():
    return sum(9 * cout(l) * cin(l) + cout(l) for l in range(NL))


def unpack(p):
    out, o = [], 0
    for l in range(NL):
        n = 9 * cout(l) * cin(l)
        W = p[o:o + n].reshape(9, cout(l), cin(l))
        o += n
        b = p[o:o + cout(l)]
        o += cout(l)
        out.append((W, b))
    return out


def init_params(seed=0, cin0=CREAL):
    """He-style initialisation; padded (unused) entries are zero.  cin0 input channels are used
    by the first layer (4, or 5 with the amplitude channel of the per-cell normalisation)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(param_count(), np.float32)
    o = 0
    for l in range(NL):
        ci_r = cin0 if l == 0 else CH
        co_r = CREAL if l == NL - 1 else CH
        W = np.zeros((9, cout(l), cin(l)), np.float32)
        W[:, :co_r, :ci_r] = rng.standard_normal((9, co_r, ci_r)) * np.sqrt(2.0 / (9 * ci_r))
        if l == NL - 1:
            W *= 0.1
        n = W.size
        p[o:o + n] = W.ravel()
        o += n + cout(l)
    return p


def bf16(x):
    """Round to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def gelu(z):
    c = np.sqrt(2 / np.pi)
    return 0.5 * z * (1 + np.tanh(c * (z + 0.044715 * z ** 3)))


def gelu_grad(z):
    c = np.sqrt(2 / np.pi)
    t = np.tanh(c * (z + 0.044715 * z ** 3))
    return 0.5 * (1 + t) + 0.5 * z * (1 - t * t) * c * (1 + 3 * 0.044715 * z * z)


def _shift(x, dy, dx):
    """x[b, y + dy, x + dx, c] at [b, y, x, c] (periodic)."""
    return np.roll(x, shift=(-dy, -dx), axis=(1, 2))


def conv(x, W, d=1):
    """y[b, p, o] = sum_t sum_c W[t, o, c] x[b, p + d off_t, c]."""
    y = 0.0
    for t in range(9):
        ky, kx = divmod(t, 3)
        y = y + np.einsum("byxc,oc->byxo", _shift(x, d * (ky - 1), d * (kx - 1)), W[t], optimize=True)
    return y


def conv_data_grad(g, W, d=1):
    """dx[b, q, c] = sum_t sum_o W[t, o, c] g[b, q - d off_t, o]."""
    r = 0.0
    for t in range(9):
        ky, kx = divmod(t, 3)
        r = r + np.einsum("byxo,oc->byxc", _shift(g, -d * (ky - 1), -d * (kx - 1)), W[t], optimize=True)
    return r


def conv_weight_grad(g, x, d=1):
    """dW[t, o, c] = sum_p g[p, o] x[p + d off_t, c]."""
    out = np.empty((9, g.shape[-1], x.shape[-1]))
    for t in range(9):
        ky, kx = divmod(t, 3)
        out[t] = np.einsum("byxo,byxc->oc", g, _shift(x, d * (ky - 1), d * (kx - 1)), optimize=True)
    return out


def cell_amp(x):
    """a_b = max |x[b, 0]| in float32, as the GPU computes it."""
    return np.abs(np.asarray(x, np.float32)[:, 0]).max(axis=(1, 2)).astype(np.float64)


def _inputs(x, iscale, norm, rnd):
    B, _, S, _ = x.shape
    xf = np.asarray(x, np.float32).astype(np.float64)
    x0 = np.zeros((B, S, S, CIO))
    if norm:
        a = cell_amp(x)
        inv = (np.float32(iscale) / a.astype(np.float32)).astype(np.float64)   # fp32 factor, as on the GPU
        x0[..., :CREAL] = np.transpose(xf * inv[:, None, None, None], (0, 2, 3, 1))
        x0[..., CREAL] = (np.float32(iscale) * a.astype(np.float32)).astype(np.float64)[:, None, None]
    else:
        x0[..., :CREAL] = np.transpose(iscale * xf, (0, 2, 3, 1))
    return rnd(x0)


def forward(x, params, iscale=1.0, oscale=1.0, bf16_round=True, dil=(1,) * NL, norm=False):
    """x: planar [B, 4, S, S] -> (output planar [B, 4, S, S], cache).  iscale multiplies the
    network input (after the per-cell division when norm), oscale (times a_b^2 when norm) its output."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    layers = unpack(np.asarray(params, np.float32))
    xs = [_inputs(x, iscale, norm, rnd)]
    zs = []
    for l in range(NL - 1):
        W, b = layers[l]
        z = conv(xs[l], rnd(W), dil[l]) + b.astype(np.float64)
        zs.append(z)
        xs.append(rnd(gelu(z)))
    W, b = layers[NL - 1]
    y = conv(xs[NL - 1], rnd(W), dil[NL - 1]) + b.astype(np.float64)
    net = np.transpose(y[..., :CREAL], (0, 3, 1, 2))
    sc = oscale * (cell_amp(x) ** 2 if norm else np.ones(x.shape[0]))
    return sc[:, None, None, None] * net, (xs, zs, layers, net, sc)


def gradients(x, target, params, iscale=1.0, oscale=1.0, bf16_round=True, dil=(1,) * NL, norm=False):
    """Loss and its gradient, flat in the parameter layout.  Without norm the loss is
    mean((out - target)^2); with norm it is mean((net - target / (oscale a_b^2))^2)."""
    rnd = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    out, (xs, zs, layers, net, sc) = forward(x, params, iscale, oscale, bf16_round, dil, norm)
    B, _, S, _ = x.shape
    n = out.size
    if norm:
        d = net - target / sc[:, None, None, None]
        gy = 2 * d / n
    else:
        d = out - target
        gy = 2 * d / n * oscale
    loss = float((d * d).sum() / n)
    g = np.zeros((B, S, S, CIO))
    g[..., :CREAL] = np.transpose(gy, (0, 2, 3, 1))
    g = rnd(g)
    grads = [None] * NL
    for l in range(NL - 1, -1, -1):
        W, _ = layers[l]
        grads[l] = (conv_weight_grad(g, xs[l], dil[l]), g.sum(axis=(0, 1, 2)))
        if l > 0:
            g = rnd(conv_data_grad(g, rnd(W), dil[l]) * gelu_grad(zs[l - 1]))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    # unused (padded) entries never receive updates
    o = 0
    for l in range(NL):
        ci_r = (CREAL + 1 if norm else CREAL) if l == 0 else CH
        co_r = CREAL if l == NL - 1 else CH
        gw = flat[o:o + 9 * cout(l) * cin(l)].reshape(9, cout(l), cin(l))
        gw[:, co_r:, :] = 0
        gw[:, :, ci_r:] = 0
        o += gw.size
        flat[o + co_r:o + cout(l)] = 0
        o += cout(l)
    return loss, flat, out
