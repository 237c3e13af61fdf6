"""Pins the CPU oracle of the CNN surrogate (oracle/cnn_ref.py) by independent mathematics:
circular convolution via the FFT, adjoint identities, finite differences of the loss, the
exact erf-GELU, and bfloat16 rounding of hand-picked values."""
import numpy as np
import pytest
from scipy.special import erf

from oracle import cnn_ref as R


def _rand_conv(rng, B=2, S=8, ci=5, co=3):
    x = rng.standard_normal((B, S, S, ci))
    W = rng.standard_normal((9, co, ci))
    return x, W


def test_conv_matches_fft_circular_correlation():
    rng = np.random.default_rng(0)
    x, W = _rand_conv(rng)
    B, S, _, ci = x.shape
    co = W.shape[1]
    # y[p] = sum_t W_t x[p + off_t]: a circular cross-correlation with the kernel K[dy, dx] = W[t]
    K = np.zeros((S, S, co, ci))
    for t in range(9):
        ky, kx = divmod(t, 3)
        K[(ky - 1) % S, (kx - 1) % S] = W[t]
    Xf = np.fft.fft2(x, axes=(1, 2))
    Kf = np.fft.fft2(K, axes=(0, 1))
    yf = np.einsum("byxc,yxoc->byxo", Xf, np.conj(Kf))
    y = np.fft.ifft2(yf, axes=(1, 2)).real
    np.testing.assert_allclose(R.conv(x, W), y, rtol=0, atol=1e-10)


def test_conv_delta_response():
    """A unit impulse at pixel q in channel c produces W[t][:, c] at q - off_t."""
    S = 6
    x = np.zeros((1, S, S, 2))
    x[0, 2, 3, 1] = 1.0
    W = np.arange(9 * 3 * 2, dtype=float).reshape(9, 3, 2)
    y = R.conv(x, W)
    for t in range(9):
        ky, kx = divmod(t, 3)
        np.testing.assert_array_equal(y[0, 2 - (ky - 1), 3 - (kx - 1)], W[t][:, 1])


@pytest.mark.parametrize("d", [1, 3])
def test_data_and_weight_gradients_are_adjoints(d):
    rng = np.random.default_rng(1)
    x, W = _rand_conv(rng)
    g = rng.standard_normal(R.conv(x, W, d).shape)
    lhs = (R.conv(x, W, d) * g).sum()
    assert abs(lhs - (x * R.conv_data_grad(g, W, d)).sum()) < 1e-9 * abs(lhs) + 1e-9
    assert abs(lhs - (W * R.conv_weight_grad(g, x, d)).sum()) < 1e-9 * abs(lhs) + 1e-9


def test_dilated_conv_matches_direct_sum():
    """Direct loop over pixels and taps with offsets d (ky-1, kx-1), periodic."""
    rng = np.random.default_rng(2)
    x, W = _rand_conv(rng, B=1, S=6, ci=2, co=2)
    d = 2
    y = np.zeros((1, 6, 6, 2))
    for py in range(6):
        for px in range(6):
            for t in range(9):
                ky, kx = divmod(t, 3)
                y[0, py, px] += W[t] @ x[0, (py + d * (ky - 1)) % 6, (px + d * (kx - 1)) % 6]
    np.testing.assert_allclose(R.conv(x, W, d), y, atol=1e-12)


def test_gelu_close_to_erf_gelu_and_its_derivative():
    z = np.linspace(-6, 6, 2001)
    exact = 0.5 * z * (1 + erf(z / np.sqrt(2)))
    assert np.abs(R.gelu(z) - exact).max() < 1e-3        # tanh approximation error bound
    h = 1e-6
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(R.gelu_grad(z), fd, atol=1e-8)


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0], np.float32)
    # 8 significant bits: ties to even
    want = np.array([1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, 1.0 + 2 ** -7, -2.5, 0.0])
    np.testing.assert_array_equal(R.bf16(vals), want)


def test_param_layout():
    assert R.param_count() == (9 * 64 * 32 + 64) + 4 * (9 * 64 * 64 + 64) + (9 * 32 * 64 + 32)
    p = R.init_params(3)
    (W0, b0), *_, (W5, b5) = R.unpack(p)
    assert np.all(W0[:, :, 4:] == 0) and np.all(W5[:, 4:, :] == 0)
    assert np.all(b0 == 0)


@pytest.mark.parametrize("seed,dil,norm", [(0, (1,) * 6, False), (1, (1,) * 6, False),
                                           (2, (1, 2, 4, 8, 4, 1), True)])
def test_backward_matches_finite_differences(seed, dil, norm):
    """float64 mode (no rounding): the analytic gradient equals central differences of the loss
    for weights and biases of every layer (plain, and dilated with per-cell normalisation)."""
    rng = np.random.default_rng(seed)
    S, B = (8, 2) if not norm else (16, 2)
    p = R.init_params(seed, cin0=5 if norm else 4).astype(np.float64)
    # nonzero biases and a larger last layer so that every term contributes
    layers_off = []
    o = 0
    for l in range(R.NL):
        nw = 9 * R.cout(l) * R.cin(l)
        layers_off.append((o, nw))
        o += nw
        p[o:o + (4 if l == R.NL - 1 else 64)] = 0.1 * rng.standard_normal(4 if l == R.NL - 1 else 64)
        o += R.cout(l)
    x = rng.standard_normal((B, 4, S, S)) * 0.5
    if norm:
        x[1] *= 0.3                     # cells of different amplitude
    target = rng.standard_normal((B, 4, S, S)) * 0.1

    # the oracle casts parameters through float32: perturb in float32-representable steps
    p = p.astype(np.float32).astype(np.float64)
    kw = dict(iscale=1.3, oscale=0.7, bf16_round=False, dil=dil, norm=norm)
    loss, g, _ = R.gradients(x, target, p, **kw)

    def L(q):
        return R.gradients(x, target, q, **kw)[0]

    idx = []
    for l in range(R.NL):
        o, nw = layers_off[l]
        W = np.zeros((9, R.cout(l), R.cin(l)), bool)
        W[:, :(4 if l == R.NL - 1 else 64), :((5 if norm else 4) if l == 0 else 64)] = True
        live = np.flatnonzero(W.ravel())
        idx += list(o + rng.choice(live, 3, replace=False))
        idx.append(o + nw + int(rng.integers(0, 4 if l == R.NL - 1 else 64)))
    for i in idx:
        h = 2.0 ** -10
        qp, qm = p.copy(), p.copy()
        qp[i] += h
        qm[i] -= h
        fd = (L(qp) - L(qm)) / (2 * h)
        assert abs(fd - g[i]) <= 2e-4 * max(abs(fd), 1e-3 * abs(g).max()), (i, fd, g[i])
