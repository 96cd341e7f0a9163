"""Pins the FP64 denoiser oracle (no GPU): the convolution against a direct
numpy sum, its two backward passes by the adjoint identity
<conv(x,w), dy> = <x, dX(dy,w)> = <w, dW(dy,x)>, and the whole network's
gradient against central finite differences."""
import numpy as np
import pytest

from oracle import oracle as O


def conv_numpy(x, w, b):
    B, H, W, Cin = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((B, H, W, w.shape[0])) + b
    for ky in range(3):
        for kx in range(3):
            y += np.einsum("bhwc,oc->bhwo", xp[:, ky:ky + H, kx:kx + W, :], w[:, ky, kx, :])
    return y


def test_conv_forward_matches_numpy():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 7, 9, 5))
    w = rng.standard_normal((4, 3, 3, 5))
    b = rng.standard_normal(4)
    assert np.allclose(O.conv3x3_fwd(x, w, b), conv_numpy(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_backward_adjoints():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 6, 5, 3))
    w = rng.standard_normal((4, 3, 3, 3))
    dy = rng.standard_normal((2, 6, 5, 4))
    y = O.conv3x3_fwd(x, w, np.zeros(4))
    lhs = np.sum(y * dy)
    assert np.sum(x * O.conv3x3_bwd_x(dy, w)) == pytest.approx(lhs, rel=1e-12)
    assert np.sum(w * O.conv3x3_bwd_w(dy, x)) == pytest.approx(lhs, rel=1e-12)


def small_batch(seed, b=1, h=32, w=32):
    rng = np.random.default_rng(seed)
    in10 = np.empty((b, h, w, 10), np.float32)
    in10[..., :3] = rng.gamma(1.0, 0.5, (b, h, w, 3))
    in10[..., 3:6] = rng.uniform(0.2, 0.8, (b, h, w, 3))
    n = rng.standard_normal((b, h, w, 3))
    in10[..., 6:9] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    in10[..., 9] = rng.uniform(1, 3, (b, h, w))
    target = (in10[..., :3] * rng.uniform(0.8, 1.2, (b, h, w, 3))).astype(np.float32)
    return in10, target


def init_params(seed):
    n = int(O.lib().orc_unet_num_params())
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) * 0.05).astype(np.float32)


def test_unet_gradient_finite_differences():
    in10, target = small_batch(3)
    p = init_params(4)
    loss, g, _ = O.unet_grad(p, in10, target, emulate=False)
    assert loss > 0
    rng = np.random.default_rng(5)
    idx = rng.choice(p.size, 24, replace=False)
    idx = np.concatenate([idx, [p.size - 1, p.size - 20]])  # output-layer biases
    checked = 0
    for i in idx:
        # float32 parameters: use a step representable in float32
        h = np.float32(1e-3) * max(1.0, abs(p[i]))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        hh = (np.float64(pp[i]) - np.float64(pm[i])) / 2
        lp = O.unet_grad(pp, in10, target, emulate=False)[0]
        lm = O.unet_grad(pm, in10, target, emulate=False)[0]
        fd = (lp - lm) / (2 * hh)
        if abs(g[i]) < 1e-9 and abs(fd) < 1e-9:
            continue
        checked += 1
        assert fd == pytest.approx(g[i], rel=5e-2, abs=1e-7), i
    assert checked >= 10
