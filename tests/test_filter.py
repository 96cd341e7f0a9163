"""Joint bilateral filter (Render.cu:353-398): the numpy oracle against
closed forms (no GPU), then the HIP kernel against the oracle (GPU)."""
import numpy as np
import pytest

from oracle import oracle_filter as F


def gbuffer(h, w, seed=0):
    rng = np.random.default_rng(seed)
    aov = np.zeros((h, w, 7))
    aov[..., 5] = 1.0
    aov[..., 6] = 2.0
    rgb = rng.uniform(0, 1, (h, w, 3))
    return rgb, aov


def test_oracle_constant_image_is_fixed_point():
    rgb, aov = gbuffer(20, 30)
    rgb[:] = 0.37
    assert np.allclose(F.bilateral(rgb, aov), 0.37, rtol=1e-14)


def test_oracle_reduces_to_clamped_gaussian():
    from scipy.ndimage import convolve1d
    rgb, aov = gbuffer(24, 40, 1)
    out = F.bilateral(rgb, aov, sG=3.0, sR=1e9, sN=1e9, sD=1e9)
    i = np.arange(-16, 17)
    k = np.exp(-i * i / 18.0)
    k /= k.sum()
    want = convolve1d(convolve1d(rgb, k, axis=1, mode="nearest"), k, axis=0, mode="nearest")
    assert np.allclose(out, want, rtol=1e-12, atol=1e-12)


def test_oracle_does_not_mix_across_normals():
    rgb, aov = gbuffer(16, 32, 2)
    aov[:, 16:, 3:6] = (1.0, 0.0, 0.0)
    rgb[:, :16] = 0.2
    rgb[:, 16:] = 0.9
    out = F.bilateral(rgb, aov, sR=1e9, sN=0.1)
    assert np.allclose(out[:, :16], 0.2, atol=1e-12) and np.allclose(out[:, 16:], 0.9, atol=1e-12)


@pytest.mark.gpu
def test_gpu_bilateral_matches_oracle(ctx):
    print("[gpu] bilateral filter", flush=True)
    import ctypes as C
    from cudapathtracing_amd import _lib as L
    rng = np.random.default_rng(3)
    h, w = 45, 70
    rgb = rng.gamma(1.0, 0.4, (h, w, 3)).astype(np.float32)
    aov = np.zeros((h, w, 7), np.float32)
    n = rng.standard_normal((h, w, 3))
    n[:, : w // 2] = (0, 0, 1)
    aov[..., 3:6] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    aov[..., 6] = rng.uniform(1.5, 2.5, (h, w))
    out = np.empty((h, w, 3), np.float32)
    rc = L.lib().cpt_bilateral(ctx._h, rgb.ctypes.data, aov.ctypes.data, out.ctypes.data, w, h, 8.0, 5.0, 0.1, 0.2, 0)
    assert rc == 0
    want = F.bilateral(rgb, aov)
    assert np.max(np.abs(out - want) / np.maximum(np.abs(want), 1e-3)) < 1e-4
