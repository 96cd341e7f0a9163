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


# ---- temporal reprojection (addPrevious, Render.cu:400-421) ----

def _cam(frm, at, w=64, h=48, fov=40.0):
    from cudapathtracing_amd import _lib as L
    from cudapathtracing_amd.context import camera
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3(frm), L.d3(at), fov, w / h, w, 0)
    return d, camera(d)


def _wall_points(cam, z=-0.5):
    """Pixel-centre rays of `cam` intersected with the plane z = const."""
    lf = np.array(cam.look_from[:])
    jj, ii = np.meshgrid(np.arange(cam.height), np.arange(cam.width), indexing="ij")
    d = (np.array(cam.pixel_center[:]) + ii[..., None] * np.array(cam.delta_h[:])
         + jj[..., None] * np.array(cam.delta_v[:]) - lf)
    t = (z - lf[2]) / d[..., 2]
    return lf + t[..., None] * d


def test_oracle_static_camera_reprojects_to_same_pixel():
    desc, cam = _cam((0.1, 0.05, 1.9), (0.0, 0.0, 0.0))
    i, j = F.reproject(_wall_points(cam), desc.look_from[:], desc.look_at[:], 40.0, cam.width, cam.height)
    jj, ii = np.meshgrid(np.arange(cam.height), np.arange(cam.width), indexing="ij")
    assert np.array_equal(i, ii) and np.array_equal(j, jj)


@pytest.mark.parametrize("k", [3, -5])
def test_oracle_translation_shifts_by_whole_pixels(k):
    # A camera translated by dx parallel to a wall at distance D sees the wall
    # shifted by dx / D * h / (2 tan(fov/2)) pixels; pick dx for k pixels.
    w, h, fov, D = 64, 48, 40.0, 2.4
    dx = 2 * k * D * np.tan(np.radians(fov) / 2) / h
    prev_from, prev_at = (-dx, 0.0, 1.9), (-dx, 0.0, 0.0)
    desc, cam = _cam((0.0, 0.0, 1.9), (0.0, 0.0, 0.0), w, h, fov)
    pos = _wall_points(cam, z=1.9 - D)
    i, j = F.reproject(pos, prev_from, prev_at, fov, w, h)
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    inside = (ii + k >= 0) & (ii + k < w)
    assert np.array_equal(i[inside], ii[inside] + k) and np.array_equal(j[inside], jj[inside])
    assert np.all(i[~inside] == -1)
    rng = np.random.default_rng(0)
    cur, prev = rng.uniform(0, 1, (h, w, 3)), rng.uniform(0, 1, (h, w, 3))
    out = F.temporal_blend(cur, prev, pos, prev_from, prev_at, fov, moving=True)
    src = np.where(inside, ii + k, ii)
    assert np.allclose(out, 0.4 * cur + 0.6 * prev[jj, src], rtol=0, atol=1e-15)


def test_oracle_misses_keep_the_same_pixel():
    desc, cam = _cam((0.0, 0.0, 1.9), (0.0, 0.0, 0.0))
    pos = np.full((cam.height, cam.width, 3), F.INF)
    rng = np.random.default_rng(1)
    cur, prev = rng.uniform(0, 1, (2, cam.height, cam.width, 3))
    out = F.temporal_blend(cur, prev, pos, (0.3, 0, 1.9), (0.3, 0, 0), 40.0)
    assert np.allclose(out, 0.1 * cur + 0.9 * prev, rtol=0, atol=1e-15)


def test_oracle_depth_positions_lie_on_pixel_rays():
    desc, cam = _cam((0.2, 0.1, 1.9), (0.0, 0.0, 0.0))
    pos = _wall_points(cam)
    dist = np.linalg.norm(pos - np.array(cam.look_from[:]), axis=-1)
    dist[0, :5] = 0.0
    got = F.positions_from_depth(dist, cam.look_from[:], cam.pixel_center[:], cam.delta_h[:], cam.delta_v[:])
    assert np.allclose(got[1:], pos[1:], atol=1e-12) and np.all(got[0, :5] == F.INF)


@pytest.mark.gpu
@pytest.mark.parametrize("moving", [False, True])
def test_gpu_temporal_blend_matches_oracle(ctx, moving):
    print(f"[gpu] temporal blend moving={moving}", flush=True)
    w, h, fov = 96, 64, 40.0
    desc, cam = _cam((0.05, 0.02, 1.9), (0.0, 0.0, 0.0), w, h, fov)
    prev_from, prev_at = (-0.04, 0.03, 1.85), (0.01, -0.01, 0.0)
    prev_desc, _ = _cam(prev_from, prev_at, w, h, fov)
    rng = np.random.default_rng(5)
    pos = _wall_points(cam, z=-0.5) + rng.normal(0, 1e-3, (h, w, 3))
    pos[:3] = F.INF
    cur = rng.gamma(1.0, 0.4, (h, w, 3)).astype(np.float32)
    prev = rng.gamma(1.0, 0.4, (h, w, 3)).astype(np.float32)
    got = ctx.temporal_blend(cur, prev, desc, prev_desc, position=pos.astype(np.float32), moving=moving)
    want = F.temporal_blend(cur, prev, pos.astype(np.float32), prev_from, prev_at, fov, moving)
    # (i, j) may differ where x*w or y*h lands within rounding of an integer;
    # the positions are float32 on both sides, so this is rare
    diff = np.abs(got - want).max(-1) > 1e-5 * np.maximum(np.abs(want).max(-1), 1)
    assert diff.mean() < 2e-3
    # depth path: positions rebuilt from the distance channel
    aov = np.zeros((h, w, 7), np.float32)
    aov[..., 6] = np.linalg.norm(pos - np.array(cam.look_from[:]), axis=-1)
    aov[:3, :, 6] = 0.0
    got = ctx.temporal_blend(cur, prev, desc, prev_desc, aov=aov, moving=moving)
    p2 = F.positions_from_depth(aov[..., 6].astype(np.float64), cam.look_from[:], cam.pixel_center[:],
                                cam.delta_h[:], cam.delta_v[:])
    want = F.temporal_blend(cur, prev, p2, prev_from, prev_at, fov, moving)
    diff = np.abs(got - want).max(-1) > 1e-5 * np.maximum(np.abs(want).max(-1), 1)
    assert diff.mean() < 2e-3
