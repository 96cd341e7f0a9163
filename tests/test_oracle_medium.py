"""Pins the oracle's heterogeneous medium (no GPU): delta tracking and ratio
tracking against Beer-Lambert on piecewise-constant grids, and the
integrator's coupling of the medium with emission against a closed form.

* Ratio tracking is unbiased: the mean of the estimates is exp(-tau), tau =
  sigma * sum(density * length) over the voxels crossed.  An anisotropic grid
  and rays along each axis and an oblique one catch a transposed or
  mis-scaled voxel index.
* Delta tracking: the collision probability over a segment is 1 - exp(-tau)
  and, in a homogeneous grid, collision distances follow the truncated
  exponential (mean checked).
* Integrator: a light seen through a purely absorbing slab (albedo 0) has
  expected radiance Le * exp(-tau) per pixel.
"""
import math

import numpy as np

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import camera, scenes
from oracle import oracle as O


def grid_scene(dens, bmin=(0.0, 0.0, 0.0), bmax=(1.0, 1.0, 1.0), sigma=3.0, albedo=(0.5, 0.5, 0.5)):
    s = scenes.SceneData()
    s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.medium = scenes.Medium(np.asarray(dens, np.float32), bmin, bmax, sigma, albedo)
    return O.OracleScene(s)


def optical_depth(dens, bmin, bmax, sigma, o, d, tmin, tmax, steps=200000):
    """Midpoint quadrature of sigma * density along the ray (voxel lookups)."""
    o, d = np.asarray(o, float), np.asarray(d, float)
    nz, ny, nx = dens.shape
    ts = tmin + (np.arange(steps) + 0.5) * (tmax - tmin) / steps
    p = o[None] + ts[:, None] * d[None]
    lo, hi = np.asarray(bmin, float), np.asarray(bmax, float)
    inside = np.all((p >= lo) & (p <= hi), axis=1)
    idx = np.floor((p - lo) / (hi - lo) * np.array([nx, ny, nz])).astype(int)
    idx = np.clip(idx, 0, np.array([nx, ny, nz]) - 1)
    rho = dens[idx[:, 2], idx[:, 1], idx[:, 0]] * inside
    return sigma * rho.sum() * (tmax - tmin) / steps * np.linalg.norm(d)


def test_ratio_tracking_matches_beer_lambert():
    rng = np.random.default_rng(3)
    dens = rng.uniform(0.0, 1.0, (2, 3, 4)).astype(np.float32)  # nz=2, ny=3, nx=4
    sc = grid_scene(dens)
    rays = [((-1.0, 0.4, 0.7), (1.0, 0.0, 0.0), 0.5, 3.0),
            ((0.2, -1.0, 0.3), (0.0, 1.0, 0.0), 0.5, 3.0),
            ((0.9, 0.1, -2.0), (0.0, 0.0, 1.0), 1.0, 4.0),
            ((-0.5, -0.3, -0.2), (1.3, 1.1, 0.9), 0.0, 2.0)]
    for o, d, t0, t1 in rays:
        tau = optical_depth(dens, (0, 0, 0), (1, 1, 1), 3.0, o, d, t0, t1)
        est = sc.medium_transmittance(o, d, t0, t1, seed=5, n=200000)
        assert np.all((est >= 0) & (est <= 1))
        se = est.std() / math.sqrt(est.size)
        assert abs(est.mean() - math.exp(-tau)) < 5 * se + 1e-4, (o, d, est.mean(), math.exp(-tau))


def test_delta_tracking_collision_statistics():
    rng = np.random.default_rng(4)
    dens = rng.uniform(0.0, 1.0, (3, 2, 5)).astype(np.float32)
    sc = grid_scene(dens, sigma=2.0)
    o, d = (-0.5, 0.3, 0.6), (0.8, 0.15, -0.1)
    t = sc.medium_collide(o, d, 0.0, 3.0, seed=9, n=200000)
    tau = optical_depth(dens, (0, 0, 0), (1, 1, 1), 2.0, o, d, 0.0, 3.0)
    p = np.mean(t >= 0)
    assert abs(p - (1 - math.exp(-tau))) < 5 * math.sqrt(p * (1 - p) / t.size)
    # homogeneous grid: truncated exponential over the box segment
    sc = grid_scene(np.full((2, 2, 2), 0.5, np.float32), sigma=2.0)
    o, d = (-1.0, 0.5, 0.5), (2.0, 0.0, 0.0)  # |d| = 2: box spans t in [0.5, 1]
    t = sc.medium_collide(o, d, 0.0, 10.0, seed=2, n=200000)
    hit = t[t >= 0]
    lam = 2.0 * 0.5 * 2.0  # sigma * density * |d| per unit t
    L_ = 0.5
    assert abs(hit.size / t.size - (1 - math.exp(-lam * L_))) < 0.005
    assert hit.min() >= 0.5 and hit.max() <= 1.0
    mean_rel = 1 / lam - L_ * math.exp(-lam * L_) / (1 - math.exp(-lam * L_))
    assert abs((hit - 0.5).mean() - mean_rel) < 0.003


def test_absorbing_slab_in_front_of_light():
    """Camera on +z looking at a light (normal +z, Le 2) behind a homogeneous
    slab z in [-0.25, 0.25] of density 0.6, sigma 2, albedo 0, black sky:
    E[pixel] = Le exp(-sigma rho 0.5 |d| / |d_z|)."""
    s = scenes.SceneData()
    s.add_light((0.0, 0.0, -1.0), 4.0, 4.0, (0.0, 0.0, 1.0), (2.0, 2.0, 2.0))
    s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.medium = scenes.Medium(np.full((1, 1, 1), 0.6, np.float32), (-3, -3, -0.25), (3, 3, 0.25), 2.0,
                             (0.0, 0.0, 0.0))
    o = O.OracleScene(s)
    o.build(np.array([-3.0, -3, -1]), np.array([3.0, 3, 1]))
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0, 0, 3.0)), L.d3((0, 0, 0)), 30.0, 1.0, 48, 0)
    cam = camera(d)
    rgb = o.render(cam, L.RenderParams(17, 8, 4, 3, L.PROFILE_CONFIG, 0.7, 0, 0, 0, 0))
    # per-pixel expectation from the pixel-centre direction (jitter changes
    # |d|/|d_z| by < 1e-3 at this field of view)
    h, w = cam.height, cam.width
    i, j = np.meshgrid(np.arange(w), np.arange(h))
    pc, dh, dv, lf = (np.array(getattr(cam, f)[:]) for f in ("pixel_center", "delta_h", "delta_v", "look_from"))
    dirs = pc[None, None] + i[..., None] * dh + j[..., None] * dv - lf
    sec = np.linalg.norm(dirs, axis=-1) / np.abs(dirs[..., 2])
    expect = 2.0 * np.exp(-2.0 * 0.6 * 0.5 * sec)
    ratio = rgb[..., 0] / expect
    # each sample is Le or 0 (binomial): mean over 48*48*8 samples
    p = math.exp(-0.6)
    se = math.sqrt(p * (1 - p) / (h * w * 8)) / p
    assert abs(ratio.mean() - 1) < 5 * se
    assert np.allclose(rgb[..., 0], rgb[..., 1]) and np.allclose(rgb[..., 0], rgb[..., 2])
