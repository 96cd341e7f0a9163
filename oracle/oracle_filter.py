"""FP64 numpy restatement of the reference's joint bilateral filter
(gaussianSeparate, CudaPathTracing/Render.cu:353-398, horizontal then
vertical as main.cu:100-104).  TEST INFRASTRUCTURE ONLY."""
import numpy as np

RADIUS = 16  # KERNEL_RADIUS, Render.cuh:7


def _pass(src, center, normal, depth, sG, sR, sN, sD, axis):
    h, w = src.shape[:2]
    acc = np.zeros_like(src)
    wsum = np.zeros(src.shape[:2])
    for i in range(-RADIUS, RADIUS + 1):
        if axis == 1:
            idx = np.clip(np.arange(w) + i, 0, w - 1)
            p, n, d = src[:, idx], normal[:, idx], depth[:, idx]
        else:
            idx = np.clip(np.arange(h) + i, 0, h - 1)
            p, n, d = src[idx], normal[idx], depth[idx]
        wt = np.exp(-(i * i) / (2.0 * sG * sG)
                    - np.sum((p - center) ** 2, -1) / (2.0 * sR * sR)
                    - np.sum((normal - n) ** 2, -1) / (2.0 * sN * sN)
                    - (depth - d) ** 2 / (2.0 * sD * sD))
        wsum += wt
        acc += p * wt[..., None]
    return acc / wsum[..., None]


def bilateral(rgb, aov, sG=8.0, sR=5.0, sN=0.1, sD=0.2):
    rgb = np.asarray(rgb, np.float64)
    aov = np.asarray(aov, np.float64)
    normal, depth = aov[..., 3:6], aov[..., 6]
    mid = _pass(rgb, rgb, normal, depth, sG, sR, sN, sD, axis=1)
    return _pass(mid, mid, normal, depth, sG, sR, sN, sD, axis=0)


INF = 999999.0  # rtmath.cuh:8


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def reproject(position, prev_from, prev_at, vfov_deg, w, h):
    """Camera::getPrePositionInImg (Camera.cuh:104-132) for an (..., 3)
    array of points: (i, j) in the previous camera's image, -1 when the point
    is a miss (x + y > INF) or projects outside [0, 1]^2."""
    p = np.asarray(position, np.float64)
    lf = np.asarray(prev_from, np.float64)
    la = np.asarray(prev_at, np.float64)
    d = _unit(p - lf)
    fwd = _unit(la - lf)
    right = _unit(np.cross(fwd, [0.0, 1.0, 0.0]))
    up = np.cross(right, fwd)
    focal = np.linalg.norm(lf - la)
    vh = 2.0 * np.tan(np.radians(vfov_deg) / 2.0) * focal
    vw = w / h * vh
    px, py, pz = d @ right, d @ up, d @ fwd
    x = 0.5 + px / pz * (focal / vw)
    y = 0.5 + py / pz * (focal / vh)
    bad = (p[..., 0] + p[..., 1] > INF) | (x < 0) | (y < 0) | (x > 1) | (y > 1)
    i = np.where(bad, -1, np.floor(np.where(bad, 0, x) * w)).astype(np.int64)
    j = np.where(bad, -1, np.floor(np.where(bad, 0, y) * h)).astype(np.int64)
    return i, j


def temporal_blend(rgb, prev_rgb, position, prev_from, prev_at, vfov_deg, moving=False):
    """addPrevious (Render.cu:400-421): out = rgb*alpha + prev[i, j]*(1-alpha),
    alpha 0.4 while moving else 0.1, the same pixel when (i, j) is invalid.
    (i, j) = (w, .) / (., h) at x or y = 1 is clamped to the last pixel (the
    reference would read past the row)."""
    rgb = np.asarray(rgb, np.float64)
    prev = np.asarray(prev_rgb, np.float64)
    h, w = rgb.shape[:2]
    i, j = reproject(position, prev_from, prev_at, vfov_deg, w, h)
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    ok = (i >= 0) & (j >= 0)
    si = np.where(ok, np.minimum(i, w - 1), ii)
    sj = np.where(ok, np.minimum(j, h - 1), jj)
    a = 0.4 if moving else 0.1
    return rgb * a + prev[sj, si] * (1.0 - a)


def positions_from_depth(depth, look_from, pixel_center, delta_h, delta_v):
    """First-hit points from a per-pixel distance along the pixel-centre ray;
    distance <= 0 marks a miss (INF)."""
    h, w = depth.shape
    jj, ii = np.meshgrid(np.arange(h), np.arange(w), indexing="ij")
    pc = np.asarray(pixel_center, np.float64)
    dh, dv = np.asarray(delta_h, np.float64), np.asarray(delta_v, np.float64)
    lf = np.asarray(look_from, np.float64)
    d = pc + ii[..., None] * dh + jj[..., None] * dv - lf
    p = lf + _unit(d) * np.asarray(depth, np.float64)[..., None]
    p[depth <= 0] = INF
    return p
