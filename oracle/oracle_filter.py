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
