"""Denoiser data, training and the equal-quality comparison.

* :func:`random_scene` draws a Cornell-box scene like the 1291-primitive
  workload: displaced icosphere (level 3-5, amplitude U[0, 0.15]), Lambert or
  GGX material with random roughness/F0, random wall albedos, a random
  camera pose and, with probability 1/2, a 64^3 fBm medium (sigma_max
  U[1, 8], grey albedo U[0.6, 0.95]) - all from one integer seed.
* :func:`orbit_cameras` draws seeded views of the same kind.
* :func:`render_pair` renders the network input (low spp + albedo, normal,
  depth) and the target (high spp) of one scene with the path tracer.
* :func:`relmse` is mean((x - r)^2 / (r^2 + 1e-2)) over pixels and channels.
* :func:`equal_quality` finds the plain path-tracing spp whose relMSE matches
  the denoised image's, by rendering at doubling spp and interpolating
  log(relMSE) against log(spp).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import _lib as L
from . import scenes
from .context import Context
from .denoiser import Denoiser, pack_input


def relmse(x: np.ndarray, ref: np.ndarray) -> float:
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def relmse_two_ref(x: np.ndarray, r1: np.ndarray, r2: np.ndarray) -> float:
    """relMSE of x with the references' own noise removed: r1 and r2 are
    independent unbiased estimates of the same image, so
    E[(x - r1)(x - r2)] = E[(x - mu)^2] (x independent of both); normalised
    by the mean reference, as relmse.  An unbiased estimate: it can fall
    below zero when x is much closer to mu than the references are."""
    x = np.asarray(x, np.float64)
    r1 = np.asarray(r1, np.float64)
    r2 = np.asarray(r2, np.float64)
    rb = 0.5 * (r1 + r2)
    return float(np.mean((x - r1) * (x - r2) / (rb ** 2 + 1e-2)))


def crossing_spp(points, target: float):
    """spp at which a measured relMSE curve [(spp, relMSE), ...] (spp
    increasing) first falls to `target`, interpolating log(relMSE) linearly
    in log(spp) between the two bracketing measurements; None when no
    measurement is at or below the target (no extrapolation)."""
    for k, (n, e) in enumerate(points):
        if e <= target:
            if k == 0:
                return float(n)
            n0, e0 = points[k - 1]
            if e <= 0 or e0 <= 0 or target <= 0:
                return float(n)
            f = (math.log(target) - math.log(e0)) / (math.log(e) - math.log(e0))
            return math.exp(math.log(n0) + f * (math.log(n) - math.log(n0)))
    return None


def random_scene(seed: int, width: int = 512, height: int = 512, medium_prob: float = 0.5) -> scenes.SceneData:
    rng = np.random.default_rng(1_000_003 * seed + 17)
    s = scenes.SceneData()
    scenes.cornell_box(s, rng)
    level = int(rng.integers(3, 6))
    amp = float(rng.uniform(0.0, 0.15))
    r = float(rng.uniform(0.18, 0.32))
    c = (float(rng.uniform(-0.12, 0.12)), float(-0.5 + r * (1 + amp) + 0.01), float(rng.uniform(-0.12, 0.12)))
    tri = scenes.icosphere(level, c, r, amp=amp, seed=int(rng.integers(1, 1 << 30)))
    if rng.random() < 0.5:
        m = s.add_material(rng.uniform(0.2, 0.8, 3), L.M_LAMBERT)
    else:
        rough = float(rng.uniform(0.05, 0.8))
        f0 = float(rng.uniform(0.02, 0.9))
        m = s.add_material((f0, f0, f0), L.M_GGX, rough, rough)
    s.add_mesh(tri, m)
    # camera: in front of the open face, looking at the box interior
    th = rng.uniform(-0.35, 0.35)
    ph = rng.uniform(-0.25, 0.25)
    dist = rng.uniform(1.6, 2.2)
    frm = (dist * math.sin(th) * math.cos(ph), dist * math.sin(ph), dist * math.cos(th) * math.cos(ph))
    at = tuple(rng.uniform(-0.08, 0.08, 3).tolist())
    fov = float(rng.uniform(34.0, 46.0))
    s.camera = L.CameraDesc(L.d3((0, 0, 0)), L.d3(frm), L.d3(at), fov, width / height, width, 0)
    # heterogeneous fBm medium filling the box in some scenes (the
    # config-5 workload has one)
    if rng.random() < medium_prob:
        a = float(rng.uniform(0.6, 0.95))
        s.medium = scenes.Medium(scenes.fbm_grid(64, int(rng.integers(1, 1 << 30))), (-0.5, -0.5, -0.5),
                                 (0.5, 0.5, 0.5), float(rng.uniform(1.0, 8.0)), (a, a, a))
    return s


def orbit_cameras(n: int, width: int, height: int, seed: int = 7) -> List[L.CameraDesc]:
    """n seeded views of the Cornell-box scenes from in front of the open face."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        th = rng.uniform(-0.35, 0.35)
        ph = rng.uniform(-0.25, 0.25)
        dist = rng.uniform(1.6, 2.2)
        frm = (dist * math.sin(th) * math.cos(ph), dist * math.sin(ph), dist * math.cos(th) * math.cos(ph))
        at = tuple(rng.uniform(-0.08, 0.08, 3).tolist())
        fov = float(rng.uniform(34.0, 46.0))
        out.append(L.CameraDesc(L.d3((0, 0, 0)), L.d3(frm), L.d3(at), fov, width / height, width, 0))
    return out


@dataclass
class Pair:
    inp: np.ndarray     # (h, w, 10)
    target: np.ndarray  # (h, w, 3)


def render_pair(ctx: Context, s: scenes.SceneData, spp_in: int, spp_ref: int, seed: int) -> Pair:
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    rgb, aov, _ = ctx.render(s.camera, spp_in, seed=seed, aov=True, profile=L.PROFILE_CONFIG)
    ref, _, _ = ctx.render(s.camera, spp_ref, seed=seed + 7919, sample_base=spp_in,
                           profile=L.PROFILE_CONFIG)
    return Pair(pack_input(rgb, aov), ref)


def make_dataset(ctx: Context, n: int, size: int, spp_in: int, spp_ref: int, seed0: int = 0,
                 log=None) -> List[Pair]:
    out = []
    t0 = time.time()
    for i in range(n):
        out.append(render_pair(ctx, random_scene(seed0 + i, size, size), spp_in, spp_ref, seed=101 + i))
        if log and (i + 1) % max(1, n // 8) == 0:
            log(f"dataset {i + 1}/{n} scenes ({time.time() - t0:.1f}s)")
    return out


def augment(inp: np.ndarray, tgt: np.ndarray, rng: np.random.Generator, crop: int):
    """Random crop and horizontal flip (normal x flips sign)."""
    h, w = inp.shape[:2]
    y = int(rng.integers(0, h - crop + 1))
    x = int(rng.integers(0, w - crop + 1))
    a = inp[y:y + crop, x:x + crop].copy()
    b = tgt[y:y + crop, x:x + crop].copy()
    if rng.random() < 0.5:
        a = a[:, ::-1].copy()
        b = b[:, ::-1].copy()
        a[..., 6] = -a[..., 6]
    return a, b


def train(dn: Denoiser, data: List[Pair], steps: int, batch: int, crop: int, lr: float = 1e-3,
          seed: int = 0, log=None) -> List[float]:
    rng = np.random.default_rng(seed)
    losses = []
    t0 = time.time()
    for it in range(steps):
        idx = rng.integers(0, len(data), batch)
        xs, ys = zip(*(augment(data[i].inp, data[i].target, rng, crop) for i in idx))
        # cosine learning-rate decay
        cur = lr * 0.5 * (1 + math.cos(math.pi * it / max(1, steps)))
        loss = dn.train_step(np.stack(xs), np.stack(ys), max(cur, lr * 0.02))
        losses.append(loss)
        if log and (it + 1) % max(1, steps // 10) == 0:
            log(f"train step {it + 1}/{steps} loss {np.mean(losses[-20:]):.4f} ({time.time() - t0:.1f}s)")
    return losses


def equal_quality(ctx: Context, cam: L.CameraDesc, ref: np.ndarray, target_relmse: float, seed: int,
                  spp_start: int = 4, spp_max: int = 4096):
    """Plain-PT spp and time at which relMSE reaches target_relmse.

    Renders spp_start, 2*spp_start, ... until the relMSE falls below the
    target, then interpolates log(relMSE) linearly in log(spp) between the
    last two renders (time interpolated the same way)."""
    pts = []
    spp = spp_start
    while spp <= spp_max:
        img, _, st = ctx.render(cam, spp, seed=seed, profile=L.PROFILE_CONFIG)
        e = relmse(img, ref)
        pts.append((spp, e, st.ms))
        if e <= target_relmse:
            break
        spp *= 2
    if len(pts) == 1 or pts[-1][1] > target_relmse:
        s, e, ms = pts[-1]
        res = dict(spp=float(s), ms=float(ms), points=pts, reached=pts[-1][1] <= target_relmse)
        # not reached: relMSE against a finite-spp reference levels off at
        # the reference's own noise, relMSE(n) ~ a / n + b; fit the last four
        # points and extrapolate where the target lies above the floor b
        if len(pts) >= 4:
            n = np.array([p[0] for p in pts[-4:]], np.float64)
            e = np.array([p[1] for p in pts[-4:]], np.float64)
            A = np.stack([1.0 / n, np.ones_like(n)], 1)
            (a, b), *_ = np.linalg.lstsq(A, e, rcond=None)
            res["fit_a"], res["fit_b"] = float(a), float(b)
            if a > 0 and target_relmse > b:
                res["spp_fit"] = float(a / (target_relmse - b))
        return res
    (s0, e0, m0), (s1, e1, m1) = pts[-2], pts[-1]
    f = (math.log(target_relmse) - math.log(e0)) / (math.log(e1) - math.log(e0))
    spp_eq = math.exp(math.log(s0) + f * (math.log(s1) - math.log(s0)))
    ms_eq = m0 + (m1 - m0) * (spp_eq - s0) / (s1 - s0)
    return dict(spp=spp_eq, ms=ms_eq, points=pts, reached=True)
