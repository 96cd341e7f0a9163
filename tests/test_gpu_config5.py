"""The many-view medium workload (displaced GGX icosphere in the Cornell box
plus a heterogeneous fBm medium, several views in one cpt_render_views call,
then BF16 U-Net inference) at a size the FP64 oracle covers: level-5
icosphere, a 32^3 density grid, two 48 x 48 views.

* Each view's paths against the oracle with that view's pixel offset in the
  random stream (certified paths, as in test_gpu_render.py).
* The trained network's output on the GPU-rendered 4-spp inputs against the
  oracle's forward pass of the same network with the same BF16 rounding
  points (test_gpu_denoise.py checks the kernels and gradients).
"""
import os

import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import camera, scenes
from cudapathtracing_amd.denoiser import Denoiser, pack_input
from cudapathtracing_amd.training import orbit_cameras
from oracle import oracle as O
from test_gpu_render import certified_match, params

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cudapathtracing_amd",
                       "weights", "unet_cornell.f32")
W = H = 48


def scene():
    s = scenes.displaced_icosphere(W, H, level=5)
    s.medium = scenes.Medium(scenes.fbm_grid(32, 0), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 8.0, (0.8, 0.8, 0.8))
    return s


def test_views_with_medium_paths_and_inference(ctx):
    print("[gpu] config-5 workload at 2 x 48 x 48, level-5 icosphere, 32^3 medium", flush=True)
    s = scene()
    cams = orbit_cameras(2, W, H, seed=7)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    try:
        for b in (0, 1):
            g, _, st = ctx.render_views(cams, 1, seed=11, profile=L.PROFILE_CONFIG, sample_base=b)
            assert g.shape == (2, H, W, 3)
            for v, d in enumerate(cams):
                r, cert = o.render(camera(d), params(1, L.PROFILE_CONFIG, base=b), pixel_offset=v * W * H,
                                   cert=True)
                certified_match(g[v], r, cert, f"view {v} sample {b}", 0.8)
        # denoised 4 spp: the network on the GPU against the oracle's forward pass
        rgb, aov, _ = ctx.render_views(cams, 4, seed=3317, profile=L.PROFILE_CONFIG, aov=True)
        # the packed output (the denoiser's input layout, CPT_RENDER_PACK_AOV) is
        # the same numbers, to host memory and into a device tensor
        packed, none, _ = ctx.render_views(cams, 4, seed=3317, profile=L.PROFILE_CONFIG, aov=True, pack=True)
        assert none is None and packed.shape == (2, H, W, 10)
        import torch
        dev = torch.empty((2, H, W, 10), dtype=torch.float32, device="cuda")
        ctx.render_views(cams, 4, seed=3317, profile=L.PROFILE_CONFIG, aov=True, pack=True, out=dev)
        torch.cuda.synchronize()
        ctx.set_stream(None)
    finally:
        ctx.set_medium(None)
    in10 = np.stack([pack_input(rgb[v], aov[v]) for v in range(2)])
    assert np.array_equal(packed, in10)
    assert np.array_equal(dev.cpu().numpy(), in10)
    p = np.fromfile(WEIGHTS, dtype="<f4")
    dn = Denoiser(ctx, 2, H, W, seed=1)
    try:
        dn.set_params(p)
        out = dn(in10)
    finally:
        dn.close()
    _, _, pred = O.unet_grad(p, in10, in10[..., :3].copy(), emulate=True)
    ref = np.maximum(np.expm1(pred[..., :3].astype(np.float64)), 0.0)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    per = np.abs(out - ref) / np.maximum(ref, 1e-2)
    print(f"  inference vs oracle: relative L2 {rel:.2e}, 99th percentile per value {np.quantile(per, 0.99):.2e}")
    assert rel < 2e-2
    assert np.quantile(per, 0.99) < 5e-2
    # the network removes most of the 4-spp noise: its output is smoother
    assert np.std(np.diff(out, axis=2)) < np.std(np.diff(rgb, axis=2))
