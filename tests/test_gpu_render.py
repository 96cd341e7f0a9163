"""GPU integrator parity against the FP64 oracle on the same random stream
(render, CudaPathTracing/Render.cu:16-261).

With one sample per pixel a pixel is one path, so each image compares path
by path.  FP32 (GPU) and FP64 (oracle) take the same branch at almost every
decision; a path can still diverge where a ray grazes a geometric edge.  The
bar: >= 99% of paths within 1e-3 relative (absolute floor 1e-3), and a
closed-form scene (sky over a Lambert floor) exact to FP32 rounding.
"""
import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import camera, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-3


def params(spp, profile, seed=11, base=0, max_depth=8):
    return L.RenderParams(seed, spp, max_depth, 3, profile, 0.7, base, 0, 0, 0)


def path_match(ctx, s, desc, profile, samples=(0, 1)):
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    cam = camera(desc)
    fr = []
    for b in samples:
        g, _, st = ctx.render(desc, 1, seed=11, profile=profile, sample_base=b)
        r = o.render(cam, params(1, profile, base=b))
        err = np.abs(g.astype(np.float64) - r) / np.maximum(np.abs(r), 1e-3)
        ok = np.all(err <= REL, axis=-1)
        fr.append(ok.mean())
        assert np.isfinite(g).all()
    return float(np.mean(fr))


def test_sky_over_floor_exact(ctx):
    print("[gpu] render sky over floor", flush=True)
    s = scenes.SceneData()
    m = s.add_material((0.6, 0.6, 0.6), L.M_LAMBERT)
    s.add_floor((-50, 0, -50), (50, 0, -50), (-50, 0, 50), (50, 0, 50), m)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    d = L.CameraDesc(L.d3((0.5, 0.7, 0.9)), L.d3((0, 2, 0)), L.d3((0, 0, 0.001)), 30.0, 1.0, 32, 0)
    g, _, st = ctx.render(d, 4, profile=L.PROFILE_CONFIG)
    assert np.allclose(g, 0.6 * np.array([0.5, 0.7, 0.9]), rtol=2e-6)
    assert st.rays > 0 and st.bounces >= 2


@pytest.mark.parametrize("width,quant,qsort,persist", [(4, False, 0, 1), (4, False, 1, 1), (4, True, 1, 1),
                                                       (4, False, 1, 0), (2, False, 0, 1), (4, True, 0, 0),
                                                       (8, False, 0, 1)])
def test_cornell_lambert_paths(ctx, width, quant, qsort, persist):
    print(f"[gpu] render cornell paths (width {width}, quantized {quant}, queue sort {qsort}, "
          f"persistent {persist})", flush=True)
    s = scenes.cornell_icosphere()
    d = scenes.default_camera(48, 48)
    ctx.set_traversal_width(width)
    ctx.set_quantized(quant)
    ctx.set_persistent(bool(persist))
    if qsort:  # bounce and shadow queues both traced in key order
        ctx.set_queue_sort_min(1)
        ctx.set_shadow_sort_min(1)
    try:
        f = path_match(ctx, s, d, L.PROFILE_CONFIG, samples=(0, 1, 2))
    finally:
        ctx.set_traversal_width(4)
        ctx.set_quantized(True)
        ctx.set_persistent(True)
        ctx.set_queue_sort_min(0)
        ctx.set_shadow_sort_min(1 << 20)
    print(f"  cornell path agreement {f:.5f}")
    assert f >= 0.99


@pytest.mark.parametrize("mode", [1, 2])
def test_displaced_icosphere_cluster_paths(ctx, mode):
    # GGX patches on a displaced icosphere, whose clustered tree is chosen:
    # mode 1 builds it with the BVH, mode 2 at the first render
    print(f"[gpu] render displaced icosphere paths (cluster mode {mode})", flush=True)
    s = scenes.displaced_icosphere(32, 32, level=5)
    d = scenes.default_camera(40, 40)
    ctx.set_cluster(mode)
    try:
        f = path_match(ctx, s, d, L.PROFILE_CONFIG)
        info = ctx.cluster_info()
    finally:
        ctx.set_cluster(2)
    print(f"  {info}; path agreement {f:.5f}")
    assert info["in_use"]
    assert f >= 0.99


@pytest.mark.parametrize("ssort", [0, 1])
def test_cornell_two_lights(ctx, ssort):
    # two shadow rays per vertex: both add to the path radiance in one launch
    print(f"[gpu] render cornell two lights (shadow sort {ssort})", flush=True)
    s = scenes.cornell_icosphere()
    s.lights.append(L.Light(L.d3((0.3, 0.5 - 1e-3, -0.2)), 0.1, 0.15, L.d3((0.0, -1.0, 0.0)),
                            L.d3((6.0, 4.0, 2.0)), 1, 0))
    d = scenes.default_camera(40, 40)
    ctx.set_shadow_sort_min(1 if ssort else 0)
    try:
        f = path_match(ctx, s, d, L.PROFILE_CONFIG)
    finally:
        ctx.set_shadow_sort_min(1 << 20)
    print(f"  two-light path agreement {f:.5f}")
    assert f >= 0.99


def test_cornell_reference_profile(ctx):
    print("[gpu] render cornell reference profile", flush=True)
    s = scenes.SceneData()
    rng = np.random.default_rng(0)
    scenes.cornell_box(s, rng, mtype=L.M_OPAQUE)
    m = s.add_material((0.9, 0.7, 0.5), L.M_OPAQUE, 0.3, 0.3)
    s.add_mesh(scenes.icosphere(2, (0.0, -0.2, 0.0), 0.25), m)
    d = scenes.default_camera(40, 40)
    f = path_match(ctx, s, d, L.PROFILE_REFERENCE)
    print(f"  reference-profile path agreement {f:.5f}")
    assert f >= 0.98


def scene1_like():
    """Scene1.ini without files: glass and volume spheres, GGX spheres."""
    s = scenes.SceneData()
    s.add_light((0.0, 0.0, 10.0), 3.0, 3.0, (0.0, 0.0, -1.0), (1000 / 255,) * 3)
    m0 = s.add_material((230 / 255,) * 3, L.M_OPAQUE, 0.15, 0.15)
    m1 = s.add_material((1.0, 86 / 255, 32 / 255), L.M_OPAQUE)
    s.add_floor((-12.0, 0.0, -6.0), (12.0, 0.0, -6.0), (-12.0, 0.0, 10.0), (12.0, 0.0, 10.0), m0)
    s.add_floor((-12.0, 10.0, -6.0), (12.0, 10.0, -6.0), (-12.0, 0.0, -6.0), (12.0, 0.0, -6.0), m1)
    s.add_sphere((-2, 2, 0), 1.0, s.add_material((1, 1, 1), L.M_SPECULAR_DIELECTRIC, 0.2, 0.2))
    s.add_sphere((2, 2, 0), 1.0, s.add_material((1, 1, 1), L.M_VOLUME, 0.2, 0.2))
    s.add_sphere((-1.8, 1, 2.3), 1.0, s.add_material((1, 10 / 255, 10 / 255), L.M_OPAQUE, 0.5, 0.5))
    s.add_sphere((0.2, 1, -1.2), 1.0, s.add_material((10 / 255, 1, 10 / 255), L.M_ROUGH_DIELECTRIC, 0.1, 0.7))
    s.add_sphere((2.3, 1, 1.9), 1.0, s.add_material((10 / 255, 10 / 255, 1), L.M_OPAQUE, 0.7, 0.1))
    return s


def test_scene1_glass_volume_rough(ctx):
    print("[gpu] render scene1 (glass, volume, rough dielectric)", flush=True)
    s = scene1_like()
    d = L.CameraDesc(L.d3((0.8, 0.8, 0.8)), L.d3((0.0, 5.0, 10.0)), L.d3((0.0, 2.0, 0.0)), 45.0, 16 / 9, 64, 0)
    f = path_match(ctx, s, d, L.PROFILE_REFERENCE)
    print(f"  scene1 path agreement {f:.5f}")
    assert f >= 0.97


def test_converged_image_statistics(ctx):
    """Mean radiance of a 32-spp GPU render agrees with the oracle's."""
    print("[gpu] render cornell 32 spp statistics", flush=True)
    s = scenes.cornell_icosphere()
    d = scenes.default_camera(32, 32)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    g, aov, st = ctx.render(d, 32, seed=5, profile=L.PROFILE_CONFIG, aov=True)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    r = o.render(camera(d), params(32, L.PROFILE_CONFIG, seed=5))
    rel = np.abs(g - r) / np.maximum(r, 1e-3)
    assert np.mean(rel <= REL) >= 0.95
    assert abs(g.mean() / r.mean() - 1) < 1e-3
    # AOV (means over the pixel's samples): albedo in [0,1], positive depth;
    # normals have length <= 1 (the sky adds none, silhouettes average two),
    # and the mean normal of a pixel inside the box is close to unit length
    alb, nrm, dep = aov[..., :3], aov[..., 3:6], aov[..., 6]
    assert alb.min() >= 0 and alb.max() <= 1.0 + 1e-5
    assert np.all(dep > 0)
    nl = np.linalg.norm(nrm, axis=-1)
    assert nl.max() <= 1.0 + 1e-4
    assert np.median(nl[8:24, 8:24]) > 0.95
    assert st.rays >= 32 * 32 * 32


def test_medium_paths(ctx):
    """Heterogeneous fBm medium (delta tracking, ratio-tracked shadow rays)
    in the Cornell box: path-by-path agreement with the oracle."""
    print("[gpu] render cornell + fBm medium paths", flush=True)
    s = scenes.cornell_icosphere()
    s.medium = scenes.Medium(scenes.fbm_grid(32, 1), (-0.5, -0.5, -0.5), (0.5, 0.5, 0.5), 8.0, (0.8, 0.7, 0.6))
    d = scenes.default_camera(48, 48)
    f = path_match(ctx, s, d, L.PROFILE_CONFIG)
    print(f"  medium path agreement {f:.5f}")
    assert f >= 0.98
    ctx.set_medium(None)


def test_absorbing_slab_closed_form(ctx):
    """A light behind a purely absorbing homogeneous slab: E[pixel] =
    Le exp(-sigma rho thickness sec(theta))."""
    import math
    print("[gpu] render absorbing slab", flush=True)
    s = scenes.SceneData()
    s.add_light((0.0, 0.0, -1.0), 4.0, 4.0, (0.0, 0.0, 1.0), (2.0, 2.0, 2.0))
    s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.medium = scenes.Medium(np.full((1, 1, 1), 0.6, np.float32), (-3, -3, -0.25), (3, 3, 0.25), 2.0,
                             (0.0, 0.0, 0.0))
    ctx.upload(s)
    ctx.build(np.array([-3.0, -3, -1]), np.array([3.0, 3, 1]))
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0, 0, 3.0)), L.d3((0, 0, 0)), 30.0, 1.0, 64, 0)
    g, _, _ = ctx.render(d, 64, seed=17, profile=L.PROFILE_CONFIG, max_depth=4)
    cam = camera(d)
    h, w = cam.height, cam.width
    i, j = np.meshgrid(np.arange(w), np.arange(h))
    pc, dh, dv, lf = (np.array(getattr(cam, f)[:]) for f in ("pixel_center", "delta_h", "delta_v", "look_from"))
    dirs = pc[None, None] + i[..., None] * dh + j[..., None] * dv - lf
    sec = np.linalg.norm(dirs, axis=-1) / np.abs(dirs[..., 2])
    expect = 2.0 * np.exp(-2.0 * 0.6 * 0.5 * sec)
    p = math.exp(-0.6)
    se = math.sqrt(p * (1 - p) / (h * w * 64)) / p
    assert abs((g[..., 0] / expect).mean() - 1) < 5 * se
    ctx.set_medium(None)


def test_render_views_streams(ctx):
    """Two views in one call: each view path-by-path equal to the oracle with
    the view's pixel offset in the random stream."""
    print("[gpu] render two views", flush=True)
    s = scenes.cornell_icosphere()
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    d0 = scenes.default_camera(40, 40)
    d1 = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0.3, 0.2, 1.7)), L.d3((0.0, -0.05, 0.0)), 38.0, 1.0, 40, 0)
    g, _, st = ctx.render_views([d0, d1], 1, seed=11, profile=L.PROFILE_CONFIG)
    assert g.shape == (2, 40, 40, 3)
    for v, d in enumerate((d0, d1)):
        r = o.render(camera(d), params(1, L.PROFILE_CONFIG), pixel_offset=v * 40 * 40)
        err = np.abs(g[v].astype(np.float64) - r) / np.maximum(np.abs(r), 1e-3)
        f = np.all(err <= REL, axis=-1).mean()
        print(f"  view {v} path agreement {f:.5f}")
        assert f >= 0.99
    g1, _, _ = ctx.render(d0, 1, seed=11, profile=L.PROFILE_CONFIG)
    assert np.allclose(g1, g[0], rtol=1e-5, atol=1e-6)
