"""GPU integrator parity against the FP64 oracle on the same random stream
(render, CudaPathTracing/Render.cu:16-261), path by path.

The GPU takes each decision of a path in FP32 and the oracle in FP64.  The
oracle certifies a path when every decision it takes (each closest hit,
Fresnel branch choice, sampled-direction side, frame threshold, NEE side,
free flight against a segment end, voxel boundary) is robust to that
rounding by a margin; both sides must then follow the same branches.  The
bar: EVERY certified path agrees within 1e-3 relative (absolute floor 1e-3);
the excluded fraction and the decision classes that excluded paths are
printed, and the certified fraction has a floor per scene.  Closed-form
scenes (sky over a Lambert floor, a glass furnace, an absorbing slab, a
w x h light) are checked on the GPU images directly.
"""
import math
import os

import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import camera, scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL = 1e-3
WHY = {1: "hit", 2: "nee-side", 3: "fresnel-choice", 4: "sample-side", 5: "frame", 6: "medium-distance",
       7: "voxel", 8: "volume-distance", 9: "face", 10: "btdf-side"}
GOLD = os.path.join(os.path.dirname(__file__), "golden", "scene_ini")


def params(spp, profile, seed=11, base=0, max_depth=8, rr_depth=3):
    return L.RenderParams(seed, spp, max_depth, rr_depth, profile, 0.7, base, 0, 0, 0)


def certified_match(g, r, cert, label, floor):
    """g, r: (h, w, 3) GPU and oracle pixels (means over the same samples);
    cert: (h*w, spp) certificates.  A pixel counts when all its samples are
    certified; every such pixel must agree within REL."""
    ok = (cert == 0).all(axis=1).reshape(g.shape[:2])
    err = np.abs(g.astype(np.float64) - r) / np.maximum(np.abs(r), 1e-3)
    good = np.all(err <= REL, axis=-1)
    bad = ok & ~good
    frac = float(ok.mean())
    why = np.bincount(cert[cert > 0].ravel(), minlength=11)
    reasons = {WHY[k]: int(why[k]) for k in range(1, 11) if why[k]}
    print(f"  {label}: certified {frac:.4f} of pixels, excluded {1 - frac:.4f} {reasons}; "
          f"agreement of the excluded {float(good[~ok].mean()) if (~ok).any() else 1.0:.3f}; "
          f"certified mismatches {int(bad.sum())}", flush=True)
    if bad.any():
        idx = np.argwhere(bad)[:5]
        for j, i in idx:
            print(f"    pixel ({i},{j}) gpu {g[j, i]} oracle {r[j, i]}", flush=True)
    assert not bad.any(), f"{label}: {int(bad.sum())} certified pixels differ"
    assert frac >= floor, f"{label}: certified fraction {frac:.4f} below {floor}"
    assert np.isfinite(g).all()
    return frac


def path_match(ctx, s, desc, profile, samples=(0, 1), floor=0.9, label="paths", build=True, **kw):
    if build:
        ctx.upload(s)
        ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    cam = camera(desc)
    fr = []
    for b in samples:
        g, _, st = ctx.render(desc, 1, seed=11, profile=profile, sample_base=b, **kw)
        r, cert = o.render(cam, params(1, profile, base=b, **{k: v for k, v in kw.items()
                                                               if k in ("max_depth", "rr_depth")}), cert=True)
        fr.append(certified_match(g, r, cert, f"{label} sample {b}", floor))
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


@pytest.mark.parametrize("width,quant,qsort,persist,pshadow", [(4, False, 0, 1, 1), (4, False, 1, 1, 1),
                                                               (4, True, 1, 1, 1), (4, False, 1, 0, 1),
                                                               (2, False, 0, 1, 1), (4, True, 0, 0, 1),
                                                               (8, False, 0, 1, 1), (4, True, 1, 1, 0),
                                                               (4, True, 0, 1, 0)])
def test_cornell_lambert_paths(ctx, width, quant, qsort, persist, pshadow):
    print(f"[gpu] render cornell paths (width {width}, quantized {quant}, queue sort {qsort}, "
          f"persistent {persist}, persistent shadows {pshadow})", flush=True)
    s = scenes.cornell_icosphere()
    d = scenes.default_camera(48, 48)
    ctx.set_traversal_width(width)
    ctx.set_quantized(quant)
    ctx.set_persistent(bool(persist))
    ctx.set_persistent_shadow(bool(pshadow))
    if qsort:  # bounce and shadow queues both traced in key order
        ctx.set_queue_sort_min(1)
        ctx.set_shadow_sort_min(1)
    try:
        path_match(ctx, s, d, L.PROFILE_CONFIG, samples=(0, 1, 2), floor=0.95, label="cornell")
    finally:
        ctx.set_traversal_width(4)
        ctx.set_quantized(True)
        ctx.set_persistent(True)
        ctx.set_persistent_shadow_min()
        ctx.set_queue_sort_min(0)
        ctx.set_shadow_sort_min(1 << 20)


@pytest.mark.parametrize("medium", [False, True])
def test_occlusion_kernels_agree_bit_for_bit(ctx, medium):
    """The persistent occlusion walk (append order) and the one-lane-per-ray
    kernel (append order, and origin-cell order) decide every shadow ray the
    same way and add each unoccluded contribution to its path's own slot: one
    light, so no atomic adds, and the images are equal bit for bit; so are
    repeats of the persistent render, whose walks pair rays with lanes by
    timing (closest-hit ties go to the lowest leaf whatever the test order).
    With the fBm medium the shadow rays are ratio-tracked."""
    print(f"[gpu] occlusion kernels bit for bit (medium {medium})", flush=True)
    if medium:
        s = scenes.displaced_icosphere_medium(64, 64, level=6, grid=32)
    else:
        s = scenes.displaced_icosphere(40, 40, level=5)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    kw = dict(seed=5, profile=L.PROFILE_CONFIG)
    try:
        ctx.set_persistent_shadow(True)
        a, _, sa = ctx.render(s.camera, 8, **kw)
        a2, _, _ = ctx.render(s.camera, 8, **kw)
        ctx.set_persistent_shadow(False)
        ctx.set_shadow_sort_min(0)
        b, _, _ = ctx.render(s.camera, 8, **kw)
        ctx.set_shadow_sort_min(1)
        c, _, _ = ctx.render(s.camera, 8, **kw)
    finally:
        ctx.set_persistent_shadow_min()
        ctx.set_shadow_sort_min(1 << 20)
        ctx.set_medium(None)
    assert sa.rays > 0 and a.mean() > 0
    assert np.array_equal(a, a2)
    assert np.array_equal(a, b) and np.array_equal(a, c)


@pytest.mark.parametrize("ppb", [37 * 29, 2 * 37 * 29 + 5, 500])
def test_odd_size_multi_batch(ctx, ppb):
    # 37 x 29 (not a multiple of the 8 x 8 path tiles) at 3 spp: one whole
    # sample per batch in tile order (ppb = w*h), whole samples spread over
    # the fewest batches (room for 2 samples: batches of 2 and 1), or batches
    # that split a sample with untiled paths (ppb = 500)
    print(f"[gpu] render 37x29 x 3 spp, {ppb} paths per batch", flush=True)
    s = scenes.cornell_icosphere()
    # height = (int)(37 / 1.27) = 29 (Camera.cu:12)
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0.0, 0.0, 1.9)), L.d3((0.0, 0.0, 0.0)), 40.0, 1.27, 37, 0)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    g, _, st = ctx.render(d, 3, seed=11, profile=L.PROFILE_CONFIG, paths_per_batch=ppb)
    assert g.shape == (29, 37, 3)
    r, cert = o.render(camera(d), params(3, L.PROFILE_CONFIG), cert=True)
    certified_match(g, r, cert, f"37x29 ppb {ppb}", 0.85)
    assert st.paths == 37 * 29 * 3


@pytest.mark.parametrize("mode", [1, 2])
def test_displaced_icosphere_cluster_paths(ctx, mode):
    # GGX patches on a displaced icosphere, whose clustered tree is chosen:
    # mode 1 builds it with the BVH, mode 2 at the first render
    print(f"[gpu] render displaced icosphere paths (cluster mode {mode})", flush=True)
    s = scenes.displaced_icosphere(32, 32, level=5)
    d = scenes.default_camera(40, 40)
    ctx.set_cluster(mode)
    try:
        path_match(ctx, s, d, L.PROFILE_CONFIG, floor=0.9, label=f"icosphere cluster {mode}")
        info = ctx.cluster_info()
    finally:
        ctx.set_cluster(2)
    print(f"  {info}")
    assert info["in_use"]


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
        path_match(ctx, s, d, L.PROFILE_CONFIG, floor=0.95, label="two lights")
    finally:
        ctx.set_shadow_sort_min(1 << 20)


def opaque_cornell():
    s = scenes.SceneData()
    rng = np.random.default_rng(0)
    scenes.cornell_box(s, rng, mtype=L.M_OPAQUE)
    m = s.add_material((0.9, 0.7, 0.5), L.M_OPAQUE, 0.3, 0.3)
    s.add_mesh(scenes.icosphere(2, (0.0, -0.2, 0.0), 0.25), m)
    return s


@pytest.mark.parametrize("profile", [L.PROFILE_REFERENCE, L.PROFILE_REAL_TIME])
def test_cornell_reference_profiles(ctx, profile):
    print(f"[gpu] render cornell profile {profile}", flush=True)
    path_match(ctx, opaque_cornell(), scenes.default_camera(40, 40), profile, floor=0.9,
               label=f"profile {profile}")


@pytest.mark.parametrize("profile", [L.PROFILE_NORMAL, L.PROFILE_DEPTH])
def test_normal_depth_display(ctx, profile):
    # NORMAL / DEPTH: one sample, the first vertex's normal or distance shown
    # as the reference shows it (Render.cu:239-251); spp is ignored
    print(f"[gpu] render display profile {profile}", flush=True)
    s = opaque_cornell()
    d = scenes.default_camera(40, 40)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    g, _, st = ctx.render(d, 16, seed=11, profile=profile)
    r, cert = o.render(camera(d), params(16, profile), cert=True)
    assert st.paths == 40 * 40
    ok = (cert[:, 0] == 0).reshape(40, 40)
    # the display depends on the first vertex only: its hit and orientation
    assert ok.mean() > 0.95
    assert np.allclose(g[ok], r[ok], rtol=1e-5, atol=2e-6)


def test_glass_furnace_reference_profile(ctx):
    # colour-1 glass under a uniform sky, STATIC roulette 0.7 at every vertex
    # and no depth cap: each path is 0 or sky * 0.7^-k; the mean is the sky
    print("[gpu] render glass furnace", flush=True)
    s = scenes.glass_column(32)
    d = s.camera
    path_match(ctx, s, d, L.PROFILE_REFERENCE, samples=(0, 1, 2, 3), floor=0.9, label="glass furnace")
    g, _, _ = ctx.render(d, 256, seed=5, profile=L.PROFILE_REFERENCE)
    sky = np.array([0.5, 0.7, 0.9])
    ratio = g / sky
    assert np.allclose(ratio, ratio[..., :1], rtol=1e-5)
    assert abs(ratio.mean() - 1.0) < 0.03


def test_nee_width_for_both(ctx):
    # w x h light with w < h: the reference's NEE samples w x w (area w^2)
    print("[gpu] render w x h light", flush=True)
    a, Le, hgt, w, h = 0.5, 1000.0, 1.0, 0.01, 0.02
    s = scenes.SceneData()
    m = s.add_material((a,) * 3, L.M_LAMBERT)
    s.add_light((0, hgt, 0), w, h, (0, -1, 0), (Le,) * 3, True)
    s.add_floor((-50, 0, -50), (50, 0, -50), (-50, 0, 50), (50, 0, 50), m)
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0, 0.5, 1e-3)), L.d3((0, 0, 1e-6)), 0.01, 1.0, 4, 0)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    unit = a / math.pi * Le / hgt ** 2
    for profile, area in ((L.PROFILE_REFERENCE, w * w), (L.PROFILE_CONFIG, w * h)):
        g, _, _ = ctx.render(d, 16, seed=7, profile=profile)
        assert np.allclose(g[..., 0], unit * area, rtol=2e-3), profile


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
    path_match(ctx, s, d, L.PROFILE_REFERENCE, samples=(0, 1, 2, 3), floor=0.85, label="scene1")


@pytest.mark.parametrize("name", ["Scene1.ini", "Scene2.ini", "Scene3.ini"])
def test_scene_ini_through_the_abi(ctx, name):
    """cpt_scene_load_ini (IniParser::Parse + Scene::init, IniParser.cpp:3-96,
    Scene.cuh:72-145) on the reference's own scene files: keys and topology
    bit-exact against the oracle built from the host parse, then per-path
    radiance in the STATIC profile.  The files' OBJ meshes and sky textures
    are not available offline: their meshes add no triangles."""
    print(f"[gpu] scene ini {name}", flush=True)
    path = os.path.join(GOLD, name)
    mn, mx, cam = ctx.load_ini(path)
    ctx.build(mn, mx)
    g = ctx.export()
    s = scenes.load_ini(path)
    assert np.array_equal(mn, s.bounds_min) and np.array_equal(mx, s.bounds_max)
    o = O.OracleScene(s)
    ref = o.build(mn, mx)
    for k in ("keys", "order", "child_a", "child_b"):
        assert np.array_equal(g[k], ref[k]), k
    cam.width = 48
    path_match(ctx, s, cam, L.PROFILE_REFERENCE, samples=(0, 1), floor=0.85, label=name, build=False)


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
    r, cert = o.render(camera(d), params(32, L.PROFILE_CONFIG, seed=5), cert=True)
    certified_match(g, r, cert, "cornell 32 spp", 0.3)
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
    try:
        path_match(ctx, s, d, L.PROFILE_CONFIG, floor=0.85, label="medium")
    finally:
        ctx.set_medium(None)


def test_absorbing_slab_closed_form(ctx):
    """A light behind a purely absorbing homogeneous slab: E[pixel] =
    Le exp(-sigma rho thickness sec(theta))."""
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
        r, cert = o.render(camera(d), params(1, L.PROFILE_CONFIG), pixel_offset=v * 40 * 40, cert=True)
        certified_match(g[v], r, cert, f"view {v}", 0.95)
    g1, _, _ = ctx.render(d0, 1, seed=11, profile=L.PROFILE_CONFIG)
    assert np.allclose(g1, g[0], rtol=1e-5, atol=1e-6)


def test_abort_flag(ctx):
    """cpt_request_abort: set before the call, a trace and a render return
    CPT_ERR_ABORTED (the persistent walks take no rays); after
    cpt_clear_abort the same calls give the normal results."""
    from cudapathtracing_amd import CptError
    print("[gpu] abort flag", flush=True)
    s = scenes.cornell_icosphere()
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    rays = scenes.random_rays(1 << 16, seed=3, lo=-0.45, hi=0.45)
    t0, i0 = ctx.trace(rays)
    ctx.request_abort()
    try:
        for call in (lambda: ctx.trace(rays), lambda: ctx.render(scenes.default_camera(16, 16), 2)):
            with pytest.raises(CptError) as e:
                call()
            assert e.value.code == L.CPT_ERR_ABORTED
    finally:
        ctx.clear_abort()
    t1, i1 = ctx.trace(rays)
    assert np.array_equal(i0, i1) and np.array_equal(t0, t1)
