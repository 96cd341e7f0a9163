"""Pins the oracle's restatement of the reference behaviours that no
reference test or golden vector covers (CPU only):

* rough dielectric sampling and BTDF (CudaPathTracing/Material.cuh:45-64,
  118-162): the Fresnel branch choice, Snell's law with the reference's
  tangent sign, and identities in which D, F and the refraction Jacobian
  cancel between fr and the sampled pdf;
* the homogeneous sigma = 25 volume object walk (Render.cu:64-135): the
  unscattered and roulette-killed fractions, and the first free flight's
  truncated exponential law;
* the STATIC (REFERENCE) profile: roulette p = 0.7 at every vertex with no
  depth cap (a glass furnace), NEE over width x width (Render.cu:170-171);
* the REAL_TIME, NORMAL and DEPTH profiles (Render.cu:27-38, 189-204,
  237-251);
* the per-path certificate used by the GPU parity tests.

Every expected value below comes from closed forms derived from the
reference's formulas (or textbook definitions), not from the oracle's code.
"""
import math

import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import camera, scenes
from oracle import oracle as O

ETA = 1.51
R0 = ((ETA - 1) / (ETA + 1)) ** 2


def mat(mtype, color=(1, 1, 1), ax=0.5, ay=0.5, eta=ETA):
    return L.Material(L.d3(color), ax, ay, eta, mtype, 0)


def lam_textbook(alpha, w):
    """Smith Lambda of isotropic GGX (Heitz 2014): (sqrt(1 + a^2 tan^2) - 1) / 2."""
    c = min(max(w[2], 0.0), 1.0)
    t2 = (1.0 - c * c) / (c * c)
    return (math.sqrt(1.0 + alpha * alpha * t2) - 1.0) / 2.0


# ---- rough dielectric ------------------------------------------------------

def test_rough_dielectric_branch_choice_smooth_limit():
    # alpha -> 0 at normal incidence: the sampled microfacet is the normal,
    # so r3 < R0 reflects and r3 > R0 refracts straight through
    m = mat(L.M_ROUGH_DIELECTRIC, ax=1e-5, ay=1e-5)
    n = (0.0, 0.0, 1.0)
    out, pdf, _ = O.bsdf_sample(m, (0, 0, -1), n, True, (0.3, 0.6, R0 - 1e-4))
    assert pdf > 0 and out[2] > 0.999999
    out, pdf, _ = O.bsdf_sample(m, (0, 0, -1), n, True, (0.3, 0.6, R0 + 1e-4))
    assert pdf > 0 and np.allclose(out, (0, 0, -1), atol=1e-4)


@pytest.mark.parametrize("front,theta", [(True, 40.0), (True, 70.0), (False, 30.0)])
def test_rough_dielectric_snell_reference_sign(front, theta):
    # smooth limit: |tangential| = sin(theta) / eta_rel (Snell) but the
    # reference builds the tangent from the outgoing direction uo
    # (Material.cuh:148), so it keeps uo's sign: the mirror image, about the
    # plane of the normal, of the physical refraction
    m = mat(L.M_ROUGH_DIELECTRIC, ax=1e-5, ay=1e-5)
    th = math.radians(theta)
    uo = np.array([math.sin(th), 0.0, math.cos(th)])
    rel = ETA if front else 1.0 / ETA
    out, pdf, _ = O.bsdf_sample(m, -uo, (0, 0, 1), front, (0.3, 0.6, 0.999))
    st = math.sin(th) / rel
    assert pdf > 0
    assert np.allclose(out, (st, 0.0, -math.sqrt(1 - st * st)), atol=2e-4)


@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.7])
def test_rough_dielectric_transmission_weight_is_g2_over_g1(alpha):
    # fr * |cos| / pdf of a refracted sample: D, the Fresnel term (F = R, the
    # same dielectric Fresnel at the same microfacet) and the Jacobian
    # |Wi.Wm| / (|Wi.Wm| + Wo.Wm / eta)^2 cancel, leaving
    # colour * G2(V, -L) / G1(V) = colour (1 + Lv) / (1 + Lv + Ll)
    colour = np.array([0.9, 0.8, 0.7])
    m = mat(L.M_ROUGH_DIELECTRIC, tuple(colour), alpha, alpha)
    rng = np.random.default_rng(3)
    th = math.radians(35.0)
    ray = np.array([-math.sin(th), 0.3, -math.cos(th)])
    V = -ray / np.linalg.norm(ray)
    lv = lam_textbook(alpha, V)
    seen = 0
    for r in rng.random((400, 3)):
        r[2] = 0.5 + 0.5 * r[2]  # r3 > R for these angles: the refraction branch
        out, pdf, f = O.bsdf_sample(m, ray, (0, 0, 1), True, r)
        if not (pdf > 0) or out[2] >= 0:
            continue
        w = f * abs(out[2]) / pdf
        ll = lam_textbook(alpha, -out)
        assert np.allclose(w, colour * (1 + lv) / (1 + lv + ll), rtol=1e-9)
        seen += 1
    assert seen > 300
    # smooth limit: the weight is the colour
    m0 = mat(L.M_ROUGH_DIELECTRIC, tuple(colour), 1e-4, 1e-4)
    out, pdf, f = O.bsdf_sample(m0, ray, (0, 0, 1), True, (0.2, 0.7, 0.9))
    assert np.allclose(f * abs(out[2]) / pdf, colour, rtol=1e-5)


def test_rough_dielectric_reflection_weight_smooth_limit():
    # the reflection branch samples the GGX conductor lobe with the
    # reference's n = 1.58, k = 0.001 (Material.cuh:78-82) and divides its
    # pdf by R: at normal incidence, alpha -> 0, the weight tends to
    # F_conductor(1.58, 0.001) / R0 * 4 / (4 + 0.001) (the 0.001 of the
    # denominator 4 NV NL + 0.001)
    n, k = 1.58, 0.001
    fc = ((n - 1) ** 2 + k * k) / ((n + 1) ** 2 + k * k)
    m = mat(L.M_ROUGH_DIELECTRIC, (1, 1, 1), 1e-4, 1e-4)
    out, pdf, f = O.bsdf_sample(m, (0, 0, -1), (0, 0, 1), True, (0.2, 0.7, R0 * 0.5))
    assert out[2] > 0.999999
    assert f[0] * out[2] / pdf == pytest.approx(fc / R0 * 4.0 / 4.001, rel=1e-5)


# ---- sigma = 25 volume object -----------------------------------------------

def volume_sphere_scene(r):
    s = scenes.SceneData()
    m = s.add_material((1, 1, 1), L.M_VOLUME)
    s.add_sphere((0.0, 0.0, 0.0), r, m)
    o = O.OracleScene(s)
    o.build(np.array([-1.0, -1, -1]), np.array([1.0, 1, 1]))
    return o


@pytest.mark.parametrize("r", [0.01, 0.03])
def test_volume_walk_unscattered_and_killed_fractions(r):
    # a ray through the centre enters at z = -r; the first free flight
    # t ~ Exp(25) passes the chord 2r with probability exp(-50 r) (the path
    # continues from beyond the far side with throughput 1); otherwise the
    # first collision's roulette (0.8) kills the walk with probability 0.2,
    # leaving the throughput 1 and the ray at the collision point
    o = volume_sphere_scene(r)
    n = 200_000
    ro, rd, thr, nc = o.volume_walk((0, 0, -5.0), (0, 0, 1.0), 5.0 - r, seed=7, n=n)
    beyond = (nc == 0) & (ro[:, 2] >= r - 1e-12)
    killed = (nc == 0) & (ro[:, 2] < r)
    p_free = math.exp(-50.0 * r)
    for frac, p in ((beyond.mean(), p_free), (killed.mean(), 0.2 * (1 - p_free)),
                    ((nc >= 1).mean(), 0.8 * (1 - p_free))):
        assert abs(frac - p) < 5 * math.sqrt(p * (1 - p) / n)
    assert np.all(thr[nc == 0] == 1.0)
    assert np.all(rd[nc == 0] == np.array([0, 0, 1.0]))


def test_volume_walk_first_flight_is_truncated_exponential():
    # killed walks stop at the first collision: its depth into the sphere
    # follows Exp(25) truncated to the chord [0, 2r)
    r = 0.03
    o = volume_sphere_scene(r)
    ro, _, _, nc = o.volume_walk((0, 0, -5.0), (0, 0, 1.0), 5.0 - r, seed=11, n=200_000)
    depth = ro[(nc == 0) & (ro[:, 2] < r), 2] + r
    c = 2 * r
    cdf = lambda t: (1 - np.exp(-25 * t)) / (1 - math.exp(-25 * c))  # noqa: E731
    x = np.sort(depth)
    ks = np.max(np.abs(cdf(x) - np.arange(1, x.size + 1) / x.size))
    assert ks < 1.63 / math.sqrt(x.size)  # KS at the 1% level
    # the weight of a survived collision is exp(-25 t) / 0.8 with t the depth
    # of that collision on the entry ray
    ro, _, thr, nc = o.volume_walk((0, 0, -5.0), (0, 0, 1.0), 5.0 - r, seed=12, n=20_000)
    one = nc == 1
    assert one.sum() > 1000
    assert np.allclose(thr[one], np.exp(-25.0 * (ro[one, 2] + r)) / 0.8, rtol=1e-12)


# ---- REFERENCE profile: roulette at every vertex, no depth cap ---------------

SKY = np.array([0.5, 0.7, 0.9])


def glass_column(width=24):
    s = scenes.glass_column(width)
    return s, s.camera


def per_sample(s, d, profile, spp, max_depth=8, rr_depth=3, seed=5):
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    prm = L.RenderParams(seed, spp, max_depth, rr_depth, profile, 0.7, 0, 0, 0, 0)
    _, ps = o.render(camera(d), prm, per_sample=True)
    return ps.reshape(-1, 3)


def test_glass_furnace_reference_profile():
    # colour-1 glass: each vertex weighs F/R or (1-F)/(1-R) = 1, so a sample
    # is 0 (killed) or sky * 0.7^-k after k vertices; E = sky at every pixel
    # (roulette is unbiased and every path escapes or is killed)
    s, d = glass_column()
    ps = per_sample(s, d, L.PROFILE_REFERENCE, 16)
    ratio = ps / SKY
    assert np.allclose(ratio, ratio[:, :1], rtol=1e-12, atol=0)  # grey scaling
    r = ratio[:, 0]
    nz = r > 0
    k = np.log(r[nz]) / math.log(1 / 0.7)
    assert np.allclose(k, np.round(k), atol=1e-6)
    # no depth cap: paths with more than 8 vertices contribute
    assert np.round(k).max() >= 9
    se = r.std() / math.sqrt(r.size)
    assert abs(r.mean() - 1.0) < 4 * se
    assert ((k >= 1)).mean() > 0.2  # most light goes through the glass


def test_glass_column_capped_profiles():
    s, d = glass_column()
    # CONFIG without roulette: every sample is 0 or exactly the sky, and the
    # 8-vertex cap loses the long paths
    cfg = per_sample(s, d, L.PROFILE_CONFIG, 8, max_depth=8, rr_depth=1000)[:, 0] / SKY[0]
    assert np.all((cfg == 0) | np.isclose(cfg, 1.0, rtol=1e-12))
    # REAL_TIME: no roulette, at most 3 vertices (Render.cu:191-194)
    rt = per_sample(s, d, L.PROFILE_REAL_TIME, 8)[:, 0] / SKY[0]
    assert np.all((rt == 0) | np.isclose(rt, 1.0, rtol=1e-12))
    assert 0.2 < rt.mean() < cfg.mean() < 1.0


def floor_scene(albedo, light=None):
    s = scenes.SceneData()
    m = s.add_material((albedo,) * 3, L.M_LAMBERT)
    if light is not None:
        s.add_light(*light)
    s.add_floor((-50, 0, -50), (50, 0, -50), (-50, 0, 50), (50, 0, 50), m)
    return s


def down_camera(width=4):
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0, 0.5, 1e-3)), L.d3((0, 0, 1e-6)), 0.01, 1.0, width, 0)
    return d


@pytest.mark.parametrize("profile", [L.PROFILE_REFERENCE, L.PROFILE_REAL_TIME])
def test_nee_width_for_both_extents(profile):
    # a w x h light 1 above a Lambert floor (albedo a): the reference samples
    # its NEE points over w x w with pdf 1/w^2 (Render.cu:170-171)
    a, Le, hgt = 0.5, 1000.0, 1.0
    unit = a / math.pi * Le / hgt ** 2
    w, h = 0.01, 0.02  # every w x w point lies on the light: area w^2
    s = floor_scene(a, ((0, hgt, 0), w, h, (0, -1, 0), (Le,) * 3, True))
    for prof, area in ((profile, w * w), (L.PROFILE_CONFIG, w * h)):
        ps = per_sample(s, down_camera(), prof, 16)
        assert np.allclose(ps[:, 0], unit * area, rtol=2e-3)
    # w > h: points beyond h/2 miss the light (the shadow ray's closest hit
    # is not the light): samples are 0 or unit * w^2, half of them 0
    w, h = 0.02, 0.01
    s = floor_scene(a, ((0, hgt, 0), w, h, (0, -1, 0), (Le,) * 3, True))
    ps = per_sample(s, down_camera(), profile, 256)[:, 0]
    hit = ps > 0
    assert np.allclose(ps[hit], unit * w * w, rtol=2e-3)
    se = math.sqrt(0.25 / ps.size)
    assert abs(hit.mean() - h / w) < 5 * se


def test_normal_and_depth_display():
    # NORMAL / DEPTH (Render.cu:239-251): one sample; the first vertex's
    # face-forwarded normal shown as (n + 0.5) / 2, its distance as
    # clamp01(d / 50); pixels with no vertex show the cleared G-buffer
    s = floor_scene(0.5)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    d = L.CameraDesc(L.d3((0.1, 0.2, 0.3)), L.d3((0, 2.0, 0)), L.d3((0, 1.5, 4.0)), 60.0, 1.0, 32, 0)
    cam = camera(d)
    prm = lambda p: L.RenderParams(3, 4, 8, 3, p, 0.7, 0, 0, 0, 0)  # noqa: E731
    nrm = o.render(cam, prm(L.PROFILE_NORMAL))
    dep = o.render(cam, prm(L.PROFILE_DEPTH))
    h, w = cam.height, cam.width
    i, j = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    pc, dh, dv, lf = (np.array(getattr(cam, f)[:]) for f in ("pixel_center", "delta_h", "delta_v", "look_from"))
    dirs = pc + (i[..., None] - 0.5) * dh + (j[..., None] - 0.5) * dv - lf
    down = dirs[..., 1] < -0.02 * np.linalg.norm(dirs, axis=-1)  # clearly hits the floor
    up = dirs[..., 1] > 0.02 * np.linalg.norm(dirs, axis=-1)     # clearly the sky
    tfloor = -lf[1] / dirs[..., 1]
    dist = tfloor * np.linalg.norm(dirs, axis=-1)
    near = down & (dist < 30)  # well inside the 100 x 100 floor
    assert near.sum() > 100 and up.sum() > 100
    assert np.allclose(nrm[near], (0.25, 0.75, 0.25), atol=1e-12)
    assert np.allclose(nrm[up], 0.25, atol=0) and np.allclose(dep[up], 1.0, atol=0)
    # the jittered sample lies inside its pixel: its distance lies between
    # the distances through the pixel's corners
    corners = []
    for di in (-0.5, 0.5):
        for dj in (-0.5, 0.5):
            dc = pc + (i[..., None] - 0.5 + di) * dh + (j[..., None] - 0.5 + dj) * dv - lf
            corners.append(-lf[1] / dc[..., 1] * np.linalg.norm(dc, axis=-1))
    lo, hi = np.min(corners, axis=0), np.max(corners, axis=0)
    close = down & (hi < 20) & (lo > 0)
    assert close.sum() > 50
    got = dep[close, 0] * 50.0
    assert np.all(got >= lo[close] * (1 - 1e-9)) and np.all(got <= hi[close] * (1 + 1e-9))


# ---- certificate -------------------------------------------------------------

def test_certificate_excludes_ties_only():
    # Cornell box: almost every 1-spp path is certified
    s = scenes.cornell_icosphere()
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    cam = camera(scenes.default_camera(32, 32))
    prm = L.RenderParams(11, 1, 8, 3, L.PROFILE_CONFIG, 0.7, 0, 0, 0, 0)
    _, cert = o.render(cam, prm, cert=True)
    assert (cert == 0).mean() > 0.9
    # the same floor twice: every path that sees the floor meets an exact
    # tie between two primitives and is excluded as a hit decision
    s = floor_scene(0.5)
    s.add_floor((-50, 0, -50), (50, 0, -50), (-50, 0, 50), (50, 0, 50), 0)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    d = L.CameraDesc(L.d3((0.1, 0.2, 0.3)), L.d3((0, 2.0, 0)), L.d3((0, 0.0, 4.0)), 60.0, 1.0, 16, 0)
    _, cert = o.render(camera(d), prm, cert=True)
    assert np.any(cert == 1) and np.any(cert == 0)
    assert set(np.unique(cert)) <= {0, 1}
