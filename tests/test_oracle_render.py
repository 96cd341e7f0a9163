"""Pins the oracle's RNG, materials and integrator (no GPU) against published
known answers and closed forms.

* Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11).
* Fresnel: normal-incidence dielectric ((eta-1)/(eta+1))^2, total internal
  reflection, normal-incidence conductor ((n-1)^2+k^2)/((n+1)^2+k^2)
  (Material.cuh:287-318) inside the reference's GGX conductor fr.
* GGX: D integrates to 1 over projected half vectors; the VNDF reflection
  pdf integrates to 2*pi in E[1/pdf] (support = upper hemisphere).
* Lambert: every cosine-weighted sample carries weight = albedo.
* Integrator: a Lambert floor under a uniform sky renders albedo*sky per
  sample exactly; a small light over a Lambert floor gives the inverse-square
  irradiance (NEE geometry and pdf of Render.cu:161-187).
"""
import math

import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import camera, scenes
from oracle import oracle as O


def test_philox_known_answers():
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def mat(mtype, color=(1, 1, 1), ax=0.5, ay=0.5, eta=1.51):
    return L.Material(L.d3(color), ax, ay, eta, mtype, 0)


def test_dielectric_fresnel_limits():
    m = mat(L.M_SPECULAR_DIELECTRIC)
    n = (0, 0, 1)
    R0 = ((1.51 - 1) / (1.51 + 1)) ** 2
    _, pdf, _ = O.bsdf_sample(m, (0, 0, -1), n, True, (0.5, 0.5, R0 * 0.5))
    assert pdf == pytest.approx(R0, rel=1e-12)  # reflected with probability R
    _, pdf, _ = O.bsdf_sample(m, (0, 0, -1), n, True, (0.5, 0.5, 0.9))
    assert pdf == pytest.approx(1 - R0, rel=1e-12)
    # inside the glass at 60 degrees: sin(60)*1.51 > 1 -> total internal reflection
    d = (math.sin(math.radians(60)), 0, -math.cos(math.radians(60)))
    out, pdf, f = O.bsdf_sample(m, d, n, False, (0.5, 0.5, 0.999))
    assert pdf == 1.0 and out[2] > 0
    # the sampled weight fr*|cos|/pdf of a smooth dielectric is the colour
    w = f * abs(out[2]) / pdf
    assert np.allclose(w, 1.0, rtol=1e-12)


def test_conductor_normal_incidence():
    ax, ay = 0.3, 0.5
    m = mat(L.M_OPAQUE, (0.8, 0.6, 0.4), ax, ay)
    nn, kk = 2.63, 9.08
    F = ((nn - 1) ** 2 + kk ** 2) / ((nn + 1) ** 2 + kk ** 2)
    f = O.bsdf_eval(m, (0, 0, -1), (0, 0, 1), True, (0, 0, 1))
    want = F / (math.pi * ax * ay) / (4.0 + 0.001) * np.array([0.8, 0.6, 0.4])
    assert np.allclose(f, want, rtol=1e-12)


@pytest.mark.parametrize("ax,ay", [(0.2, 0.2), (0.5, 0.5), (0.3, 0.7)])
def test_ggx_d_normalised(ax, ay):
    m = mat(L.M_OPAQUE, ax=ax, ay=ay)
    n = (0.0, 0.0, 1.0)
    # integrate D(h) cos(theta) over the hemisphere with u = sin^2(theta):
    # dw cos = (1/2) du dphi
    nu, nphi = 4000, 64
    u = (np.arange(nu) + 0.5) / nu
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    tot = 0.0
    for uu in u:
        st, ct = math.sqrt(uu), math.sqrt(1 - uu)
        for p in phi:
            tot += O.ggx_d(m, n, (st * math.cos(p), st * math.sin(p), ct))
    tot *= 0.5 * (1.0 / nu) * (2 * np.pi / nphi)
    assert tot == pytest.approx(1.0, rel=2e-3)


def test_vndf_pdf_integrates_over_hemisphere():
    m = mat(L.M_GGX, (1, 1, 1), 0.6, 0.6)
    rng = np.random.default_rng(0)
    d = np.array([0.5, 0.0, -math.sqrt(0.75)])
    s, nvalid, N = 0.0, 0, 100000
    for r in rng.random((N, 3)):
        out, pdf, _ = O.bsdf_sample(m, d, (0, 0, 1), True, r)
        if pdf > 0:
            s += 1.0 / pdf
            nvalid += 1
    assert nvalid > 0.6 * N
    assert s / N == pytest.approx(2 * math.pi, rel=0.05)


def test_lambert_weight_is_albedo():
    m = mat(L.M_LAMBERT, (0.3, 0.5, 0.7))
    rng = np.random.default_rng(1)
    cos = []
    for r in rng.random((2000, 3)):
        out, pdf, f = O.bsdf_sample(m, (0.3, 0.2, -1.0), (0, 0, 1), True, r)
        w = f * out[2] / pdf
        assert np.allclose(w, [0.3, 0.5, 0.7], rtol=1e-12)
        cos.append(out[2])
    assert np.mean(cos) == pytest.approx(2 / 3, abs=0.02)


def floor_scene(albedo, light=None):
    s = scenes.SceneData()
    m = s.add_material((albedo,) * 3, L.M_LAMBERT)
    if light is not None:
        s.add_light(*light)
    s.add_floor((-50, 0, -50), (50, 0, -50), (-50, 0, 50), (50, 0, 50), m)
    return s


def render_params(spp, profile=L.PROFILE_CONFIG, seed=7):
    return L.RenderParams(seed, spp, 8, 3, profile, 0.7, 0, 0, 0, 0)


def test_sky_over_lambert_floor_is_exact():
    s = floor_scene(0.6)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    d = L.CameraDesc(L.d3((0.5, 0.7, 0.9)), L.d3((0, 2, 0)), L.d3((0, 0, 0.001)), 30.0, 1.0, 16, 0)
    img = o.render(camera(d), render_params(4))
    assert np.allclose(img, 0.6 * np.array([0.5, 0.7, 0.9]), rtol=1e-12)


def test_small_light_inverse_square():
    a, Le, w, hgt = 0.5, 1000.0, 0.01, 1.0
    s = floor_scene(a, ((0, hgt, 0), w, w, (0, -1, 0), (Le,) * 3, False))
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    # camera looking straight down at the origin through a tiny field of view
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0, 0.5, 0.0)), L.d3((0, 0, 1e-6)), 0.01, 1.0, 4, 0)
    d.look_from[2] = 1e-3
    cam = camera(d)
    for profile in (L.PROFILE_CONFIG, L.PROFILE_REFERENCE):
        img = o.render(cam, render_params(16, profile))
        # radiance = a/pi * Le * A * cos*cos / r^2; the invisible light is
        # never hit directly, so NEE needs the visible flag for its shadow ray
        assert np.all(img == 0.0)
    s = floor_scene(a, ((0, hgt, 0), w, w, (0, -1, 0), (Le,) * 3, True))
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    img = o.render(cam, render_params(16, L.PROFILE_CONFIG))
    want = a / math.pi * Le * w * w / hgt ** 2
    assert np.allclose(img[..., 0], want, rtol=2e-3)
