"""Pins the oracle's FP64 hit tests and reference-order traversal
(CudaPathTracing/Hittable.cuh:27-140,367-398, BVH.cuh:282-332) against
textbook rays and brute force over all primitives (no GPU)."""
import numpy as np
import pytest

from cudapathtracing_amd import scenes
from oracle import oracle as O


def ray(o, d, tmin=0.001, tmax=999999.0):
    return np.array([[*o, tmin, *d, tmax]], np.float32)


def small_scene():
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), 2)
    s.add_light((0, 3, 0), 1.0, 1.0, (0, -1, 0), (15, 15, 15), True)
    s.add_light((5, 3, 0), 1.0, 1.0, (0, -1, 0), (15, 15, 15), False)  # invisible
    s.add_sphere((0, 0, -5), 1.0, m)
    s.add_mesh(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float64) + [3, 0, 0] * 3, m)
    return s


def test_textbook_hits():
    s = small_scene()
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    cases = [
        (ray((0, 0, 0), (0, 0, -1)), 2, 4.0),          # sphere front root
        (ray((0, 0, -5), (0, 0, -1)), 2, 1.0),         # from inside: far root
        (ray((3.25, 0.25, 1), (0, 0, -1)), 3, 1.0),    # triangle interior
        (ray((3.75, 0.75, 1), (0, 0, -1)), -1, None),  # u+v > 1
        (ray((0, 0, 0), (0, 1, 0)), 0, 3.0),           # light from below (one-sided)
        (ray((0, 5, 0), (0, -1, 0)), -1, None),        # light from above: back face
        (ray((5, 0, 0), (0, 1, 0)), -1, None),         # invisible light
        (ray((3.25, 0.25, 1), (0, 0, -1), tmax=0.5), -1, None),
        (ray((0, 0, 0), (0, 0, -2)), 2, 2.0),          # unnormalised direction
    ]
    for r, want_id, want_t in cases:
        t, ids, _ = o.trace_ref(r)
        tb, idb, _ = o.trace_brute(r)
        assert ids[0] == want_id and idb[0] == want_id
        if want_t is not None:
            assert t[0] == pytest.approx(want_t, rel=1e-12)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_reference_traversal_equals_brute_force(seed):
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), 6)
    s.add_light((0.5, 1.2, 0.5), 0.4, 0.4, (0, -1, 0), (15, 15, 15))
    s.add_sphere((0.2, 0.3, 0.8), 0.15, m)
    s.add_mesh(scenes.triangle_soup(3000, seed=seed, mu=-3.0), m)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    rays = scenes.random_rays(3000, seed=seed + 10)
    t, ids, visits = o.trace_ref(rays)
    tb, idb, cert = o.trace_brute(rays, eb=1e-6, et=1e-9)
    assert cert.mean() > 0.98
    assert np.array_equal(ids[cert], idb[cert])
    assert np.array_equal(t[cert], tb[cert])
    hit = ids >= 0
    assert 0.05 < hit.mean() < 0.95
    # the unculled reference traversal visits at least one node per ray
    assert visits.min() >= 1
