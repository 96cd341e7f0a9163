"""Pins the FP64 oracle's LBVH restatement against closed forms and brute
force (no GPU).  Reference: CudaPathTracing/BVH.cuh:12-280, Device.cuh:134-195.
"""
import numpy as np
import pytest

from cudapathtracing_amd import scenes
from oracle import oracle as O


def interleave_loop(x, y, z):
    """Bit-by-bit 3D interleave (x -> bit 3b+2, y -> 3b+1, z -> 3b)."""
    c = 0
    for b in range(10):
        c |= ((x >> b) & 1) << (3 * b + 2)
        c |= ((y >> b) & 1) << (3 * b + 1)
        c |= ((z >> b) & 1) << (3 * b)
    return c


def test_expand_bits_closed_forms():
    assert O.expand_bits(0) == 0
    assert O.expand_bits(1) == 1
    assert O.expand_bits(2) == 8
    assert O.expand_bits(3) == 9
    assert O.expand_bits(1023) == 0x09249249
    assert O.expand_bits(0x200) == 1 << 27
    for v in range(0, 1024, 37):
        want = sum(((v >> b) & 1) << (3 * b) for b in range(10))
        assert O.expand_bits(v) == want


def test_morton_axes_corners_and_index():
    smin, smax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    assert O.morton3d((0, 0, 0), smin, smax, 0) == 0
    # n*1024 >= 1023 clamps to 1023 on every axis
    assert O.morton3d((1, 1, 1), smin, smax, 0) >> 32 == 0x3FFFFFFF
    assert O.morton3d((1, 0, 0), smin, smax, 0) >> 32 == 0x24924924
    assert O.morton3d((0, 1, 0), smin, smax, 0) >> 32 == 0x12492492
    assert O.morton3d((0, 0, 1), smin, smax, 0) >> 32 == 0x09249249
    # outside the box clamps to the faces; NaN maps to 0 (mMax(NaN,0) = 0)
    assert O.morton3d((-5, -5, -5), smin, smax, 7) == 7
    assert O.morton3d((float("nan"), 0, 0), smin, smax, 3) == 3
    k = O.morton3d((0.5, 0.25, 0.75), smin, smax, 123456)
    assert k & 0xFFFFFFFF == 123456
    assert k >> 32 == interleave_loop(512, 256, 768)
    rng = np.random.default_rng(5)
    for _ in range(200):
        p = rng.uniform(-0.2, 1.2, 3)
        q = [min(max(int(p[i] * 1024.0) if p[i] * 1024.0 > 0 else 0, 0), 1023) for i in range(3)]
        assert O.morton3d(p, smin, smax, 0) >> 32 == interleave_loop(*q)


def test_bounds_rule():
    a, b = O.expand_bounds((0, 0, 0), (1, 2, 3))
    assert a.tolist() == [-1, -1, -1] and b.tolist() == [2, 3, 4]
    # no mesh: Scene starts at [INF, -INF] -> fallback +-20 (x tested only)
    a, b = O.expand_bounds((999999.0,) * 3, (-999999.0,) * 3)
    assert a.tolist() == [-20] * 3 and b.tolist() == [20] * 3
    a, b = O.expand_bounds((5.0, 0, 0), (2.0, 1, 1))  # x inverted by more than 2
    assert a.tolist() == [-20] * 3


def clz64(x):
    return 64 - int(x).bit_length()


def topdown(keys):
    """Brute-force top-down split at the highest differing bit of the range."""
    n = len(keys)
    ca = np.zeros(max(n - 1, 0), np.int64)
    cb = np.zeros(max(n - 1, 0), np.int64)
    ranges = {}
    if n < 2:
        return ca, cb, ranges
    stack = [(0, 0, n - 1)]
    while stack:
        idx, f, l = stack.pop()
        ranges[idx] = (f, l)
        common = clz64(int(keys[f]) ^ int(keys[l]))
        split = f
        for i in range(f, l):
            if clz64(int(keys[f]) ^ int(keys[i])) > common:
                split = i
        ca[idx] = ~split if split == f else split
        cb[idx] = ~(split + 1) if split + 1 == l else split + 1
        if split != f:
            stack.append((split, f, split))
        if split + 1 != l:
            stack.append((split + 1, split + 1, l))
    return ca, cb, ranges


def soup_scene(n, seed, dup_every=0):
    rng = np.random.default_rng(seed)
    tri = scenes.triangle_soup(n, seed=seed, mu=-4.0)
    if dup_every:
        tri[::dup_every] = tri[0]  # identical centroids -> identical codes
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), 6)
    s.add_mesh(tri, m)
    return s


@pytest.mark.parametrize("n,seed,dup", [(2, 0, 0), (3, 1, 0), (17, 2, 0), (500, 3, 0), (700, 4, 3),
                                        (300, 5, 1)])
def test_karras_matches_topdown_and_refit_is_union(n, seed, dup):
    s = soup_scene(n, seed, dup)
    o = O.OracleScene(s)
    bmn, bmx = s.bounds_min, s.bounds_max
    b = o.build(bmn, bmx)
    keys = b["keys"]
    assert np.all(np.diff(keys.astype(object)) > 0)  # sorted, unique
    assert np.array_equal((keys & np.uint64(0xFFFFFFFF)).astype(np.uint32), b["order"])
    # keys equal the unsorted Morton keys permuted
    k0 = o.morton_keys(bmn, bmx)
    assert np.array_equal(np.sort(k0), keys)
    ca, cb, ranges = topdown(keys)
    assert np.array_equal(ca, b["child_a"]) and np.array_equal(cb, b["child_b"])
    # every node except the root has exactly one parent
    if n > 1:
        refs = np.concatenate([b["child_a"], b["child_b"]])
        internal = refs[refs >= 0]
        leaves = ~refs[refs < 0]
        assert sorted(internal.tolist()) == list(range(1, n - 1))
        assert sorted(leaves.tolist()) == list(range(n))
        assert b["max_depth"] <= 62
    # refit: child box = union of the primitive boxes of the child's range
    _, pbox = o.prims()
    for i in range(n - 1):
        for c, ref in enumerate((b["child_a"][i], b["child_b"][i])):
            f, l = (~ref, ~ref) if ref < 0 else ranges[int(ref)]
            ids = b["order"][f:l + 1]
            want = np.concatenate([pbox[ids, :3].min(0), pbox[ids, 3:].max(0)])
            assert np.array_equal(b["boxes"][i, 6 * c:6 * c + 6], want)


def test_primitive_boxes_and_centres():
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), 2)
    s.add_light((0, 2, 0), 1.0, 0.5, (0, -1, 0), (15, 15, 15))
    s.add_sphere((1, 2, 3), 0.5, m)
    s.add_floor((-1, 0, -1), (1, 0, -1), (-1, 0, 1), (1, 0, 1), m)
    o = O.OracleScene(s)
    c, b = o.prims()
    # light: +-w/2, +-h/2 around the centre in its frame, padded 1e-4
    assert np.allclose(c[0], (0, 2, 0))
    assert np.allclose(b[0], [-0.2501, 1.9999, -0.5001, 0.2501, 2.0001, 0.5001], atol=1e-12)
    # sphere: unpadded c +- r
    assert b[1].tolist() == [0.5, 1.5, 2.5, 1.5, 2.5, 3.5]
    # floor triangles (lt,lb,rt), (rt,lb,rb); centre ((p0+p1)+p2)/3
    assert np.allclose(c[2], ((-1 + -1 + 1) / 3, 0, (-1 + 1 + -1) / 3))
    assert np.allclose(b[2], [-1.0001, -0.0001, -1.0001, 1.0001, 0.0001, 1.0001])
