"""GPU LBVH parity: Morton keys, sorted order, topology bit-exact against the
FP64 oracle; FP32 node boxes equal the oracle's FP64 boxes rounded outward.
Reference: CudaPathTracing/Device.cuh:134-195, BVH.cuh:12-280."""
import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def scene3_like():
    """Scene3.ini without its file: light, 2 floors, 5 GGX spheres."""
    s = scenes.SceneData()
    s.add_light((0.0, 0.0, 10.0), 3.0, 3.0, (0.0, 0.0, -1.0), (1000 / 255,) * 3)
    m0 = s.add_material((230 / 255,) * 3, L.M_OPAQUE, 0.15, 0.15)
    m1 = s.add_material((1.0, 86 / 255, 32 / 255), L.M_OPAQUE)
    s.add_floor((-12.0, 0.0, -6.0), (12.0, 0.0, -6.0), (-12.0, 0.0, 10.0), (12.0, 0.0, 10.0), m0)
    s.add_floor((-12.0, 10.0, -6.0), (12.0, 10.0, -6.0), (-12.0, 0.0, -6.0), (12.0, 0.0, -6.0), m1)
    for c, a in [((-2, 2, 0), 0.01), ((2, 2, 0), 0.2), ((-1.8, 1, 2.3), 0.5), ((0.2, 1, -1.2), 0.1),
                 ((2.3, 1, 1.9), 0.7)]:
        s.add_sphere(c, 1.0, s.add_material((1, 1, 1), L.M_OPAQUE, a, a))
    return s


def soup(n, seed=1, dup=0):
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    t = scenes.triangle_soup(n, seed=seed)
    if dup:
        t[::dup] = t[0]
    s.add_mesh(t, m)
    return s


CASES = {
    "cornell": lambda: scenes.cornell_icosphere(),
    "scene3": scene3_like,
    "one": lambda: soup(1),
    "two": lambda: soup(2),
    "tile_edge": lambda: soup(4097, seed=3),
    "duplicates": lambda: soup(20000, seed=4, dup=2),
    "soup_1m": lambda: soup(1_000_000, seed=1),
    # above 2^22 leaves: node boxes finished by all three refit stages
    # (1024-leaf chunks, 64K- and 4M-leaf chunks, the rest across them)
    "soup_5m": lambda: soup(5_000_000, seed=5),
}


def round_outward(b64):
    b32 = b64.astype(np.float32)
    lo = np.where(b32.astype(np.float64) > b64, np.nextafter(b32, np.float32(-np.inf)), b32)
    hi = np.where(b32.astype(np.float64) < b64, np.nextafter(b32, np.float32(np.inf)), b32)
    return lo, hi


@pytest.mark.parametrize("name", list(CASES))
def test_build_bit_exact(ctx, name):
    print(f"[gpu] build parity case={name}", flush=True)
    s = CASES[name]()
    ctx.upload(s)
    o = O.OracleScene(s)
    k_gpu = ctx.morton(s.bounds_min, s.bounds_max)
    k_orc = o.morton_keys(s.bounds_min, s.bounds_max)
    assert np.array_equal(k_gpu, k_orc)
    st = ctx.build(s.bounds_min, s.bounds_max)
    assert st.n_prims == s.n_prims
    g = ctx.export()
    r = o.build(s.bounds_min, s.bounds_max)
    assert np.array_equal(g["keys"], r["keys"])
    assert np.array_equal(g["order"], r["order"])
    assert np.array_equal(g["child_a"], r["child_a"])
    assert np.array_equal(g["child_b"], r["child_b"])
    if s.n_prims > 1:
        lo, hi = round_outward(r["boxes"])
        for c in (0, 6):
            assert np.array_equal(g["boxes"][:, c:c + 3], lo[:, c:c + 3])
            assert np.array_equal(g["boxes"][:, c + 3:c + 6], hi[:, c + 3:c + 6])
    lo, hi = round_outward(r["root_box"])
    assert np.array_equal(g["root_box"][:3], lo[:3]) and np.array_equal(g["root_box"][3:], hi[3:])


def test_bounds_fallback_without_mesh(ctx):
    # no [Mesh]: bounds stay [INF,-INF] and the build uses +-20 (Device.cuh:177-181)
    s = scene3_like()
    assert s.bounds_min[0] > s.bounds_max[0]
    ctx.upload(s)
    o = O.OracleScene(s)
    assert np.array_equal(ctx.morton(s.bounds_min, s.bounds_max), o.morton_keys(s.bounds_min, s.bounds_max))


@pytest.mark.parametrize("n,bits", [(1, 32), (2, 8), (4095, 30), (4096, 32), (4097, 16),
                                    (100_000, 32), (3_000_001, 30)])
def test_radix_sort_stable(ctx, n, bits):
    print(f"[gpu] radix sort n={n} bits={bits}", flush=True)
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        keys[::3] = keys[0]  # many equal keys: stability matters
    vals = np.arange(n, dtype=np.uint32)[::-1].copy()
    k, v = keys.copy(), vals.copy()
    ctx.sort_pairs(k, v, bits)
    mask = np.uint32((1 << bits) - 1) if bits < 32 else np.uint32(0xFFFFFFFF)
    perm = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(v, vals[perm])
    assert np.array_equal(k, keys[perm])


def test_device_pointer_sort(ctx):
    import torch
    n = 50_000
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 2**30, n, dtype=np.int64).astype(np.uint32)
    kt = torch.from_numpy(keys.view(np.int32)).cuda()
    vt = torch.arange(n, dtype=torch.int32, device="cuda")
    ctx.sort_pairs(kt, vt, 30)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    perm = np.argsort(keys, kind="stable")
    assert np.array_equal(vt.cpu().numpy(), perm.astype(np.int32))


def leaf_ranges(ca, cb, n):
    """[first, last] leaf range of every internal node from the child arrays."""
    lo = np.zeros(n - 1, np.int64)
    hi = np.zeros(n - 1, np.int64)
    order = []
    st = [0]
    while st:
        i = st.pop()
        order.append(i)
        for c in (ca[i], cb[i]):
            if c >= 0:
                st.append(c)
    for i in reversed(order):
        r = []
        for c in (ca[i], cb[i]):
            r.append((~c, ~c) if c < 0 else (lo[c], hi[c]))
        lo[i] = min(r[0][0], r[1][0])
        hi[i] = max(r[0][1], r[1][1])
    return lo, hi


def test_dynamic_refit(ctx):
    print("[gpu] dynamic refit", flush=True)
    s = soup(30000, seed=8)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    g0 = ctx.export()
    rng = np.random.default_rng(9)
    moved = s.tri + rng.normal(0, 0.01, s.tri.shape)
    moved = moved.astype(np.float32).astype(np.float64)
    ctx.update_triangles(0, moved[:15000])
    ctx.update_triangles(15000, moved[15000:])
    st = ctx.refit()
    assert st.refit_ms > 0
    g = ctx.export()
    assert np.array_equal(g["child_a"], g0["child_a"]) and np.array_equal(g["order"], g0["order"])
    s2 = scenes.SceneData()
    s2.materials = s.materials
    s2.add_mesh(moved, 0)
    _, pbox = O.OracleScene(s2).prims()
    n = s2.n_prims
    lo, hi = leaf_ranges(g["child_a"], g["child_b"], n)
    order = g["order"]
    for i in rng.choice(n - 1, 400, replace=False):
        for c, ref in enumerate((g["child_a"][i], g["child_b"][i])):
            f, l = (~ref, ~ref) if ref < 0 else (lo[ref], hi[ref])
            ids = order[f:l + 1]
            want = np.concatenate([pbox[ids, :3].min(0), pbox[ids, 3:].max(0)])
            a, b = round_outward(want)
            assert np.array_equal(g["boxes"][i, 6 * c:6 * c + 3], a[:3])
            assert np.array_equal(g["boxes"][i, 6 * c + 3:6 * c + 6], b[3:])
    # traversal on the refitted tree still finds the brute-force closest hits
    rays = scenes.random_rays(3000, seed=12)
    t, ids = ctx.trace(rays)
    tb, idb, cert = O.OracleScene(s2).trace_brute(rays, eb=1e-4, et=1e-4)
    assert np.array_equal(ids[cert], idb[cert])
