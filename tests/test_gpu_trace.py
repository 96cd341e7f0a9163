"""GPU closest-hit parity (traverseIterative, CudaPathTracing/BVH.cuh:282-332).

* small scenes: brute-force FP64 oracle; on certified rays (every primitive's
  accept/reject decision robust by a margin of 1e-4, runner-up hit farther
  by more than 1e-4 relative) the hit ids are bit-exact, and the distance,
  recomputed in FP64 for the winning primitive, is within 1e-6 relative.
* large scenes: the oracle's reference-order traversal; ids equal except
  near-ties, whose FP64 distances then agree to 1e-6.
"""
import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import scenes
from oracle import oracle as O

pytestmark = pytest.mark.gpu

T_REL = 1e-6


def camera_rays(desc, w, h, jitter_seed=0):
    from cudapathtracing_amd import camera
    cam = camera(desc)
    rng = np.random.default_rng(jitter_seed)
    i, j = np.meshgrid(np.arange(cam.width), np.arange(cam.height))
    i, j = i.ravel(), j.ravel()
    rx, ry = rng.random(i.size), rng.random(i.size)
    pc, dh, dv, lf = (np.array(getattr(cam, f)[:]) for f in ("pixel_center", "delta_h", "delta_v", "look_from"))
    pix = pc + (rx - 0.5 + i)[:, None] * dh + (ry - 0.5 + j)[:, None] * dv
    d = pix - lf
    r = np.zeros((i.size, 8), np.float32)
    r[:, :3] = lf
    r[:, 3] = 0.001
    r[:, 4:7] = d
    r[:, 7] = 999999.0
    return r


def box_rays(n, seed):
    r = scenes.random_rays(n, seed=seed, lo=-0.45, hi=0.45)
    return r


def check_small(ctx, s, rays):
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    o = O.OracleScene(s)
    tb, idb, cert = o.trace_brute(rays, eb=1e-4, et=1e-4)
    print(f"  certified {cert.mean():.5f} of {len(rays)} rays, hits {np.mean(idb >= 0):.3f}")
    assert cert.mean() > 0.95
    # binary nodes; four-wide FP32 and quantized records, one lane per ray
    # and persistent, over the LBVH's boxes and over the clustered tree's
    for width, quant, persistent, cluster in ((2, False, False, False), (4, False, False, False),
                                              (4, True, False, False), (4, True, True, False),
                                              (4, False, True, True), (4, True, True, True),
                                              (4, True, False, True), (8, False, False, False)):
        ctx.set_traversal_width(width)
        ctx.set_quantized(quant)
        ctx.set_persistent(persistent)
        ctx.set_cluster(cluster)
        cfg = (width, quant, persistent, cluster)
        t, ids = ctx.trace(rays)
        if cluster:
            print(f"  cluster {ctx.cluster_info()}", flush=True)
        bad = np.where(cert & (ids != idb))[0]
        assert bad.size == 0, (cfg, bad[:10], ids[bad[:10]], idb[bad[:10]])
        h = cert & (idb >= 0)
        if h.any():
            rel = np.abs(t[h].astype(np.float64) - tb[h]) / np.abs(tb[h])
            assert rel.max() <= T_REL, (cfg, rel.max())
        miss = cert & (idb < 0)
        assert np.all(t[miss] == rays[miss, 7])
    ctx.set_traversal_width(4)
    ctx.set_quantized(True)
    ctx.set_persistent(True)
    ctx.set_cluster(2)
    return cert.mean()


def test_cornell_camera_and_interior_rays(ctx):
    print("[gpu] trace cornell", flush=True)
    s = scenes.cornell_icosphere()
    rays = np.concatenate([camera_rays(scenes.default_camera(128, 128), 128, 128), box_rays(20000, 3)])
    check_small(ctx, s, rays)


def test_scene3_spheres_floors_light(ctx):
    from test_gpu_bvh import scene3_like
    print("[gpu] trace scene3", flush=True)
    s = scene3_like()
    d = L.CameraDesc(L.d3((0.8, 0.8, 0.8)), L.d3((0.0, 5.0, 10.0)), L.d3((0.0, 2.0, 0.0)), 45.0, 16 / 9, 160, 0)
    rays = camera_rays(d, 160, 90)
    # rays towards the light from the floor
    rng = np.random.default_rng(1)
    n = 5000
    o = np.stack([rng.uniform(-5, 5, n), np.zeros(n) + 0.01, rng.uniform(-5, 8, n)], 1)
    tgt = np.stack([rng.uniform(-1.5, 1.5, n), rng.uniform(-1.5, 1.5, n), np.full(n, 10.0)], 1)
    r2 = np.zeros((n, 8), np.float32)
    r2[:, :3], r2[:, 3], r2[:, 4:7], r2[:, 7] = o, 0.001, tgt - o, 999999.0
    check_small(ctx, s, np.concatenate([rays, r2]))


def test_soup_brute_force(ctx):
    print("[gpu] trace soup 20k brute force", flush=True)
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.add_mesh(scenes.triangle_soup(20000, seed=5, mu=-4.0), m)
    rays = scenes.random_rays(20000, seed=6)
    check_small(ctx, s, rays)


def test_axis_aligned_and_degenerate_rays(ctx):
    print("[gpu] trace axis-aligned", flush=True)
    s = scenes.cornell_icosphere()
    rng = np.random.default_rng(2)
    n = 6000
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(-0.45, 0.45, (n, 3))
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    r[np.arange(n), 4 + axis] = sign
    r[:, 3], r[:, 7] = 0.001, 999999.0
    r[:100, :3] = 0.0  # origin exactly at 0 with zero direction components
    r[100:200, 7] = 0.01  # short tmax
    check_small(ctx, s, r)


def test_empty_and_tiny_scenes(ctx):
    print("[gpu] trace empty/tiny", flush=True)
    s = scenes.SceneData()
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    rays = scenes.random_rays(100, seed=1)
    t, ids = ctx.trace(rays)
    assert np.all(ids == -1) and np.all(t == rays[:, 7])
    for n in (1, 2, 3):
        s = scenes.SceneData()
        m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
        s.add_mesh(scenes.triangle_soup(n, seed=9, mu=-1.0), m)
        tgt = s.tri.reshape(-1, 3, 3).mean(1)
        o = np.array([[-1.0, -1.0, -1.0]])
        r = np.zeros((n, 8), np.float32)
        r[:, :3], r[:, 3], r[:, 4:7], r[:, 7] = o, 0.001, tgt - o, 999999.0
        check_small(ctx, s, r)


def test_soup_1m_reference_traversal(ctx):
    print("[gpu] trace soup 1M vs reference-order traversal", flush=True)
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.add_mesh(scenes.triangle_soup(1_000_000, seed=1, mu=-5.0), m)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    rays = scenes.random_rays(200_000, seed=2)
    t, ids = ctx.trace(rays)
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    tr, idr, _ = o.trace_ref(rays)
    same = ids == idr
    print(f"  id agreement {same.mean():.6f}, hits {np.mean(idr >= 0):.3f}")
    assert same.mean() > 0.999
    h = same & (idr >= 0)
    rel = np.abs(t[h].astype(np.float64) - tr[h]) / np.abs(tr[h])
    assert rel.max() <= T_REL
    # mismatches must be near-ties: the FP64 distance to the GPU's primitive
    # equals the oracle's closest distance within 1e-6
    mm = np.where(~same)[0]
    if mm.size:
        tg = o.prim_t(rays[mm], ids[mm])
        ok = (ids[mm] >= 0) & (idr[mm] >= 0) & (np.abs(tg - tr[mm]) <= 1e-6 * np.abs(tr[mm]))
        assert ok.all(), (mm[~ok][:5], ids[mm][~ok][:5], idr[mm][~ok][:5])


def test_cluster_tree_icosphere(ctx):
    """The clustered tree over a displaced icosphere (cheaper than the LBVH,
    so the records group it): hits equal the oracle's reference-order LBVH
    traversal except near-ties, with and without it, before and after a
    refit."""
    print("[gpu] trace displaced icosphere, clustered tree vs LBVH records", flush=True)
    s = scenes.displaced_icosphere(64, 64, level=6)
    ctx.upload(s)
    ctx.set_cluster(1)
    st = ctx.build(s.bounds_min, s.bounds_max)
    info = ctx.cluster_info()
    print(f"  {info}, cluster stage {st.cluster_ms:.3f} ms", flush=True)
    assert info["in_use"] and 0 < info["depth"] <= 64
    assert info["cluster_cost"] < info["lbvh_cost"]
    rng = np.random.default_rng(5)
    rays = scenes.random_rays(100_000, seed=7, lo=-0.45, hi=0.45)
    cam = camera_rays(scenes.default_camera(96, 96), 96, 96, jitter_seed=3)
    rays = np.concatenate([rays, cam[rng.permutation(len(cam))[:20_000]]])
    o = O.OracleScene(s)
    o.build(s.bounds_min, s.bounds_max)
    tr, idr, _ = o.trace_ref(rays)
    for stage in ("build", "refit"):
        if stage == "refit":
            st = ctx.refit()
            assert ctx.cluster_info()["in_use"] and st.cluster_ms > 0
        for cluster in (True, False):
            ctx.set_cluster(cluster)
            t, ids = ctx.trace(rays)
            same = ids == idr
            print(f"  {stage} cluster={cluster}: id agreement {same.mean():.6f}", flush=True)
            assert same.mean() > 0.999
            h = same & (idr >= 0)
            rel = np.abs(t[h].astype(np.float64) - tr[h]) / np.abs(tr[h])
            assert rel.max() <= T_REL
            mm = np.where(~same)[0]
            if mm.size:
                tg = o.prim_t(rays[mm], ids[mm])
                ok = (ids[mm] >= 0) & (idr[mm] >= 0) & (np.abs(tg - tr[mm]) <= 1e-6 * np.abs(tr[mm]))
                assert ok.all(), (stage, cluster, mm[~ok][:5])
        ctx.set_cluster(1)
    ctx.set_cluster(2)


def test_cluster_tree_coincident_primitives(ctx):
    """20k copies of one triangle beside a random soup: every union area of
    the copies ties, and the clustering pairs them off in about log2(n)
    rounds (one pair per round without the buddy rule: seconds).  Hits over
    either tree agree, up to which copy wins a tie."""
    print("[gpu] clustered tree over coincident primitives", flush=True)
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    one = np.array([[0.2, 0.2, 0.5, 0.8, 0.3, 0.5, 0.4, 0.9, 0.5]])
    s.add_mesh(np.repeat(one, 20_000, axis=0), m)
    s.add_mesh(scenes.triangle_soup(5_000, seed=6, mu=-3.5), m)
    ctx.upload(s)
    ctx.set_cluster(1)
    try:
        st = ctx.build(s.bounds_min, s.bounds_max)
        info = ctx.cluster_info()
        print(f"  cluster stage {st.cluster_ms:.2f} ms, {info}", flush=True)
        assert st.cluster_ms < 250 and info["depth"] <= 64
        rays = scenes.random_rays(50_000, seed=9, lo=-0.2, hi=1.2)
        t1, i1 = ctx.trace(rays)
        ctx.set_cluster(0)
        t0, i0 = ctx.trace(rays)
    finally:
        ctx.set_cluster(2)
    copies = lambda ids: (ids >= 0) & (ids < 20_000)
    same = (i1 == i0) | (copies(i1) & copies(i0))
    assert same.all(), np.where(~same)[0][:5]
    assert np.array_equal(t1, t0)
    assert copies(i0).mean() > 0.01


def test_exact_ties_resolve_to_one_primitive(ctx):
    """Three exact copies of every triangle of a soup: each hit on the soup is
    a three-way tie at one distance, which every walk resolves to the copy in
    the lowest leaf, so binary, four-wide (FP32 and quantized, one lane per
    ray and persistent) and eight-wide walks over one tree return the same
    ids, and so does every repeat of the persistent walks, whose queued leaf
    tests run in an order set by the other lanes of the wave."""
    print("[gpu] trace exact ties", flush=True)
    n = 30_000
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    soup = scenes.triangle_soup(n, seed=12, mu=-3.5)
    s.add_mesh(np.concatenate([soup, soup, soup]), m)
    ctx.upload(s)
    ctx.set_cluster(0)
    rays = scenes.random_rays(200_000, seed=13, lo=-0.2, hi=1.2)
    ref = None
    try:
        ctx.build(s.bounds_min, s.bounds_max)
        for width, quant, persistent in ((2, False, False), (4, False, False), (4, True, False),
                                         (4, True, True), (4, False, True), (8, False, False)):
            ctx.set_traversal_width(width)
            ctx.set_quantized(quant)
            ctx.set_persistent(persistent)
            for _ in range(3 if persistent else 1):
                t, ids = ctx.trace(rays)
                if ref is None:
                    ref = (t, ids)
                    continue
                cfg = (width, quant, persistent)
                assert np.array_equal(ids, ref[1]), (cfg, np.where(ids != ref[1])[0][:5])
                assert np.array_equal(t, ref[0]), cfg
    finally:
        ctx.set_traversal_width(4)
        ctx.set_quantized(True)
        ctx.set_persistent(True)
        ctx.set_cluster(2)
    assert np.mean(ref[1] >= 0) > 0.2


def test_device_pointer_trace(ctx):
    import torch
    print("[gpu] trace with device pointers", flush=True)
    s = scenes.cornell_icosphere()
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    rays = box_rays(10000, 11)
    t0, i0 = ctx.trace(rays)
    rt = torch.from_numpy(rays).cuda()
    t1, i1 = ctx.trace(rt)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    assert np.array_equal(i1.cpu().numpy(), i0) and np.array_equal(t1.cpu().numpy(), t0)
    nodes, prims = ctx.trace_count(rays)
    assert nodes.min() >= 1 and prims.max() >= 1


def test_ray_ordering_keeps_results(ctx):
    """Traced in ray-key order, every ray's (t, id) lands at its own index and
    equals the unsorted trace."""
    print("[gpu] trace ray ordering", flush=True)
    s = scenes.SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.add_mesh(scenes.triangle_soup(200_000, seed=3, mu=-4.5), m)
    ctx.upload(s)
    ctx.build(s.bounds_min, s.bounds_max)
    rays = scenes.random_rays(300_001, seed=4)
    ctx.set_ray_sort_min(0)
    ctx.set_persistent(False)
    t0, i0 = ctx.trace(rays)
    n0, p0 = ctx.trace_count(rays)
    ctx.set_persistent(True)
    for smin in (0, 1):
        ctx.set_ray_sort_min(smin)
        t1, i1 = ctx.trace(rays)
        tot, srt, ker = ctx.last_trace_stats()
        assert (srt > 0) == (smin == 1) and tot >= ker
        n1, p1 = ctx.trace_count(rays)
        assert np.array_equal(i0, i1) and np.array_equal(t0, t1), smin
        assert np.array_equal(n0, n1) and np.array_equal(p0, p1), smin
    ctx.set_ray_sort_min(0)
    ctx.set_persistent(True)
    assert np.mean(i0 >= 0) > 0.1
