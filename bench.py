#!/usr/bin/env python3
"""Benchmarks of the MI355X path on one GPU; prints ONE JSON line.

Primary workload (`value`, one step = one frame): the plain path tracer on
the displaced level-8 icosphere (1,310,720 triangles, GGX with roughness
U[0.05,0.8] and F0 U[0.02,0.9] per 64-triangle patch) inside the Cornell box,
1920x1080, 64 spp, 8 path vertices, roulette from the 4th.  `value` is
Mrays/s = (extension + shadow rays traced) / frame time.  `roofline` covers
the dominant kernel, the extension (closest-hit) kernel: algorithmic bytes
per ray = 32 (ray) + 32 (hit record) (+ 4 for the queue permutation where
bounce queues are sorted, off by default) + 64 per
four-wide quantized node record visited (112 = 7 of the 8 float4 of an FP32
record with --records fp32) + 48 per primitive tested, node/primitive counts
from a counting run of the same frame, divided by the summed HIP-event
durations of the extension launches.  The four-wide records group the
clustered tree (CPT_OPT_CLUSTER) where its surface-area cost is clearly below
the LBVH's, as on this scene; the render leg builds it with the BVH (mode 1,
so its time is reported as `cluster_ms`) rather than at the first frame
(mode 2, the default): the records are the same.

Secondary workload (`lbvh`): LBVH build of a 10M-triangle soup (centroids
U[0,1]^3, lognormal side mu=-6 sigma=0.5, random orientation, seed 1) and
closest-hit traversal of 2^25 random rays (origins U[-0.5,1.5]^3, seed 2),
with its own roofline.

`cpu_baseline`: the FP64 CPU oracle (reference semantics, OpenMP on the host
cores) on a bounded sample of the primary workload (a centred pixel window at
the same spp); `lbvh.cpu_baseline` the same for traversal.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpix/s at equal relMSE (denoised vs plain PT); Mrays/s; LBVH build ms"
HBM_PEAK_GBS = 8000.0
def node_bytes(args):
    """Bytes read per visit of a node record: a 64 B four-wide quantized
    record (the default), 7 of the 8 float4 of a four-wide FP32 record, 6
    float4 of an eight-wide quantized record, or a 64 B binary node."""
    if args.width == 8:
        return 96.0
    if args.width == 2:
        return 64.0
    return 64.0 if args.records == "quant" else 112.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cores():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def median_time(fn, reps, warm=1):
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def bench_render(cpt, L, scenes, args):
    s = scenes.displaced_icosphere(1920, 1080)
    ctx = cpt.Context(0)
    ctx.set_quantized(args.records == "quant")
    ctx.set_traversal_width(args.width)
    ctx.set_persistent(not args.no_persistent)
    ctx.set_cluster(0 if args.no_cluster else 1)
    ctx.upload(s)
    st_b = ctx.build(s.bounds_min, s.bounds_max)
    tree = ctx.cluster_info()
    log(f"scene build: LBVH {st_b.total_ms:.3f} ms, clustered tree {st_b.cluster_ms:.3f} ms, {tree}")
    kw = dict(seed=3317, profile=L.PROFILE_CONFIG, max_depth=8, rr_depth=3)
    for _ in range(args.warmup):
        ctx.render(s.camera, 1, **kw)
    frames = []
    for k in range(args.steps):
        img, _, st = ctx.render(s.camera, args.spp, sample_base=k * args.spp, **kw)
        frames.append(st)
        log(f"frame {k}: {st.ms:.1f} ms, {st.rays / 1e6:.1f} Mrays, {st.bounces} bounces")
    # counting run (untimed) for the algorithmic bytes of the extension kernel
    _, _, stc = ctx.render(s.camera, args.spp, sample_base=0, count_work=True, **kw)
    ms = float(np.mean([f.ms for f in frames]))
    rays = float(np.mean([f.rays for f in frames]))
    ext_ms = float(np.mean([f.extend_ms for f in frames]))
    ext_rays = float(np.mean([f.extend_rays for f in frames]))
    # per ray: ray 32 B, hit record 32 B (bounce queues are traced unsorted
    # by default, so no permutation is read); per record
    # visit and per primitive test their record bytes
    ext_bytes = 64.0 * stc.extend_rays + node_bytes(args) * stc.extend_nodes + 48.0 * stc.extend_prims
    ext_bytes *= ext_rays / max(stc.extend_rays, 1.0)
    achieved = ext_bytes / (ext_ms * 1e-3) / 1e9
    # occlusion walk: 48 B shadow record (ray, light sample) + record visits
    # + primitive tests (the radiance update of unoccluded rays, <= 32 B, is
    # not counted)
    sh_ms = float(np.mean([f.shadow_ms for f in frames]))
    sh_rays = float(np.mean([f.shadow_rays for f in frames]))
    sh_bytes = 48.0 * stc.shadow_rays + node_bytes(args) * stc.shadow_nodes + 48.0 * stc.shadow_prims
    sh_bytes *= sh_rays / max(stc.shadow_rays, 1.0)
    sh_achieved = sh_bytes / (sh_ms * 1e-3) / 1e9 if sh_ms > 0 else 0.0
    res = dict(ms=ms, rays=rays, mrays=rays / (ms * 1e-3) / 1e6, ext_ms=ext_ms, ext_rays=ext_rays,
               shade_ms=float(np.mean([f.shade_ms for f in frames])),
               shadow_ms=float(np.mean([f.shadow_ms for f in frames])),
               bounces=int(frames[0].bounces), launches=int(frames[0].bounces) * 4,
               nodes_per_ray=stc.extend_nodes / max(stc.extend_rays, 1),
               prims_per_ray=stc.extend_prims / max(stc.extend_rays, 1),
               lane_utilisation=stc.extend_nodes / max(stc.extend_wave_nodes, 1),
               achieved=achieved, build_ms=st_b.total_ms, cluster_ms=st_b.cluster_ms, tree=tree,
               shadow_achieved=sh_achieved, shadow_rays=sh_rays,
               shadow_nodes_per_ray=stc.shadow_nodes / max(stc.shadow_rays, 1),
               shadow_prims_per_ray=stc.shadow_prims / max(stc.shadow_rays, 1),
               shadow_lane_utilisation=stc.shadow_nodes / max(stc.shadow_wave_nodes, 1),
               mean=float(img.mean()))
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        o = O.OracleScene(s)
        # oracle LBVH build of the same 1.31M triangles: median of 5 after 1
        cpu_build = median_time(lambda: o.build(s.bounds_min, s.bounds_max), 5)
        cam = cpt.camera(s.camera)
        # centred window of the frame, same camera and spp
        win = args.cpu_window
        sub = L.Camera()
        C = __import__("ctypes")
        C.memmove(C.byref(sub), C.byref(cam), C.sizeof(L.Camera))
        i0, j0 = (cam.width - win) // 2, (cam.height - win) // 2
        for k in range(3):
            sub.pixel_center[k] = cam.pixel_center[k] + i0 * cam.delta_h[k] + j0 * cam.delta_v[k]
        sub.width = sub.height = win
        prm = L.RenderParams(3317, args.spp, 8, 3, L.PROFILE_CONFIG, 0.7, 0, 0, 0, 0)
        t0 = time.time()
        o.render(sub, prm, nthreads=cores())
        dt = time.time() - t0
        # ray count of the same window, traced on the GPU
        gw, _, gst = ctx.render(sub, args.spp, seed=3317, profile=L.PROFILE_CONFIG)
        # config 1 (Cornell box + level-3 icosphere, 256x256, 16 spp, depth 8,
        # roulette after 3) on the host cores and on the GPU
        s1 = scenes.cornell_icosphere(256, 256)
        o1 = O.OracleScene(s1)
        o1.build(s1.bounds_min, s1.bounds_max)
        prm1 = L.RenderParams(3317, 16, 8, 3, L.PROFILE_CONFIG, 0.7, 0, 0, 0, 0)
        cam1 = cpt.camera(s1.camera)
        c1 = median_time(lambda: o1.render(cam1, prm1, nthreads=cores()), 3, warm=0)
        ctx.upload(s1)
        ctx.build(s1.bounds_min, s1.bounds_max)
        g1 = median_time(lambda: ctx.render(s1.camera, 16, seed=3317, profile=L.PROFILE_CONFIG), 5)
        cpu = {"value": round(gst.rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": cores(), "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"FP64 oracle (reference-order unculled traversal, FP64 BSDFs, OpenMP) rendering the "
                         f"centred {win}x{win} window of the same frame at {args.spp} spp in {dt:.1f} s; "
                         f"rays = the GPU's ray count for that window",
               "build_ms_1p31m": round(cpu_build * 1e3, 1),
               "config1": {"cpu_mpix_per_s": round(256 * 256 / c1 / 1e6, 4), "cpu_s": round(c1, 3),
                           "gpu_mpix_per_s": round(256 * 256 / g1 / 1e6, 2), "gpu_ms": round(g1 * 1e3, 3),
                           "spp": 16, "size": [256, 256], "max_depth": 8, "rr_depth": 3,
                           "timing": "median of 3 (CPU) / 5 after 1 warm-up (GPU, wall time incl. readback)"}}
        log(f"cpu baseline {cpu}")
    ctx.close()
    return res, cpu, int(s.tri.shape[0])


WEIGHTS = os.path.join(ROOT, "cudapathtracing_amd", "weights", "unet_cornell.f32")


def bench_equal_quality(cpt, L, scenes, args):
    """Denoised 4-spp vs plain path tracing at equal relMSE.

    Scene: the displaced-icosphere workload at args.eq_width x eq_height.
    Reference: args.eq_ref_spp spp.  Denoised: 4 spp + albedo/normal/depth,
    then the U-Net (time = render + inference).  Plain: spp doubled from 4
    until relMSE falls below the denoised relMSE, time interpolated in
    log(spp).  Mpix/s = pixels / time for each method."""
    if not os.path.exists(WEIGHTS):
        return None
    from cudapathtracing_amd.denoiser import Denoiser, pack_input
    from cudapathtracing_amd.training import equal_quality, relmse
    w, h = args.eq_width, args.eq_height
    s = scenes.displaced_icosphere(w, h)
    kw = dict(profile=L.PROFILE_CONFIG, max_depth=8, rr_depth=3)
    with cpt.Context(0) as ctx:
        ctx.upload(s)
        ctx.build(s.bounds_min, s.bounds_max)
        t0 = time.time()
        ref, _, st_ref = ctx.render(s.camera, args.eq_ref_spp, seed=99991, **kw)
        log(f"equal-quality reference {args.eq_ref_spp} spp: {time.time() - t0:.1f}s")
        dn = Denoiser(ctx, 1, h, w)
        dn.set_params(np.fromfile(WEIGHTS, dtype="<f4"))
        ctx.render(s.camera, 4, seed=5, aov=True, **kw)  # warm-up
        dn(pack_input(*ctx.render(s.camera, 4, seed=5, aov=True, **kw)[:2])[None])
        rgb, aov, st4 = ctx.render(s.camera, 4, seed=7, aov=True, **kw)
        den = dn(pack_input(rgb, aov)[None])[0]
        dstat = dn.stats()
        e_den = relmse(den, ref)
        e_noisy = relmse(rgb, ref)
        bil = ctx.bilateral(rgb, aov)
        e_bil = relmse(bil, ref)
        t_den = st4.ms + dstat["fwd_ms"]
        eq = equal_quality(ctx, s.camera, ref, e_den, seed=11, spp_start=4, spp_max=args.eq_ref_spp // 2)
        dn.close()
    npix = w * h
    return {"width": w, "height": h, "reference_spp": args.eq_ref_spp, "denoised_spp": 4,
            "relmse_noisy_4spp": e_noisy, "relmse_bilateral_4spp": e_bil, "relmse_denoised_4spp": e_den,
            "denoised_ms": round(t_den, 2), "render_4spp_ms": round(st4.ms, 2),
            "denoise_ms": round(dstat["fwd_ms"], 2),
            "denoise_tflops": round(dstat["fwd_flops"] / (dstat["fwd_ms"] * 1e-3) / 1e12, 1),
            "plain_spp_at_equal_relmse": round(eq["spp"], 1), "plain_ms_at_equal_relmse": round(eq["ms"], 2),
            "reached": eq["reached"],
            "mpix_per_s_denoised": round(npix / (t_den * 1e-3) / 1e6, 3),
            "mpix_per_s_plain": round(npix / (eq["ms"] * 1e-3) / 1e6, 3),
            "speedup": round(eq["ms"] / t_den, 2),
            "plain_points": [(p[0], round(p[1], 6), round(p[2], 2)) for p in eq["points"]]}


def window_camera(L, cam, x0, y0, w, h):
    """The w x h sub-window at pixel (x0, y0) of an expanded camera."""
    import ctypes as C
    sub = L.Camera()
    C.memmove(C.byref(sub), C.byref(cam), C.sizeof(L.Camera))
    for k in range(3):
        sub.pixel_center[k] = cam.pixel_center[k] + x0 * cam.delta_h[k] + y0 * cam.delta_v[k]
    sub.width, sub.height = w, h
    return sub


def bench_config5(cpt, L, scenes, args):
    """Config 5: denoised 4 spp vs the plain tracer at equal relMSE on the
    displaced icosphere plus a grid^3 fBm medium (sigma_max 8), nv views.

    Error: relMSE with the reference noise removed (two independent
    half-spp references r1, r2: mean((x - r1)(x - r2) / (rbar^2 + 1e-2))),
    over a centred window of each of the first `c5_score_views` views.
    Denoised: all nv views at 4 spp with AOVs in one cpt_render_views call,
    then the U-Net over the views (device buffers end to end), for
    `c5_seeds` seeds; time = median wall time of the pipeline (synchronised).
    Plain: the scored windows accumulated at spp 4, 8, 16, ... (each step
    continues the sample stream) until the relMSE is below every bootstrap
    resample of the denoised mean; the equal-relMSE spp is interpolated
    between the two bracketing measurements (never extrapolated).  Plain
    time for all nv full views at that spp = the measured time of an
    all-view render at 64 spp x spp / 64 (frame time is linear in spp:
    samples are independent passes).  CI: 95% bootstrap over the seeds."""
    if not os.path.exists(WEIGHTS):
        return None
    import torch
    from cudapathtracing_amd.denoiser import Denoiser
    from cudapathtracing_amd.training import crossing_spp, orbit_cameras, relmse_two_ref
    w, h, nv = args.c5_width, args.c5_height, args.c5_views
    nsv = min(args.c5_score_views, nv)
    ww, wh = min(args.c5_window[0], w), min(args.c5_window[1], h)
    x0, y0 = (w - ww) // 2, (h - wh) // 2
    s = scenes.displaced_icosphere_medium(w, h, grid=args.c5_grid)
    cams = orbit_cameras(nv, w, h, seed=7)
    wins = [window_camera(L, cpt.camera(c), x0, y0, ww, wh) for c in cams[:nsv]]
    kw = dict(profile=L.PROFILE_CONFIG, max_depth=8, rr_depth=3)
    chunk = min(nv, 8)
    half = args.c5_ref_spp // 2
    with cpt.Context(0) as ctx:
        ctx.upload(s)
        ctx.build(s.bounds_min, s.bounds_max)
        t0 = time.time()
        r1 = ctx.render_views(wins, half, seed=99991, **kw)[0].astype(np.float64)
        r2 = ctx.render_views(wins, half, seed=77773, **kw)[0].astype(np.float64)
        log(f"config-5 references ({nsv} windows {ww}x{wh}, 2 x {half} spp): {time.time() - t0:.1f}s")

        def score(x):  # mean over the scored windows
            return float(np.mean([relmse_two_ref(x[v], r1[v], r2[v]) for v in range(nsv)]))

        dn = Denoiser(ctx, chunk, h, w)
        dn.set_params(np.fromfile(WEIGHTS, dtype="<f4"))
        # the render writes the denoiser's input layout (radiance + 7 AOVs per pixel)
        x = torch.empty((nv, h, w, 10), dtype=torch.float32, device="cuda")
        rgb = x[..., :3]
        den = torch.empty((nv, h, w, 3), dtype=torch.float32, device="cuda")

        def denoised_pipeline(seed):
            ctx.render_views(cams, 4, seed=seed, aov=True, pack=True, out=x, **kw)
            for c0 in range(0, nv, chunk):
                dn(x[c0:c0 + chunk], out=den[c0:c0 + chunk])

        denoised_pipeline(5)  # warm-up
        e_den, e_noisy, t_den = [], [], []
        for k in range(args.c5_seeds):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            denoised_pipeline(1000 + 7 * k)
            torch.cuda.synchronize()
            t_den.append((time.perf_counter() - t0) * 1e3)
            crop = lambda a: a[:nsv, y0:y0 + wh, x0:x0 + ww].cpu().numpy()
            e_den.append(score(crop(den)))
            e_noisy.append(score(crop(rgb)))
        dn.close()
        e_den = np.array(e_den)
        rng = np.random.default_rng(0)
        boot = rng.choice(e_den, size=(2000, e_den.size), replace=True).mean(axis=1)
        lo, hi = np.quantile(boot, [0.025, 0.975])
        target = float(e_den.mean())
        # plain sweep on the scored windows, continuing the sample stream
        acc = np.zeros((nsv, wh, ww, 3), np.float64)
        pts, n = [], 0
        step = 4
        t0 = time.time()
        while n + step <= args.c5_spp_max:
            img = ctx.render_views(wins, step, seed=4242, sample_base=n, **kw)[0]
            acc += img.astype(np.float64) * step
            n += step
            e = score(acc / n)
            pts.append((n, e))
            log(f"config-5 plain {n} spp: relMSE {e:.3e} (denoised {target:.3e}, CI [{lo:.3e}, {hi:.3e}])")
            if e <= min(lo, target):
                break
            step = n  # doubles the accumulated spp
        log(f"config-5 plain sweep: {time.time() - t0:.1f}s")
        # all-view plain frame time at 64 spp
        plain = torch.empty((nv, h, w, 3), dtype=torch.float32, device="cuda")
        ctx.render_views(cams, 16, seed=13, out=plain, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.render_views(cams, 64, seed=17, out=plain, **kw)
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) * 1e3
    spp_eq = crossing_spp(pts, target)
    spp_ci = [crossing_spp(pts, float(hi)), crossing_spp(pts, float(lo))]  # larger error -> fewer spp
    npix = nv * w * h
    td = float(np.median(t_den))
    out = {"views": nv, "width": w, "height": h, "medium_grid": args.c5_grid, "medium_sigma_max": 8.0,
           "scored_views": nsv, "scored_window": [ww, wh], "reference_spp": [half, half], "seeds": int(e_den.size),
           "estimator": "two independent references: mean((x-r1)(x-r2)/(rbar^2+1e-2))",
           "denoised_spp": 4, "relmse_noisy_4spp": float(np.mean(e_noisy)),
           "relmse_denoised_4spp": target, "relmse_denoised_ci95": [float(lo), float(hi)],
           "denoised_ms": round(td, 2), "mpix_per_s_denoised": round(npix / (td * 1e-3) / 1e6, 3),
           "plain_ms_64spp_all_views": round(t64, 2),
           "plain_points": [(p[0], float(f"{p[1]:.4e}")) for p in pts],
           "reached": spp_eq is not None}
    if spp_eq is not None:
        t_plain = t64 * spp_eq / 64.0
        out.update({"plain_spp_at_equal_relmse": round(spp_eq, 1),
                    "plain_spp_ci95": [None if v is None else round(v, 1) for v in spp_ci],
                    "plain_ms_at_equal_relmse": round(t_plain, 1),
                    "mpix_per_s_plain": round(npix / (t_plain * 1e-3) / 1e6, 4),
                    "speedup": round(t_plain / td, 2),
                    "speedup_ci95": [None if v is None else round(t64 * v / 64.0 / td, 2) for v in spp_ci]})
    else:
        out.update({"plain_spp_at_equal_relmse": None, "speedup": None})
    return out


def bench_denoiser(cpt, args):
    """BF16 U-Net on MFMA (random weights, synthetic inputs): inference of a
    batch of 8 views at 1920x1080 (config 5's chunk of views) and one Adam
    training step on a batch of 8 512x512 inputs (config 4), beside the step
    on 8 128x128 crops (the shipped weights' training shape).  Times: HIP
    events around each pass on the library's
    stream (medians).  FLOPs: 2 * 9 * Cin * Cout per padded pixel per
    convolution; a training step = 3 x forward - the first layer's data
    gradient.  Roofline: dense BF16 MFMA peak 2.5 PFLOP/s; the pass time
    includes its elementwise kernels (input preparation, pooling, output,
    split sums, Adam), so `frac` is a lower bound for the convolution
    kernels themselves; `mfma_busy` is the counters' MFMA-busy share of the
    convolution kernels (committed summary, null if the kernels changed)."""
    import torch
    from cudapathtracing_amd.denoiser import Denoiser
    B, H, W = 8, 1080, 1920
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    with cpt.Context(0) as ctx:
        dn = Denoiser(ctx, B, H, W, seed=1)
        x = torch.rand((B, H, W, 10), generator=g, device="cuda")
        y = torch.empty((B, H, W, 3), device="cuda")
        for _ in range(2):
            dn(x, out=y)
        t = []
        for _ in range(10):
            dn(x, out=y)
            t.append(dn.stats()["fwd_ms"])
        fl = dn.stats()["fwd_flops"]
        inf_ms = float(np.median(t))
        rng = np.random.default_rng(0)
        L0 = dn.layers()[0]
        peak = 2500.0
        train = {}
        for S in (512, 128):
            Bt = 8
            inp = torch.rand((Bt, S, S, 10), generator=g, device="cuda")
            tgt = torch.rand((Bt, S, S, 3), generator=g, device="cuda")
            for _ in range(2):
                dn.train_step(inp, tgt, 1e-4)
            ts = []
            for _ in range(8):
                dn.train_step(inp, tgt, 1e-4)
                ts.append(dn.stats()["step_ms"])
            dn(inp)
            f1 = dn.stats()["fwd_flops"]
            f_train = 3 * f1 - 2.0 * Bt * S * S * L0["cin"] * L0["cout"] * 9
            tr_ms = float(np.median(ts))
            tf_tr = f_train / (tr_ms * 1e-3) / 1e12
            train[S] = {"batch": Bt, "size": S, "ms": round(tr_ms, 3), "gflop": round(f_train / 1e9, 1),
                        "tflops": round(tf_tr, 1), "frac": round(tf_tr / peak, 4)}
        dn.close()
    tf_inf = fl / (inf_ms * 1e-3) / 1e12
    log(f"denoiser: inference {inf_ms:.3f} ms ({tf_inf:.1f} TFLOP/s), training step 512^2 {train[512]['ms']} ms "
        f"({train[512]['tflops']}), 128^2 {train[128]['ms']} ms ({train[128]['tflops']})")
    return {"inference": {"batch": B, "width": W, "height": H, "ms": round(inf_ms, 3), "gflop": round(fl / 1e9, 1),
                          "tflops": round(tf_inf, 1), "mpix_per_s": round(B * W * H / (inf_ms * 1e-3) / 1e6, 1)},
            "training": {**train[512], "workload": "config 4: batch 8 of 512x512 inputs"},
            "training_128_crops": train[128],
            "roofline": {"bound": "mfma", "achieved": round(tf_inf, 1), "peak": peak, "unit": "TFLOP/s",
                         "frac": round(tf_inf / peak, 4), "traffic": None, "kernel": "k_conv_fwd",
                         "scope": "whole inference pass",
                         "mfma_busy": profile_value("mfma_denoiser_convs.md", "mfma_busy_k_conv_fwd", "k_conv_fwd")}}


def bench_lbvh(cpt, scenes, args):
    import torch
    scene = scenes.triangle_soup_scene(args.tris, seed=1)
    rays_np = scenes.random_rays(args.rays, seed=2)
    ctx = cpt.Context(0)
    ctx.set_quantized(args.records == "quant")
    ctx.set_traversal_width(args.width)
    ctx.set_persistent(not args.no_persistent)
    ctx.upload(scene)
    rays = torch.from_numpy(rays_np).cuda()
    t_out = torch.empty(args.rays, dtype=torch.float32, device="cuda")
    id_out = torch.empty(args.rays, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    build_ms, trace_ms, kern_ms, sort_ms, stage = [], [], [], [], []
    for it in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        st = ctx.build(scene.bounds_min, scene.bounds_max)
        ctx.trace(rays, t_out, id_out)
        tr, srt, ker = ctx.last_trace_stats()
        torch.cuda.synchronize()
        if it >= args.warmup:
            build_ms.append(st.total_ms)
            trace_ms.append(tr)
            kern_ms.append(ker)
            sort_ms.append(srt)
            stage.append((st.morton_ms, st.sort_ms, st.emit_ms, st.refit_ms))
        log(f"lbvh step {it}: build {st.total_ms:.3f} ms, trace {tr:.3f} ms (ray ordering {srt:.3f}, "
            f"kernel {ker:.3f})")
    nodes, prims = ctx.trace_count(rays)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    n_nodes = float(nodes.double().sum().item())
    n_prims = float(prims.double().sum().item())
    hits = float((id_out >= 0).double().mean().item())
    alg = 40.0 * args.rays + node_bytes(args) * n_nodes + 48.0 * n_prims
    tr_avg, bd_avg = float(np.mean(trace_ms)), float(np.mean(build_ms))
    k_avg, s_avg = float(np.mean(kern_ms)), float(np.mean(sort_ms))
    sa = np.mean(np.array(stage), axis=0).tolist()
    # the kernel reads each ray through the ordering permutation (+4 B)
    alg += 4.0 * args.rays if s_avg > 0 else 0.0
    achieved = alg / (k_avg * 1e-3) / 1e9
    # the persistent walk (four-wide formats) and its t pass, or one lane per ray
    tr_kernel = "k_trace_persist" if (args.width == 4 and not args.no_persistent) else "k_trace"
    out = {"triangles": args.tris, "rays": args.rays, "build_ms": round(bd_avg, 4),
           "build_stage_ms": {"morton": round(sa[0], 4), "sort": round(sa[1], 4), "emit": round(sa[2], 4),
                              "refit": round(sa[3], 4)},
           "trace_ms": round(tr_avg, 4), "ray_ordering_ms": round(s_avg, 4), "trace_kernel_ms": round(k_avg, 4),
           "mrays_per_s": round(args.rays / (tr_avg * 1e-3) / 1e6, 1),
           "hit_fraction": round(hits, 4), "nodes_per_ray": round(n_nodes / args.rays, 2),
           "prims_per_ray": round(n_prims / args.rays, 2),
           "roofline": roofline_hbm(tr_kernel, achieved, k_avg)}
    if not args.no_cpu:
        from oracle import oracle as O
        o = O.OracleScene(scene)
        cb = median_time(lambda: o.build(scene.bounds_min, scene.bounds_max), args.cpu_build_reps)
        sample = rays_np[:args.cpu_rays]
        tc = time.time()
        o.trace_ref(sample)
        dt = time.time() - tc
        out["cpu_baseline"] = {"value": round(sample.shape[0] / dt / 1e6, 4), "unit": "Mrays/s",
                               "cores": cores(), "kind": "port", "cpu_model": cpu_model(),
                               "sample": f"FP64 oracle, reference traversal order (no t-culling), first "
                                         f"{sample.shape[0]} of the {args.rays} rays",
                               "build_ms": round(cb * 1e3, 1),
                               "build_timing": f"oracle LBVH build of the same {args.tris} triangles, median of "
                                               f"{args.cpu_build_reps} after 1 warm-up"}
    ctx.close()
    return out


def copy_probe(nbytes):
    """A plain device copy of known size: its FETCH_SIZE / WRITE_SIZE per
    dispatch calibrate the counters (bytes read and written = nbytes)."""
    import torch
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 4}), flush=True)


# Sources whose code decides each profiled kernel's traffic and issue
# counts: a committed counter summary holds the digest of these files at the
# time it was measured, and its figures are reported only while the digest
# still matches (null otherwise: the kernel has changed since).
KERNEL_SOURCES = {
    "render": ["render.hip", "bsdf.hpp", "cpt_device.hpp", "internal.hpp", "bvh.hip", "cluster.hip"],
    "trace": ["trace.hip", "cpt_device.hpp", "internal.hpp", "bvh.hip", "cluster.hip"],
    "denoise": ["denoise.hip", "denoise_kernels.hpp"],
}
KERNEL_GROUP = {"k_extend_persist": "render", "k_extend": "render", "k_shadow": "render", "k_shadow_persist": "render",
                "k_shade": "render",
                "k_trace_persist": "trace", "k_trace": "trace", "k_conv_fwd": "denoise", "k_conv_bwd_w": "denoise"}


def _code_only(path):
    """Source text without comments and blank lines, so that a comment edit
    keeps the digest."""
    import re
    t = open(path, encoding="utf-8").read()
    t = re.sub(r"/\*.*?\*/", "", t, flags=re.S)
    t = re.sub(r"//[^\n]*", "", t)
    return "\n".join(l.rstrip() for l in t.splitlines() if l.strip()).encode()


def source_digest(group):
    """Digest of the code (comments removed) that decides a group's kernels."""
    import hashlib
    h = hashlib.sha256()
    for f in KERNEL_SOURCES[group]:
        h.update(f.encode())
        h.update(_code_only(os.path.join(ROOT, "cudapathtracing_amd", "csrc", f)))
    h.update(_code_only(os.path.join(ROOT, "include", "cpt.h")))
    return h.hexdigest()[:16]


def profile_value(fname, key, kernel):
    """`key: value` from the committed summary profiles/<fname>, when its
    source_digest line matches the current sources of `kernel`; else None."""
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and not line.startswith(("|", " ", "`", "#")):
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
    if vals.get("source_digest") != source_digest(KERNEL_GROUP[kernel]) or key not in vals:
        return None
    return float(vals[key])


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by a copy of known
    size), or None when none exists for the current sources."""
    return profile_value(f"traffic_{kernel}.md", "hbm_bytes_per_launch", kernel)


def roofline_hbm(kernel, achieved, launch_ms):
    """HBM roofline of a traversal kernel: algorithmic bytes / launch time
    against 8 TB/s, beside what the counters say -- the measured HBM traffic
    and its fraction of peak, and the share of issue slots spent on VALU
    instructions (the actual limiter of the walks, whose records are served
    largely from L2 and the MALL)."""
    tr = measured_traffic(kernel)
    issue = profile_value(f"valu_issue_{kernel}.md", f"valu_issue_{kernel}", kernel)
    if issue is None:
        issue = profile_value("valu_issue_traversal_kernels.md", f"valu_issue_{kernel}", kernel)
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr, "kernel": kernel,
           "avg_launch_ms": round(launch_ms, 4)}
    if achieved > HBM_PEAK_GBS:
        # every record visit and primitive test counts as algorithmic bytes,
        # also where the L2 serves it, so this rate can pass the HBM peak;
        # traffic / traffic_frac give the bytes that came from HBM
        out["note"] = "algorithmic bytes include L2-served record reads; see traffic_frac"
    if tr is not None and launch_ms > 0:
        out["traffic_frac"] = round(tr / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    if issue is not None:
        out["valu_issue_frac"] = issue
        out["limiter"] = "valu_issue" if issue >= 0.6 else "memory"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--tris", type=int, default=10_000_000)
    ap.add_argument("--rays", type=int, default=1 << 25)
    ap.add_argument("--cpu-rays", type=int, default=1 << 20)
    ap.add_argument("--cpu-build-reps", type=int, default=5)
    ap.add_argument("--cpu-window", type=int, default=320)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lbvh", action="store_true")
    ap.add_argument("--no-eq", action="store_true")
    ap.add_argument("--eq-width", type=int, default=960)
    ap.add_argument("--eq-height", type=int, default=540)
    ap.add_argument("--eq-ref-spp", type=int, default=4096)
    ap.add_argument("--width", type=int, choices=[2, 4, 8], default=4, help="traversal width")
    ap.add_argument("--records", choices=["quant", "fp32"], default="quant",
                    help="four-wide node records read by the traversal")
    ap.add_argument("--no-c5", action="store_true", help="skip the many-view medium equal-relMSE leg")
    ap.add_argument("--no-cluster", action="store_true",
                    help="four-wide records from the LBVH's boxes even where the clustered tree is cheaper")
    ap.add_argument("--no-persistent", action="store_true",
                    help="one lane per ray instead of the persistent lane-refilling walks")
    ap.add_argument("--c5-views", type=int, default=32)
    ap.add_argument("--c5-width", type=int, default=1920)
    ap.add_argument("--c5-height", type=int, default=1080)
    ap.add_argument("--c5-grid", type=int, default=256)
    ap.add_argument("--c5-ref-spp", type=int, default=4096, help="total over the two independent references")
    ap.add_argument("--c5-score-views", type=int, default=8)
    ap.add_argument("--c5-window", type=int, nargs=2, default=[640, 360], help="scored centred window (w h)")
    ap.add_argument("--c5-seeds", type=int, default=8)
    ap.add_argument("--c5-spp-max", type=int, default=65536)
    ap.add_argument("--no-dn", action="store_true", help="skip the denoiser kernels leg")
    ap.add_argument("--only", choices=["render", "lbvh", "eq", "c5", "dn"], default=None,
                    help="run one workload and print its record (profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="copy this many bytes device-to-device 5 times and exit (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    import cudapathtracing_amd as cpt
    from cudapathtracing_amd import _lib as L
    from cudapathtracing_amd import scenes
    if args.only == "lbvh":
        print(json.dumps({"lbvh": bench_lbvh(cpt, scenes, args)}), flush=True)
        return
    if args.only == "eq":
        print(json.dumps({"equal_relmse": bench_equal_quality(cpt, L, scenes, args)}), flush=True)
        return
    if args.only == "c5":
        print(json.dumps({"config5": bench_config5(cpt, L, scenes, args)}), flush=True)
        return
    if args.only == "dn":
        print(json.dumps({"denoiser": bench_denoiser(cpt, args)}), flush=True)
        return
    if args.only == "render":
        args.no_lbvh = args.no_eq = args.no_c5 = args.no_dn = True

    r, cpu, ntri = bench_render(cpt, L, scenes, args)
    lb = None if args.no_lbvh else bench_lbvh(cpt, scenes, args)
    eq = None if args.no_eq else bench_equal_quality(cpt, L, scenes, args)
    c5 = None if args.no_c5 else bench_config5(cpt, L, scenes, args)
    dnr = None if args.no_dn else bench_denoiser(cpt, args)
    ext_kernel = "k_extend_persist" if (args.width == 4 and not args.no_persistent) else "k_extend"
    sh_kernel = "k_shadow_persist" if (args.width == 4 and not args.no_persistent) else "k_shadow"
    out = {
        "metric": METRIC,
        "value": round(r["mrays"], 1),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain_path_tracer_displaced_icosphere", "width_bvh": args.width,
                   "records": args.records,
                   "traversal_tree": "clustered" if r["tree"]["in_use"] else "lbvh",
                   "triangles": ntri, "width": 1920,
                   "height": 1080, "spp": args.spp, "max_depth": 8, "rr_depth": 3},
        "mpix_per_s": round(1920 * 1080 / (r["ms"] * 1e-3) / 1e6, 3),
        "rays_per_frame": r["rays"],
        "kernel_ms": {"extend": round(r["ext_ms"], 2), "shade": round(r["shade_ms"], 2),
                      "shadow": round(r["shadow_ms"], 2)},
        "bounces": r["bounces"],
        "extend_nodes_per_ray": round(r["nodes_per_ray"], 2),
        "extend_prims_per_ray": round(r["prims_per_ray"], 2),
        "extend_lane_utilisation": round(r["lane_utilisation"], 3),
        "scene_lbvh_build_ms": round(r["build_ms"], 3),
        "scene_cluster_ms": round(r["cluster_ms"], 3),
        "tree_cost": {"lbvh": round(r["tree"]["lbvh_cost"], 2), "clustered": round(r["tree"]["cluster_cost"], 2),
                      "clustered_depth": r["tree"]["depth"]},
        "roofline": roofline_hbm(ext_kernel, r["achieved"], r["ext_ms"] / max(1, r["bounces"])),
        "shadow": {"rays_per_frame": r["shadow_rays"], "nodes_per_ray": round(r["shadow_nodes_per_ray"], 2),
                   "prims_per_ray": round(r["shadow_prims_per_ray"], 2),
                   "lane_utilisation": round(r["shadow_lane_utilisation"], 3),
                   "kernels": "k_shadow_persist for a bounce with at least 2^22 shadow rays (no medium), "
                              "k_shadow below (origin-cell order from 2^20 rays)",
                   "roofline": roofline_hbm(sh_kernel, r["shadow_achieved"], r["shadow_ms"] / max(1, r["bounces"]))},
        "cpu_baseline": cpu,
        "lbvh": lb,
        "equal_relmse": eq,
        "config5": c5,
        "denoiser": dnr,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
