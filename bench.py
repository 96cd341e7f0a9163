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
    res = dict(ms=ms, rays=rays, mrays=rays / (ms * 1e-3) / 1e6, ext_ms=ext_ms, ext_rays=ext_rays,
               shade_ms=float(np.mean([f.shade_ms for f in frames])),
               shadow_ms=float(np.mean([f.shadow_ms for f in frames])),
               bounces=int(frames[0].bounces), launches=int(frames[0].bounces) * 4,
               nodes_per_ray=stc.extend_nodes / max(stc.extend_rays, 1),
               prims_per_ray=stc.extend_prims / max(stc.extend_rays, 1),
               lane_utilisation=stc.extend_nodes / max(stc.extend_wave_nodes, 1),
               achieved=achieved, build_ms=st_b.total_ms, cluster_ms=st_b.cluster_ms, tree=tree,
               mean=float(img.mean()))
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        o = O.OracleScene(s)
        o.build(s.bounds_min, s.bounds_max)
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
        cpu = {"value": round(gst.rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": cores(), "kind": "port",
               "sample": f"FP64 oracle (reference-order unculled traversal, FP64 BSDFs, OpenMP) rendering the "
                         f"centred {win}x{win} window of the same frame at {args.spp} spp in {dt:.1f} s; "
                         f"rays = the GPU's ray count for that window"}
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


def bench_config5(cpt, L, scenes, args):
    """Denoised inference vs plain tracer at equal relMSE on the displaced
    icosphere plus a grid^3 fBm medium (sigma_max 8), many views at once.

    Denoised: all views rendered at 4 spp with AOVs in one cpt_render_views
    call, then the U-Net over the views (device buffers end to end; time =
    wall time of both, synchronised).  Error: relMSE of the first
    `c5_ref_views` views against their `c5_ref_spp` renders.  Plain: spp
    doubled from 4 on those views until relMSE falls below the denoised
    one (log-log interpolation); its time for all views = the measured time
    of an all-view render at 16 spp scaled linearly to that spp."""
    if not os.path.exists(WEIGHTS):
        return None
    import torch
    from cudapathtracing_amd.denoiser import Denoiser
    from cudapathtracing_amd.training import equal_quality, orbit_cameras, relmse
    w, h, nv = args.c5_width, args.c5_height, args.c5_views
    s = scenes.displaced_icosphere_medium(w, h, grid=args.c5_grid)
    cams = orbit_cameras(nv, w, h, seed=7)
    kw = dict(profile=L.PROFILE_CONFIG, max_depth=8, rr_depth=3)
    chunk = min(nv, 8)
    with cpt.Context(0) as ctx:
        ctx.upload(s)
        ctx.build(s.bounds_min, s.bounds_max)
        t0 = time.time()
        refs = [ctx.render(cams[v], args.c5_ref_spp, seed=99991 + v, **kw)[0] for v in range(args.c5_ref_views)]
        log(f"config-5 references ({args.c5_ref_views} views, {args.c5_ref_spp} spp): {time.time() - t0:.1f}s")
        dn = Denoiser(ctx, chunk, h, w)
        dn.set_params(np.fromfile(WEIGHTS, dtype="<f4"))
        rgb = torch.empty((nv, h, w, 3), dtype=torch.float32, device="cuda")
        aov = torch.empty((nv, h, w, 7), dtype=torch.float32, device="cuda")
        den = torch.empty((nv, h, w, 3), dtype=torch.float32, device="cuda")

        def denoised_pipeline(seed):
            ctx.render_views(cams, 4, seed=seed, aov=True, out=rgb, aov_out=aov, **kw)
            x = torch.cat([rgb, aov], dim=-1).contiguous()
            for c0 in range(0, nv, chunk):
                dn(x[c0:c0 + chunk], out=den[c0:c0 + chunk])

        denoised_pipeline(5)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        denoised_pipeline(7)
        torch.cuda.synchronize()
        t_den = (time.perf_counter() - t0) * 1e3
        d_np = den[:args.c5_ref_views].cpu().numpy()
        e_den = float(np.mean([relmse(d_np[v], refs[v]) for v in range(args.c5_ref_views)]))
        e_noisy = float(np.mean([relmse(rgb[v].cpu().numpy(), refs[v]) for v in range(args.c5_ref_views)]))
        eq = equal_quality(ctx, cams[0], refs[0], e_den, seed=11, spp_start=4, spp_max=args.c5_ref_spp // 2)
        plain16 = torch.empty((nv, h, w, 3), dtype=torch.float32, device="cuda")
        ctx.render_views(cams, 16, seed=13, out=plain16, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.render_views(cams, 16, seed=17, out=plain16, **kw)
        torch.cuda.synchronize()
        t16 = (time.perf_counter() - t0) * 1e3
        t_plain = t16 * eq["spp"] / 16.0
        dn.close()
    npix = nv * w * h
    fit = {}
    if not eq["reached"]:
        # measured plain spp is only a lower bound; the a/n + b fit of relMSE
        # (b = noise floor of the finite-spp reference) gives an estimate
        fit = {"plain_fit_a": eq.get("fit_a"), "plain_fit_b": eq.get("fit_b")}
        if "spp_fit" in eq:
            t_fit = t16 * eq["spp_fit"] / 16.0
            fit.update({"plain_spp_fit": round(eq["spp_fit"], 1), "plain_ms_fit": round(t_fit, 1),
                        "mpix_per_s_plain_fit": round(nv * w * h / (t_fit * 1e-3) / 1e6, 4),
                        "speedup_fit": round(t_fit / t_den, 1)})
    return {**fit, "views": nv, "width": w, "height": h, "medium_grid": args.c5_grid, "medium_sigma_max": 8.0,
            "reference_spp": args.c5_ref_spp, "reference_views": args.c5_ref_views, "denoised_spp": 4,
            "relmse_noisy_4spp": e_noisy, "relmse_denoised_4spp": e_den, "denoised_ms": round(t_den, 2),
            "plain_spp_at_equal_relmse": round(eq["spp"], 1), "plain_ms_at_equal_relmse": round(t_plain, 2),
            "plain_ms_16spp_all_views": round(t16, 2), "reached": eq["reached"],
            "mpix_per_s_denoised": round(npix / (t_den * 1e-3) / 1e6, 3),
            "mpix_per_s_plain": round(npix / (t_plain * 1e-3) / 1e6, 3),
            "speedup": round(t_plain / t_den, 2),
            "speedup_is_lower_bound": not eq["reached"],
            "plain_points": [(p[0], round(p[1], 6), round(p[2], 2)) for p in eq["points"]]}


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
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(tr_kernel),
                        "kernel": tr_kernel}}
    if not args.no_cpu:
        from oracle import oracle as O
        o = O.OracleScene(scene)
        tb = time.time()
        o.build(scene.bounds_min, scene.bounds_max)
        cb = time.time() - tb
        sample = rays_np[:args.cpu_rays]
        tc = time.time()
        o.trace_ref(sample)
        dt = time.time() - tc
        out["cpu_baseline"] = {"value": round(sample.shape[0] / dt / 1e6, 4), "unit": "Mrays/s",
                               "cores": cores(), "kind": "port",
                               "sample": f"FP64 oracle, reference traversal order (no t-culling), first "
                                         f"{sample.shape[0]} rays; oracle LBVH build of the same "
                                         f"{args.tris} triangles {cb * 1e3:.0f} ms"}
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


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by a copy of known
    size), or None when no summary for it exists."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--tris", type=int, default=10_000_000)
    ap.add_argument("--rays", type=int, default=1 << 25)
    ap.add_argument("--cpu-rays", type=int, default=1 << 17)
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
    ap.add_argument("--c5-ref-spp", type=int, default=4096)
    ap.add_argument("--c5-ref-views", type=int, default=1)
    ap.add_argument("--only", choices=["render", "lbvh", "eq", "c5"], default=None,
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
    if args.only == "render":
        args.no_lbvh = args.no_eq = args.no_c5 = True

    r, cpu, ntri = bench_render(cpt, L, scenes, args)
    lb = None if args.no_lbvh else bench_lbvh(cpt, scenes, args)
    eq = None if args.no_eq else bench_equal_quality(cpt, L, scenes, args)
    c5 = None if args.no_c5 else bench_config5(cpt, L, scenes, args)
    ext_kernel = "k_extend_persist" if (args.width == 4 and not args.no_persistent) else "k_extend"
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
        "roofline": {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["achieved"] / HBM_PEAK_GBS, 4), "traffic": measured_traffic(ext_kernel),
                     "kernel": ext_kernel,
                     "avg_launch_ms": round(r["ext_ms"] / max(1, r["bounces"]), 4)},
        "cpu_baseline": cpu,
        "lbvh": lb,
        "equal_relmse": eq,
        "config5": c5,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
