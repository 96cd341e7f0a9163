#!/usr/bin/env python3
"""Benchmark of the LBVH build + closest-hit traversal path on one MI355X.

Workload (one step): LBVH build of a 10M-triangle soup (centroids U[0,1]^3,
lognormal side mu=-6 sigma=0.5, random orientation, seed 1; 30-bit Morton
codes) followed by closest-hit traversal of 2^25 rays (origins
U[-0.5,1.5]^3, uniform directions, seed 2; tmin 0.001, tmax 999999).
Inputs are resident in HBM before the timed region.

Prints one JSON line (rank 0).  `value` is traversal throughput in Mrays/s;
`lbvh_build_ms` is the build time of the same step; `roofline` covers the
traversal kernel (algorithmic bytes: 32 B ray + 8 B result + 64 B per node
visited + 48 B per primitive tested, counts from the counting variant of the
same kernel); `cpu_baseline` is the FP64 CPU oracle (reference traversal
order, OpenMP) on a 2^18-ray sample of the same rays.
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=10_000_000)
    ap.add_argument("--rays", type=int, default=1 << 25)
    ap.add_argument("--cpu-rays", type=int, default=1 << 18)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-render", action="store_true",
                    help="skip the path-tracing workload (1.31M-triangle displaced icosphere, 1080p 64 spp)")
    ap.add_argument("--render-spp", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU figures are not produced",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch
    import cudapathtracing_amd as cpt
    from cudapathtracing_amd import scenes

    t0 = time.time()
    scene = scenes.triangle_soup_scene(args.tris, seed=1)
    rays_np = scenes.random_rays(args.rays, seed=2)
    log(f"generated {args.tris} triangles, {args.rays} rays in {time.time() - t0:.1f}s")

    ctx = cpt.Context(0)
    ctx.upload(scene)
    rays = torch.from_numpy(rays_np).cuda()
    t_out = torch.empty(args.rays, dtype=torch.float32, device="cuda")
    id_out = torch.empty(args.rays, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    log("scene and rays resident in HBM")

    build_ms, trace_ms, stage = [], [], []
    for it in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        st = ctx.build(scene.bounds_min, scene.bounds_max)
        ctx.trace(rays, t_out, id_out)
        tr = ctx.last_trace_ms()
        torch.cuda.synchronize()
        if it >= args.warmup:
            build_ms.append(st.total_ms)
            trace_ms.append(tr)
            stage.append((st.morton_ms, st.sort_ms, st.emit_ms, st.refit_ms))
        log(f"step {it}: build {st.total_ms:.3f} ms, trace {tr:.3f} ms")
    ctx.set_stream(None)

    # algorithmic bytes of the traversal kernel (counting variant, untimed)
    nodes, prims = ctx.trace_count(rays)
    torch.cuda.synchronize()
    ctx.set_stream(None)
    n_nodes = float(nodes.double().sum().item())
    n_prims = float(prims.double().sum().item())
    alg_bytes = 40.0 * args.rays + 64.0 * n_nodes + 48.0 * n_prims
    hits = float((id_out >= 0).double().mean().item())

    tr_avg = float(np.mean(trace_ms))
    bd_avg = float(np.mean(build_ms))
    mrays = args.rays / (tr_avg * 1e-3) / 1e6
    achieved = alg_bytes / (tr_avg * 1e-3) / 1e9
    stage_avg = np.mean(np.array(stage), axis=0).tolist()

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        o = O.OracleScene(scene)
        tb = time.time()
        o.build(scene.bounds_min, scene.bounds_max)
        cpu_build_s = time.time() - tb
        sample = rays_np[:args.cpu_rays]
        tc = time.time()
        o.trace_ref(sample)
        cpu_s = time.time() - tc
        cpu = {"value": round(sample.shape[0] / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": cores,
               "kind": "port",
               "sample": f"FP64 oracle, reference traversal order (no t-culling), first "
                         f"{sample.shape[0]} of the {args.rays} rays on the same {args.tris}-triangle "
                         f"LBVH; oracle LBVH build {cpu_build_s * 1e3:.0f} ms"}
        log(f"cpu baseline {cpu['value']} Mrays/s on {cores} threads, build {cpu_build_s:.2f}s")

    render = None
    if not args.no_render:
        render = bench_render(cpt, scenes, args.render_spp)
        log(f"render: {render}")

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(tr_avg + bd_avg, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "lbvh_build_plus_closest_hit", "triangles": args.tris,
                   "rays": args.rays, "morton_bits": 30, "tmin": 0.001, "tmax": 999999.0},
        "lbvh_build_ms": round(bd_avg, 4),
        "lbvh_stage_ms": {"morton": round(stage_avg[0], 4), "sort": round(stage_avg[1], 4),
                          "emit": round(stage_avg[2], 4), "refit": round(stage_avg[3], 4)},
        "trace_ms": round(tr_avg, 4),
        "hit_fraction": round(hits, 4),
        "nodes_per_ray": round(n_nodes / args.rays, 3),
        "prims_per_ray": round(n_prims / args.rays, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_trace"},
        "cpu_baseline": cpu,
        "path_tracer": render,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_render(cpt, scenes, spp):
    """Plain path tracer on the displaced level-8 icosphere (1,310,720 GGX
    triangles in the Cornell box), 1920x1080, `spp` samples, depth 8,
    roulette after depth 3.  Returns Mrays/s (extension + shadow rays),
    Mpix/s (pixels per second at this spp) and the kernel time split."""
    import time
    t0 = time.time()
    s = scenes.displaced_icosphere(1920, 1080)
    with cpt.Context(0) as ctx:
        ctx.upload(s)
        st_b = ctx.build(s.bounds_min, s.bounds_max)
        ctx.render(s.camera, 1, seed=1)  # warm-up
        img, _, st = ctx.render(s.camera, spp, seed=3317)
    return {"workload": "displaced_icosphere_1080p", "triangles": int(s.tri.shape[0]), "spp": spp,
            "max_depth": 8, "rr_depth": 3, "ms": round(st.ms, 2), "rays": st.rays,
            "mrays_per_s": round(st.rays / (st.ms * 1e-3) / 1e6, 1),
            "mpix_per_s": round(1920 * 1080 / (st.ms * 1e-3) / 1e6, 3),
            "extend_ms": round(st.extend_ms, 2), "shade_ms": round(st.shade_ms, 2),
            "shadow_ms": round(st.shadow_ms, 2), "bounces": st.bounces,
            "lbvh_build_ms": round(st_b.total_ms, 3), "mean_radiance": float(img.mean()),
            "wall_s": round(time.time() - t0, 1)}


if __name__ == "__main__":
    main()
