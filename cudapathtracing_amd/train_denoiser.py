"""Train the denoiser on path-traced data and evaluate it.

    python -m cudapathtracing_amd.train_denoiser --scenes 64 --size 256 \
        --spp-in 4 --spp-ref 256 --steps 1500 --out weights.f32

1. Renders ``--scenes`` seeded Cornell scenes (``training.random_scene``):
   input = ``--spp-in`` spp radiance + albedo/normal/depth, target =
   ``--spp-ref`` spp radiance.
2. Trains the U-Net with Adam (batch ``--batch``, random crops and flips).
3. Evaluates on ``--eval`` held-out scenes: relMSE of the noisy input and of
   the denoised output against the target.
4. Writes the packed FP32 parameters (little-endian float32) and a JSON
   summary line.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import Context
from .denoiser import Denoiser
from .training import make_dataset, relmse, train


def log(msg):
    print(f"[train] {msg}", file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=64)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--spp-in", type=int, default=4)
    ap.add_argument("--spp-ref", type=int, default=256)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", type=str, default="")
    ap.add_argument("--out", type=str, default="")
    a = ap.parse_args(argv)
    t0 = time.time()
    with Context(0) as ctx:
        data = make_dataset(ctx, a.scenes, a.size, a.spp_in, a.spp_ref, seed0=a.seed, log=log)
        held = make_dataset(ctx, a.eval, a.size, a.spp_in, a.spp_ref, seed0=a.seed + 100000, log=log)
        t_data = time.time() - t0
        dn = Denoiser(ctx, a.batch, max(a.crop, a.size), max(a.crop, a.size), seed=a.seed + 1)
        if a.init:
            dn.set_params(np.fromfile(a.init, dtype="<f4"))
        losses = train(dn, data, a.steps, a.batch, a.crop, lr=a.lr, seed=a.seed, log=log)
        st = dn.stats()
        e_noisy, e_den = [], []
        for p in held:
            out = dn(p.inp[None])[0]
            e_noisy.append(relmse(p.inp[..., :3], p.target))
            e_den.append(relmse(out, p.target))
        ev = dn.stats()
        if a.out:
            dn.get_params().astype("<f4").tofile(a.out)
        res = {
            "scenes": a.scenes, "size": a.size, "spp_in": a.spp_in, "spp_ref": a.spp_ref,
            "steps": a.steps, "batch": a.batch, "crop": a.crop, "params": dn.n_params,
            "loss_first": float(np.mean(losses[:20])), "loss_last": float(np.mean(losses[-20:])),
            "relmse_noisy": float(np.mean(e_noisy)), "relmse_denoised": float(np.mean(e_den)),
            "gain": float(np.mean(e_noisy) / max(np.mean(e_den), 1e-12)),
            "train_step_ms": st["step_ms"],
            "train_tflops": 3 * st["fwd_flops"] / max(st["step_ms"], 1e-9) / 1e9,
            "eval_fwd_ms": ev["fwd_ms"], "eval_fwd_flops": ev["fwd_flops"],
            "data_s": round(t_data, 1), "total_s": round(time.time() - t0, 1),
        }
        dn.close()
    print(json.dumps(res), flush=True)
    return res


if __name__ == "__main__":
    main()
