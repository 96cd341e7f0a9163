"""cpt_trace_closest_noalloc (no reference counterpart; for callers that
cannot allocate): called through the C ABI on device buffers, its outputs
equal cpt_trace_closest's on two calls in a row, the second smaller, with
and without ray ordering; sizes beyond CPT_NOALLOC_MAX_RAYS are refused."""
import ctypes as C

import numpy as np
import pytest

import cudapathtracing_amd as cpt
from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import scenes

pytestmark = pytest.mark.gpu


def noalloc(ctx, rays, t, ids):
    return ctx._lib.cpt_trace_closest_noalloc(ctx._h, rays.data_ptr(), int(rays.shape[0]), t.data_ptr(),
                                              ids.data_ptr())


@pytest.mark.parametrize("sort_min", [0, 4096])
def test_noalloc_matches_trace_closest(sort_min):
    import torch
    print(f"[gpu] no-alloc trace (ray ordering from {sort_min or 'never'})", flush=True)
    dev = torch.device("cuda:0")
    with cpt.Context(0) as ctx:
        s = scenes.displaced_icosphere(32, 32, level=5)
        ctx.upload(s)
        ctx.build(s.bounds_min, s.bounds_max)
        ctx.set_ray_sort_min(sort_min)
        for n, seed in ((200_000, 3), (70_001, 4)):
            rays = torch.from_numpy(scenes.random_rays(n, seed=seed, lo=-0.45, hi=0.45)).to(dev)
            t0, i0 = ctx.trace(rays)
            t1 = torch.full((n,), -7.0, dtype=torch.float32, device=dev)
            i1 = torch.full((n,), -7, dtype=torch.int32, device=dev)
            ctx._torch_stream()
            rc = noalloc(ctx, rays, t1, i1)
            assert rc == 0, ctx._lib.cpt_last_error(ctx._h).decode()
            torch.cuda.synchronize()
            assert torch.equal(i0, i1) and torch.equal(t0, t1), n
            assert (i1 >= 0).float().mean().item() > 0.5
        big = torch.zeros((L.NOALLOC_MAX_RAYS + 1, 8), dtype=torch.float32, device=dev)
        tb = torch.empty(big.shape[0], dtype=torch.float32, device=dev)
        ib = torch.empty(big.shape[0], dtype=torch.int32, device=dev)
        assert noalloc(ctx, big, tb, ib) == L.CPT_ERR_ARG
