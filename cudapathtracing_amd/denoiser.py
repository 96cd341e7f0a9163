"""The learned denoiser (BF16 MFMA U-Net) over the C ABI.

Inputs per pixel are 10 floats: noisy radiance rgb, albedo rgb, normal xyz,
depth - what :meth:`Context.render` returns as ``rgb`` and ``aov``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .context import CptError, Context, _is_torch, _ptr


def to_bf16_bits(a: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns (round to nearest even)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


def pack_input(rgb: np.ndarray, aov: np.ndarray) -> np.ndarray:
    """(H,W,3) radiance + (H,W,7) albedo/normal/depth -> (H,W,10)."""
    return np.ascontiguousarray(np.concatenate([rgb, aov], axis=-1), np.float32)


class Denoiser:
    def __init__(self, ctx: Context, max_batch: int, max_h: int, max_w: int, seed: int = 1):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        rc = self._lib.cpt_denoiser_create(ctx._h, max_batch, max_h, max_w, seed, C.byref(h))
        if rc:
            raise CptError(rc, self._lib.cpt_last_error(ctx._h).decode())
        self._h = h
        self.n_params = int(self._lib.cpt_denoiser_num_params(h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.cpt_denoiser_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            h = getattr(self.ctx, "_h", None)
            raise CptError(rc, self._lib.cpt_last_error(h).decode() if h else "the context is closed")

    def layers(self):
        n = 32
        cin, cout = (C.c_int32 * n)(), (C.c_int32 * n)()
        wo, bo = (C.c_int64 * n)(), (C.c_int64 * n)()
        k = self._lib.cpt_denoiser_layers(self._h, cin, cout, wo, bo)
        return [dict(cin=cin[i], cout=cout[i], w_off=wo[i], b_off=bo[i]) for i in range(k)]

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        self._check(self._lib.cpt_denoiser_get_params(self._h, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (self.n_params,)
        self._check(self._lib.cpt_denoiser_set_params(self._h, p.ctypes.data))

    def train_step(self, in10, target, lr: float, want_grads: bool = False):
        b, h, w = in10.shape[:3]
        dev = _is_torch(in10) and in10.is_cuda
        if dev:
            self.ctx._torch_stream()
        else:
            in10 = np.ascontiguousarray(in10, np.float32)
            target = np.ascontiguousarray(target, np.float32)
        loss = C.c_float()
        g = None
        if want_grads:
            if dev:  # device inputs: the gradients land in a device buffer too
                import torch
                g = torch.empty(self.n_params, dtype=torch.float32, device=in10.device)
            else:
                g = np.empty(self.n_params, np.float32)
        self._check(self._lib.cpt_denoiser_train_step(self._h, _ptr(in10), _ptr(target), b, h, w, lr,
                                                      C.byref(loss), _ptr(g), L.CPT_DEVICE_PTRS if dev else 0))
        if not want_grads:
            return float(loss.value)
        return float(loss.value), (g.cpu().numpy() if dev else g)

    def __call__(self, in10, out=None):
        b, h, w = in10.shape[:3]
        dev = _is_torch(in10) and in10.is_cuda
        if dev:
            import torch
            self.ctx._torch_stream()
            if out is None:
                out = torch.empty((b, h, w, 3), dtype=torch.float32, device=in10.device)
        else:
            in10 = np.ascontiguousarray(in10, np.float32)
            if out is None:
                out = np.empty((b, h, w, 3), np.float32)
        self._check(self._lib.cpt_denoise(self._h, _ptr(in10), _ptr(out), b, h, w,
                                          L.CPT_DEVICE_PTRS if dev else 0))
        return out

    def stats(self):
        f, s, fl = C.c_float(), C.c_float(), C.c_double()
        self._check(self._lib.cpt_denoiser_stats(self._h, C.byref(f), C.byref(s), C.byref(fl)))
        return dict(fwd_ms=f.value, step_ms=s.value, fwd_flops=fl.value)


def conv3x3(ctx: Context, mode: int, x_bits, w_bits, bias, dy_bits, B, H, W, Cin, Cout) -> np.ndarray:
    """Single 3x3 convolution on bf16 bit-pattern arrays (see include/cpt.h)."""
    shape = {0: (B, H, W, Cout), 1: (Cout, 3, 3, Cin), 2: (B, H, W, Cin)}[mode]
    out = np.empty(shape, np.float32)
    keep = [None if a is None else np.ascontiguousarray(a) for a in (x_bits, w_bits, bias, dy_bits)]
    ptr = [None if a is None else a.ctypes.data for a in keep]
    rc = ctx._lib.cpt_conv3x3(ctx._h, mode, ptr[0], ptr[1], ptr[2], ptr[3], out.ctypes.data, B, H, W, Cin, Cout)
    if rc:
        raise CptError(rc, ctx._lib.cpt_last_error(ctx._h).decode())
    return out
