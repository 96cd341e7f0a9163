"""ctypes wrapper of liboracle.so, the FP64 CPU restatement of the reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / reported CPU baseline.  The
library (cudapathtracing_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

P = C.c_void_p
dp = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            build()
        L = C.CDLL(ORACLE_PATH)
        sig = {
            "orc_scene_create": (P, [P, C.c_int, P, C.c_int, P, C.c_int, P, P, C.c_int64]),
            "orc_scene_destroy": (None, [P]),
            "orc_scene_count": (C.c_int64, [P]),
            "orc_expand_bits": (C.c_uint, [C.c_uint]),
            "orc_morton3d": (C.c_uint64, [dp, dp, dp, C.c_int]),
            "orc_expand_bounds": (None, [dp, dp, dp, dp]),
            "orc_scene_prims": (None, [P, P, P]),
            "orc_morton_keys": (None, [P, dp, dp, P]),
            "orc_build": (C.c_int, [P, dp, dp, P, P, P, P, P, P, P]),
            "orc_trace_ref": (C.c_int, [P, P, C.c_int64, P, P, P]),
            "orc_trace_brute": (C.c_int, [P, P, C.c_int64, C.c_double, C.c_double, P, P, P]),
            "orc_prim_t": (C.c_int, [P, P, P, C.c_int64, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        _register_render(L)
    return _lib


def _register_render(L):
    for name, (res, args) in RENDER_SIGS.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


RENDER_SIGS = {
    "orc_philox4x32_10": (None, [P, C.c_uint32, C.c_uint32, P]),
    "orc_render": (C.c_int, [P, P, P, P, P, C.c_int]),
    "orc_render_view": (C.c_int, [P, P, P, P, P, C.c_int, C.c_int64]),
    "orc_render_view_cert": (C.c_int, [P, P, P, P, P, P, C.c_int, C.c_int64]),
    "orc_volume_walk": (C.c_int, [P, P, P, C.c_double, C.c_uint64, C.c_int, P, P, P, P]),
    "orc_scene_set_medium": (C.c_int, [P, P, C.c_int, C.c_int, C.c_int, P, P, C.c_double, P]),
    "orc_medium_transmittance": (C.c_int, [P, P, P, C.c_double, C.c_double, C.c_uint64, C.c_int, P]),
    "orc_medium_collide": (C.c_int, [P, P, P, C.c_double, C.c_double, C.c_uint64, C.c_int, P]),
    "orc_bsdf": (C.c_int, [P, dp, dp, C.c_int, dp, dp, dp, dp]),
    "orc_bsdf_eval": (C.c_int, [P, dp, dp, C.c_int, dp, dp]),
    "orc_ggx_d": (C.c_double, [P, dp, dp]),
    "orc_conv3x3_fwd": (None, [P, P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_conv3x3_bwd_w": (None, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_conv3x3_bwd_x": (None, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "orc_unet_num_params": (C.c_int64, []),
    "orc_unet_grad": (C.c_double, [P, P, P, C.c_int, C.c_int, C.c_int, C.c_int, P, P]),
}


def conv3x3_fwd(x, w, bias):
    B, H, W, Cin = x.shape
    Cout = w.shape[0]
    x = np.ascontiguousarray(x, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    b = np.ascontiguousarray(bias, np.float64)
    y = np.empty((B, H, W, Cout), np.float64)
    lib().orc_conv3x3_fwd(x.ctypes.data, w.ctypes.data, b.ctypes.data, y.ctypes.data, B, H, W, Cin, Cout)
    return y


def conv3x3_bwd_w(dy, x):
    B, H, W, Cin = x.shape
    Cout = dy.shape[-1]
    dw = np.empty((Cout, 3, 3, Cin), np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    lib().orc_conv3x3_bwd_w(dy.ctypes.data, x.ctypes.data, dw.ctypes.data, B, H, W, Cin, Cout)
    return dw


def conv3x3_bwd_x(dy, w):
    B, H, W, Cout = dy.shape
    Cin = w.shape[-1]
    dx = np.empty((B, H, W, Cin), np.float64)
    dy = np.ascontiguousarray(dy, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    lib().orc_conv3x3_bwd_x(dy.ctypes.data, w.ctypes.data, dx.ctypes.data, B, H, W, Cin, Cout)
    return dx


def unet_grad(params, in10, target, emulate=True):
    """FP64 loss and parameter gradient of the denoiser network."""
    params = np.ascontiguousarray(params, np.float32)
    in10 = np.ascontiguousarray(in10, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    B, h, w, _ = in10.shape
    g = np.empty(params.shape[0], np.float64)
    pred = np.empty((B, h, w, 32), np.float32)
    loss = lib().orc_unet_grad(params.ctypes.data, in10.ctypes.data, target.ctypes.data, B, h, w,
                               1 if emulate else 0, g.ctypes.data, pred.ctypes.data)
    return loss, g, pred


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, key[0], key[1], o)
    return list(o)


def bsdf_sample(mat, ray_dir, normal, front, r):
    d, pdf, f = (C.c_double * 3)(), C.c_double(), (C.c_double * 3)()
    lib().orc_bsdf(C.byref(mat), d3(ray_dir), d3(normal), int(front), d3(r), d, C.byref(pdf), f)
    return np.array(d[:]), pdf.value, np.array(f[:])


def bsdf_eval(mat, ray_dir, normal, front, direction):
    f = (C.c_double * 3)()
    lib().orc_bsdf_eval(C.byref(mat), d3(ray_dir), d3(normal), int(front), d3(direction), f)
    return np.array(f[:])


def ggx_d(mat, n, h):
    return lib().orc_ggx_d(C.byref(mat), d3(n), d3(h))


def d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def expand_bits(v: int) -> int:
    return int(lib().orc_expand_bits(v))


def morton3d(c, smin, smax, index: int) -> int:
    return int(lib().orc_morton3d(d3(c), d3(smin), d3(smax), index))


def expand_bounds(mn, mx):
    a, b = (C.c_double * 3)(), (C.c_double * 3)()
    lib().orc_expand_bounds(d3(mn), d3(mx), a, b)
    return np.array(a[:]), np.array(b[:])


class OracleScene:
    """FP64 scene + LBVH + traversal over the same inputs as the HIP path."""

    def __init__(self, s):
        L = lib()
        self._keep = [(type(s.materials[0]) * len(s.materials))(*s.materials) if s.materials else None,
                      (type(s.lights[0]) * len(s.lights))(*s.lights) if s.lights else None,
                      (type(s.spheres[0]) * len(s.spheres))(*s.spheres) if s.spheres else None,
                      np.ascontiguousarray(s.tri, np.float64),
                      np.ascontiguousarray(s.tri_mat, np.int32)]
        m, l, sp, tri, tm = self._keep
        self._h = L.orc_scene_create(C.cast(m, P) if m else None, len(s.materials),
                                     C.cast(l, P) if l else None, len(s.lights),
                                     C.cast(sp, P) if sp else None, len(s.spheres),
                                     tri.ctypes.data, tm.ctypes.data, tri.shape[0])
        self.n = int(L.orc_scene_count(self._h))
        self.scene = s
        m = getattr(s, "medium", None)
        if m is not None:
            dens = np.ascontiguousarray(m.density, np.float32)
            self._keep.append(dens)
            nz, ny, nx = dens.shape
            d3 = C.c_double * 3
            L.orc_scene_set_medium(self._h, dens.ctypes.data, nx, ny, nz, d3(*m.bmin), d3(*m.bmax),
                                   float(m.sigma), d3(*m.albedo))

    def __del__(self):
        try:
            lib().orc_scene_destroy(self._h)
        except Exception:
            pass

    def prims(self):
        c = np.empty((self.n, 3), np.float64)
        b = np.empty((self.n, 6), np.float64)
        lib().orc_scene_prims(self._h, c.ctypes.data, b.ctypes.data)
        return c, b

    def morton_keys(self, mn, mx):
        k = np.empty(self.n, np.uint64)
        lib().orc_morton_keys(self._h, d3(mn), d3(mx), k.ctypes.data)
        return k

    def build(self, mn, mx) -> dict:
        n = self.n
        keys = np.empty(n, np.uint64)
        order = np.empty(n, np.uint32)
        ca = np.empty(max(n - 1, 0), np.int32)
        cb = np.empty(max(n - 1, 0), np.int32)
        boxes = np.empty((max(n - 1, 0), 12), np.float64)
        root = np.empty(6, np.float64)
        depth = C.c_int()
        lib().orc_build(self._h, d3(mn), d3(mx), keys.ctypes.data, order.ctypes.data, ca.ctypes.data,
                        cb.ctypes.data, boxes.ctypes.data, root.ctypes.data, C.byref(depth))
        return dict(keys=keys, order=order, child_a=ca, child_b=cb, boxes=boxes, root_box=root,
                    max_depth=depth.value)

    def trace_ref(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        t = np.empty(n, np.float64)
        ids = np.empty(n, np.int32)
        v = np.empty(n, np.int64)
        lib().orc_trace_ref(self._h, rays.ctypes.data, n, t.ctypes.data, ids.ctypes.data, v.ctypes.data)
        return t, ids, v

    def trace_brute(self, rays: np.ndarray, eb=1e-3, et=1e-4):
        rays = np.ascontiguousarray(rays, np.float32)
        n = rays.shape[0]
        t = np.empty(n, np.float64)
        ids = np.empty(n, np.int32)
        cert = np.empty(n, np.uint8)
        lib().orc_trace_brute(self._h, rays.ctypes.data, n, eb, et, t.ctypes.data, ids.ctypes.data,
                              cert.ctypes.data)
        return t, ids, cert.astype(bool)

    def render(self, cam, params, per_sample=False, nthreads=0, pixel_offset=0, cert=False):
        """Mean radiance per pixel (h, w, 3) float64 (display values for the
        NORMAL / DEPTH profiles); per-sample radiance (npix, spp, 3) when
        per_sample; the per-sample path certificates (npix, spp) uint8 when
        cert (0 = certified, else the class of the first decision that was
        not robust).  pixel_offset: the view's first pixel index in a
        multi-view stream."""
        h, w = cam.height, cam.width
        spp = 1 if params.profile in (3, 4) else params.spp
        rgb = np.empty((h, w, 3), np.float64)
        ps = np.empty((h * w, spp, 3), np.float64) if per_sample else None
        ct = np.empty((h * w, spp), np.uint8) if cert else None
        rc = lib().orc_render_view_cert(self._h, C.byref(cam), C.byref(params), rgb.ctypes.data,
                                        ps.ctypes.data if per_sample else None,
                                        ct.ctypes.data if cert else None, nthreads, pixel_offset)
        if rc:
            raise RuntimeError("orc_render: build the oracle BVH first")
        out = (rgb,)
        if per_sample:
            out += (ps,)
        if cert:
            out += (ct,)
        return out if len(out) > 1 else rgb

    def volume_walk(self, o, d, t_hit, seed, n):
        """n walks of the volume object hit at o + t_hit d: continuing ray
        origins (n, 3), directions (n, 3), throughput (n,), collisions (n,)."""
        ro = np.empty((n, 3), np.float64)
        rd = np.empty((n, 3), np.float64)
        thr = np.empty(n, np.float64)
        nc = np.empty(n, np.int32)
        d3_ = C.c_double * 3
        if lib().orc_volume_walk(self._h, d3_(*o), d3_(*d), t_hit, seed, n, ro.ctypes.data, rd.ctypes.data,
                                 thr.ctypes.data, nc.ctypes.data):
            raise RuntimeError("orc_volume_walk: build the oracle BVH first")
        return ro, rd, thr, nc

    def medium_transmittance(self, o, d, tmin, tmax, seed, n):
        """n ratio-tracking transmittance estimates along o + t d."""
        out = np.empty(n, np.float64)
        d3 = C.c_double * 3
        if lib().orc_medium_transmittance(self._h, d3(*o), d3(*d), tmin, tmax, seed, n, out.ctypes.data):
            raise RuntimeError("no medium")
        return out

    def medium_collide(self, o, d, tmin, tmax, seed, n):
        """n delta-tracking runs: collision t or -1."""
        out = np.empty(n, np.float64)
        d3 = C.c_double * 3
        if lib().orc_medium_collide(self._h, d3(*o), d3(*d), tmin, tmax, seed, n, out.ctypes.data):
            raise RuntimeError("no medium")
        return out

    def prim_t(self, rays: np.ndarray, ids: np.ndarray):
        rays = np.ascontiguousarray(rays, np.float32)
        ids = np.ascontiguousarray(ids, np.int32)
        t = np.empty(rays.shape[0], np.float64)
        lib().orc_prim_t(self._h, rays.ctypes.data, ids.ctypes.data, rays.shape[0], t.ctypes.data)
        return t
