"""Host-side object over the C ABI: one ``cpt_ctx`` (HIP stream + device
buffers) per :class:`Context`.

Mirrors the reference's call sequence: ``Scene::init`` uploads primitives in
insertion order and calls ``Device::buildBVH`` (Scene.cuh:72-145,
Device.cuh:169-195); rendering launches ``render`` (Render.cu:16-261).
Errors raise :class:`CptError` with the library's message.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _lib as L
from .scenes import SceneData


class CptError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cpt error {code}: {msg}")
        self.code = code


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _ptr(a) -> int:
    if a is None:
        return None
    if _is_torch(a):
        return a.data_ptr()
    return a.ctypes.data


class Context:
    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        rc = self._lib.cpt_create(C.byref(h), device)
        if rc:
            raise CptError(rc, f"cpt_create(device={device}) failed (no usable HIP device?)")
        self._h = h
        self.n_prims = 0

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.cpt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int):
        if rc != 0:
            raise CptError(rc, self._lib.cpt_last_error(self._h).decode())

    def set_stream(self, stream_handle: Optional[int]):
        self._check(self._lib.cpt_set_stream(self._h, stream_handle))

    def set_traversal_width(self, width: int):
        """4 (default): four-wide records; 2: binary LBVH nodes."""
        self._check(self._lib.cpt_set_option(self._h, 1, width))

    def set_quantized(self, on: bool):
        """Four-wide walks on 64 B quantized records (default) or 128 B FP32 records."""
        self._check(self._lib.cpt_set_option(self._h, 4, 1 if on else 0))

    def set_persistent(self, on: bool):
        """Persistent lane-refilling traversal kernels (default on)."""
        self._check(self._lib.cpt_set_option(self._h, 3, 1 if on else 0))

    def set_persistent_shadow(self, on: bool):
        """cpt_render's four-wide occlusion walks always as persistent
        lane-refilling kernels in append order (on) or never (off: one lane
        per shadow ray, ordered by origin cell for long queues);
        set_persistent_shadow_min sets the automatic choice."""
        self.set_persistent_shadow_min(1 if on else 0)

    def set_persistent_shadow_min(self, n: int = L.PERSISTENT_SHADOW_MIN_DEFAULT):
        """The persistent occlusion walk for a bounce with at least n shadow
        rays in a scene without a medium, the one-lane kernel otherwise (1:
        always, 0: never; without an argument the default, 2^22)."""
        self._check(self._lib.cpt_set_option(self._h, 8, int(n)))

    def set_queue_sort_min(self, n: int):
        """cpt_render bounce queues of at least n paths are traced in ray-key order (0: never, the default)."""
        self._check(self._lib.cpt_set_option(self._h, 5, n))

    def set_shadow_sort_min(self, n: int):
        """cpt_render shadow queues of at least n rays are traced in origin-cell order (0: never)."""
        self._check(self._lib.cpt_set_option(self._h, 6, n))

    def set_cluster(self, mode: int):
        """Four-wide records over a clustered (PLOC) tree of the same leaves
        where its surface-area cost is clearly lower: 2 (default) built at the
        first render after a build or refit, 1 after every build or refit,
        0 never (True/False mean 1/0)."""
        self._check(self._lib.cpt_set_option(self._h, 7, int(mode)))

    def set_ray_sort_min(self, n: int):
        """Trace batches of at least n rays in ray-key order (0: never)."""
        self._check(self._lib.cpt_set_option(self._h, 2, n))

    def synchronize(self):
        self._check(self._lib.cpt_synchronize(self._h))

    def request_abort(self):
        """Stops the running or next call (it raises CptError with code
        CPT_ERR_ABORTED) until clear_abort; callable from another thread."""
        self._check(self._lib.cpt_request_abort(self._h))

    def clear_abort(self):
        self._check(self._lib.cpt_clear_abort(self._h))

    def _torch_stream(self, *tensors):
        """Runs the library on torch's current stream when given device tensors."""
        import torch
        self.set_stream(torch.cuda.current_stream().cuda_stream)

    # -- scene --------------------------------------------------------------
    def upload(self, s: SceneData):
        mats = (L.Material * max(1, len(s.materials)))(*s.materials)
        lights = (L.Light * max(1, len(s.lights)))(*s.lights)
        sph = (L.Sphere * max(1, len(s.spheres)))(*s.spheres)
        tri = np.ascontiguousarray(s.tri, np.float64)
        tm = np.ascontiguousarray(s.tri_mat, np.int32)
        self._check(self._lib.cpt_scene_upload(self._h, mats, len(s.materials), lights, len(s.lights),
                                               sph, len(s.spheres), tri.ctypes.data, tm.ctypes.data,
                                               tri.shape[0]))
        self.n_prims = s.n_prims
        self.scene = s
        if s.medium is not None:
            self.set_medium(s.medium)

    def set_medium(self, m):
        """Heterogeneous medium (scenes.Medium) of the uploaded scene; None removes it."""
        if m is None:
            self._check(self._lib.cpt_scene_set_medium(self._h, None, 0, 0, 0, None, None, 0.0, None))
            return
        dens = np.ascontiguousarray(m.density, np.float32)
        nz, ny, nx = dens.shape
        self._check(self._lib.cpt_scene_set_medium(self._h, dens.ctypes.data, nx, ny, nz, L.d3(m.bmin),
                                                   L.d3(m.bmax), float(m.sigma), L.d3(m.albedo)))

    def load_ini(self, path: str):
        mn, mx = (C.c_double * 3)(), (C.c_double * 3)()
        cam = L.CameraDesc()
        self._check(self._lib.cpt_scene_load_ini(self._h, path.encode(), mn, mx, C.byref(cam)))
        self.n_prims = int(self._lib.cpt_scene_count(self._h))
        return np.array(mn[:]), np.array(mx[:]), cam

    # -- LBVH ---------------------------------------------------------------
    def build(self, bounds_min, bounds_max) -> L.BuildStats:
        st = L.BuildStats()
        self._check(self._lib.cpt_bvh_build(self._h, L.d3(bounds_min), L.d3(bounds_max), C.byref(st)))
        return st

    def update_triangles(self, first: int, tri: np.ndarray):
        """Replaces triangles [first, first+n) (indices after lights and spheres)."""
        tri = np.ascontiguousarray(tri, np.float64).reshape(-1, 9)
        self._check(self._lib.cpt_scene_update_triangles(self._h, first, tri.ctypes.data, tri.shape[0]))

    def refit(self) -> L.BuildStats:
        st = L.BuildStats()
        self._check(self._lib.cpt_bvh_refit(self._h, C.byref(st)))
        return st

    def cluster_info(self) -> dict:
        """Which tree the four-wide records group after the last build or
        refit, the clustered tree's depth and both trees' surface-area costs."""
        u, d = C.c_int32(), C.c_int32()
        a, b = C.c_double(), C.c_double()
        self._check(self._lib.cpt_bvh_cluster_info(self._h, C.byref(u), C.byref(d), C.byref(a), C.byref(b)))
        return {"in_use": bool(u.value), "depth": d.value, "lbvh_cost": a.value, "cluster_cost": b.value}

    def bilateral(self, rgb, aov, sigma_g=8.0, sigma_r=5.0, sigma_n=0.1, sigma_d=0.2):
        """Joint bilateral filter of Render.cu:353-398 (host arrays)."""
        h, w = rgb.shape[:2]
        rgb = np.ascontiguousarray(rgb, np.float32)
        aov = np.ascontiguousarray(aov, np.float32)
        out = np.empty((h, w, 3), np.float32)
        self._check(self._lib.cpt_bilateral(self._h, rgb.ctypes.data, aov.ctypes.data, out.ctypes.data, w, h,
                                            sigma_g, sigma_r, sigma_n, sigma_d, 0))
        return out

    def temporal_blend(self, rgb, prev_rgb, camera: L.CameraDesc, prev_camera: L.CameraDesc, position=None,
                       aov=None, moving=False):
        """addPrevious (Render.cu:400-421): blends ``rgb`` with ``prev_rgb``
        reprojected through ``prev_camera`` (host arrays, h x w x 3).
        ``position`` (h x w x 3 first-hit points, x + y > 999999 = miss) or
        ``aov`` (its depth channel) locates each pixel's surface."""
        if position is None and aov is None:
            raise ValueError("temporal_blend needs position or aov")
        h, w = rgb.shape[:2]
        cam = L.Camera()
        self._check(self._lib.cpt_camera_from_desc(C.byref(camera), C.byref(cam)))
        if (cam.width, cam.height) != (w, h):
            raise ValueError(f"image is {w}x{h}, camera is {cam.width}x{cam.height}")
        rgb = np.ascontiguousarray(rgb, np.float32)
        prev_rgb = np.ascontiguousarray(prev_rgb, np.float32)
        if prev_rgb.shape != rgb.shape:
            raise ValueError("prev_rgb and rgb differ in shape")
        pos = None if position is None else np.ascontiguousarray(position, np.float32).reshape(h, w, 3)
        av = None if aov is None else np.ascontiguousarray(aov, np.float32).reshape(h, w, 7)
        out = np.empty((h, w, 3), np.float32)
        self._check(self._lib.cpt_temporal_blend(self._h, rgb.ctypes.data, prev_rgb.ctypes.data,
                                                 None if pos is None else pos.ctypes.data,
                                                 None if av is None else av.ctypes.data, C.byref(cam),
                                                 C.byref(prev_camera), 1 if moving else 0, out.ctypes.data, 0))
        return out

    def export(self) -> dict:
        n = self.n_prims
        keys = np.empty(n, np.uint64)
        order = np.empty(n, np.uint32)
        ca = np.empty(max(n - 1, 0), np.int32)
        cb = np.empty(max(n - 1, 0), np.int32)
        boxes = np.empty((max(n - 1, 0), 12), np.float32)
        root = np.empty(6, np.float32)
        self._check(self._lib.cpt_bvh_export(self._h, keys.ctypes.data, order.ctypes.data,
                                             ca.ctypes.data, cb.ctypes.data, boxes.ctypes.data,
                                             root.ctypes.data))
        return dict(keys=keys, order=order, child_a=ca, child_b=cb, boxes=boxes, root_box=root)

    def morton(self, bounds_min, bounds_max) -> np.ndarray:
        keys = np.empty(self.n_prims, np.uint64)
        self._check(self._lib.cpt_bvh_morton(self._h, L.d3(bounds_min), L.d3(bounds_max),
                                             keys.ctypes.data, 0))
        return keys

    def sort_pairs(self, keys, vals, bits: int = 32):
        """Stable radix sort in place (numpy arrays or device tensors)."""
        dev = _is_torch(keys) and keys.is_cuda
        if dev:
            self._torch_stream()
        n = int(keys.shape[0])
        self._check(self._lib.cpt_radix_sort_pairs(self._h, _ptr(keys), _ptr(vals), n, bits,
                                                   L.CPT_DEVICE_PTRS if dev else 0))

    # -- traversal ------------------------------------------------------------
    def trace(self, rays, t_out=None, id_out=None) -> Tuple:
        """Closest hit for rays (n x 8 float32: o, tmin, d, tmax)."""
        n = int(rays.shape[0])
        dev = _is_torch(rays) and rays.is_cuda
        if dev:
            import torch
            self._torch_stream()
            if t_out is None:
                t_out = torch.empty(n, dtype=torch.float32, device=rays.device)
            if id_out is None:
                id_out = torch.empty(n, dtype=torch.int32, device=rays.device)
        else:
            rays = np.ascontiguousarray(rays, np.float32)
            if t_out is None:
                t_out = np.empty(n, np.float32)
            if id_out is None:
                id_out = np.empty(n, np.int32)
        self._check(self._lib.cpt_trace_closest(self._h, _ptr(rays), n, _ptr(t_out), _ptr(id_out),
                                                L.CPT_DEVICE_PTRS if dev else 0))
        return t_out, id_out

    def trace_count(self, rays):
        n = int(rays.shape[0])
        dev = _is_torch(rays) and rays.is_cuda
        if dev:
            import torch
            self._torch_stream()
            a = torch.empty(n, dtype=torch.int32, device=rays.device)
            b = torch.empty(n, dtype=torch.int32, device=rays.device)
        else:
            rays = np.ascontiguousarray(rays, np.float32)
            a = np.empty(n, np.int32)
            b = np.empty(n, np.int32)
        self._check(self._lib.cpt_trace_count(self._h, _ptr(rays), n, _ptr(a), _ptr(b),
                                              L.CPT_DEVICE_PTRS if dev else 0))
        return a, b

    def last_trace_ms(self) -> float:
        return float(self._lib.cpt_last_trace_ms(self._h))

    def last_trace_stats(self):
        """(total, ray ordering, traversal kernel) ms of the last trace call."""
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self._check(self._lib.cpt_last_trace_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    # -- path tracing -----------------------------------------------------------
    def render(self, camera_desc: L.CameraDesc, spp: int, seed: int = 3317,
               profile: int = L.PROFILE_CONFIG, max_depth: int = 8, rr_depth: int = 3,
               rr_prob: float = 0.7, sample_base: int = 0, aov: bool = False,
               paths_per_batch: int = 0, out=None, aov_out=None, count_work: bool = False):
        cam = camera_desc if isinstance(camera_desc, L.Camera) else camera(camera_desc)
        p = L.RenderParams(seed, spp, max_depth, rr_depth, profile, rr_prob, sample_base,
                           1 if aov else 0, paths_per_batch, L.RENDER_COUNT_WORK if count_work else 0)
        h, w = cam.height, cam.width
        dev = out is not None and _is_torch(out) and out.is_cuda
        if dev:
            self._torch_stream()
        if out is None:
            out = np.empty((h, w, 3), np.float32)
        if aov and aov_out is None:
            aov_out = np.empty((h, w, 7), np.float32)
        st = L.RenderStats()
        self._check(self._lib.cpt_render(self._h, C.byref(cam), C.byref(p), _ptr(out),
                                         _ptr(aov_out) if aov else None, C.byref(st),
                                         L.CPT_DEVICE_PTRS if dev else 0))
        return out, (aov_out if aov else None), st


def _render_views_impl(self, cams, spp, seed=3317, profile=L.PROFILE_CONFIG, max_depth=8, rr_depth=3,
                       rr_prob=0.7, sample_base=0, aov=False, paths_per_batch=0, out=None, aov_out=None,
                       pack=False):
    cl = [c if isinstance(c, L.Camera) else camera(c) for c in cams]
    n = len(cl)
    arr = (L.Camera * n)(*cl)
    if pack and (not aov or aov_out is not None):
        raise ValueError("pack=True writes the AOVs into out: it needs aov=True and no aov_out")
    p = L.RenderParams(seed, spp, max_depth, rr_depth, profile, rr_prob, sample_base, 1 if aov else 0,
                       paths_per_batch, L.RENDER_PACK_AOV if pack else 0)
    h, w = cl[0].height, cl[0].width
    dev = out is not None and _is_torch(out) and out.is_cuda
    if dev:
        self._torch_stream()
    if out is None:
        out = np.empty((n, h, w, 10 if pack else 3), np.float32)
    if aov and not pack and aov_out is None:
        aov_out = np.empty((n, h, w, 7), np.float32)
    st = L.RenderStats()
    self._check(self._lib.cpt_render_views(self._h, arr, n, C.byref(p), _ptr(out),
                                           _ptr(aov_out) if aov and not pack else None,
                                           C.byref(st), L.CPT_DEVICE_PTRS if dev else 0))
    return out, (aov_out if aov and not pack else None), st


Context.render_views = _render_views_impl
Context.render_views.__doc__ = """Renders several same-size views in one wavefront pipeline;
returns (nviews, h, w, 3) radiance, optional AOVs and the stats.  pack=True
(with aov=True) writes (nviews, h, w, 10) into out: radiance then the 7 AOVs
per pixel, the learned denoiser's input layout."""


def camera(desc: L.CameraDesc) -> L.Camera:
    cam = L.Camera()
    rc = L.lib().cpt_camera_from_desc(C.byref(desc), C.byref(cam))
    if rc:
        raise CptError(rc, "invalid camera")
    return cam


def tonemap_rgba8(rgb: np.ndarray) -> np.ndarray:
    h, w = rgb.shape[:2]
    out = np.empty((h, w, 4), np.uint8)
    rgb = np.ascontiguousarray(rgb, np.float32)
    rc = L.lib().cpt_tonemap_rgba8(rgb.ctypes.data, w, h, out.ctypes.data)
    if rc:
        raise CptError(rc, "tonemap")
    return out


def write_pfm(path: str, rgb: np.ndarray):
    """Float framebuffer (h x w x 3, row 0 = bottom) as a PFM file."""
    h, w = rgb.shape[:2]
    rgb = np.ascontiguousarray(rgb, np.float32)
    rc = L.lib().cpt_write_pfm(path.encode(), rgb.ctypes.data, w, h)
    if rc:
        raise CptError(rc, f"cannot write {path}")


def write_png(path: str, rgb: np.ndarray, gamma: bool = True):
    """The reference's displayed bytes (pic2RGBW with gamma, the NORMAL /
    DEPTH display without) as an 8-bit RGBA PNG."""
    h, w = rgb.shape[:2]
    rgb = np.ascontiguousarray(rgb, np.float32)
    rc = L.lib().cpt_write_png(path.encode(), rgb.ctypes.data, w, h, 1 if gamma else 0)
    if rc:
        raise CptError(rc, f"cannot write {path}")
