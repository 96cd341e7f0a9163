"""Scene descriptions in the reference's terms and the seeded workloads.

A :class:`SceneData` is what ``Scene::init`` (CudaPathTracing/Scene.cuh:72-145)
assembles: lights, spheres and triangles in insertion order, a material per
object, the Morton bounds taken from the *untranslated* mesh vertices
(ObjLoader.cuh:107-112) and a camera.  Floors are split into the triangles
(lt, lb, rt) and (rt, lb, rb) as allocateFloorOnDevice does
(Device.cuh:40-44).

Workloads (sizes fixed here, used by tests and bench.py):

* ``cornell_icosphere``  - Cornell box of 10 wall triangles + a 0.25 x 0.25
  quad light (radiance 15) + a level-3 icosphere (1280 triangles, r = 0.3),
  Lambert albedo U[0.2, 0.8] per object (seed 0); 256 x 256, 16 spp, depth 8.
* ``displaced_icosphere`` - level-8 icosphere (1,310,720 triangles) displaced
  by 4-octave value noise (amplitude 0.1) in the same box; GGX with roughness
  U[0.05, 0.8] and F0 U[0.02, 0.9] per 64-triangle patch; 1920 x 1080, 64 spp.
* ``triangle_soup`` - 10M triangles, centroids U[0,1]^3, lognormal side
  (mu -6, sigma 0.5), random orientation (seed 1); 2^25 rays with origins
  U[-0.5, 1.5]^3 and uniform directions (seed 2).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L


@dataclass
class Medium:
    """Heterogeneous medium: density grid in [0,1] of shape (nz, ny, nx),
    sigma_t = sigma * density of the nearest voxel, over [bmin, bmax]."""
    density: np.ndarray
    bmin: tuple = (-0.5, -0.5, -0.5)
    bmax: tuple = (0.5, 0.5, 0.5)
    sigma: float = 8.0
    albedo: tuple = (0.8, 0.8, 0.8)


def fbm_grid(n: int, seed: int = 0) -> np.ndarray:
    """Seeded 4-octave value-noise fBm density grid, shape (n, n, n)."""
    out = np.empty((n, n, n), np.float32)
    rc = L.lib().cpt_gen_fbm_grid(n, seed, out.ctypes.data)
    if rc:
        raise ValueError("cpt_gen_fbm_grid: bad size")
    return out


@dataclass
class SceneData:
    materials: List[L.Material] = field(default_factory=list)
    lights: List[L.Light] = field(default_factory=list)
    spheres: List[L.Sphere] = field(default_factory=list)
    tri: np.ndarray = field(default_factory=lambda: np.zeros((0, 9), np.float64))
    tri_mat: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    bounds_min: np.ndarray = field(default_factory=lambda: np.full(3, 999999.0))
    bounds_max: np.ndarray = field(default_factory=lambda: np.full(3, -999999.0))
    camera: Optional[L.CameraDesc] = None
    medium: Optional[Medium] = None

    @property
    def n_prims(self) -> int:
        return len(self.lights) + len(self.spheres) + int(self.tri.shape[0])

    def add_material(self, color, mtype, alpha_x=0.5, alpha_y=0.5, eta=1.51) -> int:
        self.materials.append(L.Material(L.d3(color), alpha_x, alpha_y, eta, mtype, 0))
        return len(self.materials) - 1

    def add_light(self, center, width, height, normal, color, visible=True):
        # Scene::addOneLight (Scene.cuh:175-187)
        self.lights.append(L.Light(L.d3(center), width, height, L.d3(normal), L.d3(color),
                                   1 if visible else 0, 0))

    def add_sphere(self, center, radius, material: int):
        # Scene::addOneSphere (Scene.cuh:189-200)
        self.spheres.append(L.Sphere(L.d3(center), radius, material, 0))

    def add_floor(self, lt, rt, lb, rb, material: int):
        # Scene::addFloor -> allocateFloorOnDevice (Device.cuh:40-44)
        t = np.array([[*lt, *lb, *rt], [*rt, *lb, *rb]], np.float64)
        self._append_tris(t, material)

    def add_mesh(self, tri: np.ndarray, material, untranslated: Optional[np.ndarray] = None):
        """Appends a triangle mesh.  ``untranslated`` are the mesh vertices
        before the [Mesh] centre offset (ObjLoader.cuh:107-112 bounds); by
        default the triangles themselves."""
        src = tri if untranslated is None else untranslated
        v = np.asarray(src, np.float64).reshape(-1, 3)
        if v.shape[0]:
            self.bounds_min = np.minimum(self.bounds_min, v.min(axis=0))
            self.bounds_max = np.maximum(self.bounds_max, v.max(axis=0))
        self._append_tris(np.asarray(tri, np.float64).reshape(-1, 9), material)

    def _append_tris(self, t: np.ndarray, material):
        m = np.broadcast_to(np.asarray(material, np.int32), (t.shape[0],)).astype(np.int32)
        self.tri = np.ascontiguousarray(np.concatenate([self.tri, t], axis=0))
        self.tri_mat = np.ascontiguousarray(np.concatenate([self.tri_mat, m]))


def load_ini(path: str) -> SceneData:
    """IniParser::Parse + Scene::init (IniParser.cpp:3-96, Scene.cuh:72-145)
    on the host: objects in the reference's order, [Mesh] OBJ files loaded
    with the reference's rotate/scale/bounds/translate rule."""
    lib = L.lib()
    cnt = (C.c_int64 * 4)()
    rc = lib.cpt_ini_counts(path.encode(), cnt)
    if rc:
        raise OSError(f"cannot read scene file {path} ({rc})")
    nm, nl, ns, nt = (int(x) for x in cnt)
    mats = (L.Material * max(nm, 1))()
    lights = (L.Light * max(nl, 1))()
    sph = (L.Sphere * max(ns, 1))()
    tri = np.empty((nt, 9), np.float64)
    tm = np.empty(nt, np.int32)
    mn, mx = (C.c_double * 3)(), (C.c_double * 3)()
    cam = L.CameraDesc()
    rc = lib.cpt_ini_read(path.encode(), mats, lights, sph, tri.ctypes.data, tm.ctypes.data, mn, mx,
                          C.byref(cam))
    if rc:
        raise OSError(f"cannot read scene file {path} ({rc})")
    s = SceneData()
    s.materials = list(mats[:nm])
    s.lights = list(lights[:nl])
    s.spheres = list(sph[:ns])
    s.tri, s.tri_mat = tri, tm
    s.bounds_min, s.bounds_max = np.array(mn[:]), np.array(mx[:])
    s.camera = cam
    return s


def icosphere(level: int, center=(0.0, 0.0, 0.0), radius=1.0, amp=0.0, seed=0) -> np.ndarray:
    lib = L.lib()
    n = lib.cpt_gen_icosphere_count(level)
    out = np.empty((n, 9), np.float64)
    c = (C.c_double * 3)(*center)
    rc = lib.cpt_gen_icosphere(level, c, radius, amp, seed, out.ctypes.data)
    if rc:
        raise ValueError(f"cpt_gen_icosphere failed ({rc})")
    return out


def triangle_soup(n: int, seed=1, mu=-6.0, sigma=0.5) -> np.ndarray:
    out = np.empty((n, 9), np.float64)
    rc = L.lib().cpt_gen_triangle_soup(n, seed, mu, sigma, out.ctypes.data)
    if rc:
        raise ValueError(f"cpt_gen_triangle_soup failed ({rc})")
    return out


def random_rays(n: int, seed=2, lo=-0.5, hi=1.5, tmin=0.001, tmax=999999.0) -> np.ndarray:
    out = np.empty((n, 8), np.float32)
    rc = L.lib().cpt_gen_random_rays(n, seed, lo, hi, tmin, tmax, out.ctypes.data)
    if rc:
        raise ValueError(f"cpt_gen_random_rays failed ({rc})")
    return out


def cornell_box(scene: SceneData, albedo_rng: np.random.Generator, mtype=L.M_LAMBERT):
    """Box [-0.5, 0.5]^3 open towards +z, 0.25 x 0.25 light under the ceiling."""
    scene.add_light((0.0, 0.5 - 1e-3, 0.0), 0.25, 0.25, (0.0, -1.0, 0.0), (15.0, 15.0, 15.0), True)
    h = 0.5
    walls = [  # (lt, rt, lb, rb), each facing the box interior
        ((-h, h, h), (-h, h, -h), (-h, -h, h), (-h, -h, -h)),  # left   x=-h
        ((-h, h, -h), (h, h, -h), (-h, -h, -h), (h, -h, -h)),  # back   z=-h
        ((h, h, -h), (h, h, h), (h, -h, -h), (h, -h, h)),  # right  x=+h
        ((-h, h, h), (h, h, h), (-h, h, -h), (h, h, -h)),  # top    y=+h
        ((-h, -h, -h), (h, -h, -h), (-h, -h, h), (h, -h, h)),  # bottom y=-h
    ]
    for w in walls:
        m = scene.add_material(albedo_rng.uniform(0.2, 0.8, 3), mtype)
        scene.add_floor(*w, material=m)


def box_tris(lo, hi) -> np.ndarray:
    """12 triangles (n x 9) of an axis-aligned box, wound so that every
    geometric normal (p1 - p0) x (p2 - p0) points out of the box (the front
    face of a dielectric solid is its outside)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = (lo + hi) / 2
    v = np.array([[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])])
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in quads:
        for t in ((a, b, cc), (a, cc, d)):
            p = v[list(t)]
            if np.dot(np.cross(p[1] - p[0], p[2] - p[0]), p.mean(axis=0) - c) < 0:
                p = p[[0, 2, 1]]
            tris.append(p.reshape(9))
    return np.array(tris)


def glass_column(width=24) -> SceneData:
    """A tall colour-1 specular-dielectric box (0.2 x 1 x 0.2) under a uniform
    sky (0.5, 0.7, 0.9), seen from above: light entering the top is trapped
    by total internal reflection for many vertices (a furnace test of the
    STATIC profile's roulette with no depth cap)."""
    s = SceneData()
    m = s.add_material((1, 1, 1), L.M_SPECULAR_DIELECTRIC)
    s.add_mesh(box_tris((-0.1, 0.0, -0.1), (0.1, 1.0, 0.1)), m)
    s.camera = L.CameraDesc(L.d3((0.5, 0.7, 0.9)), L.d3((0.35, 1.6, 0.45)), L.d3((0.0, 0.6, 0.0)), 30.0, 1.0,
                            width, 0)
    return s


def default_camera(width: int, height: int, fov=40.0) -> L.CameraDesc:
    return L.CameraDesc(L.d3((0, 0, 0)), L.d3((0.0, 0.0, 1.9)), L.d3((0.0, 0.0, 0.0)), fov,
                        width / height, width, 0)


def cornell_icosphere(width=256, height=256, seed=0) -> SceneData:
    rng = np.random.default_rng(seed)
    s = SceneData()
    cornell_box(s, rng)
    m = s.add_material(rng.uniform(0.2, 0.8, 3), L.M_LAMBERT)
    s.add_mesh(icosphere(3, (0.0, 0.0, 0.0), 0.3), m)
    s.camera = default_camera(width, height)
    return s


def displaced_icosphere(width=1920, height=1080, seed=0, level=8) -> SceneData:
    rng = np.random.default_rng(seed)
    s = SceneData()
    cornell_box(s, rng)
    tri = icosphere(level, (0.0, 0.0, 0.0), 0.3, amp=0.1, seed=seed)
    n = tri.shape[0]
    patch = 64
    npatch = (n + patch - 1) // patch
    base = len(s.materials)
    rough = rng.uniform(0.05, 0.8, npatch)
    f0 = rng.uniform(0.02, 0.9, npatch)
    for k in range(npatch):
        s.add_material((f0[k], f0[k], f0[k]), L.M_GGX, float(rough[k]), float(rough[k]))
    mats = base + (np.arange(n) // patch)
    s.add_mesh(tri, mats.astype(np.int32))
    s.camera = default_camera(width, height)
    return s


def displaced_icosphere_medium(width=1920, height=1080, seed=0, level=8, grid=256) -> SceneData:
    """The displaced-icosphere scene with a grid^3 fBm density medium
    (sigma_t max 8, albedo 0.8) filling the box."""
    s = displaced_icosphere(width, height, seed, level)
    s.medium = Medium(fbm_grid(grid, seed))
    return s


def triangle_soup_scene(n=10_000_000, seed=1) -> SceneData:
    s = SceneData()
    m = s.add_material((0.5, 0.5, 0.5), L.M_LAMBERT)
    s.add_mesh(triangle_soup(n, seed), m)
    return s
