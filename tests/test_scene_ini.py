"""Scene ingestion on the host (IniParser.cpp:3-127, Scene.cuh:72-270,
ObjLoader.cuh:57-163) against the reference's own scene files and small
hand-made OBJ meshes (no GPU)."""
import os

import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import scenes

GOLD = os.path.join(os.path.dirname(__file__), "golden", "scene_ini")


def test_scene2_cornell_without_meshes():
    s = scenes.load_ini(os.path.join(GOLD, "Scene2.ini"))
    assert len(s.lights) == 1 and len(s.spheres) == 0
    assert s.tri.shape == (10, 9)  # 5 floors, 2 triangles each
    assert s.lights[0].color[:] == [15.0, 15.0, 15.0]  # (3825,...)/255
    assert s.lights[0].visible == 1  # IniParser.h:33 default
    assert [m.type for m in s.materials] == [L.M_OPAQUE] * 5
    assert np.allclose(s.materials[0].color[:], np.array([31, 115, 39]) / 255.0)
    # floor (lt, rt, lb, rb) -> (lt, lb, rt), (rt, lb, rb)
    assert s.tri[0].tolist() == [-5, 10, 10, -5, 0, 10, -5, 10, 0]
    assert s.tri[1].tolist() == [-5, 10, 0, -5, 0, 10, -5, 0, 0]
    # mesh files are missing: bounds keep their INF start values
    assert s.bounds_min.tolist() == [999999.0] * 3 and s.bounds_max.tolist() == [-999999.0] * 3
    c = s.camera
    assert c.width == 512 and c.aspect == 1.0 and c.vfov_deg == 40.0
    assert c.look_from[:] == [0.0, 5.0, 22.0]


def test_scene1_material_strings_and_order():
    s = scenes.load_ini(os.path.join(GOLD, "Scene1.ini"))
    # spheres are added before floors, so their materials come first
    assert [m.type for m in s.materials] == [L.M_SPECULAR_DIELECTRIC, L.M_VOLUME, L.M_OPAQUE,
                                             L.M_OPAQUE, L.M_OPAQUE, L.M_OPAQUE, L.M_OPAQUE]
    assert len(s.spheres) == 5 and s.tri.shape == (4, 9)
    assert s.camera.aspect == pytest.approx(16 / 9)  # default aspectRatio
    assert s.materials[2].alpha_x == 0.5 and s.materials[3].alpha_y == 0.7


def test_mesh_rotation_scale_bounds_translation(tmp_path):
    obj = tmp_path / "quad.obj"
    obj.write_text("v 1 0 0\nv 1 1 0\nv 0 1 0\nv 0 0 0\nvt 0 0\nf 1/1 2/1 3/1 4/1\nf -4 -3 -2\n")
    ini = tmp_path / "s.ini"
    ini.write_text(f"[Camera]\nwidth=64\n[Mesh]\npath={obj}\ncenter=(10.0, 20.0, 30.0)\n"
                   "rotation=90.0\nscale=2.0\ncolor=(255, 0, 0)\ntype=Lambert\n"
                   "; comment\n  ; not a comment at column 0 but has no '='\n")
    s = scenes.load_ini(str(ini))
    assert s.tri.shape == (3, 9)  # quad -> 2 fan triangles + 1 triangle
    v = s.tri.reshape(-1, 3) - [10.0, 20.0, 30.0]
    th = np.radians(90.0)
    p = np.array([1.0, 0.0, 0.0])
    want = np.array([p[0] * np.cos(th) + p[2] * np.sin(th), p[1], -p[0] * np.sin(th) + p[2] * np.cos(th)]) * 2
    assert np.allclose(v[0], want, atol=1e-12)
    # bounds come from the rotated+scaled, untranslated vertices
    assert np.allclose(s.bounds_min, v.min(0)) and np.allclose(s.bounds_max, v.max(0))
    assert s.materials[0].type == L.M_LAMBERT
    assert s.camera.width == 64


def test_malformed_and_missing_files(tmp_path):
    with pytest.raises(OSError):
        scenes.load_ini(str(tmp_path / "nope.ini"))
    bad = tmp_path / "bad.ini"
    bad.write_text("[Camera]\nwidth=abc\n")
    with pytest.raises(OSError):
        scenes.load_ini(str(bad))
    # unknown material strings and keys with spaces keep the defaults
    odd = tmp_path / "odd.ini"
    odd.write_text("[Sphere]\ntype=Chrome\nradius = 3\n")
    s = scenes.load_ini(str(odd))
    assert s.materials[0].type == L.M_OPAQUE and s.spheres[0].radius == 0.0
