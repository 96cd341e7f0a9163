"""The C-ABI library loads without a GPU and exports every symbol that
include/cpt.h declares; host-only entry points work (no compute calls)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import cudapathtracing_amd as cpt
from cudapathtracing_amd import _lib as L
from cudapathtracing_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "cpt.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cpt_[a-z0-9_]+)\s*\(", src)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(L.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in L.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def header_defines():
    src = open(os.path.join(ROOT, "include", "cpt.h")).read()
    out = {}
    for name, val in re.findall(r"^#define\s+(CPT_[A-Z0-9_]+)\s+(.+?)\s*(?:/\*.*)?$", src, flags=re.M):
        try:
            out[name] = eval(val, {})
        except Exception:
            pass
    return out


def test_python_constants_match_header():
    """The Python binding's copies of the header's constants (status codes,
    flags, option defaults and limits) hold the header's values."""
    d = header_defines()
    pairs = {"CPT_ERR_ARG": L.CPT_ERR_ARG, "CPT_ERR_HIP": L.CPT_ERR_HIP, "CPT_ERR_STATE": L.CPT_ERR_STATE,
             "CPT_ERR_OOM": L.CPT_ERR_OOM, "CPT_ERR_IO": L.CPT_ERR_IO, "CPT_ERR_ABORTED": L.CPT_ERR_ABORTED,
             "CPT_DEVICE_PTRS": L.CPT_DEVICE_PTRS, "CPT_NOALLOC_MAX_RAYS": L.NOALLOC_MAX_RAYS,
             "CPT_PERSISTENT_SHADOW_MIN_DEFAULT": L.PERSISTENT_SHADOW_MIN_DEFAULT,
             "CPT_M_LAMBERT": L.M_LAMBERT, "CPT_M_GGX": L.M_GGX, "CPT_OBJ_TRIANGLE": L.OBJ_TRIANGLE}
    for k, v in pairs.items():
        assert d[k] == v, (k, d[k], v)


def test_struct_layouts_match_header(tmp_path):
    """ctypes layouts equal what a C compiler makes of include/cpt.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"cpt_material": L.Material, "cpt_light": L.Light, "cpt_sphere": L.Sphere,
               "cpt_camera_desc": L.CameraDesc, "cpt_camera": L.Camera,
               "cpt_build_stats": L.BuildStats, "cpt_render_params": L.RenderParams,
               "cpt_render_stats": L.RenderStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "cpt.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                  check=True).stdout.splitlines())
    for cname, py in structs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(got[f"{cname}.{f}"]) == getattr(py, f).offset, (cname, f)


def test_host_entry_points():
    lib = L.lib()
    assert b"gfx950" in lib.cpt_version()
    assert lib.cpt_gen_icosphere_count(3) == 1280
    assert lib.cpt_gen_icosphere_count(8) == 1310720
    assert lib.cpt_create(None, 0) == L.CPT_ERR_ARG


def test_camera_matches_reference_formulas():
    d = L.CameraDesc(L.d3((0, 0, 0)), L.d3((0.0, 5.0, 22.0)), L.d3((0.0, 5.0, 0.0)), 40.0, 1.0, 512, 0)
    cam = cpt.camera(d)
    assert cam.width == 512 and cam.height == 512
    h = np.tan(np.radians(40.0) / 2) * 22.0 * 2.0
    assert cam.delta_v[1] == pytest.approx(h / 512, rel=1e-14)
    # pixel (0,0) is the bottom-left pixel centre
    pc = np.array(cam.pixel_center[:])
    assert pc[0] == pytest.approx(-h / 2 + h / 1024, rel=1e-12)
    assert pc[1] == pytest.approx(5.0 - h / 2 + h / 1024, rel=1e-12)
    d.aspect = 16 / 9
    d.width = 800
    assert cpt.camera(d).height == 450


def test_tonemap():
    rgb = np.array([[[0.0, 0.25, 4.0]]], np.float32)
    out = cpt.tonemap_rgba8(rgb)
    assert out.tolist() == [[[0, 127, 254, 255]]]


def test_generators_deterministic_and_sized():
    a = scenes.icosphere(2, (0, 0, 0), 2.0)
    assert a.shape == (320, 9)
    r = np.linalg.norm(a.reshape(-1, 3), axis=1)
    assert np.allclose(r, 2.0, rtol=1e-12)
    b = scenes.icosphere(2, (1, 1, 1), 0.3, amp=0.1, seed=4)
    rb = np.linalg.norm(b.reshape(-1, 3) - 1, axis=1)
    assert rb.min() >= 0.27 - 1e-9 and rb.max() <= 0.33 + 1e-9
    s1 = scenes.triangle_soup(1000, seed=1)
    assert np.array_equal(s1, scenes.triangle_soup(1000, seed=1))
    c = s1.reshape(-1, 3, 3).mean(axis=1)
    assert c.min() > -0.05 and c.max() < 1.05
    e = np.linalg.norm(s1[:, 0:3] - s1[:, 3:6], axis=1)
    assert 0.5 * np.exp(-6) < np.median(e) < 2 * np.exp(-6)
    rays = scenes.random_rays(10000, seed=2)
    assert np.allclose(np.linalg.norm(rays[:, 4:7], axis=1), 1.0, atol=1e-6)
    assert rays[:, :3].min() >= -0.5 and rays[:, :3].max() <= 1.5
    assert abs(rays[:, 4:7].mean()) < 0.02


def test_dropin_program_links_and_reports_missing_device():
    """The C++ drop-in binding (cudapathtracing_amd/dropin/scene_dropin.cpp)
    is built by the library Makefile, links libcpt.so by soname through its
    run path, and without a GPU fails cleanly at cpt_create."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    binary = os.path.join(root, "cudapathtracing_amd", "dropin", "scene_dropin")
    assert os.path.exists(binary), "build the library first (its Makefile builds the program)"
    dyn = subprocess.run(["readelf", "-d", binary], capture_output=True, text=True, check=True).stdout
    assert "[libcpt.so]" in dyn and "$ORIGIN/.." in dyn
    ini = os.path.join(root, "tests", "golden", "scene_ini", "Scene1.ini")
    r = subprocess.run([binary], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
    import torch
    if not torch.cuda.is_available():
        r = subprocess.run([binary, ini, "1", "16", os.devnull, os.devnull], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 1 and "no HIP device" in r.stderr
