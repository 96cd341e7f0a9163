"""The reference-side C++ binding run as its own program.

cudapathtracing_amd/dropin/scene_dropin.cpp is the Scene class a reference
maintainer adds in place of Scene.cuh's device members (Scene.cuh:72-145)
and main.cu's STATIC frame (Render.cu:16-261, pic2RGBW Render.cu:462-481):
plain C++ against include/cpt.h, linked to libcpt.so by the library's
Makefile.  Run on the reference's own scene files, its PFM must equal the
Python binding's render of the same file and camera bit for bit, and its PNG
the displayed bytes of that render.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from cudapathtracing_amd import _lib as L
from cudapathtracing_amd.context import tonemap_rgba8
from test_image_io import read_png

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cudapathtracing_amd", "dropin", "scene_dropin")
GOLD = os.path.join(ROOT, "tests", "golden", "scene_ini")


def read_pfm(path):
    data = open(path, "rb").read()
    head = data.split(b"\n", 3)
    w, h = map(int, head[1].split())
    assert head[0] == b"PF" and float(head[2]) < 0
    return np.frombuffer(head[3], "<f4").reshape(h, w, 3)


@pytest.mark.parametrize("name", ["Scene1.ini", "Scene3.ini"])
def test_cpp_scene_binding_matches_python(ctx, tmp_path, name):
    print(f"[gpu] C++ drop-in program on {name}", flush=True)
    assert os.path.exists(BIN), f"{BIN} is not built (the library Makefile builds it)"
    ini = os.path.join(GOLD, name)
    pfm, png = str(tmp_path / "a.pfm"), str(tmp_path / "a.png")
    r = subprocess.run([BIN, ini, "2", "48", pfm, png], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"  {info}")
    img = read_pfm(pfm)
    # the same file through the Python binding
    mn, mx, cam = ctx.load_ini(ini)
    assert np.allclose(info["min"], mn, rtol=0, atol=0) and np.allclose(info["max"], mx, rtol=0, atol=0)
    ctx.build(mn, mx)
    cam.width = 48
    ref, _, st = ctx.render(cam, 2, seed=3317, profile=L.PROFILE_REFERENCE)
    assert img.shape == ref.shape == (info["height"], info["width"], 3)
    assert np.array_equal(img, ref)
    assert info["rays"] == st.rays
    assert np.array_equal(read_png(png), tonemap_rgba8(ref)[::-1])
    assert ref.mean() > 0
