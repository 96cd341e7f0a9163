"""MI355X (gfx950) LBVH build, closest-hit traversal and path tracer.

Python is only the host-side convenience layer over ``libcpt.so`` (the C ABI
in ``include/cpt.h``); every computation runs in the HIP library.
"""
from . import _lib
from ._lib import lib
from .context import Context, CptError, camera, tonemap_rgba8, write_pfm, write_png
from .scenes import SceneData

__all__ = ["Context", "CptError", "SceneData", "camera", "tonemap_rgba8", "write_pfm", "write_png", "lib", "_lib"]
