"""MI355X-native single-rank frequency-domain solve path of
demo_matrixfree_csem: HIP kernels for gfx950 behind the reference's C entry
points (include/csem.h, include/csem_dev.h), built into libcsem.so.

    python -m demo_matrixfree_csem_amd.build     # compile in-tree
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["LIB_PATH", "lib"]
