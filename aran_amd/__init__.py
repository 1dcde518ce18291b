"""aran_amd -- MI355X-native 3D Newton FMM behind LibAran's AranSolver3d C API.

The compute path is the HIP library aran_amd/libaran_amd.so (gfx950); this
package only loads it and mirrors its C entry points for Python callers.
Importing the package does not touch the GPU.
"""
from ._lib import LIB_PATH, lib  # noqa: F401

__all__ = ["lib", "LIB_PATH"]
