"""MI355X MM-PIHM hydrology: host-side mirror of the reference's RHS interface.

The compute path is libpihm_amd.so (HIP kernels for gfx950 behind a C ABI); this package
builds meshes and forcing in the reference's conventions and calls that library.
"""
from .mesh import Mesh  # noqa: F401
