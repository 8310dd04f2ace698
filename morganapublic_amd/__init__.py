"""MI355X-native FE assembly + Jacobi-PCG for Morgana meshes.

The compute path is libmgfe.so (hand-written HIP for gfx950 behind the C ABI of
include/mgfe.h); this package is its Python face: the ctypes binding (_lib), a
mirror of the reference's builder/solver interfaces (fe) and the synthetic
workloads (workloads).
"""
from ._lib import Context, MgfeError  # noqa: F401
