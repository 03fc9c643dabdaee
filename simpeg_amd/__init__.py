"""simpeg_amd: SimPEG's 3D nodal DC-resistivity path (``Simulation3DNodal``) on AMD Instinct
MI355X (gfx950), with hand-written HIP kernels behind a C ABI (libsimpeg_dc.so)."""

from . import maps  # noqa: F401
from .mesh import TensorMesh  # noqa: F401

__version__ = "0.1.0"
