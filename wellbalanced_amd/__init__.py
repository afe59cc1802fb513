"""wellbalanced_amd: the ideal-MHD step of carafe (max-models/wellbalanced) on MI355X.

Hand-written HIP kernels for gfx950 behind the C ABI of ``include/wbmhd.h``
(``libwbmhd.so``, loaded with ctypes); PyTorch-ROCm only holds the device
buffers.  ``MHD2D`` and ``MHD3D`` keep the reference's model API.
"""

from .hierarchy import AMRHierarchy, Patch
from .mhd import CONSERVED, MHD2D
from .mhd3d import CONSERVED_3D, MHD3D

__all__ = ["AMRHierarchy", "CONSERVED", "CONSERVED_3D", "MHD2D", "MHD3D", "Patch"]
