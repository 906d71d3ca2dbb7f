"""calculixpp_amd — MI355X (gfx950) element integration and global assembly for
CalculiX++'s fem:: hot path: C3D8/C3D20/C3D20R/C3D6 stiffness and internal force,
J2 return map with consistent tangent, deterministic CSR assembly and the adjoint
of r = K(x, rho) u. The compute runs in libcxg.so (hand-written HIP, C ABI in
include/cxg.h); this package is its host-side interface.
"""
from . import _lib
from ._lib import (C3D6, C3D8, C3D20, C3D20R, CxgError, CxgLibraryError)
from .model import J2Params, Model

__all__ = ["_lib", "C3D6", "C3D8", "C3D20", "C3D20R", "CxgError", "CxgLibraryError",
           "J2Params", "Model"]
