"""lmtpp_amd: MI355X-native (gfx950) assemble + PCG path for LMT++-style linear elasticity.

The compute lives in libfemgpu.so (hand-written HIP kernels behind the C ABI of
include/femgpu.h); this package is its host-side mirror of the reference Formulation
interface.  There is no CPU fallback: a missing library raises at first use.
"""
from . import _capi  # noqa: F401
from ._capi import (ABS_MAXNORM, BLOCK_IC0, BLOCK_IC0_COLOURED, BLOCK_JACOBI, CHEBYSHEV_BLOCK_JACOBI, FP64,  # noqa: F401
                    HEXA, JACOBI, MIXED, REL_L2, ROW_ORDER_AUTO, ROW_ORDER_BY_LENGTH, ROW_ORDER_NATURAL, TETRA, TETRA10,
                    FemGpuError, load, row_order)
from .batch import Batch  # noqa: F401
from .formulation import BlockSystem, Formulation  # noqa: F401
from .mesh import clamp_dofs, faces_on_plane, make_rect, nodes_on_plane, skin_faces  # noqa: F401
from .problem import Problem  # noqa: F401

__all__ = ["Formulation", "Batch", "BlockSystem", "Problem", "make_rect", "skin_faces", "faces_on_plane", "nodes_on_plane",
           "clamp_dofs", "FemGpuError", "load", "TETRA", "TETRA10", "HEXA", "FP64", "MIXED", "JACOBI",
           "BLOCK_JACOBI", "BLOCK_IC0", "BLOCK_IC0_COLOURED", "CHEBYSHEV_BLOCK_JACOBI", "REL_L2", "ABS_MAXNORM", "ROW_ORDER_AUTO",
           "ROW_ORDER_NATURAL", "ROW_ORDER_BY_LENGTH", "row_order"]
