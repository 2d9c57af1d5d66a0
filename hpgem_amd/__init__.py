"""MI355X-native explicit DG operator with hpGEM's RHS/mass/RK interface.

The compute path is libhgdg.so (hand-written HIP for gfx950) behind the C ABI
in include/hgdg.h; this package is the host-side mirror of hpGEM's API.
"""
from ._lib import (HgdgError, OP_ACOUSTIC_CENTRAL, OP_ADVECTION_UPWIND, OP_MATRIX, PREC_FP64, PREC_MIXED_FP32,
                   SHAPE_CUBE, SHAPE_SQUARE, SHAPE_TETRAHEDRON, load)
from .api import HpgemAPIDevice, HpgemAPIMatrixDevice, Mesh, basis_eval, basis_size, write_vtk
from .tableaux import (BOGACKI_SHAMPINE, DORMAND_PRINCE, FORWARD_EULER, MIDPOINT, RK2_TVD, RK3_TVD, RK4,
                       ButcherTableau, get_rule)

__all__ = ["HgdgError", "HpgemAPIDevice", "HpgemAPIMatrixDevice", "OP_MATRIX", "Mesh", "basis_eval", "basis_size", "write_vtk", "load", "RK4", "DORMAND_PRINCE",
           "FORWARD_EULER", "MIDPOINT", "RK2_TVD", "RK3_TVD", "BOGACKI_SHAMPINE", "ButcherTableau", "get_rule", "SHAPE_SQUARE", "SHAPE_CUBE", "SHAPE_TETRAHEDRON",
           "OP_ADVECTION_UPWIND", "OP_ACOUSTIC_CENTRAL", "PREC_FP64", "PREC_MIXED_FP32"]
