"""MI355X Galerkin BEM hot path: a drop-in for the operator assembly and
dense solve of GeCao/taichi-BEM, with the reference's manager/layer class
names over hand-written HIP kernels (libbem_hip.so, C ABI in
include/bem_hip.h)."""
from . import _lib, adjoint, meshes, sweep
from ._lib import BEMError, Context, SingularMatrixError, build, convert, matrix_axpy
from .bem_functions.utils import (AssembleType, CellFluxType, KernelType, PanelsRelation, ScopeType,
                                  VertAttachType, get_gaussion_integration_points_and_weights,
                                  unwarp_tensor, warp_tensor)
from .managers import BEMManager, CoreManager, MeshManager

__all__ = [
    "_lib", "adjoint", "meshes", "sweep", "BEMError", "Context", "SingularMatrixError", "build", "convert", "matrix_axpy",
    "AssembleType", "CellFluxType", "KernelType", "PanelsRelation", "ScopeType", "VertAttachType",
    "get_gaussion_integration_points_and_weights", "warp_tensor", "unwarp_tensor",
    "BEMManager", "CoreManager", "MeshManager",
]
