"""Shared plumbing of the operator layers: each layer owns its matrix as a
device tensor (the reference owns a Taichi field, e.g.
src/BEM_functions/single_layer/single_layer_3d.py:29-36) and calls the C ABI
for the work its Taichi kernels did."""
from __future__ import annotations

from .. import _lib
from .utils import KernelType


class Layer3d:
    rank = 3
    op_code = -1

    def __init__(self, BEM_manager):
        self._BEM_manager = BEM_manager
        self._GaussQR = BEM_manager.get_GaussQR()
        self._dim = 3
        self._n = BEM_manager._n
        self.N_Neumann = BEM_manager.get_N_Neumann()
        self._Q_Neumann = BEM_manager.get_shape_func_degree_Neumann()
        self.M_Dirichlet = BEM_manager.get_M_Dirichlet()
        self._Q_Dirichlet = BEM_manager.get_shape_func_degree_Dirichlet()
        self.num_of_vertices = BEM_manager.get_num_of_vertices()
        self.num_of_panels = BEM_manager.get_num_of_panels()
        self.num_of_panels_Neumann = BEM_manager.get_num_of_panels_Neumann()
        self.num_of_panels_Dirichlet = BEM_manager.get_num_of_panels_Dirichlet()

    def _desc(self, deg_test: int, deg_trial: int, k: float, sqrt_n: float, precision: int | None = None):
        bm = self._BEM_manager
        return _lib.Context.op_desc(
            self.op_code, bm._kernel_type, k, sqrt_n, deg_test, deg_trial,
            bm._precision if precision is None else precision, _lib.C128, bm._quadrature_flags,
        )

    def _apply(self, deg_test: int, deg_trial: int, k: float, sqrt_n: float, x, multiplier: float,
               row_offset: int, rows: int) -> None:
        """rhs_vec[row_offset : row_offset+rows] += multiplier * Op . conj(x)."""
        bm = self._BEM_manager
        y = bm.rhs_constructor.get_rhs_vec()[row_offset: row_offset + rows]
        desc = self._desc(deg_test, deg_trial, k, sqrt_n, precision=bm._rhs_precision)
        xx = x.to(y.dtype)
        conj = bm._n == 2
        bm.ctx.apply(desc, xx, y, multiplier, conj_x=conj)

    def kill(self):
        self.N_Neumann = 0
        self.M_Dirichlet = 0
        self._Q_Neumann = -1
        self._Q_Dirichlet = -1
        self.num_of_vertices = 0
        self.num_of_panels = 0
        self.num_of_panels_Neumann = 0
        self.num_of_panels_Dirichlet = 0
