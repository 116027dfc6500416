"""Hypersingular W (src/BEM_functions/hypersingular_layer/hypersingular_layer_3d.py)."""
from __future__ import annotations

from .. import _lib
from ._layer import Layer3d


class HypersingularLayer3d(Layer3d):
    op_code = _lib.OP_W

    def __init__(self, BEM_manager, *args, **kwargs):
        super().__init__(BEM_manager)
        self._Wmat = None
        if self.M_Dirichlet > 0:
            self._Wmat = BEM_manager.new_matrix(self.M_Dirichlet, self.M_Dirichlet)

    def get_W_mat(self):
        return self._Wmat

    def forward(self, scope_type: int, k: float, sqrt_n: float):
        """W_mat <- W (hypersingular_layer_3d.py:437-500)."""
        if self._Wmat is not None:
            self._Wmat.zero_()
            self._BEM_manager.ctx.assemble(
                self._desc(self._Q_Dirichlet, self._Q_Dirichlet, k, sqrt_n), self._Wmat, 1.0)

    def apply_W_dot_D_boundary(self, scope_type: int, k: float, sqrt_n: float, multiplier: float):
        """rhs[Dirichlet rows] += multiplier W conj(g), trial loop over the
        Neumann panels (hypersingular_layer_3d.py:502-565)."""
        bm = self._BEM_manager
        if self.num_of_panels_Dirichlet > 0 and self.num_of_panels_Neumann > 0:
            g = bm.rhs_constructor._Dirichlet_boundary
            self._apply(self._Q_Dirichlet, self._Q_Dirichlet, k, sqrt_n, g, multiplier,
                        bm.get_Dirichlet_offset_i(), self.M_Dirichlet)

    def kill(self):
        super().kill()
        self._Wmat = None
