"""Single layer V (src/BEM_functions/single_layer/single_layer_3d.py)."""
from __future__ import annotations

from .. import _lib
from ._layer import Layer3d


class SingleLayer3d(Layer3d):
    op_code = _lib.OP_V

    def __init__(self, BEM_manager, *args, **kwargs):
        super().__init__(BEM_manager)
        assert self.N_Neumann + self.M_Dirichlet > 0
        self._Vmat = None
        if self.N_Neumann > 0:
            self._Vmat = BEM_manager.new_matrix(self.N_Neumann, self.N_Neumann)

    def get_V_mat(self):
        return self._Vmat

    def forward(self, scope_type: int, k: float, sqrt_n: float):
        """V_mat <- V (single_layer_3d.py:371-433)."""
        if self.N_Neumann > 0:
            self._Vmat.zero_()
            self._BEM_manager.ctx.assemble(
                self._desc(self._Q_Neumann, self._Q_Neumann, k, sqrt_n), self._Vmat, 1.0)

    def apply_V_dot_N_boundary(self, scope_type: int, k: float, sqrt_n: float, multiplier: float):
        """rhs[Neumann rows] += multiplier V conj(f) over Dirichlet trial panels
        (single_layer_3d.py:435-498)."""
        bm = self._BEM_manager
        if self.num_of_panels_Neumann > 0 and self.num_of_panels_Dirichlet > 0:
            f = bm.rhs_constructor._Neumann_boundary
            self._apply(self._Q_Neumann, self._Q_Neumann, k, sqrt_n, f, multiplier,
                        bm.get_Neumann_offset_i(), self.N_Neumann)

    def kill(self):
        super().kill()
        self._Vmat = None
