"""Double layer K (src/BEM_functions/double_layer/double_layer_3d.py)."""
from __future__ import annotations

from .. import _lib
from ._layer import Layer3d


class DoubleLayer3d(Layer3d):
    op_code = _lib.OP_K

    def __init__(self, BEM_manager, *args, **kwargs):
        super().__init__(BEM_manager)
        self._Kmat = None
        if self.N_Neumann > 0 and self.M_Dirichlet > 0:
            self._Kmat = BEM_manager.new_matrix(self.N_Neumann, self.M_Dirichlet)

    def get_K_mat(self):
        return self._Kmat

    def forward(self, scope_type: int, k: float, sqrt_n: float):
        """K_mat <- K, P0/P1 test x Dirichlet trial (double_layer_3d.py:376-438)."""
        if self._Kmat is not None:
            self._Kmat.zero_()
            self._BEM_manager.ctx.assemble(
                self._desc(self._Q_Neumann, self._Q_Dirichlet, k, sqrt_n), self._Kmat, 1.0)

    def apply_K_dot_D_boundary(self, scope_type: int, k: float, sqrt_n: float, multiplier: float):
        """rhs[Neumann rows] += multiplier K conj(g); the reference runs its trial
        loop over the Neumann panels (double_layer_3d.py:440-501)."""
        bm = self._BEM_manager
        if self.num_of_panels_Neumann > 0:
            g = bm.rhs_constructor._Dirichlet_boundary
            self._apply(self._Q_Neumann, self._Q_Dirichlet, k, sqrt_n, g, multiplier,
                        bm.get_Neumann_offset_i(), self.N_Neumann)

    def kill(self):
        super().kill()
        self._Kmat = None
