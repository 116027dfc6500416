"""Adjoint double layer K' (src/BEM_functions/adj_double_layer/adj_double_layer_3d.py)."""
from __future__ import annotations

from .. import _lib
from ._layer import Layer3d


class AdjDoubleLayer3d(Layer3d):
    op_code = _lib.OP_KADJ

    def __init__(self, BEM_manager, *args, **kwargs):
        super().__init__(BEM_manager)
        # The reference allocates _Kmat but never fills it in 3D and uses -K^T
        # in the block system (BEM_manager.py:619-622).
        self._Kmat = None

    def forward(self, scope_type: int, k: float, sqrt_n: float):
        """No-op, as in the reference (adj_double_layer_3d.py:378-383)."""
        return

    def assemble(self, k: float, sqrt_n: float, out, alpha: complex = 1.0):
        """out += alpha K' (Dirichlet test x Neumann trial), assembled directly."""
        self._BEM_manager.ctx.assemble(self._desc(self._Q_Dirichlet, self._Q_Neumann, k, sqrt_n), out, alpha)

    def apply_K_dot_N_boundary(self, scope_type: int, k: float, sqrt_n: float, multiplier: float):
        """rhs[Dirichlet rows] += multiplier K' conj(f), trial loop over the
        Dirichlet panels (adj_double_layer_3d.py:443-506)."""
        bm = self._BEM_manager
        if self.num_of_panels_Dirichlet > 0:
            f = bm.rhs_constructor._Neumann_boundary
            self._apply(self._Q_Dirichlet, self._Q_Neumann, k, sqrt_n, f, multiplier,
                        bm.get_Dirichlet_offset_i(), self.M_Dirichlet)
