"""Mass matrix M (src/BEM_functions/identity_layer/identity_layer_3d.py)."""
from __future__ import annotations

from ._layer import Layer3d


class IdentityLayer3d(Layer3d):
    def __init__(self, BEM_manager, *args, **kwargs):
        super().__init__(BEM_manager)
        self._Neumann_offset_i = BEM_manager.get_Neumann_offset_i()
        self._Dirichlet_offset_j = BEM_manager.get_Dirichlet_offset_j()
        self._Dirichlet_offset_i = BEM_manager.get_Dirichlet_offset_i()
        self._Neumann_offset_j = BEM_manager.get_Neumann_offset_j()

    def matA_add_global_Identity_matrix(self, mutiplier: float):
        """A[N+i, D+v] += m M and A[D+v, N+i] += m M, transmission only
        (identity_layer_3d.py:148-196)."""
        bm = self._BEM_manager
        if not bm._is_transmission:
            return
        A = bm.get_mat_A()
        blk = A[self._Neumann_offset_i:, self._Dirichlet_offset_j:]
        bm.ctx.mass_assemble(self._Q_Neumann, self._Q_Dirichlet, blk, mutiplier, transpose=False)
        blk_t = A[self._Dirichlet_offset_i:, self._Neumann_offset_j:]
        bm.ctx.mass_assemble(self._Q_Neumann, self._Q_Dirichlet, blk_t, mutiplier, transpose=True)
