"""Boundary data and right-hand side
(src/BEM_functions/rhs_constructor/rhs_constructor_3d.py)."""
from __future__ import annotations

import torch

from .utils import ScopeType, VertAttachType


class RHSConstructor3d:
    rank = 3

    def __init__(self, BEM_manager, *args, **kwargs):
        bm = BEM_manager
        self._BEM_manager = bm
        self._GaussQR = bm.get_GaussQR()
        self._dim = 3
        self._n = bm._n
        self.N_Neumann = bm.get_N_Neumann()
        self._Q_Neumann = bm.get_shape_func_degree_Neumann()
        self.M_Dirichlet = bm.get_M_Dirichlet()
        self._Q_Dirichlet = bm.get_shape_func_degree_Dirichlet()
        self.num_of_vertices = bm.get_num_of_vertices()
        self.num_of_panels = bm.get_num_of_panels()
        self.num_of_panels_Neumann = bm.get_num_of_panels_Neumann()
        self.num_of_panels_Dirichlet = bm.get_num_of_panels_Dirichlet()
        self.analytical_function_Dirichlet = bm.analytical_function_Dirichlet
        self.analytical_function_Neumann = bm.analytical_function_Neumann
        nN = self.num_of_panels if self._Q_Neumann == 0 else self.num_of_vertices
        nD = self.num_of_panels if self._Q_Dirichlet == 0 else self.num_of_vertices
        self._Neumann_boundary = bm.new_vector(nN)
        self._Dirichlet_boundary = bm.new_vector(nD)
        self._f_boundary = bm.new_vector(self.N_Neumann + self.M_Dirichlet)
        self._rhs_vec = bm.new_vector(self.N_Neumann + self.M_Dirichlet)

    # -- data sampling (set_boundaries, rhs_constructor_3d.py:50-152) --------
    def _as_field(self, v: torch.Tensor) -> torch.Tensor:
        bm = self._BEM_manager
        v = v.to(bm.device)
        if v.dim() == 2 and v.shape[-1] == 1:
            v = v[:, 0]
        if v.dim() == 2 and v.shape[-1] == 2:
            v = torch.complex(v[:, 0], v[:, 1])
        return v.to(self._rhs_vec.dtype)

    def set_boundaries(self, scope_type: int, sqrt_ni: float, sqrt_no: float):
        bm = self._BEM_manager
        mm = bm._core_manager._mesh_manager
        self._Dirichlet_boundary.zero_()
        self._Neumann_boundary.zero_()
        self._f_boundary.zero_()
        Doff, Noff = bm.get_Dirichlet_offset_j(), bm.get_Neumann_offset_j()
        cen, pn = mm.panel_centroids(), mm.get_panel_normals()
        vx, vn = mm.vertices, mm.get_vert_normals()
        fD, fN = self.analytical_function_Dirichlet, self.analytical_function_Neumann
        vtype = mm.get_vertice_types()
        if bm._is_transmission:
            # interior minus exterior traces (:59-114); normals are not
            # flipped by scope (BEM_manager.py:329-343)
            if self._Q_Neumann == 0:
                fi = self._as_field(fN(cen, pn, sqrt_ni))
                fo = self._as_field(fN(cen, pn, sqrt_no))
                self._Neumann_boundary += fi
                self._f_boundary[Noff: Noff + self.num_of_panels] += fi - fo
            else:
                sel = (vtype == int(VertAttachType.BOTH_TOBESOLVED)) | (vtype == int(VertAttachType.NEUMANN_TOBESOLVED))
                fi = self._as_field(fN(vx, vn, sqrt_ni)) * sel
                fo = self._as_field(fN(vx, vn, sqrt_no)) * sel
                self._Neumann_boundary += fi
                self._f_boundary[Noff: Noff + self.num_of_vertices] += fi - fo
            if self._Q_Dirichlet == 0:
                gi = self._as_field(fD(cen, sqrt_ni))
                go = self._as_field(fD(cen, sqrt_no))
                self._Dirichlet_boundary += gi
                self._f_boundary[Doff: Doff + self.num_of_panels] += gi - go
            else:
                sel = (vtype == int(VertAttachType.BOTH_TOBESOLVED)) | (vtype == int(VertAttachType.DIRICHLET_TOBESOLVED))
                gi = self._as_field(fD(vx, sqrt_ni)) * sel
                go = self._as_field(fD(vx, sqrt_no)) * sel
                self._Dirichlet_boundary += gi
                self._f_boundary[Doff: Doff + self.num_of_vertices] += gi - go
        else:
            if self._Q_Neumann == 0:
                if self.num_of_panels_Dirichlet > 0:
                    self._Neumann_boundary += self._as_field(fN(cen, pn))
            else:
                sel = vtype == int(VertAttachType.DIRICHLET_TOBESOLVED)
                self._Neumann_boundary += self._as_field(fN(vx, vn)) * sel
            if self._Q_Dirichlet == 0:
                if self.num_of_panels_Neumann > 0:
                    self._Dirichlet_boundary += self._as_field(fD(cen))
            else:
                sel = vtype == int(VertAttachType.NEUMANN_TOBESOLVED)
                self._Dirichlet_boundary += self._as_field(fD(vx)) * sel

    def get_rhs_vec(self):
        return self._rhs_vec

    def get_f_boundary(self, local_index=None):
        return self._f_boundary if local_index is None else self._f_boundary[local_index]

    def get_Dirichlet_boundary(self, panel_index=None, basis_function_index=-1):
        return self._Dirichlet_boundary

    def get_Neumann_boundary(self, panel_index=None, basis_function_index=-1):
        return self._Neumann_boundary

    # -- RHS kernels ---------------------------------------------------------
    def multiply_identity_to_vector(self, multiplier: float):
        """rhs += multiplier M conj(data) (rhs_constructor_3d.py:256-359)."""
        bm = self._BEM_manager
        conj = self._n == 2
        if self.N_Neumann > 0:
            y = self._rhs_vec[bm.get_Neumann_offset_i(): bm.get_Neumann_offset_i() + self.N_Neumann]
            bm.ctx.mass_apply(self._Q_Neumann, self._Q_Dirichlet, self._Dirichlet_boundary, y, multiplier,
                              conj_x=conj)
        if self.M_Dirichlet > 0:
            y = self._rhs_vec[bm.get_Dirichlet_offset_i(): bm.get_Dirichlet_offset_i() + self.M_Dirichlet]
            bm.ctx.mass_apply(self._Q_Dirichlet, self._Q_Neumann, self._Neumann_boundary, y, multiplier,
                              conj_x=conj)

    def multiply_M_to_vector(self, scope_type: int, k: float, sqrt_n: float, multiplier: float):
        """rhs += K g - V f - K' f - W g, matrix-free (rhs_constructor_3d.py:361-372)."""
        bm = self._BEM_manager
        if self.N_Neumann > 0:
            bm.double_layer.apply_K_dot_D_boundary(scope_type=scope_type, k=k, sqrt_n=sqrt_n, multiplier=multiplier)
            bm.single_layer.apply_V_dot_N_boundary(scope_type=scope_type, k=k, sqrt_n=sqrt_n, multiplier=-multiplier)
        if self.M_Dirichlet > 0:
            bm.adj_double_layer.apply_K_dot_N_boundary(scope_type=scope_type, k=k, sqrt_n=sqrt_n, multiplier=-multiplier)
            bm.hypersingular_layer.apply_W_dot_D_boundary(scope_type=scope_type, k=k, sqrt_n=sqrt_n, multiplier=-multiplier)

    def forward(self, scope_type: int, k: float, sqrt_n: float):
        """rhs = (+/-)1/2 M g + K g - ... (rhs_constructor_3d.py:374-386)."""
        self._rhs_vec.zero_()
        bm = self._BEM_manager
        if not bm._is_transmission:
            scope_type = bm._scope_type
            if scope_type == int(ScopeType.INTERIOR):
                self.multiply_identity_to_vector(multiplier=0.5)
            elif scope_type == int(ScopeType.EXTERIOR):
                self.multiply_identity_to_vector(multiplier=-0.5)
            else:
                raise RuntimeError("You have to indicate a concrete scope type to construct rhs")
            self.multiply_M_to_vector(scope_type=scope_type, k=k, sqrt_n=sqrt_n, multiplier=1.0)

    def kill(self):
        self._rhs_vec = None
        self._Dirichlet_boundary = None
        self._Neumann_boundary = None
        self._f_boundary = None
