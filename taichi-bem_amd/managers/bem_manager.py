"""BEM manager (src/managers/BEM_manager.py): orchestrates operator
assembly, block placement, right-hand side and the dense solve.  Method names,
argument order, enum values and the block layout (Dirichlet unknowns first,
Neumann unknowns at offset M_Dirichlet, BEM_manager.py:104-111) follow the
reference; the work runs in libbem_hip.so."""
from __future__ import annotations

import math

import torch

from .. import _lib
from ..bem_functions import (AdjDoubleLayer3d, DoubleLayer3d, HypersingularLayer3d, IdentityLayer3d,
                             RHSConstructor3d, SingleLayer3d)
from ..bem_functions.utils import (AssembleType, CellFluxType, KernelType, ScopeType, VertAttachType,
                                   get_gaussion_integration_points_and_weights)

_PRECISION = {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32}
_QUADRATURE = {"reference": 0, "corrected": _lib.FLAG_ROTATE_NEAR | _lib.FLAG_SYM_COINCIDE}


class BEMManager:
    def __init__(self, core_manager):
        self._core_manager = core_manager
        p = core_manager._simulation_parameters
        self._simulation_parameters = p
        self._is_transmission = core_manager._is_transmission
        kernel = p["kernel"]
        if kernel == "Laplace":
            self._kernel_type, self._n = int(KernelType.LAPLACE), 1
        elif kernel == "Helmholtz":
            self._kernel_type, self._n = int(KernelType.HELMHOLTZ), 2
        elif kernel == "Helmholtz_Transmission":
            self._kernel_type, self._n = int(KernelType.HELMHOLTZ_TRANSMISSION), 2
        else:
            raise RuntimeError("Kernel Type only support Laplace and Helmholtz for now")
        self._k = 0.0
        if self._kernel_type != int(KernelType.LAPLACE):
            self._k = float(p["k"])
        self._sqrt_ni = 1.0
        self._sqrt_no = 1.0
        if self._kernel_type == int(KernelType.HELMHOLTZ_TRANSMISSION):
            self._sqrt_ni = float(math.sqrt(p["n_i"]))
            self._sqrt_no = float(math.sqrt(p["n_o"]))
        self._dim = p.get("dim", 3)
        self._GaussQR = p["GaussQR"]
        # the reference raises NotImplementedError for unsupported GaussQR
        get_gaussion_integration_points_and_weights(self._GaussQR)
        self._scope_type = int(ScopeType.NOT_DEFINED)
        if p.get("scope") == "Interior":
            self._scope_type = int(ScopeType.INTERIOR)
        elif p.get("scope") == "Exterior":
            self._scope_type = int(ScopeType.EXTERIOR)
        prec = p.get("precision", "fp64")
        if prec not in _PRECISION:
            raise RuntimeError(f"precision must be one of {sorted(_PRECISION)}")
        self._precision = _PRECISION[prec]
        self._rhs_precision = self._precision
        quad = p.get("quadrature", "reference")
        if quad not in _QUADRATURE:
            raise RuntimeError(f"quadrature must be one of {sorted(_QUADRATURE)}")
        self._quadrature_flags = _QUADRATURE[quad]
        if self._n == 1:
            self._dtype = torch.float64 if prec == "fp64" else torch.float32
        else:
            self._dtype = torch.complex128 if prec == "fp64" else torch.complex64
        self.num_of_vertices = 0
        self.num_of_panels = 0
        self._L1_inv_H = None
        self._L2_inv_H = None
        self.initialized = False
        # one fused V/K/W pass per wave number where the layout allows it
        self.fuse_layers = True

    @property
    def ctx(self) -> _lib.Context:
        return self._core_manager.ctx

    @property
    def device(self):
        return self._core_manager.device

    def new_matrix(self, rows: int, cols: int) -> torch.Tensor:
        return torch.zeros(rows, cols, dtype=self._dtype, device=self.device)

    def new_vector(self, n: int) -> torch.Tensor:
        return torch.zeros(n, dtype=self._dtype, device=self.device)

    def initialization(self, analytical_function_Dirichlet, analytical_function_Neumann):
        mm = self._core_manager._mesh_manager
        if not mm.initialized:
            raise RuntimeError("The initialization of Mesh Manager should always before the BEM Manager")
        self.num_of_vertices = mm.get_num_of_vertices()
        self.num_of_panels = mm.get_num_of_panels()
        self.N_Neumann = mm.get_N_Neumann()
        self.shape_func_degree_Neumann = mm.get_shape_func_degree_Neumann()
        self.M_Dirichlet = mm.get_M_Dirichlet()
        self.shape_func_degree_Dirichlet = mm.get_shape_func_degree_Dirichlet()
        self.num_of_panels_Neumann = mm.get_num_of_panels_Neumann()
        self.num_of_panels_Dirichlet = mm.get_num_of_panels_Dirichlet()
        self._Dirichlet_offset_i = 0
        self._Dirichlet_offset_j = 0
        self._Neumann_offset_i = self.M_Dirichlet
        self._Neumann_offset_j = self.M_Dirichlet
        self.analytical_function_Dirichlet = analytical_function_Dirichlet
        self.analytical_function_Neumann = analytical_function_Neumann
        n = self.N_Neumann + self.M_Dirichlet
        self.raw_solved = self.new_vector(n)
        self.raw_analytical_solved = self.new_vector(n)
        self._mat_A = self.new_matrix(n, n)
        self.rhs = self.new_vector(2 * n if self._is_transmission else n)
        self.solved = self.new_vector(self.num_of_vertices)
        self.analytical_solved = self.new_vector(self.num_of_vertices)
        self.identity_layer = IdentityLayer3d(self)
        self.single_layer = SingleLayer3d(self)
        self.double_layer = DoubleLayer3d(self)
        self.adj_double_layer = AdjDoubleLayer3d(self)
        self.hypersingular_layer = HypersingularLayer3d(self)
        self.rhs_constructor = RHSConstructor3d(self)
        self.initialized = True

    # -- getters (BEM_manager.py:197-403) -----------------------------------------
    def get_GaussQR(self):
        return self._GaussQR

    def get_num_of_vertices(self):
        return self.num_of_vertices

    def get_num_of_panels(self):
        return self.num_of_panels

    def get_M_Dirichlet(self):
        return self.M_Dirichlet

    def get_N_Neumann(self):
        return self.N_Neumann

    def get_num_of_panels_Neumann(self):
        return self.num_of_panels_Neumann

    def get_num_of_panels_Dirichlet(self):
        return self.num_of_panels_Dirichlet

    def get_shape_func_degree_Neumann(self):
        return self.shape_func_degree_Neumann

    def get_shape_func_degree_Dirichlet(self):
        return self.shape_func_degree_Dirichlet

    def get_mat_A(self):
        return self._mat_A

    def get_Dirichlet_offset_i(self):
        return self._Dirichlet_offset_i

    def get_Dirichlet_offset_j(self):
        return self._Dirichlet_offset_j

    def get_Neumann_offset_i(self):
        return self._Neumann_offset_i

    def get_Neumann_offset_j(self):
        return self._Neumann_offset_j

    def get_scope_type(self):
        return self._scope_type

    def set_scope_type(self, scope_type):
        self._scope_type = scope_type

    # -- assembly -------------------------------------------------------------------
    def matrix_layer_init(self):
        for m in (self.single_layer._Vmat, self.double_layer._Kmat, self.hypersingular_layer._Wmat):
            if m is not None:
                m.zero_()

    def matrix_layer_forward(self, k: float, sqrt_n: float = 1.0, scope_type: int = int(ScopeType.NOT_DEFINED)):
        """BEM_manager.py:829-844; an undefined scope raises instead of exit()."""
        if scope_type == int(ScopeType.NOT_DEFINED):
            raise RuntimeError("Did not indicate a correct scope type!")
        self.set_scope_type(scope_type=scope_type)
        self.matrix_layer_init()
        if self._can_fuse_layers():
            # V, K and W of every panel pair from one kernel evaluation
            V, K, W = self.single_layer._Vmat, self.double_layer._Kmat, self.hypersingular_layer._Wmat
            desc = self.single_layer._desc(0, 0, k, sqrt_n)
            self.ctx.assemble_vkw(desc, V, K, W)
            self.adj_double_layer.forward(scope_type, k, sqrt_n)
            return
        self.single_layer.forward(scope_type, k, sqrt_n)
        self.double_layer.forward(scope_type, k, sqrt_n)
        self.adj_double_layer.forward(scope_type, k, sqrt_n)
        self.hypersingular_layer.forward(scope_type, k, sqrt_n)

    def _can_fuse_layers(self) -> bool:
        """The fused V/K/W pass covers P0 Neumann unknowns on every panel and P1
        Dirichlet unknowns on every vertex (the transmission layout);
        `fuse_layers = False` selects the per-operator passes."""
        mm = self._core_manager._mesh_manager
        return (self.fuse_layers and self.shape_func_degree_Neumann == 0 and self.shape_func_degree_Dirichlet == 1
                and self.N_Neumann == mm.get_num_of_panels() and self.M_Dirichlet == mm.get_num_of_vertices()
                and self.single_layer._Vmat is not None and self.double_layer._Kmat is not None
                and self.hypersingular_layer._Wmat is not None)

    def rhs_layer_forward(self, scope_type, k: float, sqrt_n: float = 1.0):
        self.rhs_constructor.forward(scope_type=scope_type, k=k, sqrt_n=sqrt_n)

    def _block(self, r0: int, c0: int, rows: int, cols: int):
        return self._mat_A[r0: r0 + rows, c0: c0 + cols]

    def matA_add_M(self, multiplier: float):
        """A += -m V at (N,N), +m K at (N,D), -m K^T at (D,N), -m W at (D,D)
        (BEM_manager.py:608-628)."""
        N, M = self.N_Neumann, self.M_Dirichlet
        if N > 0:
            _lib.matrix_axpy(self.single_layer._Vmat, self._block(self._Neumann_offset_i, self._Neumann_offset_j, N, N),
                             -multiplier)
        if N * M > 0:
            K = self.double_layer._Kmat
            _lib.matrix_axpy(K, self._block(self._Neumann_offset_i, self._Dirichlet_offset_j, N, M), multiplier)
            _lib.matrix_axpy(K, self._block(self._Dirichlet_offset_i, self._Neumann_offset_j, M, N), -multiplier,
                             transpose=True)
        if M > 0:
            _lib.matrix_axpy(self.hypersingular_layer._Wmat,
                             self._block(self._Dirichlet_offset_i, self._Dirichlet_offset_j, M, M), -multiplier)

    def assemble_matA(self, assemble_type: int, multiplier: float):
        """BEM_manager.py:630-645."""
        if assemble_type == int(AssembleType.ADD_M):
            self.matA_add_M(multiplier)
        elif assemble_type == int(AssembleType.ADD_HALF_IDENTITY):
            self.identity_layer.matA_add_global_Identity_matrix(0.5 * multiplier)
        elif assemble_type == int(AssembleType.ADD_P_MINUS):
            self.identity_layer.matA_add_global_Identity_matrix(0.5 * multiplier)
            self.matA_add_M(-multiplier)
        elif assemble_type == int(AssembleType.ADD_P_PLUS):
            self.identity_layer.matA_add_global_Identity_matrix(0.5 * multiplier)
            self.matA_add_M(multiplier)

    def assemble_rhs(self):
        """BEM_manager.py:647-662: transmission rhs = A conj(f) over the first
        N+M rows, otherwise a copy of the constructor's vector."""
        n = self.N_Neumann + self.M_Dirichlet
        self.rhs.zero_()
        if self._is_transmission:
            f = self.rhs_constructor.get_f_boundary()
            A = self._mat_A
            A64 = A if A.dtype in (torch.complex128, torch.float64) else A.to(
                torch.complex128 if A.is_complex() else torch.float64)
            # conj() is a lazy view; the library reads storage, so the
            # conjugate is materialised before the pointer is taken
            x = f.to(A64.dtype).conj().resolve_conj().contiguous()
            zero = torch.zeros_like(x)
            r = torch.empty_like(x)
            self.ctx.residual(A64, x, zero, r)  # r = 0 - A x
            self.rhs[:n] = (-r).to(self.rhs.dtype)
        else:
            self.rhs.copy_(self.rhs_constructor.get_rhs_vec())

    # -- solve ------------------------------------------------------------------------
    def solve(self):
        """BEM_manager.py:1032-1073 with the HIP LU.  For transmission the
        reference solves the stacked [A; P+_i] system by least squares; this
        build solves the square block system A x = rhs[:n] (documented
        difference)."""
        n = self.N_Neumann + self.M_Dirichlet
        LU = self._mat_A.clone()
        ipiv = torch.empty(n, dtype=torch.int32, device=self.device)
        self.ctx.lu_factor(LU, ipiv)
        b = self.rhs[:n].clone()
        self.ctx.lu_solve(LU, ipiv, b)
        self.raw_solved.copy_(b)
        self._LU, self._ipiv = LU, ipiv
        return self.raw_solved

    def run(self):
        """BEM_manager.py:990-1030 (without the GUI colouring)."""
        self._mat_A.zero_()
        if self._is_transmission:
            self.matrix_layer_forward(k=self._k, sqrt_n=self._sqrt_no, scope_type=int(ScopeType.EXTERIOR))
            self.assemble_matA(assemble_type=int(AssembleType.ADD_P_MINUS), multiplier=1.0)
            self.rhs_constructor.set_boundaries(scope_type=int(ScopeType.INTERIOR), sqrt_ni=self._sqrt_ni,
                                                sqrt_no=self._sqrt_no)
            self.assemble_rhs()
            self.matrix_layer_forward(k=self._k, sqrt_n=self._sqrt_ni, scope_type=int(ScopeType.INTERIOR))
            self.assemble_matA(assemble_type=int(AssembleType.ADD_P_PLUS), multiplier=1.0)
        else:
            self.matrix_layer_forward(k=self._k, scope_type=self._scope_type)
            self.assemble_matA(assemble_type=int(AssembleType.ADD_M), multiplier=-1.0)
            self.rhs_constructor.set_boundaries(scope_type=self._scope_type, sqrt_ni=self._sqrt_ni,
                                                sqrt_no=self._sqrt_no)
            self.rhs_layer_forward(k=self._k, scope_type=self._scope_type)
            self.assemble_rhs()
        self.solve()
        self.compute_analytical_raw_solved()
        self.splat_u_from_raw_solved_to_vertices()

    # -- norm analysis (BEM_manager.py:850-988) ---------------------------------------
    # Blocks of an (N+M) x (N+M) matrix in the block layout of matA_add_M.
    def get_V_part(self, torch_input: torch.Tensor):
        N0, N = self._Neumann_offset_i, self.N_Neumann
        return torch_input[N0: N0 + N, self._Neumann_offset_j: self._Neumann_offset_j + N]

    def get_K_part(self, torch_input: torch.Tensor):
        return torch_input[self._Neumann_offset_i: self._Neumann_offset_i + self.N_Neumann,
                           self._Dirichlet_offset_j: self._Dirichlet_offset_j + self.M_Dirichlet]

    def get_adj_K_part(self, torch_input: torch.Tensor):
        return torch_input[self._Dirichlet_offset_i: self._Dirichlet_offset_i + self.M_Dirichlet,
                           self._Neumann_offset_j: self._Neumann_offset_j + self.N_Neumann]

    def get_W_part(self, torch_input: torch.Tensor):
        D0, M = self._Dirichlet_offset_i, self.M_Dirichlet
        return torch_input[D0: D0 + M, self._Dirichlet_offset_j: self._Dirichlet_offset_j + M]

    def set_V_part(self, old_tensor: torch.Tensor, new_tensor: torch.Tensor):
        self.get_V_part(old_tensor).copy_(new_tensor)

    def set_K_part(self, old_tensor: torch.Tensor, new_tensor: torch.Tensor):
        self.get_K_part(old_tensor).copy_(new_tensor)

    def set_adj_K_part(self, old_tensor: torch.Tensor, new_tensor: torch.Tensor):
        self.get_adj_K_part(old_tensor).copy_(new_tensor)

    def set_W_part(self, old_tensor: torch.Tensor, new_tensor: torch.Tensor):
        self.get_W_part(old_tensor).copy_(new_tensor)

    def compute_inv_norm(self, A: torch.Tensor, A_type: int) -> float:
        """max 1 / sigma_i of L^-1 A L^-H (BEM_manager.py:874-919): the
        inverse's norm in the energy norm of the Cholesky factor L of
        blkdiag(W, V) of the first matrix seen (A_type 1; (MV + MV^T)/2 as
        there) or of blkdiag(K', K) (A_type 2), kept for later calls.  The
        dense factorisation, inverse and singular values are the library
        routines the reference calls (torch.linalg, here on the device);
        the per-wave-number estimate of the batched sweep is
        bem_batch_condest."""
        if self._L1_inv_H is None and A_type == 1:
            MV = torch.zeros_like(A)
            self.set_W_part(MV, self.get_W_part(A))
            self.set_V_part(MV, self.get_V_part(A))
            MV = (MV + MV.T) / 2.0
            self._L1_inv_H = torch.linalg.inv(torch.linalg.cholesky(MV, upper=False)).T.conj()
            self._L2_inv_H = None
        elif self._L2_inv_H is None and A_type == 2:
            MV = torch.block_diag(self.get_adj_K_part(A), self.get_K_part(A))
            self._L2_inv_H = torch.linalg.inv(torch.linalg.cholesky(MV, upper=False)).T.conj()
            self._L1_inv_H = None
        LH = self._L1_inv_H if A_type == 1 else self._L2_inv_H
        if LH is None:
            raise RuntimeError("A_type must be 1 or 2")
        n = self.N_Neumann + self.M_Dirichlet
        A = A @ LH
        if A.shape[0] == n:
            A = LH.T.conj() @ A
        elif A.shape[0] == 2 * n:
            A = torch.block_diag(LH.T.conj(), LH.T.conj()) @ A
        else:
            raise NotImplementedError("We only support Non-Augmentd matrix and Augmented Matrix for Norm analysis")
        s_val = torch.linalg.svdvals(A)
        return (1.0 / s_val).max().item()

    def _assemble_norm_matrix(self, k: float, interior_multiplier: float):
        """_mat_A = P_MINUS(exterior) + m P_PLUS(interior) at wave number k
        (BEM_manager.py:921-927, 935-941)."""
        self._mat_A.zero_()
        self.matrix_layer_forward(k=k, sqrt_n=self._sqrt_no, scope_type=int(ScopeType.EXTERIOR))
        self.assemble_matA(assemble_type=int(AssembleType.ADD_P_MINUS), multiplier=1.0)
        self.matrix_layer_forward(k=k, sqrt_n=self._sqrt_ni, scope_type=int(ScopeType.INTERIOR))
        self.assemble_matA(assemble_type=int(AssembleType.ADD_P_PLUS), multiplier=interior_multiplier)

    def get_mat_A1_inv_norm(self, k: float) -> float:
        """A1 = P-(exterior) - P+(interior) (BEM_manager.py:921-933)."""
        self._assemble_norm_matrix(k, -1.0)
        return self.compute_inv_norm(self._mat_A, A_type=1)

    def get_mat_A2_inv_norm(self, k: float) -> float:
        """A2 = P-(exterior) + P+(interior) (BEM_manager.py:935-947)."""
        self._assemble_norm_matrix(k, 1.0)
        return self.compute_inv_norm(self._mat_A, A_type=2)

    def _augmented(self, k: float, need_recompute_matrix_layers: bool) -> torch.Tensor:
        """[_mat_A; P+(interior)] (BEM_manager.py:949-966)."""
        A = self._mat_A.clone()
        self._mat_A.zero_()
        if need_recompute_matrix_layers:
            self.matrix_layer_forward(k=k, sqrt_n=self._sqrt_ni, scope_type=int(ScopeType.INTERIOR))
        self.assemble_matA(assemble_type=int(AssembleType.ADD_P_PLUS), multiplier=1.0)
        return torch.cat((A, self._mat_A.clone()), 0)

    def get_augment_mat_A1_inv_norm(self, k: float, need_recompute_matrix_layers: bool) -> float:
        return self.compute_inv_norm(self._augmented(k, need_recompute_matrix_layers), A_type=1)

    def get_augment_mat_A2_inv_norm(self, k: float, need_recompute_matrix_layers: bool) -> float:
        return self.compute_inv_norm(self._augmented(k, need_recompute_matrix_layers), A_type=2)

    # -- post-processing (BEM_manager.py:664-801) ---------------------------------------
    def compute_analytical_raw_solved(self):
        mm = self._core_manager._mesh_manager
        rc = self.rhs_constructor
        self.raw_analytical_solved.zero_()
        if self._is_transmission:
            D, N = self._Dirichlet_offset_j, self._Neumann_offset_j
            self.raw_analytical_solved[D: D + self.M_Dirichlet] = rc._Dirichlet_boundary[: self.M_Dirichlet]
            self.raw_analytical_solved[N: N + self.N_Neumann] = rc._Neumann_boundary[: self.N_Neumann]
            return
        if self.N_Neumann > 0:
            off = self._Neumann_offset_j
            if self.shape_func_degree_Neumann == 0:
                v = self.analytical_function_Neumann(mm.panel_centroids(), mm.get_panel_normals())
            else:
                v = self.analytical_function_Neumann(mm.vertices, mm.get_vert_normals())
            self.raw_analytical_solved[off: off + self.N_Neumann] = rc._as_field(v)
        if self.M_Dirichlet > 0:
            off = self._Dirichlet_offset_j
            if self.shape_func_degree_Dirichlet == 0:
                v = self.analytical_function_Dirichlet(mm.panel_centroids())
            else:
                v = self.analytical_function_Dirichlet(mm.vertices)
            self.raw_analytical_solved[off: off + self.M_Dirichlet] = rc._as_field(v)

    def splat_u_from_raw_solved_to_vertices(self):
        """BEM_manager.py:753-801: the Neumann unknowns are added into the
        vertex fields unless only the Dirichlet field is shown, and the
        Dirichlet unknowns unless only the Neumann field is shown (`_show` is
        -1, both, except for transmission, where it is the Neumann field,
        BEM_manager.py:28,52-53).  P0 values are spread area-weighted over a
        panel's vertices, P1 values are added as they are; panels and vertices
        of the other unknown type are skipped."""
        self.solved.zero_()
        self.analytical_solved.zero_()
        show = int(VertAttachType.NEUMANN_TOBESOLVED) if self._is_transmission else -1
        if show != int(VertAttachType.DIRICHLET_TOBESOLVED) and self.N_Neumann > 0:
            self._splat(self._Neumann_offset_j, self.N_Neumann, self.shape_func_degree_Neumann,
                        int(CellFluxType.DIRICHLET_TOBESOLVED), int(VertAttachType.DIRICHLET_TOBESOLVED))
        if show != int(VertAttachType.NEUMANN_TOBESOLVED) and self.M_Dirichlet > 0:
            self._splat(self._Dirichlet_offset_j, self.M_Dirichlet, self.shape_func_degree_Dirichlet,
                        int(CellFluxType.NEUMANN_TOBESOLVED), int(VertAttachType.NEUMANN_TOBESOLVED))

    def _splat(self, off: int, n: int, degree: int, skip_panel: int, skip_vertex: int):
        mm = self._core_manager._mesh_manager
        sol = self.raw_solved[off: off + n]
        ana = self.raw_analytical_solved[off: off + n]
        if degree == 0:
            # the dof maps are the identity: local panel I is panel I
            tris = mm.panels.long()
            keep = (mm.get_panel_types() != skip_panel).to(self.solved.real.dtype)
            w = (mm.get_panel_areas() / 3.0).to(self.solved.real.dtype) * keep
            va = mm.get_vert_areas().to(self.solved.real.dtype)
            for s in range(3):
                self.solved.index_add_(0, tris[:, s], sol * w / va[tris[:, s]])
                self.analytical_solved.index_add_(0, tris[:, s], ana * w / va[tris[:, s]])
        else:
            keep = (mm.get_vertice_types() != skip_vertex).to(self.solved.real.dtype)
            self.solved += sol * keep
            self.analytical_solved += ana * keep

    def kill(self):
        self.initialized = False
        self.num_of_vertices = 0
        self.num_of_panels = 0
        self.solved = None
        self.analytical_solved = None
        self.raw_solved = None
        self.raw_analytical_solved = None
        self._mat_A = None
        self._LU = self._ipiv = None
        self.rhs = None
        for layer in ("single_layer", "double_layer", "adj_double_layer", "hypersingular_layer",
                      "identity_layer", "rhs_constructor"):
            if hasattr(self, layer):
                getattr(self, layer).kill()
