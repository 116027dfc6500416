"""Mesh manager (src/managers/mesh_manager.py): loads or generates the
surface, uploads it to the device, where the BEM library computes areas,
normals and the near-field list, and classifies the boundary-value problem
into dof maps.  Only boundary-value problems whose panels are all of one type
(Dirichlet, Neumann, Full) are supported: their local<->global maps are the
identity (mesh_manager.py:166-264)."""
from __future__ import annotations

import os

import numpy as np
import torch

from .. import meshes
from ..bem_functions.utils import CellFluxType, VertAttachType


class MeshManager:
    def __init__(self, core_manager):
        self._core_manager = core_manager
        p = core_manager._simulation_parameters
        self._dim = p.get("dim", 3)
        self._boundary_type = p["boundary"]
        self._is_transmission = core_manager._is_transmission
        self.vertices = None
        self.panels = None
        self.panel_areas = None
        self.vert_areas = None
        self.panel_normals = None
        self.vert_normals = None
        self.panel_types = None
        self.vertice_types = None
        self.num_of_vertices = 0
        self.num_of_panels = 0
        self.num_of_panels_Dirichlet = 0
        self.num_of_panels_Neumann = 0
        self.M_Dirichlet = 0
        self.N_Neumann = 0
        Q_Neumann = p.get("Q_Neumann", 0)
        Q_Dirichlet = p.get("Q_Dirichlet", 1)
        assert Q_Neumann in (0, 1)
        assert Q_Dirichlet in (0, 1)
        self.shape_func_degree_Neumann = Q_Neumann
        self.shape_func_degree_Dirichlet = Q_Dirichlet
        self.initialized = False

    # -- loading ----------------------------------------------------------------
    def load_asset(self):
        """Vertices/panels from the parameters: explicit arrays, a generated
        (perturbed) icosphere, or an OBJ file normalised like the reference
        (mesh_manager.py:479-494)."""
        p = self._core_manager._simulation_parameters
        if p.get("vertices") is not None:
            return np.asarray(p["vertices"], dtype=np.float64), np.asarray(p["panels"], dtype=np.int32)
        obj = p.get("object", "icosphere")
        if obj == "icosphere":
            return meshes.icosphere(int(p.get("level", 3)))
        if obj == "perturbed_icosphere":
            return meshes.perturbed_icosphere(int(p.get("level", 5)), seed=int(p.get("seed", 0)))
        path = obj if obj.endswith(".obj") else os.path.join(p.get("asset_path", "."), obj + ".obj")
        return meshes.load_obj(path)

    def initialization(self):
        if self._dim != 3:
            raise RuntimeError("The dimension of our scene should only be 2D or 3D (this build: 3D)")
        np_vertices, np_faces = self.load_asset()
        ctx = self._core_manager.ctx
        dev = self._core_manager.device
        self.vertices = torch.as_tensor(np_vertices, dtype=torch.float64, device=dev).contiguous()
        self.panels = torch.as_tensor(np_faces, dtype=torch.int32, device=dev).reshape(-1, 3).contiguous()
        self.num_of_vertices = int(self.vertices.shape[0])
        self.num_of_panels = int(self.panels.shape[0])
        ctx.set_mesh(self.vertices, self.panels, self._core_manager._simulation_parameters["GaussQR"])
        self.panel_areas, self.panel_normals, self.vert_areas, self.vert_normals = ctx.geometry()
        self.panel_types = torch.zeros(self.num_of_panels, dtype=torch.int32, device=dev)
        self.vertice_types = torch.zeros(self.num_of_vertices, dtype=torch.int32, device=dev)
        if self._boundary_type == "Full" or self._is_transmission > 0:
            self.set_full_bvp()
        elif self._boundary_type == "Dirichlet":
            self.set_Dirichlet_bvp()
        elif self._boundary_type == "Neumann":
            self.set_Neumann_bvp()
        elif self._boundary_type == "Mix":
            raise NotImplementedError("the Mix boundary-value problem is outside this hot path")
        else:
            raise RuntimeError("The Boundary Type can only be Dirichlet/Neumann/Mix")
        self.initialized = True

    def _n_dofs(self, q: int) -> int:
        return self.num_of_panels if q == 0 else self.num_of_vertices

    def set_full_bvp(self):
        self.N_Neumann = self._n_dofs(self.shape_func_degree_Neumann)
        self.M_Dirichlet = self._n_dofs(self.shape_func_degree_Dirichlet)
        self.num_of_panels_Dirichlet = self.num_of_panels
        self.num_of_panels_Neumann = self.num_of_panels
        self.panel_types.fill_(int(CellFluxType.BOTH_TOBESOLVED))
        self.vertice_types.fill_(int(VertAttachType.BOTH_TOBESOLVED))

    def set_Dirichlet_bvp(self):
        self.N_Neumann = self._n_dofs(self.shape_func_degree_Neumann)
        self.M_Dirichlet = 0
        self.num_of_panels_Dirichlet = 0
        self.num_of_panels_Neumann = self.num_of_panels
        self.panel_types.fill_(int(CellFluxType.NEUMANN_TOBESOLVED))
        self.vertice_types.fill_(int(VertAttachType.NEUMANN_TOBESOLVED))

    def set_Neumann_bvp(self):
        self.N_Neumann = 0
        self.M_Dirichlet = self._n_dofs(self.shape_func_degree_Dirichlet)
        self.num_of_panels_Dirichlet = self.num_of_panels
        self.num_of_panels_Neumann = 0
        self.panel_types.fill_(int(CellFluxType.DIRICHLET_TOBESOLVED))
        self.vertice_types.fill_(int(VertAttachType.DIRICHLET_TOBESOLVED))

    # -- getters (mesh_manager.py:391-477) ---------------------------------------
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

    def get_vertices(self):
        return self.vertices

    def get_panels(self):
        return self.panels

    def get_panel_areas(self):
        return self.panel_areas

    def get_vert_areas(self):
        return self.vert_areas

    def get_vert_normals(self):
        return self.vert_normals

    def get_panel_normals(self):
        return self.panel_normals

    def get_panel_types(self):
        return self.panel_types

    def get_vertice_types(self):
        return self.vertice_types

    def panel_centroids(self):
        return self.vertices[self.panels.long()].mean(dim=1)

    def kill(self):
        self.initialized = False
        self.vertices = None
        self.panels = None
        self.panel_areas = None
        self.vert_areas = None
        self.panel_normals = None
        self.vert_normals = None
        self.panel_types = None
        self.vertice_types = None
