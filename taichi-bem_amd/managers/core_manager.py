"""Core manager (src/managers/core_manager.py:11-148) without the GUI: turns
the simulation-parameter dict into the kernel/transmission flags, owns the
device context and the mesh and BEM managers."""
from __future__ import annotations

import torch

from .. import _lib
from .bem_manager import BEMManager
from .mesh_manager import MeshManager


class CoreManager:
    def __init__(self, simulation_parameters: dict, device: int = 0):
        self._simulation_parameters = dict(simulation_parameters)
        p = self._simulation_parameters
        p.setdefault("dim", 3)
        if "Helmholtz" not in p["kernel"]:
            p["k"] = 0  # core_manager.py:21-22
        # core_manager.py:24
        self._is_transmission = 1 if ("Full" in p["boundary"] or "Transmission" in p["kernel"]) else 0
        self.device = torch.device("cuda", device)
        self.ctx = _lib.Context(device)
        self._mesh_manager = MeshManager(self)
        self._BEM_manager = BEMManager(self)
        self.initialized = False

    def initialization(self, analytical_function_Dirichlet, analytical_function_Neumann):
        self._mesh_manager.initialization()
        self._BEM_manager.initialization(analytical_function_Dirichlet, analytical_function_Neumann)
        self.initialized = True

    def run(self):
        self._BEM_manager.run()

    def kill(self):
        self.initialized = False
        self._BEM_manager.kill()
        self._mesh_manager.kill()
        self.ctx.close()
