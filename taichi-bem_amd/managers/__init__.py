from .mesh_manager import MeshManager
from .bem_manager import BEMManager
from .core_manager import CoreManager
