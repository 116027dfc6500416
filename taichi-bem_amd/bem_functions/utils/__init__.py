from .types import CellFluxType, ScopeType, VertAttachType, KernelType, PanelsRelation, AssembleType
from .functions import get_gaussion_integration_points_and_weights
from .io import warp_tensor, unwarp_tensor
