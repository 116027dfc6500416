from .single_layer import SingleLayer3d
from .double_layer import DoubleLayer3d
from .adj_double_layer import AdjDoubleLayer3d
from .hypersingular_layer import HypersingularLayer3d
from .identity_layer import IdentityLayer3d
from .rhs_constructor import RHSConstructor3d
