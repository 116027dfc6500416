"""Enumerations of the BEM path with the reference's integer values, so that
host code passes the same numbers across the ABI
(src/BEM_functions/utils/types.py:4-58)."""
from enum import Enum


class _IntEnum(Enum):
    def __int__(self):
        return self.value


class CellFluxType(_IntEnum):
    NEUMANN_TOBESOLVED = 0
    DIRICHLET_TOBESOLVED = 1
    MIX_NONRESOLVED = 2
    BOTH_TOBESOLVED = 3


class ScopeType(_IntEnum):
    NOT_DEFINED = 0
    INTERIOR = 1
    EXTERIOR = 2


class VertAttachType(_IntEnum):
    NEUMANN_TOBESOLVED = 0
    DIRICHLET_TOBESOLVED = 1
    BOTH_TOBESOLVED = 2


class KernelType(_IntEnum):
    LAPLACE = 0
    HELMHOLTZ = 1
    HELMHOLTZ_TRANSMISSION = 2


class PanelsRelation(_IntEnum):
    SEPARATE = 0
    COINCIDE = 1
    COMMON_VERTEX = 2
    COMMON_EDGE = 3


class AssembleType(_IntEnum):
    NOTHING = 0
    ADD_HALF_IDENTITY = 1
    ADD_M = 2
    ADD_P_PLUS = 3
    ADD_P_MINUS = 4
