"""Experimental constraints of the pair path: G(r) and S(Q)."""
from .PairDistributionConstraints import PairDistributionConstraint  # noqa: F401
from .StructureFactorConstraints import StructureFactorConstraint  # noqa: F401
