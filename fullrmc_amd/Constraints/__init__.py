"""Experimental constraints of the pair path: G(r), g(r), S(Q) and reduced S(Q)."""
from .PairCorrelationConstraints import PairCorrelationConstraint  # noqa: F401
from .PairDistributionConstraints import PairDistributionConstraint  # noqa: F401
from .StructureFactorConstraints import ReducedStructureFactorConstraint, StructureFactorConstraint  # noqa: F401
