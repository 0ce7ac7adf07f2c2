"""Constraints of the pair path: G(r), g(r), S(Q), reduced S(Q), and the
inter/intra-molecular distance constraints."""
from .DistanceConstraints import InterMolecularDistanceConstraint, IntraMolecularDistanceConstraint  # noqa: F401
from .PairCorrelationConstraints import PairCorrelationConstraint  # noqa: F401
from .PairDistributionConstraints import PairDistributionConstraint  # noqa: F401
from .StructureFactorConstraints import ReducedStructureFactorConstraint, StructureFactorConstraint  # noqa: F401
