"""PairCorrelationConstraint with the reference's float32 semantics
(Constraints/PairCorrelationConstraints.py): the PairDistributionConstraint
grid and histograms, with g(r) as the model instead of G(r).
"""
from ..Globals import FLOAT_TYPE, PI
from .PairDistributionConstraints import PairDistributionConstraint


class PairCorrelationConstraint(PairDistributionConstraint):
    """g(r) constraint.  ``experimentalData`` is (n,2): r and g(r).

    g(r) = sum_ij w_ij n_ij(r) V / (N_ij v(r)), the G(r) sum before its
    4*PI*r*rho0*(g-1) transform; G(r) = 4*PI*r*rho0*(g(r)-1)
    (PairCorrelationConstraints.py:26-60)."""

    kind = 2

    def _model(self, data, rho0=None, counts=None, weights=None):
        """__get_total_gr (PairCorrelationConstraints.py:126-169) without
        shape function, scale-factor fitting, multiframe prior or window:
        a scale factor acts on G(r) and is mapped back to g(r)."""
        rho0 = self.engine.numberDensity if rho0 is None else rho0
        gr = self._gr_sum(data, cast_density=False, counts=counts, weights=weights)
        gr /= self.shellVolumes
        sf = FLOAT_TYPE(self.scaleFactor)
        if sf != 1:
            Gr = (4. * PI * self.shellCenters * rho0) * (gr - 1)
            Gr *= sf
            gr = 1. + Gr / (4. * PI * self.shellCenters * rho0)
        return gr
