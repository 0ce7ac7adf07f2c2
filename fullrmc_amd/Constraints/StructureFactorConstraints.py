"""StructureFactorConstraint with the reference's float32 semantics
(Constraints/StructureFactorConstraints.py): its own histogram grid
(rmin, rmax, dr), G(r) on that grid, S(Q) = sum_r G(r) * dr sin(rQ)/Q + 1;
ReducedStructureFactorConstraint: the same sum without the +1.
"""
import numpy as np

from ..Globals import FLOAT_TYPE, INT_TYPE, PI
from .PairDistributionConstraints import _HistogramConstraint


class StructureFactorConstraint(_HistogramConstraint):
    """S(Q) constraint.  ``experimentalData`` is (n,2): Q (>0) and S(Q)."""

    kind = 1

    def __init__(self, experimentalData, dataWeights=None, weighting="atomicNumber", atomsWeight=None,
                 rmin=None, rmax=None, dr=None, scaleFactor=1.0, adjustScaleFactor=(0, 0.8, 1.2),
                 windowFunction=None, limits=None):
        if limits is not None:
            raise NotImplementedError("limits are not part of this path")
        super().__init__(experimentalData, dataWeights, weighting, atomsWeight, scaleFactor, adjustScaleFactor,
                         windowFunction)
        x = self.experimentalData
        # set_limits (StructureFactorConstraints.py:686-705)
        self.experimentalQValues = x[:, 0].astype(FLOAT_TYPE)
        self.experimentalSF = x[:, 1].astype(FLOAT_TYPE)
        qmin, qmax = self.experimentalQValues[0], self.experimentalQValues[-1]
        if not qmin > 0:
            raise ValueError("qmin must be bigger than 0")
        # set_rmin / set_rmax / set_dr (:502-575)
        self.minimumDistance = FLOAT_TYPE(2. * PI / qmax) if rmin is None else FLOAT_TYPE(rmin)
        if rmax is None:
            dq = self.experimentalQValues[1] - self.experimentalQValues[0]
            self.maximumDistance = FLOAT_TYPE(2. * PI / dq)
        else:
            self.maximumDistance = FLOAT_TYPE(rmax)
        if dr is None:
            b = 2. * PI / qmax
            rb = round(b, 1)
            if rb > b:
                rb -= 0.1
            self.bin = FLOAT_TYPE(rb)
        else:
            self.bin = FLOAT_TYPE(dr)
        # __set_histogram (:330-360)
        self.edges = np.arange(self.minimumDistance, self.maximumDistance + self.bin, self.bin).astype(FLOAT_TYPE)
        self.minimumDistance = self.edges[0]
        self.maximumDistance = self.edges[-1]
        self.shellCenters = (self.edges[0:-1] + self.edges[1:]) / FLOAT_TYPE(2.)
        self.histogramSize = INT_TYPE(len(self.edges) - 1)
        self.shellVolumes = FLOAT_TYPE(4.0 / 3.) * PI * ((self.edges[1:]) ** 3 - self.edges[0:-1] ** 3)
        # __set_Gr_2_Sq_matrix (:302-312)
        Qs, Rs = self.experimentalQValues, self.shellCenters
        dR = self.shellCenters[1] - self.shellCenters[0]
        qr = Rs.reshape((-1, 1)) * (np.ones((len(Rs), 1), dtype=FLOAT_TYPE) * Qs)
        self.Gr2SqMatrix = dR * (np.sin(qr) / Qs)
        self._experimental = self.experimentalSF
        self._usedDataWeights = self._used_data_weights(x.shape[0])

    def _prefactor(self):
        return FLOAT_TYPE(4.) * PI * self.shellCenters * self.engine.numberDensity

    def _get_Sq_from_Gr(self, Gr):
        """StructureFactorConstraints.py:772-773"""
        return np.sum(Gr.reshape((-1, 1)) * self.Gr2SqMatrix, axis=0) + 1

    def _apply_scale_factor(self, Sq, scaleFactor):
        """StructureFactorConstraints.py:775-778"""
        if scaleFactor != 1:
            Sq = scaleFactor * (Sq - 1) + 1
        return Sq

    def _model(self, data, rho0=None, counts=None, weights=None):
        """__get_total_Sq (StructureFactorConstraints.py:780-822) without
        scale-factor fitting, multiframe prior or window."""
        rho0 = self.engine.numberDensity if rho0 is None else rho0
        Gr = self._gr_sum(data, cast_density=False, counts=counts, weights=weights)
        Gr /= self.shellVolumes
        Gr = (FLOAT_TYPE(4.) * PI * self.shellCenters * rho0) * (Gr - 1)
        Sq = self._get_Sq_from_Gr(Gr)
        return self._apply_scale_factor(Sq, FLOAT_TYPE(self.scaleFactor))

    def _device_arrays(self):
        e = self.engine
        T = len(self.elementsPairs)
        w = np.empty(T, FLOAT_TYPE)
        Dd = np.empty(T, FLOAT_TYPE)
        for t, (a, b) in enumerate(self.elementsPairs):
            wij = self.weightingScheme.get(a + "-" + b, self.weightingScheme.get(b + "-" + a))
            ni, nj = e.numberOfAtomsPerElement[a], e.numberOfAtomsPerElement[b]
            Nij = ni * (ni - 1) / 2.0 if a == b else ni * nj
            w[t] = wij
            Dd[t] = Nij / e.volume
        return dict(kind=self.kind, H=int(self.histogramSize), minD=self.minimumDistance, maxD=self.maximumDistance,
                    bin=self.bin, nQ=len(self.experimentalQValues), w=w, D=Dd,
                    shellVolumes=np.ascontiguousarray(self.shellVolumes),
                    prefactor=np.ascontiguousarray(self._prefactor(), dtype=FLOAT_TYPE),
                    experimental=np.ascontiguousarray(self._experimental), dataWeights=self._usedDataWeights,
                    sqMatrix=np.ascontiguousarray(self.Gr2SqMatrix, dtype=FLOAT_TYPE),
                    scaleFactor=FLOAT_TYPE(self.scaleFactor))


class ReducedStructureFactorConstraint(StructureFactorConstraint):
    """Reduced S(Q) constraint (StructureFactorConstraints.py:1237-1261):
    S(Q) normalised to 0, the sine transform of G(r) without the +1, and a
    scale factor that multiplies S(Q) itself.  ``experimentalData`` is (n,2):
    Q (>0) and S(Q)-1."""

    kind = 3

    def _get_Sq_from_Gr(self, Gr):
        """StructureFactorConstraints.py:1255-1256"""
        return np.sum(Gr.reshape((-1, 1)) * self.Gr2SqMatrix, axis=0)

    def _apply_scale_factor(self, Sq, scaleFactor):
        """StructureFactorConstraints.py:1258-1261"""
        if scaleFactor != 1:
            Sq = scaleFactor * Sq
        return Sq
