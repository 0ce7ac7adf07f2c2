"""PairDistributionConstraint with the reference's float32 semantics
(Constraints/PairDistributionConstraints.py), histograms through the gfx950
drop-in ``fullrmc_amd.Core.pairs_histograms``.  PairCorrelationConstraint
(Constraints/PairCorrelationConstraints.py) lives in PairCorrelationConstraints.

Supported options: experimental data (n,2), data weights, weighting scheme
or explicit atoms weights, constant scale factor.  Shape functions, window
functions, scale-factor fitting and multiframe priors (all no-ops in the
benchmark configurations) are not part of this path and raise.
"""
import itertools

import numpy as np

from ..Core.pairs_histograms import full_pairs_histograms_coords, multiple_pairs_histograms_coords
from ..Globals import FLOAT_TYPE, INT_TYPE, PI, PRECISION

# Atomic numbers for the 'atomicNumber' weighting.  The reference reads them
# from pdbparser's element database (not available here).
ATOMIC_NUMBER = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10,
                 "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P": 15, "S": 16, "Cl": 17, "Ar": 18, "K": 19,
                 "Ca": 20, "Ti": 22, "Cr": 24, "Mn": 25, "Fe": 26, "Co": 27, "Ni": 28, "Cu": 29, "Zn": 30,
                 "Ge": 32, "Se": 34, "Br": 35, "Kr": 36, "Zr": 40, "Ag": 47, "Sn": 50, "Xe": 54, "Au": 79,
                 "Pb": 82}


def elements_weight(elements, atomsWeight, weighting):
    """Per-element weights: explicit ``atomsWeight`` first, else the
    weighting scheme (Core/Collection.py:91-120)."""
    out = {}
    for el in elements:
        w = (atomsWeight or {}).get(el)
        if w is None:
            key = el.capitalize()  # element symbols are matched without regard to case
            if weighting != "atomicNumber" or key not in ATOMIC_NUMBER:
                raise ValueError("no '%s' weight for element '%s'; pass atomsWeight" % (weighting, el))
            w = ATOMIC_NUMBER[key]
        out[el] = FLOAT_TYPE(w)
    return out


def normalized_weighting(numbers, weights):
    """Pair weights c_i c_j b_i b_j / (sum_k c_k b_k)^2, doubled for unlike
    pairs (the normalisation pdbparser's get_normalized_weighting provides to
    the reference), as float32 keyed 'A-B' for sorted pairs."""
    els = sorted(numbers)
    total = float(sum(numbers.values()))
    c = {e: numbers[e] / total for e in els}
    mean = sum(c[e] * float(weights[e]) for e in els)
    out = {}
    for a, b in itertools.combinations_with_replacement(els, 2):
        w = c[a] * c[b] * float(weights[a]) * float(weights[b]) / mean ** 2
        if a != b:
            w *= 2.0
        out[a + "-" + b] = FLOAT_TYPE(w)
    return out


class _HistogramConstraint:
    """State machine shared by PDC and SFC (Core/Constraint.py:27-160)."""

    def __init__(self, experimentalData, dataWeights, weighting, atomsWeight, scaleFactor, adjustScaleFactor,
                 windowFunction, shapeFuncParams=None):
        if windowFunction is not None or shapeFuncParams is not None:
            raise NotImplementedError("window and shape functions are not part of this path")
        if adjustScaleFactor is not None and adjustScaleFactor[0]:
            raise NotImplementedError("scale-factor fitting is not part of this path")
        # set_experimental_data (Core/Constraint.py:1240-1279): a path is read
        # by np.loadtxt as float32; lists and arrays are cast
        if isinstance(experimentalData, str):
            data = np.loadtxt(experimentalData, dtype=FLOAT_TYPE)
        else:
            data = np.asarray(experimentalData)
        self.experimentalData = data.astype(FLOAT_TYPE)
        self._check_data(self.experimentalData)
        self.dataWeights = None if dataWeights is None else np.array(dataWeights, dtype=FLOAT_TYPE)
        self.weighting = weighting
        self.atomsWeight = atomsWeight
        self.scaleFactor = scaleFactor
        self.engine = None
        self.data = None
        self.standardError = None
        self.afterMoveStandardError = None
        self.activeAtomsDataBeforeMove = None
        self.activeAtomsDataAfterMove = None
        self.tried = 0
        self.accepted = 0
        self.varianceSquared = 1

    @staticmethod
    def _check_data(x):
        if x.ndim != 2 or x.shape[1] != 2:
            raise ValueError("experimentalData must be an (n,2) array")
        if not np.all((np.array(sorted(x[:, 0]), dtype=FLOAT_TYPE) - x[:, 0]) <= PRECISION):
            raise ValueError("experimentalData abscissae are not sorted")
        if x[0, 0] < 0:
            raise ValueError("experimentalData abscissae must be non-negative")

    def _used_data_weights(self, n):
        """_set_used_data_weights (Core/Constraint.py:1332-1349)."""
        if self.dataWeights is None:
            return None
        w = np.copy(self.dataWeights[:n])
        if not np.sum(w):
            raise ValueError("dataWeights are all zero")
        w /= FLOAT_TYPE(np.sum(w))
        w *= FLOAT_TYPE(len(w))
        return w

    def _attach(self, engine):
        self.engine = engine
        self._elementsWeight = elements_weight(engine.elements, self.atomsWeight, self.weighting)
        self.weightingScheme = normalized_weighting(engine.numberOfAtomsPerElement, self._elementsWeight)
        self.elementsPairs = sorted(itertools.combinations_with_replacement(engine.elements, 2))

    # --- histograms (the drop-in calls, exactly as the reference makes them)
    def _full(self):
        e = self.engine
        return full_pairs_histograms_coords(e.boxCoordinates, e.basisVectors, e.isPBC, e.moleculesIndex,
                                            e.elementsIndex, e.numberOfElements, self.minimumDistance,
                                            self.maximumDistance, self.bin, self.histogramSize)

    def _group(self, relativeIndexes):
        e = self.engine
        iM, eM = multiple_pairs_histograms_coords(relativeIndexes, e.boxCoordinates, e.basisVectors, e.isPBC,
                                                  e.moleculesIndex, e.elementsIndex, e.numberOfElements,
                                                  self.minimumDistance, self.maximumDistance, self.bin,
                                                  self.histogramSize, allAtoms=True)
        iF, eF = full_pairs_histograms_coords(e.boxCoordinates[relativeIndexes], e.basisVectors, e.isPBC,
                                              e.moleculesIndex[relativeIndexes], e.elementsIndex[relativeIndexes],
                                              e.numberOfElements, self.minimumDistance, self.maximumDistance,
                                              self.bin, self.histogramSize)
        return {"intra": iM - iF, "inter": eM - eF}

    def _gr_sum(self, data, cast_density, counts=None, weights=None):
        """sum over element pairs of w_ij n_ij / D_ij; counts and weights
        default to the engine's atom counts and the weighting scheme."""
        e = self.engine
        counts = e.numberOfAtomsPerElement if counts is None else counts
        weights = self.weightingScheme if weights is None else weights
        Gr = np.zeros(self.histogramSize, dtype=FLOAT_TYPE)
        for a, b in self.elementsPairs:
            wij = weights.get(a + "-" + b)
            if wij is None:
                wij = weights[b + "-" + a]
            ni, nj = counts[a], counts[b]
            i, j = e.elements.index(a), e.elements.index(b)
            if i == j:
                Nij = ni * (ni - 1) / 2.0
                nij = data["intra"][i, j, :] + data["inter"][i, j, :]
            else:
                Nij = ni * nj
                nij = data["intra"][i, j, :] + data["intra"][j, i, :] + data["inter"][i, j, :] + \
                    data["inter"][j, i, :]
            Dij = Nij / e.volume
            if cast_density:
                Dij = FLOAT_TYPE(Dij)
            Gr += wij * nij / Dij
        return Gr

    # --- constraint protocol
    def compute_data(self):
        intra, inter = self._full()
        self.data = {"intra": intra, "inter": inter}
        self.standardError = self.compute_standard_error(self._model(self.data))
        self.activeAtomsDataBeforeMove = None
        self.activeAtomsDataAfterMove = None
        return self.data, self.standardError

    def compute_before_move(self, realIndexes, relativeIndexes):
        self.activeAtomsDataBeforeMove = self._group(relativeIndexes)
        self.activeAtomsDataAfterMove = None

    def compute_after_move(self, realIndexes, relativeIndexes, movedBoxCoordinates):
        e = self.engine
        saved = np.array(e.boxCoordinates[relativeIndexes], dtype=FLOAT_TYPE)
        e.boxCoordinates[relativeIndexes] = movedBoxCoordinates
        self.activeAtomsDataAfterMove = self._group(relativeIndexes)
        e.boxCoordinates[relativeIndexes] = saved
        b, a = self.activeAtomsDataBeforeMove, self.activeAtomsDataAfterMove
        self._newData = {"intra": self.data["intra"] - b["intra"] + a["intra"],
                         "inter": self.data["inter"] - b["inter"] + a["inter"]}
        self.afterMoveStandardError = self.compute_standard_error(self._model(self._newData))
        self.tried += 1

    def accept_move(self, realIndexes, relativeIndexes):
        self.data = self._newData
        self.activeAtomsDataBeforeMove = self.activeAtomsDataAfterMove = None
        self.standardError = self.afterMoveStandardError
        self.afterMoveStandardError = None
        self.accepted += 1

    def reject_move(self, realIndexes, relativeIndexes):
        self.activeAtomsDataBeforeMove = self.activeAtomsDataAfterMove = None
        self.afterMoveStandardError = None

    # --- amputation (remove moves), PairDistributionConstraints.py:1168-1233,
    #     StructureFactorConstraints.py:1098-1163, PairCorrelationConstraints.py:394-436
    def compute_as_if_amputated(self, realIndex, relativeIndex):
        """Data and standard error as if the atom were removed: its pairs
        taken out, the atom counts, weighting scheme and number density of
        the system without it."""
        e = self.engine
        self.compute_before_move(realIndexes=realIndex, relativeIndexes=relativeIndex)
        b = self.activeAtomsDataBeforeMove
        data = {"intra": self.data["intra"] - b["intra"], "inter": self.data["inter"] - b["inter"]}
        el = e.allElements[relativeIndex[0]]
        counts = dict(e.numberOfAtomsPerElement)
        counts[el] -= 1
        weights = normalized_weighting(counts, self._elementsWeight)
        rho0 = FLOAT_TYPE((e.numberOfAtoms - 1) / e.volume)
        self.amputationStandardError = self.compute_standard_error(self._model(data, rho0, counts, weights))
        self.amputationData = {"data": data, "weightingScheme": weights}
        self.activeAtomsDataBeforeMove = None

    def accept_amputation(self, realIndex, relativeIndex):
        self.data = self.amputationData["data"]
        self.weightingScheme = self.amputationData["weightingScheme"]
        self.standardError = self.amputationStandardError
        self.amputationData = None
        self.amputationStandardError = None

    def reject_amputation(self, realIndex, relativeIndex):
        self.amputationData = None
        self.amputationStandardError = None

    def _on_collector_collect_atom(self, realIndex, relativeIndex):
        pass

    def compute_standard_error(self, modelData):
        diff = self._experimental - modelData
        if self._usedDataWeights is None:
            return np.add.reduce((diff) ** 2)
        return np.add.reduce(self._usedDataWeights * ((diff) ** 2))

    def get_constraint_value(self):
        return {"total": self._model(self.data)}


class PairDistributionConstraint(_HistogramConstraint):
    """G(r) constraint.  ``experimentalData`` is (n,2): r and G(r)."""

    kind = 0

    def __init__(self, experimentalData, dataWeights=None, weighting="atomicNumber", atomsWeight=None,
                 scaleFactor=1.0, adjustScaleFactor=(0, 0.8, 1.2), shapeFuncParams=None, windowFunction=None,
                 limits=None):
        if limits is not None:
            raise NotImplementedError("limits are not part of this path")
        super().__init__(experimentalData, dataWeights, weighting, atomsWeight, scaleFactor, adjustScaleFactor,
                         windowFunction, shapeFuncParams)
        x = self.experimentalData
        # set_experimental_data / set_limits (PairDistributionConstraints.py:722, 748-758)
        self.bin = FLOAT_TYPE(x[1, 0] - x[0, 0])
        self.minimumDistance = FLOAT_TYPE(x[0, 0] - self.bin / 2.)
        self.maximumDistance = FLOAT_TYPE(x[-1, 0] + self.bin / 2.)
        self.shellCenters = np.array([x[i, 0] for i in range(x.shape[0])], dtype=FLOAT_TYPE)
        edges = [x[i, 0] - self.bin / 2. for i in range(x.shape[0])]
        edges.append(x[-1, 0] + self.bin / 2.)
        self.edges = np.array(edges, dtype=FLOAT_TYPE)
        self.histogramSize = INT_TYPE(len(self.edges) - 1)
        self.shellVolumes = FLOAT_TYPE(4.0 / 3.) * PI * ((self.edges[1:]) ** 3 - self.edges[0:-1] ** 3)
        self._experimental = x[:, 1]
        self._usedDataWeights = self._used_data_weights(x.shape[0])

    def _prefactor(self):
        return 4. * PI * self.shellCenters * self.engine.numberDensity

    def _model(self, data, rho0=None, counts=None, weights=None):
        """__get_total_Gr (PairDistributionConstraints.py:847-895) without
        shape function, scale-factor fitting, multiframe prior or window."""
        rho0 = self.engine.numberDensity if rho0 is None else rho0
        Gr = self._gr_sum(data, cast_density=True, counts=counts, weights=weights)
        Gr /= self.shellVolumes
        Gr = (4. * PI * self.shellCenters * rho0) * (Gr - 1)
        if self.scaleFactor != 1:
            Gr *= FLOAT_TYPE(self.scaleFactor)
        return Gr

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
            Dd[t] = FLOAT_TYPE(Nij / e.volume)
        return dict(kind=self.kind, H=int(self.histogramSize), minD=self.minimumDistance, maxD=self.maximumDistance,
                    bin=self.bin, nQ=0, w=w, D=Dd, shellVolumes=np.ascontiguousarray(self.shellVolumes),
                    prefactor=np.ascontiguousarray(self._prefactor(), dtype=FLOAT_TYPE),
                    experimental=np.ascontiguousarray(self._experimental), dataWeights=self._usedDataWeights,
                    sqMatrix=None, scaleFactor=FLOAT_TYPE(self.scaleFactor))
