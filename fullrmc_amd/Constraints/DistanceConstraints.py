"""Inter- and intra-molecular distance constraints with the reference's
float32 semantics (Constraints/DistanceConstraints.py:23-1210), counted
through the gfx950 drop-in ``fullrmc_amd.Core.atomic_distances``.

They are rigid constraints: the engine evaluates them before the
experimental constraints and a move they reject is never tried
(Engine.py:3278-3296).  The standard error is the sum over type pairs of
the mean distance by which pairs closer than the pair's minimum distance
fall short of it (reduceDistanceToUpper).
"""
import itertools
import random
import warnings

import numpy as np

from .. import _device as D
from .._lib import check, lib
from ..Core.atomic_distances import _flags as atomic_distance_flags
from ..Core.atomic_distances import full_atomic_distances_coords, multiple_atomic_distances_coords
from ..Globals import FLOAT_TYPE, INT_TYPE


class _MolecularDistanceConstraint:
    """Shared state machine (DistanceConstraints.py:23-470 and 470-820)."""

    rigid = True
    _interMolecular = True
    _intraMolecular = False
    # fixed flags of the molecular distance constraints (:479-483)
    _countWithinLimits = True
    _reduceDistance = False
    _reduceDistanceToUpper = True
    _reduceDistanceToLower = False

    def __init__(self, defaultDistance, typeDefinition, pairsDistanceDefinition, flexible, rejectProbability):
        # set_default_distance (:181-194), set_flexible (:163-179), set_reject_probability
        # (Core/Constraint.py:2064-2081)
        d = FLOAT_TYPE(defaultDistance)
        if not d >= 0:
            raise ValueError("defaultDistance must be positive")
        if not isinstance(flexible, bool):
            raise ValueError("flexible must be boolean")
        rp = FLOAT_TYPE(rejectProbability)
        if not 0 <= rp <= 1:
            raise ValueError("rejectProbability must be between 0 and 1")
        if typeDefinition not in ("name", "element"):
            raise ValueError("typeDefinition must be either 'name' or 'element'")
        self.defaultDistance = d
        self.flexible = flexible
        self.rejectProbability = rp
        self.typeDefinition = typeDefinition
        self.pairsDistanceDefinition = pairsDistanceDefinition
        self.engine = None
        self.data = None
        self.standardError = None
        self.afterMoveStandardError = None
        self.activeAtomsDataBeforeMove = None
        self.activeAtomsDataAfterMove = None
        self.tried = 0
        self.accepted = 0
        self._res = None  # resident device copies while the device engine's windows run

    # ------------------------------------------------------------- setup
    def _attach(self, engine):
        """set_type_definition (:195-263) and set_pairs_distance (:265-359)."""
        self.engine = engine
        if self.typeDefinition == "name":
            self.types, self.typesIndex = list(engine.names), np.array(engine.namesIndex)
            self.numberOfAtomsPerType = dict(engine.numberOfAtomsPerName)
        else:
            self.types, self.typesIndex = list(engine.elements), np.array(engine.elementsIndex)
            self.numberOfAtomsPerType = dict(engine.numberOfAtomsPerElement)
        self.numberOfTypes = len(self.types)
        self.typePairs = sorted(itertools.combinations_with_replacement(self.types, 2))
        self.typePairsIndex = np.array([[self.types.index(a), self.types.index(b)] for a, b in self.typePairs],
                                       dtype=INT_TYPE).reshape(-1, 2)
        self._set_pairs_distance(self.pairsDistanceDefinition)

    def _set_pairs_distance(self, definition):
        types = self.types
        pd = {}
        if definition is not None:
            if not isinstance(definition, (list, set, tuple)):
                raise ValueError("pairsDistanceDefinition must be a list")
            for pair in definition:
                pair = list(pair)
                if len(pair) != 3:
                    raise ValueError("pairsDistanceDefinition list pair item list must have three items")
                if pair[0] not in types or pair[1] not in types:
                    warnings.warn("pairsDistanceDefinition item %r names a type not in the engine; omitted" % (pair,))
                    continue
                pd.setdefault(pair[0], {})
                pd.setdefault(pair[1], {})
                dist = FLOAT_TYPE(pair[2])
                if not dist >= 0:
                    raise ValueError("pairsDistanceDefinition distances must be >= 0")
                # the first definition of a pair wins
                if pair[1] not in pd[pair[0]]:
                    pd[pair[0]][pair[1]] = dist
                if not (pair[0] in pd[pair[1]] and pair[0] != pair[1]):
                    pd[pair[1]][pair[0]] = dist
        for a in types:
            pd.setdefault(a, {})
            for b in types:
                pd.setdefault(b, {})
                if b not in pd[a]:
                    if a in pd[b]:
                        pd[a][b] = pd[b][a]
                    else:
                        pd[a][b] = self.defaultDistance
                        pd[b][a] = self.defaultDistance
                if pd[a][b] != pd[b][a]:
                    raise ValueError("types '%s' and '%s' pair distance definitions are in conflict" % (a, b))
        self.pairsDistance = pd
        T = self.numberOfTypes
        self.lowerLimitArray = np.zeros((T, T, 1), dtype=FLOAT_TYPE)
        self.upperLimitArray = np.zeros((T, T, 1), dtype=FLOAT_TYPE)
        for i1, a in enumerate(types):
            for i2, b in enumerate(types):
                self.upperLimitArray[i1, i2, 0] = FLOAT_TYPE(pd[a][b])
                self.upperLimitArray[i2, i1, 0] = FLOAT_TYPE(pd[a][b])

    # ------------------------------------------------------ computations
    def _kw(self):
        return dict(interMolecular=self._interMolecular, intraMolecular=self._intraMolecular,
                    reduceDistance=self._reduceDistance, reduceDistanceToUpper=self._reduceDistanceToUpper,
                    reduceDistanceToLower=self._reduceDistanceToLower, countWithinLimits=self._countWithinLimits)

    def _pick(self, nintra, dintra, ninter, dinter):
        if self._interMolecular:
            return ninter, dinter
        return nintra, dintra

    def _full(self, box, mol, types):
        e = self.engine
        return self._pick(*full_atomic_distances_coords(box, e.basisVectors, e.isPBC, mol, types, self.numberOfTypes,
                                                        self.lowerLimitArray, self.upperLimitArray, **self._kw()))

    def _group(self, rel):
        if self._res is not None:
            return self._group_resident(rel)
        e = self.engine
        kw = self._kw()
        nM, dM = self._pick(*multiple_atomic_distances_coords(
            rel, e.boxCoordinates, e.basisVectors, e.isPBC, e.moleculesIndex, self.typesIndex, self.numberOfTypes,
            self.lowerLimitArray, self.upperLimitArray, interMolecular=kw["interMolecular"],
            intraMolecular=kw["intraMolecular"], countWithinLimits=kw["countWithinLimits"],
            reduceDistanceToUpper=kw["reduceDistanceToUpper"], reduceDistanceToLower=kw["reduceDistanceToLower"],
            reduceDistance=kw["reduceDistance"], allAtoms=True))
        nF, dF = self._full(e.boxCoordinates[rel], e.moleculesIndex[rel], self.typesIndex[rel])
        return {"number": nM - nF, "distanceSum": dM - dF}

    # ------------------------------------------- resident device copies
    def bind_resident(self):
        """Device copies of the engine's box coordinates, molecule and type
        indexes, basis and limits, held while the device engine's windows
        run: an evaluation then uploads only the group's rows.  It runs the
        drop-in calls' kernels on the same inputs, so its counts and float32
        sums are theirs."""
        e = self.engine
        self._res = {"box": D.upload(np.asarray(e.boxCoordinates, np.float32)),
                     "mol": D.upload(np.asarray(e.moleculesIndex, np.int32)),
                     "types": D.upload(np.asarray(self.typesIndex, np.int32)),
                     "basis": D.upload(np.asarray(e.basisVectors, np.float32)),
                     "lo": D.upload(self.lowerLimitArray), "up": D.upload(self.upperLimitArray),
                     "flags": atomic_distance_flags(**self._kw()), "pbc": int(bool(e.isPBC))}

    def unbind_resident(self):
        self._res = None

    def resident_commit(self, rel, boxRows):
        """An accepted move's new box coordinates into the device copy."""
        if self._res is not None:
            idx = D.upload(np.asarray(rel, np.int32)).long()
            self._res["box"][idx] = D.upload(np.asarray(boxRows, np.float32))

    def _group_resident(self, rel):
        """_group on the resident copies, at the group's current host rows
        (the moved ones inside compute_after_move)."""
        e = self.engine
        rel = np.asarray(rel, np.int32)
        return self._groups_resident(rel, None, np.asarray(e.boxCoordinates[rel], np.float32))[0]

    def _groups_resident(self, rel, moved, rows=None):
        """_group before the move (the resident rows, or rows) and, with
        moved, after it: for each state the group-vs-all call with the
        group's rows patched in (restored after it) and the group-alone call
        on those rows, outputs from zeros as the drop-in calls take them,
        downloaded once."""
        t = D.torch()
        r = self._res
        rel = np.asarray(rel, np.int32)
        g, T, n = len(rel), self.numberOfTypes, r["box"].shape[0]
        idx = D.upload(rel)
        idx64 = idx.long()
        resident = r["box"][idx64]
        states = [resident if rows is None else D.upload(rows)]
        if moved is not None:
            states.append(D.upload(np.asarray(moved, np.float32)))
        cnt = D.zeros((4 * len(states), T, T), np.int32)
        dst = D.zeros((4 * len(states), T, T), np.float32)
        mol_g, types_g = r["mol"][idx64], r["types"][idx64]
        ar = t.arange(g, dtype=t.int32, device="cuda")
        st = D.stream()
        L = lib()
        for si, rows_d in enumerate(states):
            o = 4 * si
            if rows_d is not resident:
                r["box"][idx64] = rows_d
            check(L.frmc_multiple_atomic_distances_coords(
                idx.data_ptr(), g, r["box"].data_ptr(), n, r["basis"].data_ptr(), r["pbc"], r["mol"].data_ptr(),
                r["types"].data_ptr(), T, r["lo"].data_ptr(), r["up"].data_ptr(), r["flags"], 1,
                cnt[o].data_ptr(), dst[o].data_ptr(), cnt[o + 1].data_ptr(), dst[o + 1].data_ptr(), st))
            if rows_d is not resident:
                r["box"][idx64] = resident
            check(L.frmc_multiple_atomic_distances_coords(
                ar.data_ptr(), g, rows_d.data_ptr(), g, r["basis"].data_ptr(), r["pbc"], mol_g.data_ptr(),
                types_g.data_ptr(), T, r["lo"].data_ptr(), r["up"].data_ptr(), r["flags"], 0, cnt[o + 2].data_ptr(),
                dst[o + 2].data_ptr(), cnt[o + 3].data_ptr(), dst[o + 3].data_ptr(), st))
        c = D.download(cnt).reshape(-1, T, T, 1)
        d = D.download(dst).reshape(-1, T, T, 1)
        out = []
        for si in range(len(states)):
            o = 4 * si
            nM, dM = self._pick(c[o], d[o], c[o + 1], d[o + 1])
            nF, dF = self._pick(c[o + 2], d[o + 2], c[o + 3], d[o + 3])
            out.append({"number": nM - nF, "distanceSum": dM - dF})
        return out

    def _get_constraint_value(self, data=None):
        """Mean counted distance per type pair (:415-428)."""
        data = self.data if data is None else data
        idi = self.typePairsIndex[:, 0]
        idj = self.typePairsIndex[:, 1]
        numbers = (data["number"][idi, idj] + data["number"][idj, idi]).reshape(-1)
        distances = (data["distanceSum"][idi, idj] + data["distanceSum"][idj, idi]).reshape(-1)
        nonZero = np.where(numbers)
        distances[nonZero] /= numbers[nonZero]
        return distances

    def _compute_standard_error(self, distances):
        """:385-386"""
        return FLOAT_TYPE(np.sum(distances))

    def get_constraint_value(self):
        d = self._get_constraint_value()
        return {"md_%s-%s" % pair: d[i] for i, pair in enumerate(self.typePairs)}

    # ------------------------------------------------ constraint protocol
    def compute_data(self):
        """:552-604"""
        e = self.engine
        number, distanceSum = self._full(e.boxCoordinates, e.moleculesIndex, self.typesIndex)
        self.data = {"number": number, "distanceSum": distanceSum}
        self.standardError = self._compute_standard_error(self._get_constraint_value(self.data))
        self.activeAtomsDataBeforeMove = None
        self.activeAtomsDataAfterMove = None
        return self.data, self.standardError

    def compute_before_move(self, realIndexes, relativeIndexes):
        """:606-662"""
        self.activeAtomsDataBeforeMove = self._group(relativeIndexes)
        self.activeAtomsDataAfterMove = None

    def compute_after_move(self, realIndexes, relativeIndexes, movedBoxCoordinates):
        """:664-737"""
        e = self.engine
        saved = np.array(e.boxCoordinates[relativeIndexes], dtype=FLOAT_TYPE)
        e.boxCoordinates[relativeIndexes] = movedBoxCoordinates
        self.activeAtomsDataAfterMove = self._group(relativeIndexes)
        e.boxCoordinates[relativeIndexes] = saved
        self._after_move_totals()

    def _after_move_totals(self):
        b, a = self.activeAtomsDataBeforeMove, self.activeAtomsDataAfterMove
        self._newData = {"number": self.data["number"] - b["number"] + a["number"],
                         "distanceSum": self.data["distanceSum"] - b["distanceSum"] + a["distanceSum"]}
        self.afterMoveStandardError = self._compute_standard_error(self._get_constraint_value(self._newData))
        self.tried += 1

    def compute_move(self, realIndexes, relativeIndexes, movedBoxCoordinates):
        """compute_before_move then compute_after_move; on the resident
        copies both in one pass with one download."""
        if self._res is None:
            self.compute_before_move(realIndexes, relativeIndexes)
            self.compute_after_move(realIndexes, relativeIndexes, movedBoxCoordinates)
            return
        self.activeAtomsDataBeforeMove, self.activeAtomsDataAfterMove = self._groups_resident(
            relativeIndexes, movedBoxCoordinates)
        self._after_move_totals()

    def should_step_get_rejected(self, standardError):
        """:361-383; flexible: Core/Constraint.py:2083-2099 (one Python
        random draw when the standard error grows)."""
        if self.flexible:
            if standardError <= self.standardError:
                return False
            return random.random() < self.rejectProbability
        return bool(np.any(self.activeAtomsDataAfterMove["number"] > self.activeAtomsDataBeforeMove["number"]))

    def after_move_state(self):
        """What accept_move takes from the last compute_after_move."""
        return self._newData, self.afterMoveStandardError

    def set_after_move_state(self, state):
        self._newData, self.afterMoveStandardError = state

    def rejection_draws(self, standardError):
        """The Python random draws should_step_get_rejected makes for this
        standard error: one when a flexible constraint's error grows."""
        return 1 if self.flexible and not standardError <= self.standardError else 0

    def accept_move(self, realIndexes, relativeIndexes):
        """:739-758"""
        self.data = self._newData
        self.activeAtomsDataBeforeMove = self.activeAtomsDataAfterMove = None
        self.standardError = self.afterMoveStandardError
        self.afterMoveStandardError = None
        self.accepted += 1

    def reject_move(self, realIndexes, relativeIndexes):
        """:760-772"""
        self.activeAtomsDataBeforeMove = self.activeAtomsDataAfterMove = None
        self.afterMoveStandardError = None

    # --- amputation (remove moves), :774-819 and :451-464
    def compute_as_if_amputated(self, realIndex, relativeIndex):
        pass

    def accept_amputation(self, realIndex, relativeIndex):
        """The removed atom's counted pairs leave the data."""
        self.compute_before_move(realIndexes=realIndex, relativeIndexes=relativeIndex)
        b = self.activeAtomsDataBeforeMove
        self.data = {"number": self.data["number"] - b["number"],
                     "distanceSum": self.data["distanceSum"] - b["distanceSum"]}
        self.activeAtomsDataBeforeMove = None
        self.standardError = self._compute_standard_error(self._get_constraint_value())

    def reject_amputation(self, realIndex, relativeIndex):
        pass

    def _on_collector_collect_atom(self, realIndex, relativeIndex):
        t = self.types[int(self.typesIndex[relativeIndex])]
        self.typesIndex = np.delete(self.typesIndex, relativeIndex, axis=0)
        self.numberOfAtomsPerType[t] -= 1


class InterMolecularDistanceConstraint(_MolecularDistanceConstraint):
    """Minimum distances between atoms of different molecules
    (DistanceConstraints.py:966-1090)."""

    _interMolecular = True
    _intraMolecular = False

    def __init__(self, defaultDistance=1.5, typeDefinition="element", pairsDistanceDefinition=None, flexible=True,
                 rejectProbability=1):
        super().__init__(defaultDistance, typeDefinition, pairsDistanceDefinition, flexible, rejectProbability)


class IntraMolecularDistanceConstraint(_MolecularDistanceConstraint):
    """Minimum distances between atoms of the same molecule
    (DistanceConstraints.py:1093-1210)."""

    _interMolecular = False
    _intraMolecular = True

    def __init__(self, defaultDistance=1.5, typeDefinition="name", pairsDistanceDefinition=None, flexible=True,
                 rejectProbability=1):
        super().__init__(defaultDistance, typeDefinition, pairsDistanceDefinition, flexible, rejectProbability)
