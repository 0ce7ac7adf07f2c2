"""Engine: the stochastic run loop of fullrmc's Engine.run for the pair path.

Two ways to run the same trajectory:

* :meth:`Engine.run` -- the reference's loop structure (Engine.py:3377-3516):
  every move selects a group, generates the move, and asks each constraint
  for its before/after histograms through the gfx950 drop-in extension
  calls, then applies the accept rule.  One move per kernel launch.
* :meth:`Engine.run_device` -- the state lives in HBM and k proposals are
  evaluated per window against it (histogram deltas, G(r), S(Q),
  chi-squares on the GPU); the first accepted proposal is committed and the
  random streams are rewound to just after it, so the accepted-move
  sequence, histograms and chi-squares equal :meth:`run`'s.

Both use fullrmc's two global streams, Python's ``random`` and NumPy's
legacy ``numpy.random``, exactly where the reference draws from them.
"""
import collections
import ctypes
import random
import time

import numpy as np

from . import _lib
from .Core.boundary_conditions_collection import transform_coordinates
from .Generators import RandomSelector, TranslationGenerator
from .Globals import FLOAT_TYPE, INT_TYPE
from .systems import box_coordinates, reciprocal, volume


class Engine:
    """Atomic configuration, groups, constraints and the run loop."""

    def __init__(self):
        self.realCoordinates = None
        self.boxCoordinates = None
        self.groups = []
        self.generators = []
        self.constraints = []
        self.rigidConstraints = []
        self.selector = RandomSelector()
        self.tolerance = FLOAT_TYPE(0)
        self.generated = self.tried = self.accepted = self.tolerated = 0
        self.removed = [0, 0]  # removals tried, accepted
        self._last_accepted = False
        self.collectedAtoms = []
        self._collected_set = set()
        self.totalStandardError = None
        self._dev = None
        self._dev_kmax = 0
        self._dev_synced = False

    # ------------------------------------------------------------ system
    def set_system(self, realCoordinates, allElements, moleculesIndex=None, basisVectors=None,
                   reciprocalBasisVectors=None, allNames=None):
        """Coordinates (n,3), element name per atom, molecule index per atom
        and periodic cell vectors as rows (number = cube edge); atom names
        (default: the elements) for constraints typed by name.  Replaces
        Engine.set_pdb + set_boundary_conditions (Engine.py:2134-2298)."""
        real = np.array(realCoordinates, dtype=FLOAT_TYPE)
        if real.ndim != 2 or real.shape[1] != 3:
            raise ValueError("realCoordinates must be (n,3)")
        n = real.shape[0]
        if len(allElements) != n:
            raise ValueError("one element per atom is required")
        if basisVectors is None:
            raise NotImplementedError("infinite boundary conditions are not part of this path")
        if np.isscalar(basisVectors):
            basisVectors = np.diag([basisVectors] * 3)
        self.basisVectors = np.array(basisVectors, dtype=FLOAT_TYPE)
        self.reciprocalBasisVectors = (reciprocal(self.basisVectors) if reciprocalBasisVectors is None
                                       else np.array(reciprocalBasisVectors, dtype=FLOAT_TYPE))
        self.volume = volume(self.basisVectors)
        self.isPBC = True
        self.realCoordinates = real
        self.boxCoordinates = box_coordinates(self.reciprocalBasisVectors, real)
        self.allElements = list(allElements)
        self.elements = sorted(set(self.allElements))
        lut = {e: i for i, e in enumerate(self.elements)}
        self.elementsIndex = np.array([lut[e] for e in self.allElements], dtype=INT_TYPE)
        self.numberOfElements = len(self.elements)
        self.numberOfAtomsPerElement = {}
        for e in self.allElements:
            self.numberOfAtomsPerElement[e] = self.numberOfAtomsPerElement.get(e, 0) + 1
        self.moleculesIndex = (np.arange(n, dtype=INT_TYPE) if moleculesIndex is None
                               else np.array(moleculesIndex, dtype=INT_TYPE))
        self.allNames = list(self.allElements if allNames is None else allNames)
        if len(self.allNames) != n:
            raise ValueError("one name per atom is required")
        self.names = sorted(set(self.allNames))
        nlut = {e: i for i, e in enumerate(self.names)}
        self.namesIndex = np.array([nlut[e] for e in self.allNames], dtype=INT_TYPE)
        self.numberOfAtomsPerName = {}
        for e in self.allNames:
            self.numberOfAtomsPerName[e] = self.numberOfAtomsPerName.get(e, 0) + 1
        self.numberOfAtoms = n
        self.originalNumberOfAtoms = n
        self.numberDensity = FLOAT_TYPE(n) / FLOAT_TYPE(self.volume)
        self._realIndex = np.arange(n, dtype=INT_TYPE)
        self._collected_set = set()
        self.collectedAtoms = []
        self.set_groups_as_atoms()
        self._drop_device()

    def set_pdb(self, pdb, boundaryConditions=None, names=None, elements=None, moleculesIndex=None):
        """Engine.set_pdb (Engine.py:2134-2208): a PDB path (``.gz`` allowed)
        or a :func:`fullrmc_amd.Core.pdb.read_pdb` dict.  The cell comes from
        ``boundaryConditions`` (number, 3 or 9 values) if given, else from
        the file; names, elements and molecules from the file unless given."""
        from .Core.pdb import _bc_vectors, read_pdb
        data = read_pdb(pdb) if isinstance(pdb, str) else pdb
        bc = data["basis"] if boundaryConditions is None else _bc_vectors(np.ravel(boundaryConditions))
        if bc is None:
            raise NotImplementedError("infinite boundary conditions are not part of this path")
        self.set_system(data["coordinates"], list(data["elements"] if elements is None else elements),
                        data["moleculesIndex"] if moleculesIndex is None else moleculesIndex,
                        basisVectors=np.array(bc, dtype=FLOAT_TYPE),
                        allNames=list(data["names"] if names is None else names))

    def set_groups_as_atoms(self):
        self.set_groups([np.array([i], dtype=INT_TYPE) for i in range(self.numberOfAtoms)])

    def set_groups_as_molecules(self):
        order = {}
        for i, m in enumerate(self.moleculesIndex):
            order.setdefault(int(m), []).append(i)
        self.set_groups([np.array(v, dtype=INT_TYPE) for v in order.values()])

    def set_groups(self, groups, generator=None):
        """Groups of atom indexes; ``generator`` is one move generator for
        every group or a list with one per group (default: translations of
        0.2 A, as the reference's groups get)."""
        self.groups = [np.array(g, dtype=INT_TYPE) for g in groups]
        if isinstance(generator, (list, tuple)):
            if len(generator) != len(self.groups):
                raise ValueError("one move generator per group is required")
            self.generators = list(generator)
        else:
            gen = generator if generator is not None else TranslationGenerator(amplitude=0.2)
            self.generators = [gen] * len(self.groups)
        self._check_generators()
        self._drop_device()

    def set_move_generator(self, generator):
        """Use one move generator for every group."""
        self.generators = [generator] * len(self.groups)
        self._check_generators()

    def _check_generators(self):
        """Each group's generator must accept the group: the reference's
        MoveGenerator.set_group raises Exception with the check's message
        (Core/MoveGenerator.py:56-61), e.g. for a rotation on a one-atom group
        (Generators/Rotations.py:133-143)."""
        for grp, gen in zip(self.groups, self.generators):
            if hasattr(gen, "check_group"):
                ok, msg = gen.check_group(grp)
                if not ok:
                    raise Exception(msg)

    def _distinct_generators(self):
        """The distinct move generators of the groups (one C-level pass when
        every group shares one)."""
        gens = self.generators
        if gens and gens.count(gens[0]) == len(gens):
            return [gens[0]]
        seen = {}
        for g in gens:
            seen.setdefault(id(g), g)
        return list(seen.values())

    def set_tolerance(self, tolerance):
        """Percentage [0,100] of worse moves allowed (Engine.py:1826-1842)."""
        t = FLOAT_TYPE(tolerance)
        if not 0 <= t <= 100:
            raise ValueError("tolerance must be in [0,100]")
        self.tolerance = FLOAT_TYPE(t / 100.)

    def add_constraints(self, constraints):
        """Experimental constraints enter the total standard error; rigid
        ones (distance constraints) only veto moves before the experimental
        constraints are evaluated (Engine.py:3040-3082, 3278-3296)."""
        for c in constraints if isinstance(constraints, (list, tuple)) else [constraints]:
            c._attach(self)
            (self.rigidConstraints if getattr(c, "rigid", False) else self.constraints).append(c)
        self._drop_device()

    # ----------------------------------------------------- standard error
    def compute_total_standard_error(self, constraints, current="standardError"):
        """np.sum of standard errors over variance (Engine.py:3000-3029)."""
        return np.sum([getattr(c, current) / c.varianceSquared for c in constraints])

    def initialize_used_constraints(self, force=False):
        for c in self.constraints + self.rigidConstraints:
            if c.data is None or force:
                c.compute_data()
        self.totalStandardError = self.compute_total_standard_error(self.constraints)

    # ------------------------------------------------------------- plain
    def run(self, numberOfSteps=100, log=None):
        """The reference step loop (Engine.py:3377-3516), one move per
        drop-in call sequence: group selection, move or removal, rigid
        constraints, experimental constraints, the accept rule."""
        self._pull_device()
        self.initialize_used_constraints()
        for _ in range(int(numberOfSteps)):
            self.generated += 1
            gi = self.selector.select_index(len(self.groups))
            gen = self.generators[gi]
            if hasattr(gen, "pick_from_list"):
                real = gen.pick_from_list(self)
                if len(real):
                    self._try_remove(real, self.relative_indexes(real))
                if log is not None:
                    log.append((int(gi), bool(len(real)) and self._last_accepted, float(self.totalStandardError)))
                continue
            real = self.groups[gi]
            if self.collectedAtoms:
                real = real[[not self.is_collected(int(i)) for i in real]]
            if hasattr(gen, "get_ready_for_move") and len(real):
                # swap moves act on the group plus the drawn partners (Engine.py:3192-3201)
                if len(real) == gen.swapLength:
                    real = gen.get_ready_for_move(real, self)
                else:
                    real = real[:0]
            if not len(real):
                # every atom of the group is gone: the move cannot be tried
                if log is not None:
                    log.append((int(gi), False, float(self.totalStandardError)))
                continue
            rel = self.relative_indexes(real)
            before = np.array(self.realCoordinates[rel], dtype=FLOAT_TYPE)
            moved_real = gen.move(before)
            moved_box = transform_coordinates(self.reciprocalBasisVectors, moved_real)
            self._try_move(real, rel, moved_real, moved_box)
            if log is not None:
                log.append((int(gi), self._last_accepted, float(self.totalStandardError)))
        self._dev_synced = False

    def _accept_rule(self, new):
        """Engine.py:3300-3310: a worse total is kept with probability
        `tolerance` (one Python random draw)."""
        if new > self.totalStandardError:
            if random.random() > self.tolerance:
                return True
            self.tolerated += 1
        self.totalStandardError = new
        return False

    def _try_move(self, real, rel, moved_real, moved_box):
        """Engine.py:3278-3344"""
        reject = False
        for c in self.rigidConstraints:
            c.compute_before_move(realIndexes=real, relativeIndexes=rel)
            c.compute_after_move(realIndexes=real, relativeIndexes=rel, movedBoxCoordinates=moved_box)
            reject = c.should_step_get_rejected(c.afterMoveStandardError)
            if reject:
                break
        tried = not reject
        if reject:
            for c in self.rigidConstraints:
                c.reject_move(realIndexes=real, relativeIndexes=rel)
        else:
            self.tried += 1
            for c in self.constraints:
                c.compute_before_move(realIndexes=real, relativeIndexes=rel)
                c.compute_after_move(realIndexes=real, relativeIndexes=rel, movedBoxCoordinates=moved_box)
            reject = self._accept_rule(self.compute_total_standard_error(self.constraints,
                                                                         current="afterMoveStandardError"))
        if reject:
            if tried:
                for c in self.constraints:
                    c.reject_move(realIndexes=real, relativeIndexes=rel)
        else:
            self.accepted += 1
            for c in self.constraints + self.rigidConstraints:
                c.accept_move(realIndexes=real, relativeIndexes=rel)
            self.realCoordinates[rel] = moved_real
            self.boxCoordinates[rel] = moved_box
        self._last_accepted = not reject

    def _try_remove(self, real, rel):
        """Engine.py:3231-3275: the constraints as if the atom were gone,
        the accept rule, then the atom leaves the configuration."""
        self.tried += 1
        self.removed[0] += 1
        for c in self.constraints:
            c.compute_as_if_amputated(realIndex=real, relativeIndex=rel)
        reject = self._accept_rule(self.compute_total_standard_error(self.constraints,
                                                                     current="amputationStandardError"))
        if reject:
            for c in self.constraints:
                c.reject_amputation(realIndex=real, relativeIndex=rel)
        else:
            self.accepted += 1
            self.removed[1] += 1
            for c in self.constraints + self.rigidConstraints:
                c.accept_amputation(realIndex=real, relativeIndex=rel)
            self._collect_atom(int(real[0]), int(rel[0]))
        self._last_accepted = not reject

    # ---------------------------------------------------- atoms collector
    def is_collected(self, realIndex):
        return realIndex in self._collected_set

    def relative_indexes(self, realIndexes):
        """Positions of atoms (original indexes) in the current arrays."""
        return np.searchsorted(self._realIndex, np.asarray(realIndexes)).astype(INT_TYPE)

    def _collect_atom(self, real, rel):
        """Engine._on_collector_collect_atom (Engine.py:758-794): the atom
        leaves every per-atom array; counts and number density follow."""
        el, name = self.allElements[rel], self.allNames[rel]
        if self.numberOfAtomsPerElement[el] - 1 <= 0:
            raise ValueError("removing the last '%s' atom is not allowed" % el)
        for c in self.constraints + self.rigidConstraints:
            c._on_collector_collect_atom(real, rel)
        self.realCoordinates = np.delete(self.realCoordinates, rel, axis=0)
        self.boxCoordinates = np.delete(self.boxCoordinates, rel, axis=0)
        self.moleculesIndex = np.delete(self.moleculesIndex, rel, axis=0)
        self.elementsIndex = np.delete(self.elementsIndex, rel, axis=0)
        self.namesIndex = np.delete(self.namesIndex, rel, axis=0)
        self.allElements.pop(rel)
        self.allNames.pop(rel)
        self.numberOfAtomsPerElement[el] -= 1
        self.numberOfAtomsPerName[name] -= 1
        self.numberOfAtoms -= 1
        self.numberDensity = FLOAT_TYPE(self.numberOfAtoms) / FLOAT_TYPE(self.volume)
        self._realIndex = np.delete(self._realIndex, rel)
        self._collected_set.add(real)
        self.collectedAtoms.append(real)
        self._drop_device()

    # ------------------------------------------------------------ device
    def _drop_device(self):
        if self._dev is not None:
            _lib.lib().frmc_engine_destroy(self._dev)
        self._dev = None
        self._dev_synced = False

    def __del__(self):
        try:
            self._drop_device()
        except Exception:
            pass

    def _device_engine(self, kmax):
        """Create (or reuse) the device engine holding the current state."""
        from . import _device
        _device.require_gpu()
        distinct = self._distinct_generators()
        swaps = any(hasattr(g, "get_ready_for_move") for g in distinct)
        for grp, gen in zip(self.groups, self.generators) if swaps else ():
            if hasattr(gen, "get_ready_for_move") and (len(grp) != gen.swapLength or not len(gen.swapList)):
                # Engine.run skips such a move without a draw (Engine.py:3192-3201), and
                # with an empty swap list the group would be swapped with itself
                raise NotImplementedError("a swap group whose size is not the swap length, or with an empty "
                                          "swap list, runs in Engine.run")
        if self._dev is not None and self._dev_kmax >= kmax and self._dev_synced:
            return self._dev
        self._pull_device()
        self._drop_device()
        if not self.constraints:
            raise ValueError("no constraints")
        if len(self.constraints) > 4:
            raise ValueError("at most 4 constraints run on the device")
        self.initialize_used_constraints()
        L = _lib.lib()
        keep = []

        def f32p(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(a)
            return a.ctypes.data_as(_lib.c_f32p)

        def i32p(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            keep.append(a)
            return a.ctypes.data_as(_lib.c_i32p)

        d = _lib.EngineDesc()
        d.n, d.nel, d.isPBC = self.numberOfAtoms, self.numberOfElements, 1
        d.box, d.real = f32p(self.boxCoordinates), f32p(self.realCoordinates)
        d.mol, d.elem = i32p(self.moleculesIndex), i32p(self.elementsIndex)
        for q in range(9):
            d.basis[q] = float(self.basisVectors.reshape(-1)[q])
            d.recip[q] = float(self.reciprocalBasisVectors.reshape(-1)[q])
        groups = self.groups
        if self.collectedAtoms:
            # the remaining atoms of each group, by position in the arrays; the
            # windows of such an engine name their atoms, so emptied groups
            # leave the table
            groups = [self.relative_indexes(g[[not self.is_collected(int(i)) for i in g]]) for g in groups]
            groups = [g for g in groups if len(g)] or [np.zeros(1, np.int32)]
        offs = np.zeros(len(groups) + 1, dtype=np.int32)
        offs[1:] = np.cumsum([len(g) for g in groups])
        d.ngroups = len(groups)
        d.group_offsets, d.group_atoms = i32p(offs), i32p(np.concatenate(groups))
        d.nconstraints = len(self.constraints)
        for ci, c in enumerate(self.constraints):
            a = c._device_arrays()
            cd = d.constraints[ci]
            cd.kind, cd.histSize, cd.nQ = a["kind"], a["H"], a["nQ"]
            cd.minDistance, cd.maxDistance, cd.bin = float(a["minD"]), float(a["maxD"]), float(a["bin"])
            cd.pairWeights, cd.pairDensity = f32p(a["w"]), f32p(a["D"])
            cd.shellVolumes, cd.prefactor = f32p(a["shellVolumes"]), f32p(a["prefactor"])
            cd.experimental, cd.dataWeights = f32p(a["experimental"]), f32p(a["dataWeights"])
            cd.sqMatrix, cd.scaleFactor = f32p(a["sqMatrix"]), float(a["scaleFactor"])
        d.kmax = int(kmax)
        d.stream = None
        h = ctypes.c_void_p()
        _lib.check(L.frmc_engine_create(ctypes.byref(h), ctypes.byref(d)))
        self._dev, self._dev_kmax = h, int(kmax)
        # device state from the constraints' current (slot-exact) histograms
        for ci, c in enumerate(self.constraints):
            slot = np.concatenate([c.data["intra"].reshape(-1), c.data["inter"].reshape(-1), [0]])
            if np.any(slot != np.rint(slot)):
                raise ValueError("constraint histograms are not integral")
            s = np.ascontiguousarray(slot.astype(np.int32))
            _lib.check(L.frmc_engine_set_counts(h, ci, s.ctypes.data_as(_lib.c_i32p)))
        chi2 = (ctypes.c_float * len(self.constraints))()
        tse = ctypes.c_float()
        _lib.check(L.frmc_engine_refresh(h, chi2, ctypes.byref(tse)))
        for ci, c in enumerate(self.constraints):
            if np.float32(chi2[ci]).tobytes() != np.float32(c.standardError).tobytes():
                raise _lib.FrmcError("device chi-square of constraint %d (%r) differs from the host chain (%r)"
                                     % (ci, chi2[ci], c.standardError))
        if np.float32(tse.value).tobytes() != np.float32(self.totalStandardError).tobytes():
            raise _lib.FrmcError("device total %r differs from the host total %r"
                                 % (tse.value, self.totalStandardError))
        self._dev_synced = True
        return h

    def _pull_device(self):
        """Copy the device state (coordinates, histograms, chi-squares) back
        into the host mirror after device runs."""
        if self._dev is None or not self._dev_synced or not getattr(self, "_dev_dirty", False):
            return
        L = _lib.lib()
        box = np.empty((self.numberOfAtoms, 3), np.float32)
        real = np.empty((self.numberOfAtoms, 3), np.float32)
        chi2 = (ctypes.c_float * len(self.constraints))()
        tse = ctypes.c_float()
        for ci, c in enumerate(self.constraints):
            nel, H = self.numberOfElements, int(c.histogramSize)
            slot = np.empty(2 * nel * nel * H + 1, np.int32)
            _lib.check(L.frmc_engine_get_state(self._dev, box.ctypes.data_as(_lib.c_f32p), ci,
                                               slot.ctypes.data_as(_lib.c_i32p), None, chi2, ctypes.byref(tse)))
            size = nel * nel * H
            c.data = {"intra": slot[:size].reshape(nel, nel, H).astype(np.float32),
                      "inter": slot[size:2 * size].reshape(nel, nel, H).astype(np.float32)}
            c.standardError = np.float32(chi2[ci])
        _lib.check(L.frmc_engine_get_real(self._dev, real.ctypes.data_as(_lib.c_f32p)))
        self.boxCoordinates = box
        self.realCoordinates = real
        self.totalStandardError = np.float32(tse.value)
        self._dev_dirty = False

    def run_device(self, numberOfSteps=100, k=64, log=None, native=True, pull=True):
        """Device-resident run with k speculative proposals per window.

        With translation moves on every group (the default generator) the
        whole loop, random streams included, runs in the native library;
        otherwise, or with rigid (distance) constraints or removed atoms,
        proposals come from the Python generators (the reference's own NumPy
        calls), the rigid constraints veto them in Python, removals are
        evaluated on the host at the end of their window, and the moves are
        evaluated on the device.
        ``pull=False`` leaves the state on the device (the host mirror is
        refreshed by the next :meth:`run`, :meth:`sync_from_device` or a
        pulling call).  Returns a dict of counters and timings."""
        h = self._device_engine(max(k, 1))
        distinct = self._distinct_generators()
        gen0 = distinct[0]
        # rigid constraints veto moves with draws of their own, so their runs
        # take the Python-drawn windows
        nat = (native and not self.rigidConstraints and not self.collectedAtoms and len(distinct) == 1
               and getattr(gen0, "native_kind", None) == "translation")
        t0 = time.perf_counter()
        if nat:
            stats = self._run_native(h, int(numberOfSteps), int(k), gen0, log)
        else:
            stats = self._run_python(h, int(numberOfSteps), int(k), log)
        stats["seconds"] = time.perf_counter() - t0
        self._dev_dirty = True
        if pull:
            self._pull_device()
        return stats

    def sync_from_device(self):
        """Refresh coordinates, histograms and chi-squares from the device."""
        self._pull_device()

    def device_handle(self):
        """The frmc_engine handle of the device state (None before run_device)."""
        return self._dev

    def _run_native(self, h, nsteps, k, gen, log):
        L = _lib.lib()
        pys = random.getstate()
        nps = np.random.get_state()
        py, npst = _mt_from_python(pys), _mt_from_numpy(nps)
        st = _lib.RunStats()
        glog = np.empty(nsteps, np.int32) if log is not None else None
        alog = np.empty(nsteps, np.uint8) if log is not None else None
        tlog = np.empty(nsteps, np.float32) if log is not None else None
        _lib.check(L.frmc_engine_run_translation(
            h, nsteps, k, float(gen.amplitude[0]), float(gen.amplitude[1]), float(self.tolerance),
            ctypes.byref(py), ctypes.byref(npst), ctypes.byref(st),
            glog.ctypes.data_as(_lib.c_i32p) if glog is not None else None,
            alog.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if alog is not None else None,
            tlog.ctypes.data_as(_lib.c_f32p) if tlog is not None else None))
        random.setstate(_mt_to_python(py, pys))
        np.random.set_state(_mt_to_numpy(npst))
        self.generated += st.generated
        self.tried += st.tried
        self.accepted += st.accepted
        self.tolerated += st.tolerated
        if log is not None:
            log.extend((int(g), bool(a), float(t)) for g, a, t in zip(glog, alog, tlog))
        return dict(generated=st.generated, accepted=st.accepted, windows=st.windows, evaluated=st.evaluated)

    def _run_python(self, h, nsteps, k, log):
        """Speculative windows with proposals from the Python generators.

        Each proposal reserves the acceptance draw a rejected worse move
        makes (Engine.py:3300-3310).  Rigid constraints veto a proposal
        before it is tried (Engine.py:3278-3296): the veto is decided while
        the window is drawn, in order and on the live streams, since it
        depends only on the configuration the window starts from; a vetoed
        proposal reserves no acceptance draw, and the draws a flexible veto
        makes are replayed with the proposal's own.  A removal
        (Engine.py:3231-3275) ends its window: it is evaluated on the host
        with the constraints' amputation calls once the moves before it are
        walked, and an accepted one rebuilds the device engine on the
        remaining atoms."""
        self._pull_device()
        rigid = self.rigidConstraints
        distinct = self._distinct_generators()
        removals = any(hasattr(g, "pick_from_list") for g in distinct)
        # swap moves act on the group plus drawn partners (Engine.py:3192-3201),
        # and with removals the device's group table no longer matches the
        # groups: windows then name each proposal's atoms
        lists = removals or bool(self.collectedAtoms) or any(hasattr(g, "get_ready_for_move") for g in distinct)

        def propose():
            """Engine.run's selection and move (Engine.py:3160-3210)."""
            gi = self.selector.select_index(len(self.groups))
            gen = self.generators[gi]
            if hasattr(gen, "pick_from_list"):
                real = gen.pick_from_list(self)
                if not len(real):
                    return _Proposal(gi, "none", None, None, None)
                return _Proposal(gi, "remove", real, self.relative_indexes(real), None)
            real = self.groups[gi]
            if self.collectedAtoms:
                real = real[[not self.is_collected(int(i)) for i in real]]
            if hasattr(gen, "get_ready_for_move") and len(real):
                real = gen.get_ready_for_move(real, self) if len(real) == gen.swapLength else real[:0]
            if not len(real):
                return _Proposal(gi, "none", None, None, None)
            rel = self.relative_indexes(real)
            before = np.array(self.realCoordinates[rel], dtype=FLOAT_TYPE)
            return _Proposal(gi, "move", real, rel, gen.move(before))

        def veto(real, rel, mb):
            """The rigid constraints in turn (a _Veto)."""
            draws = 0
            for ci, c in enumerate(rigid):
                if hasattr(c, "compute_move"):
                    c.compute_move(realIndexes=real, relativeIndexes=rel, movedBoxCoordinates=mb)
                else:
                    c.compute_before_move(realIndexes=real, relativeIndexes=rel)
                    c.compute_after_move(realIndexes=real, relativeIndexes=rel, movedBoxCoordinates=mb)
                draws += c.rejection_draws(c.afterMoveStandardError)
                if c.should_step_get_rejected(c.afterMoveStandardError):
                    for c2 in rigid:
                        c2.reject_move(realIndexes=real, relativeIndexes=rel)
                    return _Veto(True, ci + 1, draws, None)
            return _Veto(False, len(rigid), draws, [c.after_move_state() for c in rigid])

        resident = [c for c in rigid if hasattr(c, "bind_resident")]
        for c in resident:
            c.bind_resident()
        try:
            stats = self._python_windows(h, nsteps, k, log, rigid, resident, lists, propose, veto)
        finally:
            for c in resident:
                c.unbind_resident()
        # the native mirror of real coordinates is refreshed from this loop
        self._push_real()
        return stats

    def _push_real(self):
        if self._dev is not None:
            r = np.ascontiguousarray(self.realCoordinates, dtype=np.float32)
            _lib.check(_lib.lib().frmc_engine_set_real(self._dev, r.ctypes.data_as(_lib.c_f32p)))

    def _python_windows(self, h, nsteps, k, log, rigid, resident, lists, propose, veto):
        """The window loop of _run_python."""
        L = _lib.lib()
        step = 0
        windows = evaluated = accepted = 0
        gen0 = acc0 = 0
        while step < nsteps:
            if h is None:
                # after an accepted removal: the engine on the remaining atoms
                h = self._device_engine(k)
            # proposals drawn by the Python generators cost more than a window
            # on the device: about one expected acceptance per window
            p_acc = (acc0 + 1.0) / (gen0 + 2.0)
            kw = min(k, nsteps - step, max(1, int(np.ceil(1.0 / p_acc))))
            if rigid:
                # a veto in Python (four drop-in calls per constraint) costs
                # more than a device window, so a window ends at its first
                # proposal that passes them: no veto is evaluated for a
                # proposal past the window's acceptance
                kw = min(k, nsteps - step)
            props, boxes, u, outcome = [], [], [], []
            # the streams are saved once per window; an accepted proposal's
            # position is reached again by replaying the window's draws
            st0 = (random.getstate(), np.random.get_state())
            tried0 = [c.tried for c in rigid]
            for p in range(kw):
                pr = propose()
                v = _NO_VETO
                if pr.kind == "move":
                    mb = box_coordinates(self.reciprocalBasisVectors, pr.moved)
                    if rigid:
                        v = veto(pr.real, pr.rel, mb)
                else:
                    # no move for the device: atom 0 where it is
                    mb = np.array(self.boxCoordinates[:1], dtype=np.float32)
                props.append(pr)
                outcome.append(v)
                u.append(None if v.vetoed or pr.kind == "none" else random.random())
                boxes.append(mb)
                # a removal is evaluated on the host after the moves before it
                if pr.kind == "remove" or (rigid and pr.kind == "move" and not v.vetoed):
                    break
            kw = len(props)
            tse = np.empty(kw, np.float32)
            moves = [p for p in range(kw) if props[p].kind == "move" and not outcome[p].vetoed]
            if moves:
                nb = np.ascontiguousarray(np.concatenate(boxes), dtype=np.float32)
                if lists:
                    atoms = [pr.rel if pr.kind == "move" else np.zeros(1, np.int32) for pr in props]
                    counts = np.array([len(a) for a in atoms], dtype=np.int32)
                    flat = np.ascontiguousarray(np.concatenate(atoms), dtype=np.int32)
                    _lib.check(L.frmc_engine_eval_atoms(h, kw, counts.ctypes.data_as(_lib.c_i32p),
                                                        flat.ctypes.data_as(_lib.c_i32p),
                                                        nb.ctypes.data_as(_lib.c_f32p),
                                                        tse.ctypes.data_as(_lib.c_f32p), None))
                else:
                    gids = np.array([pr.gi for pr in props], dtype=np.int32)
                    _lib.check(L.frmc_engine_eval(h, kw, gids.ctypes.data_as(_lib.c_i32p),
                                                  nb.ctypes.data_as(_lib.c_f32p), tse.ctypes.data_as(_lib.c_f32p),
                                                  None))
            windows += 1
            evaluated += len(moves)
            cur = np.float32(self.totalStandardError)
            acc = last = -1
            amputated = None
            for p in range(kw):
                pr = props[p]
                self.generated += 1
                gen0 += 1
                step += 1
                last = p
                if pr.kind == "none" or outcome[p].vetoed:
                    # nothing to try, or never tried (Engine.py:3290-3296)
                    if log is not None:
                        log.append((int(pr.gi), False, float(cur)))
                    continue
                self.tried += 1
                if pr.kind == "remove":
                    # the constraints as if the atom were gone, from the state
                    # the window started from (Engine.py:3231-3246)
                    self.removed[0] += 1
                    self._push_real()
                    self._dev_dirty = True
                    self._pull_device()
                    for c in self.constraints:
                        c.compute_as_if_amputated(realIndex=pr.real, relativeIndex=pr.rel)
                    amputated = self.compute_total_standard_error(self.constraints, current="amputationStandardError")
                    val = amputated
                else:
                    val = tse[p]
                ok, drew = True, False
                if val > cur:
                    drew = True
                    if u[p] > self.tolerance:
                        ok = False
                    else:
                        self.tolerated += 1
                if pr.kind == "remove" and not ok:
                    for c in self.constraints:
                        c.reject_amputation(realIndex=pr.real, relativeIndex=pr.rel)
                if log is not None:
                    log.append((int(pr.gi), ok, float(val if ok else cur)))
                if ok:
                    acc = p
                    acc0 += 1
                    break
            # each rigid constraint counts the proposals it evaluated up to
            # the last one walked, as the per-move loop would
            for ci, c in enumerate(rigid):
                c.tried = tried0[ci] + sum(1 for p in range(last + 1) if outcome[p].evaluated > ci)
            if acc < 0:
                continue
            random.setstate(st0[0])
            np.random.set_state(st0[1])
            for p in range(acc + 1):
                propose()
                for _ in range(outcome[p].draws):
                    random.random()
                if (p < acc and u[p] is not None) or (p == acc and drew):
                    random.random()
            pr = props[acc]
            self.accepted += 1
            accepted += 1
            if pr.kind == "remove":
                # Engine.py:3260-3275: the atom leaves the configuration, and
                # the device engine is rebuilt on the remaining atoms
                self.removed[1] += 1
                self.totalStandardError = amputated
                for c in self.constraints + rigid:
                    c.accept_amputation(realIndex=pr.real, relativeIndex=pr.rel)
                for c in resident:
                    c.unbind_resident()
                self._collect_atom(int(pr.real[0]), int(pr.rel[0]))
                for c in resident:
                    c.bind_resident()
                h = None
                continue
            for c, state in zip(rigid, outcome[acc].states or ()):
                # the accepted move's after-move data, kept from its veto
                c.set_after_move_state(state)
                c.accept_move(realIndexes=pr.real, relativeIndexes=pr.rel)
            _lib.check(L.frmc_engine_commit(h, acc))
            self.totalStandardError = tse[acc]
            self.realCoordinates[pr.rel] = pr.moved
            self.boxCoordinates[pr.rel] = boxes[acc]
            for c in resident:
                c.resident_commit(pr.rel, boxes[acc])
        if h is None:
            self._device_engine(k)
        return dict(generated=step, accepted=accepted, windows=windows, evaluated=evaluated)


_Proposal = collections.namedtuple("_Proposal", "gi kind real rel moved")
# A proposal's outcome under the rigid constraints: vetoed, how many of them
# evaluated it, the Python random draws they made, and (when it passed) their
# after-move states for its acceptance.
_Veto = collections.namedtuple("_Veto", "vetoed evaluated draws states")
_NO_VETO = _Veto(False, 0, 0, None)


# ---------------------------------------------------------------- replicas

def _mt_words(m):
    """The 624 state words of an frmc_mt19937, as a uint32 view."""
    return np.ctypeslib.as_array(m.mt)


def _mt_from_python(state):
    """frmc_mt19937 from a random.getstate() tuple."""
    m = _lib.MT19937()
    _mt_words(m)[:] = np.array(state[1][:624], dtype=np.uint32)
    m.pos = state[1][624]
    return m


def _mt_from_numpy(state):
    """frmc_mt19937 from a numpy.random.get_state() tuple."""
    if state[3]:
        raise ValueError("numpy.random holds a cached Gaussian; the native generator cannot carry it")
    m = _lib.MT19937()
    _mt_words(m)[:] = np.asarray(state[1], dtype=np.uint32)
    m.pos = int(state[2])
    return m


def _mt_to_python(m, like):
    return (like[0], tuple(_mt_words(m).tolist()) + (int(m.pos),), like[2])


def _mt_to_numpy(m):
    return ("MT19937", _mt_words(m).copy(), int(m.pos), 0, 0.0)


def run_replicas(engines, numberOfSteps=100, k=32, seeds=None, batched=True):
    """Advance independent engines (replicas) by numberOfSteps generated
    moves each, with every replica's windows evaluated concurrently on the
    GPU.  Each replica runs exactly as ``random.seed(s); numpy.random.seed(s);
    engine.run_device(numberOfSteps, k)`` would run it alone, but on its own
    pair of random streams: ``seeds`` (one per engine) starts them, and
    without seeds each engine continues the streams of its previous replica
    run.  Every engine needs translation moves on all groups (the native
    loop) and the same generator amplitude and tolerance.  ``batched``: with
    single-atom groups, all replicas' windows of a step go out in one launch
    per kernel; False gives each engine its own stream.  Returns one stats
    dict per engine."""
    engines = list(engines)
    if not engines:
        raise ValueError("no engines")
    if seeds is not None and len(seeds) != len(engines):
        raise ValueError("one seed per engine is required")
    gen = engines[0].generators[0]
    amp = (float(gen.amplitude[0]), float(gen.amplitude[1]))
    tol = float(engines[0].tolerance)
    for E in engines:
        distinct = E._distinct_generators()
        g0 = distinct[0]
        if not (len(distinct) == 1 and getattr(g0, "native_kind", None) == "translation"):
            raise ValueError("replicas need translation moves on every group")
        if E.rigidConstraints:
            raise ValueError("replicas run the native loop, which has no rigid (distance) constraints")
        if E.collectedAtoms:
            raise ValueError("replicas run the native loop, whose group table holds every original atom")
        if (float(g0.amplitude[0]), float(g0.amplitude[1])) != amp or float(E.tolerance) != tol:
            raise ValueError("replicas need the same move amplitude and tolerance")
    R = len(engines)
    handles = (ctypes.c_void_p * R)()
    pys, nps = (_lib.MT19937 * R)(), (_lib.MT19937 * R)()
    stats = (_lib.RunStats * R)()
    for r, E in enumerate(engines):
        handles[r] = E._device_engine(max(k, 1)).value
        if seeds is not None:
            E._replica_rng = (random.Random(int(seeds[r])).getstate(),
                              np.random.RandomState(int(seeds[r])).get_state())
        elif getattr(E, "_replica_rng", None) is None:
            raise ValueError("engine %d has no replica random streams: pass seeds" % r)
        pys[r] = _mt_from_python(E._replica_rng[0])
        nps[r] = _mt_from_numpy(E._replica_rng[1])
    _lib.lib().frmc_set_replica_batching(1 if batched else 0)
    try:
        _lib.check(_lib.lib().frmc_engines_run_translation(handles, R, int(numberOfSteps), int(k), amp[0], amp[1],
                                                           tol, pys, nps, stats))
    finally:
        _lib.lib().frmc_set_replica_batching(1)
    out = []
    for r, E in enumerate(engines):
        E._replica_rng = (_mt_to_python(pys[r], E._replica_rng[0]), _mt_to_numpy(nps[r]))
        st = stats[r]
        E.generated += st.generated
        E.tried += st.tried
        E.accepted += st.accepted
        E.tolerated += st.tolerated
        E._dev_dirty = True
        out.append(dict(generated=st.generated, accepted=st.accepted, windows=st.windows, evaluated=st.evaluated))
    return out

