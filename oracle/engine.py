"""ORACLE / TEST INFRASTRUCTURE ONLY -- never imported by the product path.

Sequential restatement of fullrmc's Engine.run step for the configurations
of this path, on the CPU: group selection, move generation, constraint
before/after histograms through the C restatement (refcpu), the NumPy
constraint chain (chain.py), the accept rule, and accept/reject updates.

  step loop           Engine.py:3458-3516, 3166-3228, 3278-3344
  RandomSelector      Selectors/RandomSelectors.py:87-94
  Translation move    Generators/Translations.py:171-187, Core/Collection.py:674-701
  Rotation move       Generators/Rotations.py:141-178, Core/Collection.py:437-456
  Combinator          Core/MoveGenerator.py:792-811 (no shuffle)
  Swap                Generators/Swaps.py:115-133, Core/MoveGenerator.py:460-529
  Remove + amputation Generators/Removes.py:59-86, Engine.py:3231-3275, 758-794,
                      PairDistributionConstraints.py:1168-1233,
                      StructureFactorConstraints.py:1098-1163,
                      DistanceConstraints.py:774-819
  PDC / SFC update    PairDistributionConstraints.py:1044-1166,
                      StructureFactorConstraints.py:975-1096
  PCF / reduced SFC   PairCorrelationConstraints.py:305-392 (same update,
                      g(r) model), StructureFactorConstraints.py:1237-1261
  rigid constraints   Engine.py:3278-3296; InterMolecularDistanceConstraint
                      DistanceConstraints.py:361-428, 552-772 (element types)
"""
import random

import numpy as np

from . import chain
from . import refcpu

F32 = np.float32
PRECISION = F32(1e-5)


def random_vector(minAmp, maxAmp):
    vector = np.array(1 - 2 * np.random.random(3), dtype=F32)
    norm = np.linalg.norm(vector)
    while norm == 0:
        vector = np.array(1 - 2 * np.random.random(3), dtype=F32)
        norm = np.linalg.norm(vector)
    vector /= F32(norm)
    base = F32(vector * minAmp)
    maxAmp = F32(maxAmp - minAmp)
    vector *= F32(random.random() * maxAmp)
    vector += base
    return vector


class Translation:
    def __init__(self, amplitude):
        self.amp = (F32(0), F32(float(amplitude)))

    def move(self, coords):
        return coords + random_vector(self.amp[0], self.amp[1])


class Rotation:
    def __init__(self, amplitude_deg):
        self.amp = F32(chain.PI * float(amplitude_deg) / 180.)

    def move(self, coords):
        n = 0
        while n < PRECISION:
            axis = 1 - 2 * np.random.random(3)
            n = np.linalg.norm(axis)
        axis /= n
        angle = (1 - 2 * random.random()) * self.amp
        angle = float(angle)
        ax = axis / np.sqrt(np.dot(axis, axis))
        a = np.cos(angle / 2)
        b, c, d = -ax * np.sin(angle / 2.)
        R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                      [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                      [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - b * b - c * c]], dtype=F32)
        center = np.sum(coords, 0) / coords.shape[0]
        rot = coords - center
        for i in range(rot.shape[0]):
            rot[i, :] = np.dot(R, rot[i, :])
        return np.array(rot + center, dtype=F32)


class Combined:
    def __init__(self, generators):
        self.generators = generators

    def move(self, coords):
        for g in self.generators:
            coords = g.move(coords)
        return coords


class Swap:
    """Position exchange with a random sub-list of swapList."""

    def __init__(self, swapList, swapLength=1):
        self.L = swapLength
        self.swapList = [np.array(x, dtype=np.int32) for x in swapList]

    def ready(self, idx, sysm):
        rem = [x for x in self.swapList if not any(int(i) in sysm.collected for i in x)]
        if len(rem):
            return np.concatenate((idx, rem[random.randint(0, len(rem) - 1)]))
        return np.concatenate((idx, idx))

    def move(self, coords):
        out = np.array(coords, dtype=F32)
        out[:self.L] = coords[self.L:]
        out[self.L:] = coords[:self.L]
        return out


class Remove:
    """Atom removal: from atomsList by np.random.choice, else any original
    index by random.randint; nothing once maxCollected atoms are gone or
    when the drawn atom already is."""

    def __init__(self, atomsList=None, maxCollected=None):
        self.atomsList = None if atomsList is None else np.array(atomsList, dtype=np.int32)
        self.maxCollected = maxCollected

    def pick(self, sysm):
        if self.atomsList is None:
            idx = np.array([random.randint(0, sysm.n0 - 1)], dtype=np.int32)
        else:
            idx = np.random.choice(self.atomsList, 1)
        if self.maxCollected is not None and len(sysm.collected) >= self.maxCollected:
            return idx[:0]
        if len(idx) and int(idx[0]) in sysm.collected:
            return idx[:0]
        return idx


class System:
    """Box, species and molecules: the engine attributes the constraints read."""

    def __init__(self, real, elements_of_atoms, molecules, basis, recip):
        self.real = np.array(real, dtype=F32)
        self.basis = np.array(basis, dtype=F32)
        self.recip = np.array(recip, dtype=F32)
        self.box = refcpu.transform_coordinates(self.recip, self.real)
        self.allElements = list(elements_of_atoms)
        self.elements = sorted(set(self.allElements))
        lut = {e: i for i, e in enumerate(self.elements)}
        self.elem = np.array([lut[e] for e in self.allElements], dtype=np.int32)
        self.mol = np.array(molecules, dtype=np.int32)
        self.counts = {}
        for e in self.allElements:
            self.counts[e] = self.counts.get(e, 0) + 1
        self.volume = F32(abs(np.linalg.det(self.basis.astype(np.float64))))
        self.rho0 = F32(len(self.real)) / F32(self.volume)
        self.nel = len(self.elements)
        self.n0 = len(self.real)
        self.collected = set()
        # threads of the per-move histogram loops (the reference's ncores;
        # private bins per thread, so counts do not depend on it)
        self.ncores = 1
        self.real_index = np.arange(self.n0, dtype=np.int32)

    def rel(self, real):
        return np.searchsorted(self.real_index, np.asarray(real)).astype(np.int32)

    def collect(self, real, rel):
        """Engine.py:758-794"""
        el = self.allElements[rel]
        self.real = np.delete(self.real, rel, axis=0)
        self.box = np.delete(self.box, rel, axis=0)
        self.mol = np.delete(self.mol, rel, axis=0)
        self.elem = np.delete(self.elem, rel, axis=0)
        self.allElements.pop(rel)
        self.counts[el] -= 1
        self.rho0 = F32(len(self.real)) / F32(self.volume)
        self.real_index = np.delete(self.real_index, rel)
        self.collected.add(int(real))


class PDC:
    """PairDistributionConstraint on the oracle path."""

    def __init__(self, sysm, experimentalData, weights, scaleFactor=1.0, dataWeights=None, elementWeights=None):
        self.s = sysm
        self.exp = np.asarray(experimentalData, dtype=F32)
        self.bin, self.min, self.max, self.centers, self.H, self.vols = chain.pdc_grid(self.exp)
        self.y = self.exp[:, 1]
        self.w = weights
        self.sf = scaleFactor
        self.dw = dataWeights
        self.elw = elementWeights  # per-element weights: needed to re-weight after a removal

    def _args(self, counts, w, rho0):
        s = self.s
        return (s.counts if counts is None else counts, self.w if w is None else w,
                s.rho0 if rho0 is None else rho0)

    def model(self, intra, inter, counts=None, w=None, rho0=None):
        s = self.s
        counts, w, rho0 = self._args(counts, w, rho0)
        return chain.total_gr(intra, inter, s.elements, counts, w, s.volume, rho0, self.vols, self.centers, self.sf)

    def hist_full(self, box, mol, elem):
        if len(box) > 10000:
            # large boxes: the cell-list form on every core (equal to the
            # all-pairs loop, pinned in tests/test_oracle_pins.py)
            return refcpu.full_pairs_histograms_coords_fast(box, self.s.basis, True, mol, elem, self.s.nel,
                                                            self.min, self.max, self.bin, self.H)
        return refcpu.full_pairs_histograms_coords(box, self.s.basis, True, mol, elem, self.s.nel, self.min,
                                                   self.max, self.bin, self.H)

    def hist_multi(self, idx, box):
        return refcpu.multiple_pairs_histograms_coords(idx, box, self.s.basis, True, self.s.mol, self.s.elem,
                                                       self.s.nel, self.min, self.max, self.bin, self.H,
                                                       allAtoms=True, ncores=self.s.ncores)

    def chi2(self, intra, inter, counts=None, w=None, rho0=None):
        return chain.chi_square(self.y, self.model(intra, inter, counts, w, rho0), self.dw)

    def amputate(self, idx):
        """compute_as_if_amputated: the atom's pairs out, counts, weights and
        number density of the system without it."""
        s = self.s
        self.before(idx)
        self.amp_i, self.amp_e = self.intra - self.bi, self.inter - self.be
        counts = dict(s.counts)
        counts[s.allElements[int(idx[0])]] -= 1
        self.amp_w = chain.normalized_weighting(counts, self.elw)
        rho0 = F32((len(s.real) - 1) / s.volume)
        self.amp_sd = self.chi2(self.amp_i, self.amp_e, counts, self.amp_w, rho0)

    def accept_amputation(self):
        self.intra, self.inter, self.w, self.sd = self.amp_i, self.amp_e, self.amp_w, self.amp_sd

    # reference state machine
    def compute_data(self):
        self.intra, self.inter = self.hist_full(self.s.box, self.s.mol, self.s.elem)
        self.sd = self.chi2(self.intra, self.inter)

    def _group(self, idx, box):
        iM, eM = self.hist_multi(idx, box)
        iF, eF = refcpu.full_pairs_histograms_coords(box[idx], self.s.basis, True, self.s.mol[idx],
                                                     self.s.elem[idx], self.s.nel, self.min, self.max, self.bin,
                                                     self.H)
        return iM - iF, eM - eF

    def before(self, idx):
        self.bi, self.be = self._group(idx, self.s.box)

    def after(self, idx, moved_box):
        saved = np.array(self.s.box[idx], dtype=F32)
        self.s.box[idx] = moved_box
        self.ai, self.ae = self._group(idx, self.s.box)
        self.s.box[idx] = saved
        self.new_i = self.intra - self.bi + self.ai
        self.new_e = self.inter - self.be + self.ae
        self.after_sd = self.chi2(self.new_i, self.new_e)

    def accept(self):
        self.intra, self.inter, self.sd = self.new_i, self.new_e, self.after_sd


class SFC(PDC):
    """StructureFactorConstraint on the oracle path."""

    def __init__(self, sysm, experimentalData, weights, rmin, rmax, dr, scaleFactor=1.0, dataWeights=None,
                 elementWeights=None):
        self.s = sysm
        self.elw = elementWeights
        self.exp = np.asarray(experimentalData, dtype=F32)
        self.q = self.exp[:, 0].astype(F32)
        self.y = self.exp[:, 1].astype(F32)
        self.min, self.max, self.bin, self.centers, self.H, self.vols = chain.sfc_grid(rmin, rmax, dr)
        self.M = chain.sq_matrix(self.centers, self.q)
        self.w = weights
        self.sf = scaleFactor
        self.dw = dataWeights

    def model(self, intra, inter, counts=None, w=None, rho0=None):
        s = self.s
        counts, w, rho0 = self._args(counts, w, rho0)
        return chain.total_sq(intra, inter, s.elements, counts, w, s.volume, rho0, self.vols, self.centers, self.M,
                              self.sf)


class PCF(PDC):
    """PairCorrelationConstraint on the oracle path: PDC grid, g(r) model."""

    def model(self, intra, inter, counts=None, w=None, rho0=None):
        s = self.s
        counts, w, rho0 = self._args(counts, w, rho0)
        return chain.total_gofr(intra, inter, s.elements, counts, w, s.volume, rho0, self.vols, self.centers,
                                self.sf)


class RSFC(SFC):
    """ReducedStructureFactorConstraint on the oracle path."""

    def model(self, intra, inter, counts=None, w=None, rho0=None):
        s = self.s
        counts, w, rho0 = self._args(counts, w, rho0)
        return chain.total_sq(intra, inter, s.elements, counts, w, s.volume, rho0, self.vols, self.centers, self.M,
                              self.sf, reduced=True)


class IMD:
    """InterMolecularDistanceConstraint on the oracle path, element types,
    every pair limit = defaultDistance unless given in limits[(a, b)]."""

    def __init__(self, sysm, defaultDistance=1.5, limits=None, flexible=True, rejectProbability=1.0,
                 inter=True):
        self.s = sysm
        nel = sysm.nel
        self.up = np.full((nel, nel, 1), F32(defaultDistance), F32)
        for (a, b), d in (limits or {}).items():
            i, j = sysm.elements.index(a), sysm.elements.index(b)
            self.up[i, j, 0] = self.up[j, i, 0] = F32(d)
        self.lo = np.zeros((nel, nel, 1), F32)
        self.flags = refcpu.ad_flags(inter, not inter, True, True)
        self.inter = inter
        self.flexible = flexible
        self.rp = F32(rejectProbability)
        self.pairs = [(i, j) for i in range(nel) for j in range(i, nel)]

    def _pick(self, r):
        ni, di, ne, de = r
        return (ne, de) if self.inter else (ni, di)

    def _full(self, box, mol, elem):
        return self._pick(refcpu.full_atomic_distances_coords(box, self.s.basis, True, mol, elem, self.s.nel,
                                                              self.lo, self.up, self.flags))

    def value(self, number, dsum):
        idi = np.array([p[0] for p in self.pairs])
        idj = np.array([p[1] for p in self.pairs])
        numbers = (number[idi, idj] + number[idj, idi]).reshape(-1)
        distances = (dsum[idi, idj] + dsum[idj, idi]).reshape(-1)
        nz = np.where(numbers)
        distances[nz] /= numbers[nz]
        return F32(np.sum(distances))

    def compute_data(self):
        self.number, self.dsum = self._full(self.s.box, self.s.mol, self.s.elem)
        self.sd = self.value(self.number, self.dsum)

    def _group(self, idx, box):
        nM, dM = self._pick(refcpu.multiple_atomic_distances_coords(idx, box, self.s.basis, True, self.s.mol,
                                                                    self.s.elem, self.s.nel, self.lo, self.up,
                                                                    self.flags, allAtoms=True))
        nF, dF = self._full(box[idx], self.s.mol[idx], self.s.elem[idx])
        return nM - nF, dM - dF

    def before(self, idx):
        self.bn, self.bd = self._group(idx, self.s.box)

    def after(self, idx, moved_box):
        saved = np.array(self.s.box[idx], dtype=F32)
        self.s.box[idx] = moved_box
        self.an, self.ad = self._group(idx, self.s.box)
        self.s.box[idx] = saved
        self.new_n = self.number - self.bn + self.an
        self.new_d = self.dsum - self.bd + self.ad
        self.after_sd = self.value(self.new_n, self.new_d)

    def rejects(self):
        if self.flexible:
            if self.after_sd <= self.sd:
                return False
            return random.random() < self.rp
        return bool(np.any(self.an > self.bn))

    def accept(self):
        self.number, self.dsum, self.sd = self.new_n, self.new_d, self.after_sd

    def accept_amputation(self, idx):
        self.before(idx)
        self.number, self.dsum = self.number - self.bn, self.dsum - self.bd
        self.sd = self.value(self.number, self.dsum)


class Engine:
    def remove(self, real, idx):
        """Engine.py:3231-3275"""
        self.tried += 1
        for c in self.constraints:
            c.amputate(idx)
        new = chain.total_error([c.amp_sd for c in self.constraints])
        reject = False
        if new > self.tse:
            if random.random() > self.tolerance:
                reject = True
            else:
                self.tolerated += 1
                self.tse = new
        else:
            self.tse = new
        if not reject:
            self.accepted += 1
            for c in self.constraints:
                c.accept_amputation()
            for c in self.rigid:
                c.accept_amputation(idx)
            self.s.collect(int(real[0]), int(idx[0]))
        return not reject

    def __init__(self, sysm, groups, generator, constraints, tolerance=0.0, rigid=()):
        """generator: one for every group, or a list with one per group."""
        self.s = sysm
        self.groups = [np.array(g, dtype=np.int32) for g in groups]
        self.generators = list(generator) if isinstance(generator, (list, tuple)) else [generator] * len(groups)
        self.constraints = constraints
        self.rigid = list(rigid)
        self.tolerance = F32(F32(tolerance) / 100.)
        self.generated = self.tried = self.accepted = self.tolerated = 0

    def initialize(self):
        for c in self.constraints + self.rigid:
            c.compute_data()
        self.tse = chain.total_error([c.sd for c in self.constraints])

    def run(self, nsteps, log=None):
        s = self.s
        for _ in range(nsteps):
            self.generated += 1
            gi = np.int32(random.randint(0, len(self.groups) - 1))
            gen = self.generators[gi]
            if hasattr(gen, "pick"):
                real = gen.pick(s)
                ok = False
                if len(real):
                    ok = self.remove(real, s.rel(real))
                if log is not None:
                    log.append((int(gi), ok, float(self.tse)))
                continue
            real = self.groups[gi]
            if s.collected:
                real = real[[int(i) not in s.collected for i in real]]
            if hasattr(gen, "ready") and len(real):
                real = gen.ready(real, s) if len(real) == gen.L else real[:0]
            if not len(real):
                if log is not None:
                    log.append((int(gi), False, float(self.tse)))
                continue
            idx = s.rel(real)
            before = np.array(s.real[idx], dtype=F32)
            moved_real = gen.move(before)
            moved_box = refcpu.transform_coordinates(s.recip, moved_real)
            reject = False
            for c in self.rigid:
                c.before(idx)
                c.after(idx, moved_box)
                reject = c.rejects()
                if reject:
                    break
            if not reject:
                self.tried += 1
                for c in self.constraints:
                    c.before(idx)
                    c.after(idx, moved_box)
                new = chain.total_error([c.after_sd for c in self.constraints])
                if new > self.tse:
                    if random.random() > self.tolerance:
                        reject = True
                    else:
                        self.tolerated += 1
                        self.tse = new
                else:
                    self.tse = new
            if not reject:
                self.accepted += 1
                for c in self.constraints + self.rigid:
                    c.accept()
                s.real[idx] = moved_real
                s.box[idx] = moved_box
            if log is not None:
                log.append((int(gi), not reject, float(self.tse)))
