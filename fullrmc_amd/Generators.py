"""Move proposals with the reference's random-number semantics.

Both of fullrmc's global MT19937 streams are used exactly as the reference
uses them: Python's ``random`` (group selection, translation amplitude,
rotation angle, acceptance) and NumPy's legacy ``numpy.random`` (directions).

  RandomSelector        Selectors/RandomSelectors.py:87-94
  TranslationGenerator  Generators/Translations.py:112-187 with
                        generate_random_vector, Core/Collection.py:674-701
  RotationGenerator     Generators/Rotations.py:113-178 with
                        get_rotation_matrix, Core/Collection.py:437-456
  MoveGeneratorCombinator  Core/MoveGenerator.py:792-811 (shuffle=False)
  SwapPositionsGenerator   Generators/Swaps.py:22-133 with SwapGenerator,
                        Core/MoveGenerator.py:321-530
  AtomsRemoveGenerator  Generators/Removes.py:22-86 with RemoveGenerator,
                        Core/MoveGenerator.py:107-301
"""
import random

import numpy as np

from .Globals import FLOAT_TYPE, INT_TYPE, PI, PRECISION


class RandomSelector:
    """Uniform group index: ``random.randint(0, ngroups-1)``."""

    def select_index(self, ngroups):
        return INT_TYPE(random.randint(0, ngroups - 1))


def generate_random_vector(minAmp, maxAmp):
    """Random direction (NumPy stream) scaled by an amplitude drawn from the
    Python stream in [minAmp, maxAmp); float32 throughout."""
    vector = np.array(1 - 2 * np.random.random(3), dtype=FLOAT_TYPE)
    norm = np.linalg.norm(vector)
    while norm == 0:
        vector = np.array(1 - 2 * np.random.random(3), dtype=FLOAT_TYPE)
        norm = np.linalg.norm(vector)
    vector /= FLOAT_TYPE(norm)
    base = FLOAT_TYPE(vector * minAmp)
    span = FLOAT_TYPE(maxAmp - minAmp)
    vector *= FLOAT_TYPE(random.random() * span)
    vector += base
    return vector


class TranslationGenerator:
    """Translate the whole group by one random vector of length below
    ``amplitude`` (or within ``(min, max)``)."""

    native_kind = "translation"

    def __init__(self, amplitude=0.2):
        self.set_amplitude(amplitude)

    def set_amplitude(self, amplitude):
        if isinstance(amplitude, (list, tuple)):
            mn, mx = FLOAT_TYPE(amplitude[0]), FLOAT_TYPE(amplitude[1])
            if not (mn >= 0 and mx > mn):
                raise ValueError("translation amplitude must satisfy 0 <= min < max")
            self.amplitude = (mn, mx)
        else:
            a = float(amplitude)
            if not a > 0:
                raise ValueError("translation amplitude must be bigger than 0")
            self.amplitude = (FLOAT_TYPE(0), FLOAT_TYPE(a))

    def move(self, coordinates):
        return coordinates + generate_random_vector(self.amplitude[0], self.amplitude[1])


def get_rotation_matrix(rotationVector, angle):
    angle = float(angle)
    axis = rotationVector / np.sqrt(np.dot(rotationVector, rotationVector))
    a = np.cos(angle / 2)
    b, c, d = -axis * np.sin(angle / 2.)
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - b * b - c * c]], dtype=FLOAT_TYPE)


class RotationGenerator:
    """Rotate the group about its centre by a random axis and an angle below
    ``amplitude`` degrees."""

    native_kind = None

    def __init__(self, amplitude=2):
        a = float(amplitude)
        if not 0 < a < 360:
            raise ValueError("rotation amplitude must be in (0, 360) degrees")
        self.amplitude = FLOAT_TYPE(PI * a / 180.)

    def check_group(self, group):
        """Generators/Rotations.py:133-143: a rotation needs two atoms."""
        if len(group) <= 1:
            return False, "At least two atoms needed in a group to perform rotation."
        return True, ""

    def move(self, coordinates):
        if coordinates.shape[0] <= 1:
            # The reference's fallback for a group reduced to one atom
            # (Generators/Rotations.py:155-157) passes self.__amplitude[0] to
            # generate_random_vector, but the amplitude is a float32 scalar:
            # the indexing raises IndexError before any random draw.  The
            # same error, with no draw, here.
            raise IndexError("invalid index to scalar variable: the reference's one-atom rotation fallback "
                             "indexes its scalar amplitude (a rotation needs at least two atoms)")
        n = 0
        while n < PRECISION:
            axis = 1 - 2 * np.random.random(3)
            n = np.linalg.norm(axis)
        axis /= n
        angle = (1 - 2 * random.random()) * self.amplitude
        R = get_rotation_matrix(axis, angle)
        center = np.sum(coordinates, 0) / coordinates.shape[0]
        rotated = coordinates - center
        for i in range(rotated.shape[0]):
            rotated[i, :] = np.dot(R, rotated[i, :])
        return np.array(rotated + center, dtype=FLOAT_TYPE)


class MoveGeneratorCombinator:
    """Apply several generators in list order (no shuffling)."""

    native_kind = None

    def __init__(self, combination, shuffle=False):
        if shuffle:
            raise NotImplementedError("shuffled combinations consume random.shuffle draws; not supported")
        self.combination = list(combination)

    def check_group(self, group):
        """Core/MoveGenerator.py:741-752: a combination accepts every group."""
        return True, ""

    def move(self, coordinates):
        for g in self.combination:
            coordinates = g.move(coordinates)
        return coordinates


class SwapPositionsGenerator:
    """Exchange the positions of the group's atoms with those of a sub-list
    drawn from ``swapList`` (``random.randint`` over the sub-lists).  The
    engine extends the moved group to the group plus the drawn sub-list
    (get_ready_for_move, Core/MoveGenerator.py:502-529); the move puts the
    partner's positions on the group's atoms and the group's on the partner's
    (Generators/Swaps.py:115-133)."""

    native_kind = None

    def __init__(self, group=None, swapLength=1, swapList=None):
        n = int(swapLength)
        if n < 1:
            raise ValueError("swapLength must be a positive integer")
        self.swapLength = n
        self.set_swap_list(swapList)

    def set_swap_list(self, swapList):
        """Core/MoveGenerator.py:390-436: sub-lists of swapLength distinct
        non-negative atom indexes."""
        if swapList is None:
            self.swapList = ()
            return
        if not isinstance(swapList, (list, tuple)):
            raise ValueError("swapList must be a list")
        out = []
        for sl in swapList:
            if not isinstance(sl, (list, tuple)):
                raise ValueError("swapList items must be a list")
            sub = [INT_TYPE(x) for x in sl]
            if any(x < 0 for x in sub) or len(set(int(x) for x in sub)) != len(sub) or len(sub) != self.swapLength:
                raise ValueError("swapList items must hold swapLength distinct non-negative indexes")
            out.append(np.array(sub, dtype=INT_TYPE))
        self.swapList = tuple(out)

    def get_ready_for_move(self, groupAtomsIndexes, engine=None):
        """The atoms the swap moves: the group, then a random sub-list of
        those with no removed atom (Core/MoveGenerator.py:460-529; the group
        itself when none is left)."""
        remaining = self.swapList
        if engine is not None and engine.collectedAtoms:
            remaining = [sl for sl in self.swapList if not any(engine.is_collected(int(i)) for i in sl)]
        if len(remaining):
            swap = remaining[random.randint(0, len(remaining) - 1)]
        else:
            swap = groupAtomsIndexes
        return np.concatenate((groupAtomsIndexes, swap))

    def move(self, coordinates):
        L = self.swapLength
        out = np.array(coordinates, dtype=FLOAT_TYPE)
        out[:L, :] = coordinates[L:, :]
        out[L:, :] = coordinates[:L, :]
        return out


class AtomsRemoveGenerator:
    """Remove one atom from the configuration: an atom drawn from
    ``atomsList`` (``np.random.choice``) or, without a list, any atom index
    of the original configuration (``random.randint``); nothing is removed
    once ``maximumCollected`` atoms are gone or when the drawn atom already
    is (pick_from_list, Generators/Removes.py:59-86)."""

    native_kind = None

    def __init__(self, group=None, maximumCollected=None, allowFittingScaleFactor=False, atomsList=None):
        if maximumCollected is not None:
            maximumCollected = INT_TYPE(maximumCollected)
            if not maximumCollected > 0:
                raise ValueError("maximumCollected must be bigger than 0")
        self.maximumCollected = maximumCollected
        self.allowFittingScaleFactor = bool(allowFittingScaleFactor)
        if atomsList is not None:
            # set_atoms_list (Core/MoveGenerator.py:236-270): order kept, no repeats
            atomsList = np.array([INT_TYPE(i) for i in atomsList], dtype=INT_TYPE)
            if np.any(atomsList < 0) or len(set(atomsList.tolist())) != len(atomsList):
                raise ValueError("atomsList items must be distinct indexes >= 0")
        self.atomsList = atomsList

    def pick_from_list(self, engine):
        if self.atomsList is None:
            index = np.array([random.randint(0, engine.originalNumberOfAtoms - 1)], dtype=INT_TYPE)
        else:
            index = np.random.choice(self.atomsList, 1)
        if self.maximumCollected is not None and len(engine.collectedAtoms) >= self.maximumCollected:
            index = np.array([], dtype=INT_TYPE)
        if len(index) and engine.is_collected(int(index[0])):
            index = np.array([], dtype=INT_TYPE)
        return index

    def move(self, coordinates):
        raise TypeError("a remove generator removes atoms; it does not move them")
