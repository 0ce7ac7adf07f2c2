"""Host-side pieces of the Engine mirror that need no GPU: the MT19937
state conversions between Python's random / NumPy's RandomState and the C
ABI's frmc_mt19937, and the shared-generator check of the device paths."""
import random

import numpy as np

from fullrmc_amd import Engine as eng
from fullrmc_amd.Generators import RotationGenerator, TranslationGenerator


def test_mt_state_round_trips_are_exact():
    for seed in (0, 1, 12345, 2 ** 31 - 1):
        r = random.Random(seed)
        for _ in range(seed % 7):
            r.random()
        st = r.getstate()
        m = eng._mt_from_python(st)
        assert eng._mt_to_python(m, st) == st
        rs = np.random.RandomState(seed)
        rs.random_sample(seed % 5 + 1)
        ns = rs.get_state()
        m2 = eng._mt_from_numpy(ns)
        back = eng._mt_to_numpy(m2)
        assert back[0] == ns[0] and back[2] == ns[2] and back[1].dtype == np.uint32
        assert np.array_equal(back[1], ns[1])
        # the restored stream continues exactly where the original would
        rs2 = np.random.RandomState()
        rs2.set_state(back)
        assert np.array_equal(rs2.random_sample(50), rs.random_sample(50))
        r2 = random.Random()
        r2.setstate(eng._mt_to_python(m, st))
        assert [r2.random() for _ in range(50)] == [r.random() for _ in range(50)]


def test_mt_state_rejects_a_cached_gaussian():
    rs = np.random.RandomState(3)
    rs.standard_normal()
    try:
        eng._mt_from_numpy(rs.get_state())
    except ValueError:
        return
    raise AssertionError("a cached Gaussian must be refused")


class _Stub:
    _distinct_generators = eng.Engine._distinct_generators

    def __init__(self, gens):
        self.generators = gens


def test_distinct_generators():
    t, r = TranslationGenerator(0.1), RotationGenerator(10)
    assert _Stub([t] * 1000)._distinct_generators() == [t]
    d = _Stub([t] * 500 + [r] + [t] * 10)._distinct_generators()
    assert len(d) == 2 and d[0] is t and d[1] is r
    # equal-looking but distinct generators are distinct
    t2 = TranslationGenerator(0.1)
    assert len(_Stub([t, t2])._distinct_generators()) == 2


def test_rotation_of_one_atom_raises_before_any_draw():
    """A group reduced to one atom: the reference's fallback indexes its
    float32 amplitude and raises IndexError before drawing
    (Generators/Rotations.py:155-157); assigning a rotation to a one-atom
    group is refused with the reference's message."""
    r = RotationGenerator(10)
    random.seed(1)
    np.random.seed(1)
    py, npst = random.getstate(), np.random.get_state()
    try:
        r.move(np.zeros((1, 3), np.float32))
    except IndexError:
        pass
    else:
        raise AssertionError("IndexError expected")
    assert random.getstate() == py and np.array_equal(np.random.get_state()[1], npst[1])
    E = eng.Engine()
    E.set_system(np.zeros((3, 3), np.float32) + np.arange(3, dtype=np.float32)[:, None], ["Ar"] * 3,
                 basisVectors=np.eye(3, dtype=np.float32) * 10)
    try:
        E.set_groups([[0], [1, 2]], RotationGenerator(10))
    except Exception as e:
        assert "At least two atoms" in str(e)
    else:
        raise AssertionError("a rotation on a one-atom group must be refused")
    E.set_groups([[0, 1], [1, 2]], RotationGenerator(10))


def test_replicas_refuse_rigid_constraints():
    """Replicas run the native translation loop, which has no veto: an engine
    with a distance constraint is refused before anything runs."""
    from fullrmc_amd import systems
    from fullrmc_amd.Constraints import InterMolecularDistanceConstraint

    s = systems.random_box(20, 10.0, 1.5, seed=3)
    E = eng.Engine()
    E.set_system(s["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    E.add_constraints([InterMolecularDistanceConstraint(1.5)])
    try:
        eng.run_replicas([E], 10)
    except ValueError as e:
        assert "rigid" in str(e)
    else:
        raise AssertionError("run_replicas accepted an engine with a rigid constraint")


def test_replicas_refuse_removed_atoms():
    """An engine with atoms already removed keeps its groups as original
    indexes, which the native loop's group table cannot hold: refused before
    anything runs."""
    from fullrmc_amd import systems
    from fullrmc_amd.Constraints import PairDistributionConstraint

    s = systems.random_box(20, 10.0, 1.5, seed=4)
    E = eng.Engine()
    E.set_system(s["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    x = (0.05 + 0.1 * np.arange(40)).astype(np.float32)
    E.add_constraints([PairDistributionConstraint(np.stack([x, np.zeros_like(x)], 1))])
    E.collectedAtoms.append(3)
    try:
        eng.run_replicas([E], 10)
    except ValueError as e:
        assert "removed" in str(e) or "original" in str(e)
    else:
        raise AssertionError("run_replicas accepted an engine with removed atoms")


def test_rejection_draws_counts_the_draws_of_the_veto():
    """The device engine's windows replay a veto's Python random draws from
    rejection_draws: it must equal the draws should_step_get_rejected makes
    (one for a flexible constraint whose error grows, none otherwise), for
    every outcome."""
    from fullrmc_amd.Constraints import InterMolecularDistanceConstraint

    for flexible in (True, False):
        for rp in (0.0, 0.5, 1.0):
            c = InterMolecularDistanceConstraint(2.0, flexible=flexible, rejectProbability=rp)
            c.standardError = np.float32(1.0)
            # strict vetoes read the counts of the move's pairs
            c.activeAtomsDataBeforeMove = {"number": np.zeros((1, 1, 1), np.int32)}
            c.activeAtomsDataAfterMove = {"number": np.ones((1, 1, 1), np.int32)}
            for after in (np.float32(0.5), np.float32(1.0), np.float32(1.5)):
                for seed in range(20):
                    random.seed(seed)
                    before = random.getstate()
                    for _ in range(c.rejection_draws(after)):
                        random.random()
                    replayed = random.getstate()
                    random.setstate(before)
                    c.should_step_get_rejected(after)
                    assert random.getstate() == replayed, (flexible, rp, after)
