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
