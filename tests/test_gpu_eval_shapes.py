"""Every window-evaluation shape (bins per thread 4 or 8; work units spread
over the XCDs or aligned to them; S(Q) units of 64 Q points with k_eval_b,
or pairwise-sum leaves whose sums the host combines; the S(Q) matrix in its
transposed or its four-row interleaved copy; the split chain, products in
LDS from producer waves) evaluates a window to
the same bits: forced in turn, each gives the oracle's trajectory; the tuned
engine reports a shape per window-size bucket and also matches."""
import ctypes
import os
import random
import sys

import numpy as np
import pytest

from fullrmc_amd import _lib, systems
from fullrmc_amd.Constraints import PairDistributionConstraint, StructureFactorConstraint
from fullrmc_amd.Generators import TranslationGenerator
from oracle import engine as oeng

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_engine import assert_same_state, device_engine, oracle_run, pdc_target, sfc_target  # noqa: E402

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def case():
    """1,200 atoms (Si:O 1:2), 24 A cube; G(r) on 1,500 bins of 0.008 A (two
    G(r) passes at four bins per thread, one at eight) and S(Q) on 150 Q
    points (three work units) on its own 0.01 A grid."""
    s = systems.random_box(1200, 24.0, 1.6, seed=81, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    pdc = pdc_target(s, 0.1, 82, 0.004, 0.008, 1500)
    q = np.linspace(0.5, 15.0, 150).astype(F32)
    sfc = sfc_target(s, 0.1, 82, q, 0.0, 11.0, 0.01)
    cons = lambda osys, w: [oeng.PDC(osys, pdc, w), oeng.SFC(osys, sfc, w, 0.0, 11.0, 0.01)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(1200)], oeng.Translation(0.2), cons, 300, 83)
    return s, pdc, sfc, oe, olog, pys, nps


@pytest.mark.parametrize("shape", list(range(0, 64, 4)) + [80, 84, 88, 92, -1])
def test_every_eval_shape_gives_the_oracle_trajectory(case, shape):
    s, pdc, sfc, oe, olog, pys, nps = case
    L = _lib.lib()
    _lib.check(L.frmc_set_eval_shape(shape))
    try:
        random.seed(83)
        np.random.seed(83)
        E = device_engine(s, [PairDistributionConstraint(pdc),
                              StructureFactorConstraint(sfc, rmin=0.0, rmax=11.0, dr=0.01)])
        E.set_move_generator(TranslationGenerator(0.2))
        dlog = []
        E.run_device(300, k=16, log=dlog)
        assert_same_state(E, oe, dlog, olog, (pys, nps))
        got = (ctypes.c_int32 * 6)()
        _lib.check(L.frmc_engine_eval_shapes(E.device_handle(), got, 6))
        assert all(g in list(range(0, 64, 4)) + [80, 84, 88, 92] for g in got)
        if shape >= 0:
            assert list(got) == [shape] * 6
        E._drop_device()
    finally:
        _lib.check(L.frmc_set_eval_shape(-1))


def test_eval_shape_setter_rejects_unknown_shapes():
    L = _lib.lib()
    for bad in (1, 2, 3, 5, 6, 7, -2, 9, 30, 62, 64, 68, 76, 96, 112):
        with pytest.raises(ValueError):
            _lib.check(L.frmc_set_eval_shape(bad))


# S(Q) grids whose row count ends the matrix stream with every possible tail:
# ceil(H/16) batches of 16 rows are 1, 2, 3, 4 (all of them in the tail),
# 8 (a whole prefetch round, then four), 9, 10 and 11, with partial last
# batches among them.
SQ_ROWS = [10, 20, 40, 64, 120, 130, 150, 170]


@pytest.fixture(scope="module")
def small_box():
    return systems.random_box(300, 15.0, 1.6, seed=91, species=("Si", "O"), fractions=(1 / 3, 2 / 3))


@pytest.mark.parametrize("shape", [0, 16, 32, 48, 80])
@pytest.mark.parametrize("rows", SQ_ROWS)
def test_sq_sum_over_every_row_tail(small_box, rows, shape):
    s = small_box
    dr = 0.05
    rmax = rows * dr
    pdc = pdc_target(s, 0.1, 92, 0.025, 0.05, 150)
    q = np.linspace(0.5, 12.0, 70).astype(F32)
    sfc = sfc_target(s, 0.1, 92, q, 0.0, rmax, dr)
    cons = lambda osys, w: [oeng.PDC(osys, pdc, w), oeng.SFC(osys, sfc, w, 0.0, rmax, dr)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(300)], oeng.Translation(0.2), cons, 80, 93)
    L = _lib.lib()
    _lib.check(L.frmc_set_eval_shape(shape))
    try:
        random.seed(93)
        np.random.seed(93)
        E = device_engine(s, [PairDistributionConstraint(pdc),
                              StructureFactorConstraint(sfc, rmin=0.0, rmax=rmax, dr=dr)])
        E.set_move_generator(TranslationGenerator(0.2))
        dlog = []
        E.run_device(80, k=8, log=dlog)
        assert_same_state(E, oe, dlog, olog, (pys, nps))
        E._drop_device()
    finally:
        _lib.check(L.frmc_set_eval_shape(-1))
