"""The device window loop (k_loop_step: the walk of each window's totals, the
rewind of both random streams and the next window's draws on the GPU, in
batches of windows of one size) gives the oracle's trajectory: logs, totals,
histograms, coordinates and both random streams after the run.  Covered:
windows of 8 to 128 proposals, the tolerance branch, the four kinds of
evaluation layout (64-Q units with k_eval_b, pairwise leaves combined by
k_loop_step, with the split chain), a group count whose randrange rejects
half of the words (many MT19937 blocks crossed per window), runs split over
several calls with the host loop taking the last windows, the chunk walk
with rebuilds between batches, groups out of atom order, the host loop alone
(the loop off), and batched replicas on and off the device loop."""
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


@pytest.fixture
def loop_mode():
    L = _lib.lib()

    def set_mode(m):
        _lib.check(L.frmc_set_window_loop(m))

    yield set_mode
    _lib.check(L.frmc_set_window_loop(-1))


@pytest.fixture
def eval_shape():
    L = _lib.lib()

    def force(sh):
        _lib.check(L.frmc_set_eval_shape(sh))

    yield force
    _lib.check(L.frmc_set_eval_shape(-1))


def loop_windows(E):
    n = ctypes.c_int64(-1)
    _lib.check(_lib.lib().frmc_engine_loop_windows(E.device_handle(), ctypes.byref(n)))
    return n.value


@pytest.fixture(scope="module")
def box500():
    s = systems.random_box(500, 18.0, 2.0, seed=121)
    s["exp"] = pdc_target(s, 0.1, 122, 0.025, 0.05, 180)
    return s


def _pdc(s):
    return lambda osys, w: [oeng.PDC(osys, s["exp"], w)]


@pytest.mark.parametrize("k", [8, 32, 128])
def test_window_loop_matches_oracle(box500, loop_mode, k):
    s = box500
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), _pdc(s), 2500, 123)
    loop_mode(1)
    random.seed(123)
    np.random.seed(123)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    dlog = []
    st = E.run_device(2500, k=k, log=dlog)
    assert st["generated"] == 2500
    assert loop_windows(E) > 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert 100 < sum(a for _, a, _ in dlog) < 2400


def test_window_loop_with_tolerance(box500, loop_mode):
    s = box500
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.3), _pdc(s), 1500, 124,
                                    tolerance=30.0)
    loop_mode(1)
    random.seed(124)
    np.random.seed(124)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.3))
    E.set_tolerance(30.0)
    dlog = []
    E.run_device(1500, k=32, log=dlog)
    assert loop_windows(E) > 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert E.tolerated > 0


def test_window_loop_host_loop_alone(box500, loop_mode):
    s = box500
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), _pdc(s), 800, 125)
    loop_mode(0)
    random.seed(125)
    np.random.seed(125)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    dlog = []
    E.run_device(800, k=32, log=dlog)
    assert loop_windows(E) == 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().frmc_set_window_loop(2))


@pytest.fixture(scope="module")
def alloy():
    s = systems.b2_alloy(cells=(6, 6, 6), a=3.01, sigma=0.08, seed=2)
    s["pdc"] = pdc_target(s, 0.08, 3, 0.01, 0.02, 400)
    q = np.linspace(0.5, 25, 120).astype(F32)
    s["sfc"] = sfc_target(s, 0.08, 3, q, 0.0, 8.0, 0.02)
    return s


# 0: 64-Q units and k_eval_b; 16: pairwise leaves, combined by k_loop_step;
# 48: leaves with the interleaved matrix; 88: the split chain
@pytest.mark.parametrize("shape", [0, 16, 48, 88])
def test_window_loop_pdc_plus_sfc_every_layout(alloy, loop_mode, eval_shape, shape):
    s = alloy
    n = len(s["real"])

    def cons(osys, w):
        return [oeng.PDC(osys, s["pdc"], w), oeng.SFC(osys, s["sfc"], w, 0.0, 8.0, 0.02)]

    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.1), cons, 900, 126)
    loop_mode(1)
    eval_shape(shape)
    random.seed(126)
    np.random.seed(126)
    E = device_engine(s, [PairDistributionConstraint(s["pdc"]),
                          StructureFactorConstraint(s["sfc"], rmin=0.0, rmax=8.0, dr=0.02)])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    E.run_device(900, k=16, log=dlog)
    assert loop_windows(E) > 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))


def test_window_loop_half_rejected_group_draws(loop_mode):
    """1,025 groups: randrange(1025) takes the top 11 bits of a word and
    rejects about half of them.  Windows of a fixed 128 proposals use about
    770 words of the Python stream and 768 of the NumPy stream each, so
    every window crosses state blocks of both."""
    s = systems.random_box(1025, 24.0, 1.8, seed=127)
    s["exp"] = pdc_target(s, 0.1, 128, 0.025, 0.05, 200)
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(1025)], oeng.Translation(0.2), _pdc(s), 1800, 129)
    loop_mode(1)
    random.seed(129)
    np.random.seed(129)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    h = E._device_engine(128)
    _lib.check(_lib.lib().frmc_engine_set_adaptive(h, 0))
    dlog = []
    st = E.run_device(1800, k=128, log=dlog)
    assert loop_windows(E) >= 5
    assert st["evaluated"] >= 128 * loop_windows(E)
    assert_same_state(E, oe, dlog, olog, (pys, nps))


def test_window_loop_runs_split_over_calls(box500, loop_mode):
    """Three calls (the last one too short for a window of the loop) give the
    trajectory of one run."""
    s = box500
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), _pdc(s), 1000, 130)
    loop_mode(1)
    random.seed(130)
    np.random.seed(130)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    dlog = []
    E.run_device(480, k=24, log=dlog)
    w1 = loop_windows(E)
    E.run_device(517, k=24, log=dlog)
    E.run_device(3, k=24, log=dlog)
    assert 0 < w1 < loop_windows(E)
    assert_same_state(E, oe, dlog, olog, (pys, nps))


@pytest.mark.parametrize("mode", [-1, 0])
def test_batched_replicas_on_and_off_the_device_loop(loop_mode, mode):
    """Batched replicas take the device loop by default (-1) and the host
    walks with it off (0); both end every replica in the state of its own
    single-engine run (host loop), tolerance above zero included, and
    replica 0 in the oracle's."""
    from fullrmc_amd.Engine import run_replicas
    boxes = []
    for i, n in enumerate((700, 400, 1100, 520)):
        L = (n / 0.05) ** (1 / 3)
        b = systems.random_box(n, L, 2.0, seed=140 + i)
        b["exp"] = pdc_target(b, 0.1, 150 + i, 0.025, 0.05, 150)
        boxes.append(b)

    def engine(b):
        E = device_engine(b, [PairDistributionConstraint(b["exp"])])
        E.set_move_generator(TranslationGenerator(0.2))
        E.set_tolerance(10.0)
        return E

    loop_mode(mode)
    reps = [engine(b) for b in boxes]
    seeds, nsteps = [160, 161, 162, 163], 900
    stats = run_replicas(reps, nsteps, k=24, seeds=seeds)
    assert all(st["generated"] == nsteps for st in stats)
    windows = [loop_windows(E) for E in reps]
    assert all(w > 0 for w in windows) if mode == -1 else all(w == 0 for w in windows)
    loop_mode(0)
    for E, b, seed in zip(reps, boxes, seeds):
        E.sync_from_device()
        random.seed(seed)
        np.random.seed(seed)
        S = engine(b)
        S.run_device(nsteps, k=24)
        assert E.realCoordinates.tobytes() == S.realCoordinates.tobytes()
        assert E.boxCoordinates.tobytes() == S.boxCoordinates.tobytes()
        assert np.float32(E.totalStandardError).tobytes() == np.float32(S.totalStandardError).tobytes()
        assert (E.accepted, E.tolerated) == (S.accepted, S.tolerated)
        for c, cs in zip(E.constraints, S.constraints):
            assert c.data["intra"].tobytes() == cs.data["intra"].tobytes()
            assert c.data["inter"].tobytes() == cs.data["inter"].tobytes()
        assert E._replica_rng[0] == random.getstate()
        assert np.array_equal(E._replica_rng[1][1], np.random.get_state()[1])
    b = boxes[0]
    oe, olog, pys, nps = oracle_run(b, [[i] for i in range(len(b["real"]))], oeng.Translation(0.2), _pdc(b),
                                    nsteps, seeds[0], tolerance=10.0)
    assert np.float32(reps[0].totalStandardError).tobytes() == np.float32(oe.tse).tobytes()
    assert reps[0].realCoordinates.tobytes() == oe.s.real.tobytes()


def test_window_loop_with_the_chunk_walk(loop_mode):
    """The device loop with k_delta's chunk walk (forced on for a 1,500-atom
    periodic box): commits by the next k_delta keep the sorted copy, which the
    host rebuilds between batches (every n/4 = 375 commits)."""
    s = systems.random_box(1500, 26.0, 1.8, seed=171, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    s["exp"] = pdc_target(s, 0.1, 172, 0.025, 0.05, 160)
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), _pdc(s), 2000, 173)
    L = _lib.lib()
    _lib.check(L.frmc_set_delta_cells(1))
    try:
        loop_mode(1)
        random.seed(173)
        np.random.seed(173)
        E = device_engine(s, [PairDistributionConstraint(s["exp"])])
        E.set_move_generator(TranslationGenerator(0.2))
        dlog = []
        E.run_device(2000, k=32, log=dlog)
        nch = ctypes.c_int32(0)
        _lib.check(L.frmc_engine_delta_cells(E.device_handle(), ctypes.byref(nch), None))
        assert nch.value == (n + 255) // 256
        assert loop_windows(E) > 0
        assert_same_state(E, oe, dlog, olog, (pys, nps))
        assert sum(a for _, a, _ in dlog) > 400
    finally:
        _lib.check(L.frmc_set_delta_cells(-1))


def test_window_loop_groups_out_of_atom_order(loop_mode):
    """Single-atom groups in a shuffled order over part of the atoms: the
    draw looks each group's atom up in the group table."""
    s = systems.random_box(600, 20.0, 1.8, seed=181)
    s["exp"] = pdc_target(s, 0.1, 182, 0.025, 0.05, 150)
    order = np.random.default_rng(183).permutation(600)[:450]
    groups = [[int(i)] for i in order]
    oe, olog, pys, nps = oracle_run(s, groups, oeng.Translation(0.2), _pdc(s), 1500, 184)
    loop_mode(1)
    random.seed(184)
    np.random.seed(184)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])], groups=groups, gen=TranslationGenerator(0.2))
    dlog = []
    E.run_device(1500, k=32, log=dlog)
    assert loop_windows(E) > 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))
