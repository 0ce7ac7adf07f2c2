"""k_delta's chunk walk over the spatially sorted atom copy gives the
oracle's trajectory.  Forced on for small periodic boxes (the default turns
it on from 4,096 atoms): cubic and triclinic cells, windows whose proposals
travel in the kernel arguments (k <= 64) or in the staging buffer (k > 64),
the native loop (commits deferred to the next k_delta) and windows of
Python-generated proposals decided on the host (commits by k_commit), G(r) plus S(Q) on its own longer r grid,
and runs long enough for the chunk boxes to be rebuilt (every n/4 commits)."""
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
def cells_mode():
    L = _lib.lib()

    def set_mode(m):
        _lib.check(L.frmc_set_delta_cells(m))

    yield set_mode
    _lib.check(L.frmc_set_delta_cells(-1))


def triclinic_box(n, seed):
    rng = np.random.default_rng(seed)
    basis = systems.triclinic_basis(30.0, 32.0, 34.0, 80.0, 95.0, 105.0)
    real = np.asarray(systems._rsa(n, basis, 1.8, rng), F32)
    elements = ["Si" if i % 3 == 0 else "O" for i in range(n)]
    return dict(real=real, elements=elements, molecules=np.arange(n, dtype=np.int32), basis=basis,
                recip=systems.reciprocal(basis))


def chunks_of(E):
    n = ctypes.c_int32(-1)
    w = ctypes.c_float(0.0)
    _lib.check(_lib.lib().frmc_engine_delta_cells(E.device_handle(), ctypes.byref(n), ctypes.byref(w)))
    return n.value, w.value


@pytest.fixture(scope="module")
def cubic():
    s = systems.random_box(1500, 26.0, 1.8, seed=61, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    s["exp"] = pdc_target(s, 0.1, 62, 0.025, 0.05, 160)  # r < 8 A
    return s


@pytest.fixture(scope="module")
def tric():
    s = triclinic_box(1800, 63)
    s["exp"] = pdc_target(s, 0.1, 64, 0.025, 0.05, 190)  # r < 9.5 A
    return s


@pytest.mark.parametrize("k", [8, 64, 128])
def test_chunk_walk_cubic_matches_oracle(cubic, cells_mode, k):
    s = cubic
    n = len(s["real"])
    cons = lambda osys, w: [oeng.PDC(osys, s["exp"], w)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), cons, 1500, 65)
    cells_mode(1)
    random.seed(65)
    np.random.seed(65)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    dlog = []
    E.run_device(1500, k=k, log=dlog)
    nchunk, wcut = chunks_of(E)
    assert nchunk == (n + 255) // 256
    assert 8.0 < wcut < 8.02
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    # enough commits for the chunk boxes to be rebuilt (every n/4 = 375)
    assert sum(a for _, a, _ in dlog) > 400


@pytest.mark.parametrize("k", [16, 100])
def test_chunk_walk_triclinic_matches_oracle(tric, cells_mode, k):
    s = tric
    n = len(s["real"])
    cons = lambda osys, w: [oeng.PDC(osys, s["exp"], w)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.3), cons, 1500, 66)
    cells_mode(1)
    random.seed(66)
    np.random.seed(66)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.3))
    dlog = []
    E.run_device(1500, k=k, log=dlog)
    assert chunks_of(E)[0] == (n + 255) // 256
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert sum(a for _, a, _ in dlog) > 500


def test_chunk_walk_with_sq_on_its_own_grid(tric, cells_mode):
    """G(r) on r < 9.5 A and S(Q) on its own r grid up to 12 A: the cut-off
    follows the longer grid."""
    s = tric
    n = len(s["real"])
    q = np.linspace(0.5, 12.0, 90).astype(F32)
    sfc = sfc_target(s, 0.1, 67, q, 0.0, 12.0, 0.05)
    cons = lambda osys, w: [oeng.PDC(osys, s["exp"], w), oeng.SFC(osys, sfc, w, 0.0, 12.0, 0.05)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.3), cons, 600, 68)
    cells_mode(1)
    random.seed(68)
    np.random.seed(68)
    E = device_engine(s, [PairDistributionConstraint(s["exp"]),
                          StructureFactorConstraint(sfc, rmin=0.0, rmax=12.0, dr=0.05)])
    E.set_move_generator(TranslationGenerator(0.3))
    dlog = []
    E.run_device(600, k=32, log=dlog)
    assert 12.0 < chunks_of(E)[1] < 12.02
    assert_same_state(E, oe, dlog, olog, (pys, nps))


def test_chunk_walk_with_host_decided_commits(cubic, cells_mode):
    """Proposals from the Python generators, evaluated on the device and
    decided on the host: every accepted move is committed by k_commit, which
    keeps the sorted copy."""
    s = cubic
    n = len(s["real"])
    cons = lambda osys, w: [oeng.PDC(osys, s["exp"], w)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), cons, 250, 69)
    cells_mode(1)
    random.seed(69)
    np.random.seed(69)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    dlog = []
    E.run_device(250, k=8, log=dlog, native=False)
    assert chunks_of(E)[0] > 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))


def test_chunk_walk_off_and_by_size(cubic, cells_mode):
    s = cubic
    cells_mode(0)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.run_device(10, k=4)
    assert chunks_of(E)[0] == 0
    E._drop_device()
    cells_mode(-1)  # 1,500 atoms: below the size from which it is on by default
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.run_device(10, k=4)
    assert chunks_of(E)[0] == 0
    with pytest.raises(ValueError):
        _lib.check(_lib.lib().frmc_set_delta_cells(2))


def test_chunk_walk_survives_batched_replica_runs(cubic, cells_mode):
    """Native windows, then a batched replica run (whose kernels do not keep
    the sorted copy), then native windows again: the engine rebuilds the copy
    and ends in exactly the state of a twin engine without the chunk walk
    (whose path the other engine tests check against the oracle)."""
    from fullrmc_amd.Engine import run_replicas
    s = cubic
    small = systems.random_box(400, 20.0, 1.8, seed=71, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    small["exp"] = pdc_target(small, 0.1, 72, 0.025, 0.05, 120)

    def engine(box, mode):
        cells_mode(mode)
        E = device_engine(box, [PairDistributionConstraint(box["exp"])])
        E.set_move_generator(TranslationGenerator(0.2))
        return E

    finals = []
    for mode in (1, 0):
        E = engine(s, mode)
        F = engine(small, mode)
        random.seed(73)
        np.random.seed(73)
        E.run_device(300, k=16)
        run_replicas([E, F], 200, k=16, seeds=[74, 75])
        random.seed(76)
        np.random.seed(76)
        E.run_device(300, k=16)
        assert (chunks_of(E)[0] > 0) == (mode == 1)
        finals.append((E.realCoordinates.tobytes(), E.boxCoordinates.tobytes(),
                       np.float32(E.totalStandardError).tobytes(),
                       [(c.data["intra"].tobytes(), c.data["inter"].tobytes()) for c in E.constraints],
                       E.accepted))
        E._drop_device()
        F._drop_device()
    assert finals[0] == finals[1]
    assert finals[0][4] > 100


def test_far_unwrapped_coordinates_turn_the_chunk_walk_off(cells_mode):
    """Box coordinates beyond 64 box lengths would need a larger rounding
    margin than the chunk test allows: the engine keeps no sorted copy and
    still gives the oracle's trajectory."""
    s = systems.random_box(600, 18.0, 1.8, seed=81, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    s["exp"] = pdc_target(s, 0.1, 82, 0.025, 0.05, 120)
    s = dict(s)
    s["real"] = (s["real"] + np.float32(18.0 * 100)).astype(F32)  # 100 cells away along every axis
    n = len(s["real"])
    cons = lambda osys, w: [oeng.PDC(osys, s["exp"], w)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), cons, 300, 83)
    cells_mode(1)
    random.seed(83)
    np.random.seed(83)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.2))
    dlog = []
    E.run_device(300, k=16, log=dlog)
    assert chunks_of(E)[0] == 0
    assert_same_state(E, oe, dlog, olog, (pys, nps))
