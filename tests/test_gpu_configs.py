"""GPU parity at the full sizes of the benchmark workloads:

* the 200,000-atom SiO2-like box (66,667 Si, 145 A cube, 1.5 A minimum
  distance): one O(N^2) full histogram on r in [0, 20) at 0.01 A (2,000
  bins, 64 KB of LDS bins per workgroup) in the reference's float32 binning
  and in the binary64-recompute mode, each slot for slot against the oracle
  (its cell-list form, pinned to the all-pairs loops on the CPU); then the
  device engine with G(r) on 2,000 bins and S(Q) on Q = 0.3..30 (1,000
  points) and windows of up to 128 proposals, move for move against the
  sequential oracle engine;
* the 13,500-atom B2 alloy with G(r) and S(Q): a 3,000-move trajectory
  against the oracle;
* 256 replicas of 1,000-20,000 atoms (log-uniform, 0.05 atoms per A^3),
  batched with windows of up to 32: every replica bit-identical to its own
  single-engine run, and four across the size range to the oracle.
"""
import multiprocessing
import os
import random
import sys

import numpy as np
import pytest

from fullrmc_amd import systems, tuning
from fullrmc_amd.Constraints import PairDistributionConstraint, StructureFactorConstraint
from fullrmc_amd.Core import pairs_histograms as gpu
from fullrmc_amd.Engine import run_replicas
from fullrmc_amd.Generators import TranslationGenerator
from oracle import chain, refcpu
from oracle import engine as oeng

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_engine import (WEIGHTS, assert_same_state, device_engine, element_index,  # noqa: E402
                             oracle_run, pdc_target, sfc_target)

pytestmark = pytest.mark.gpu
F32 = np.float32


def say(msg):
    print(msg, flush=True)


# ------------------------------------------------------ 200,000-atom SiO2 box

@pytest.fixture(scope="module")
def sio2():
    s = systems.sio2_box(200000, 145.0, 1.5, seed=2030)
    s["elem"] = element_index(s["elements"])
    s["box"] = systems.box_coordinates(s["recip"], s["real"])
    return s


def test_sio2_200k_full_histogram_float32_matches_oracle(sio2):
    """float32 binning on r in [0, 20) at 0.01 A, 2,000 bins, tuned and
    fixed-rule schedules: every intra and inter slot equals the oracle."""
    s = sio2
    args = (s["box"], s["basis"], True, s["molecules"], s["elem"], 2, F32(0.0), F32(20.0), F32(0.01), 2000)
    ri, re_, ov = refcpu.full_pairs_counts_cells(s["box"], s["basis"], s["molecules"], s["elem"], 2, F32(0.0),
                                                 F32(20.0), F32(0.01), 2000)
    say("sio2 f32: oracle has %d in-range pairs" % (ri.sum() + re_.sum()))
    assert re_.sum() > 2e8
    for autotune in (True, False):
        tuning.set_autotune(autotune)
        try:
            hi, he = gpu.full_pairs_histograms_coords(*args)
            sched = tuning.last_full_schedule()
        finally:
            tuning.set_autotune(True)
        say("sio2 f32: schedule %r" % sched)
        assert hi.dtype == np.float32 and hi.shape == (2, 2, 2000)
        assert np.array_equal(hi, ri.astype(F32)) and np.array_equal(he, re_.astype(F32))


def test_sio2_200k_full_histogram_float64_mode_matches_oracle(sio2):
    """The binary64-recompute mode at full size: equal to the oracle's
    binary64 enumeration slot for slot."""
    s = sio2
    hi, he, rechecked = gpu.full_pairs_histograms_coords_float64(s["box"], s["basis"], True, s["molecules"],
                                                                 s["elem"], 2, F32(0.0), F32(20.0), F32(0.01), 2000)
    ri, re_, ov = refcpu.full_pairs_counts_cells(s["box"], s["basis"], s["molecules"], s["elem"], 2, F32(0.0),
                                                 F32(20.0), F32(0.01), 2000, float64=True)
    say("sio2 f64: %d pairs re-checked in double" % rechecked)
    assert np.array_equal(hi, ri.astype(F32)) and np.array_equal(he, re_.astype(F32))
    assert 0 < rechecked < 1e9


def test_sio2_200k_engine_matches_oracle(sio2):
    """The engine on the SiO2 box with G(r) (2,000 bins) and S(Q) (its own
    r grid [0, 20) at 0.01 A, Q = 0.3..30, 1,000 points), single-atom
    translations of 0.1 A, windows of up to 128 proposals: 400 moves with
    the same groups, decisions and totals as the oracle, and the same final
    histograms, coordinates and random streams."""
    s = dict(sio2)
    nb = 2000
    pdc_exp = pdc_target(s, 0.05, 2031, 0.005, 0.01, nb)
    q = np.linspace(0.3, 30.0, 1000).astype(F32)
    sfc_exp = sfc_target(s, 0.05, 2031, q, 0.0, 20.0, 0.01)
    say("sio2 engine: targets built")
    cons = lambda osys, w: [oeng.PDC(osys, pdc_exp, w), oeng.SFC(osys, sfc_exp, w, 0.0, 20.0, 0.01)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(len(s["real"]))], oeng.Translation(0.1), cons, 400, 31)
    say("sio2 engine: oracle ran 400 moves, %d accepted" % oe.accepted)
    random.seed(31)
    np.random.seed(31)
    E = device_engine(s, [PairDistributionConstraint(pdc_exp),
                          StructureFactorConstraint(sfc_exp, rmin=0.0, rmax=20.0, dr=0.01)])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    st = E.run_device(400, k=128, log=dlog)
    assert st["generated"] == 400
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert 20 < oe.accepted < 380
    E._drop_device()


# ------------------------------------------------------- 13,500-atom B2 alloy

def test_b2_13500_trajectory_matches_oracle():
    """B2 alloy, 13,500 atoms (15 x 15 x 30 cells, a = 3.01 A, sigma 0.08 A),
    G(r) on r in [0, 20) at 0.02 A and S(Q) on Q = 0.5..25 (500 points)
    against an independently jittered lattice, translations of 0.1 A,
    windows of up to 64: 3,000 moves identical to the oracle's."""
    s = systems.b2_alloy(cells=(15, 15, 30), a=3.01, sigma=0.08, seed=2026)
    pdc_exp = pdc_target(s, 0.08, 2027, 0.01, 0.02, 1000)
    q = np.linspace(0.5, 25.0, 500).astype(F32)
    sfc_exp = sfc_target(s, 0.08, 2027, q, 0.0, 20.0, 0.02)
    cons = lambda osys, w: [oeng.PDC(osys, pdc_exp, w), oeng.SFC(osys, sfc_exp, w, 0.0, 20.0, 0.02)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(len(s["real"]))], oeng.Translation(0.1), cons, 3000, 41)
    say("b2: oracle ran 3000 moves, %d accepted" % oe.accepted)
    random.seed(41)
    np.random.seed(41)
    E = device_engine(s, [PairDistributionConstraint(pdc_exp),
                          StructureFactorConstraint(sfc_exp, rmin=0.0, rmax=20.0, dr=0.02)])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    st = E.run_device(3000, k=64, log=dlog)
    assert st["generated"] == 3000 and st["windows"] < 3000
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert 30 < oe.accepted < 2970
    E._drop_device()


# ------------------------------------------------------------- 256 replicas

def test_256_replicas_match_independent_runs_and_oracle():
    """256 single-element boxes, N log-uniform in [1000, 20000] (density
    0.05 per A^3, 2.0 A minimum distance), each with its own target G(r)
    (r in [0, 15) at 0.05 A, from a 0.1 A jittered copy), advanced together
    by the batched replica loop with windows of up to 32: every replica ends
    bit-identical to its own single-engine run with the same seed, and the
    replicas at the 0, 1/3, 2/3 and 1 size quantiles equal the oracle."""
    sizes = systems.replica_sizes(256, 1000, 20000, 2041)
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    # boxes built in fresh interpreters (this process already holds the GPU)
    with multiprocessing.get_context("spawn").Pool(workers) as pool:
        boxes = pool.starmap(systems.random_box, [(int(n), float((n / 0.05) ** (1.0 / 3.0)), 2.0, 5000 + r)
                                                  for r, n in enumerate(sizes)])
    say("replicas: %d boxes built, %d atoms in all" % (len(boxes), int(sum(sizes))))
    for r, b in enumerate(boxes):
        b["exp"] = _gpu_pdc_target(b, 0.1, 6000 + r)

    def engine(b):
        E = device_engine(b, [PairDistributionConstraint(b["exp"])])
        E.set_move_generator(TranslationGenerator(0.1))
        return E

    nsteps = 120
    seeds = [7000 + r for r in range(len(boxes))]
    reps = [engine(b) for b in boxes]
    stats = run_replicas(reps, nsteps, k=32, seeds=seeds)
    assert all(st["generated"] == nsteps for st in stats)
    assert sum(st["accepted"] for st in stats) > 256
    say("replicas: batched run done")
    for r, (E, b, seed) in enumerate(zip(reps, boxes, seeds)):
        E.sync_from_device()
        random.seed(seed)
        np.random.seed(seed)
        S = engine(b)
        S.run_device(nsteps, k=32)
        assert E.realCoordinates.tobytes() == S.realCoordinates.tobytes(), "replica %d" % r
        assert np.float32(E.totalStandardError).tobytes() == np.float32(S.totalStandardError).tobytes()
        assert E.accepted == S.accepted
        assert E.constraints[0].data["inter"].tobytes() == S.constraints[0].data["inter"].tobytes()
        assert E._replica_rng[0] == random.getstate()
        assert np.array_equal(E._replica_rng[1][1], np.random.get_state()[1])
        S._drop_device()
    say("replicas: all 256 equal their single-engine runs")
    order = np.argsort(sizes)
    for qt in (0.0, 1 / 3, 2 / 3, 1.0):
        r = int(order[int(round(qt * (len(order) - 1)))])
        b = boxes[r]
        oe, olog, pys, nps = oracle_run(b, [[i] for i in range(len(b["real"]))], oeng.Translation(0.1),
                                        lambda osys, w, b=b: [oeng.PDC(osys, b["exp"], w)], nsteps, seeds[r])
        E = reps[r]
        assert E.realCoordinates.tobytes() == oe.s.real.tobytes(), "replica %d (%d atoms)" % (r, sizes[r])
        assert np.float32(E.totalStandardError).tobytes() == np.float32(oe.tse).tobytes()
        assert E.accepted == oe.accepted
        assert E.constraints[0].data["inter"].tobytes() == oe.constraints[0].inter.tobytes()
        assert E._replica_rng[0] == pys
        assert np.array_equal(E._replica_rng[1][1], nps[1])
        say("replicas: replica %d (%d atoms) equals the oracle" % (r, sizes[r]))
    for E in reps:
        E._drop_device()


def _gpu_pdc_target(s, sigma, seed):
    """experimentalData (r in [0, 15) at 0.05 A) of a jittered copy of s: the
    oracle chain on the GPU drop-in's full histogram (the targets are inputs
    here; the engines' own histograms are what the test compares)."""
    t = systems.jitter(s, sigma, seed)
    x = (0.025 + 0.05 * np.arange(300)).astype(F32)
    exp = np.stack([x, np.zeros_like(x)], 1).astype(F32)
    osys = oeng.System(t["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    w = chain.normalized_weighting(osys.counts, {e: WEIGHTS[e] for e in osys.elements})
    p = oeng.PDC(osys, exp, w)
    hi, he = gpu.full_pairs_histograms_coords(osys.box, s["basis"], True, osys.mol, osys.elem, osys.nel, p.min,
                                              p.max, p.bin, p.H)
    exp[:, 1] = p.model(hi, he)
    return exp
