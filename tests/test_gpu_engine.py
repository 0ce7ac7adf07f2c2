"""GPU parity of the run loop: the device-resident speculative engine and the
per-move drop-in loop reproduce the sequential oracle engine's trajectory
(accepted-move sequence, totals after every step), its slot-exact
histograms, chi-squares, coordinates and both random streams."""
import random

import numpy as np
import pytest

from fullrmc_amd import systems
from fullrmc_amd.Constraints import (InterMolecularDistanceConstraint, IntraMolecularDistanceConstraint,
                                     PairCorrelationConstraint,
                                     PairDistributionConstraint, ReducedStructureFactorConstraint,
                                     StructureFactorConstraint)
from fullrmc_amd.Engine import Engine
from fullrmc_amd.Generators import (AtomsRemoveGenerator, MoveGeneratorCombinator, RotationGenerator,
                                    SwapPositionsGenerator, TranslationGenerator)
from oracle import chain, refcpu
from oracle import engine as oeng

pytestmark = pytest.mark.gpu
F32 = np.float32
WEIGHTS = {"O": 8, "Si": 14, "Fe": 26, "Al": 13, "C": 6, "H": 1, "Ar": 18}


def element_index(elements):
    lut = {e: i for i, e in enumerate(sorted(set(elements)))}
    return np.array([lut[e] for e in elements], dtype=np.int32)


def counts_of(elements):
    c = {}
    for e in elements:
        c[e] = c.get(e, 0) + 1
    return c


def pdc_target(s, sigma, seed, r_first, dr, nbins):
    """experimentalData of a jittered copy of s (G(r) by the oracle chain)."""
    t = systems.jitter(s, sigma, seed)
    x = (r_first + dr * np.arange(nbins)).astype(F32)
    exp = np.stack([x, np.zeros_like(x)], 1).astype(F32)
    osys = oeng.System(t["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    w = chain.normalized_weighting(osys.counts, {e: WEIGHTS[e] for e in osys.elements})
    p = oeng.PDC(osys, exp, w)
    hi, he = p.hist_full(osys.box, osys.mol, osys.elem)
    exp[:, 1] = p.model(hi, he)
    return exp


def sfc_target(s, sigma, seed, q, rmin, rmax, dr):
    t = systems.jitter(s, sigma, seed)
    osys = oeng.System(t["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    w = chain.normalized_weighting(osys.counts, {e: WEIGHTS[e] for e in osys.elements})
    exp = np.stack([q, np.zeros_like(q)], 1).astype(F32)
    c = oeng.SFC(osys, exp, w, rmin, rmax, dr)
    hi, he = c.hist_full(osys.box, osys.mol, osys.elem)
    exp[:, 1] = c.model(hi, he)
    return exp


def oracle_run(s, groups, gen, cons_fn, nsteps, seed, tolerance=0.0, rigid_fn=None):
    random.seed(seed)
    np.random.seed(seed)
    osys = oeng.System(s["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    w = chain.normalized_weighting(osys.counts, {e: WEIGHTS[e] for e in osys.elements})
    oe = oeng.Engine(osys, groups, gen, cons_fn(osys, w), tolerance,
                     rigid=rigid_fn(osys) if rigid_fn is not None else ())
    oe.initialize()
    log = []
    oe.run(nsteps, log)
    return oe, log, random.getstate(), np.random.get_state()


def device_engine(s, cons, groups=None, gen=None):
    E = Engine()
    E.set_system(s["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    if groups is not None:
        E.set_groups(groups, gen)
    E.add_constraints(cons)
    return E


def assert_same_state(E, oe, dlog, olog, states):
    assert len(dlog) == len(olog)
    first = next((i for i, (a, b) in enumerate(zip(dlog, olog)) if a != b), None)
    assert first is None, "trajectories diverge at step %d: %r vs %r" % (first, dlog[first], olog[first])
    assert np.float32(E.totalStandardError).tobytes() == np.float32(oe.tse).tobytes()
    for c, oc in zip(E.constraints, oe.constraints):
        assert c.data["intra"].tobytes() == oc.intra.tobytes()
        assert c.data["inter"].tobytes() == oc.inter.tobytes()
        assert np.float32(c.standardError).tobytes() == np.float32(oc.sd).tobytes()
    assert E.realCoordinates.tobytes() == oe.s.real.tobytes()
    assert E.boxCoordinates.tobytes() == oe.s.box.tobytes()
    pys, nps = states
    assert random.getstate() == pys
    d_np = np.random.get_state()
    assert np.array_equal(d_np[1], nps[1]) and d_np[2] == nps[2]


@pytest.fixture(scope="module")
def single_box():
    s = systems.random_box(500, 18.0, 2.0, seed=21)
    s["exp"] = pdc_target(s, 0.1, 22, 0.025, 0.05, 180)
    return s


def _pdc_cons(s):
    return lambda osys, w: [oeng.PDC(osys, s["exp"], w)]


@pytest.fixture(scope="module")
def plain_rmc_2000_run():
    """Plain RMC, one element, 2,000 atoms in a 30 A cube (uniform, rejection
    below 2.0 A), target G(r) from a 0.1 A Gaussian-jittered copy on
    r in [0,15) at 0.05 A, translations of 0.1 A, 10,000 steps: the oracle's
    trajectory."""
    s = systems.random_box(2000, 30.0, 2.0, seed=100)
    s["exp"] = pdc_target(s, 0.1, 101, 0.025, 0.05, 300)
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(2000)], oeng.Translation(0.1), _pdc_cons(s), 10000, 7)
    return s, oe, olog, pys, nps


@pytest.mark.parametrize("k", [1, 64])
def test_plain_rmc_2000_atoms_full_run_matches_oracle(plain_rmc_2000_run, k):
    """The whole 10,000-step run, one proposal per window and speculative
    windows of up to 64: the same group, decision and total at every step, and
    the same final state, as the oracle."""
    s, oe, olog, pys, nps = plain_rmc_2000_run
    random.seed(7)
    np.random.seed(7)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    st = E.run_device(10000, k=k, log=dlog)
    assert st["generated"] == 10000
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert 1000 < sum(a for _, a, _ in dlog) < 9000


@pytest.mark.parametrize("k", [1, 7, 64])
def test_native_translation_matches_oracle(single_box, k):
    s = single_box
    n = len(s["real"])
    nsteps = 400
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), _pdc_cons(s), nsteps, 5)
    random.seed(5)
    np.random.seed(5)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    dlog = []
    st = E.run_device(nsteps, k=k, log=dlog)
    assert st["generated"] == nsteps
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert 0 < sum(a for _, a, _ in dlog) < nsteps


def test_native_translation_with_tolerance(single_box):
    s = single_box
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.3), _pdc_cons(s), 300, 9,
                                    tolerance=30.0)
    random.seed(9)
    np.random.seed(9)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    E.set_move_generator(TranslationGenerator(0.3))
    E.set_tolerance(30.0)
    dlog = []
    E.run_device(300, k=32, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert E.tolerated > 0


def test_plain_dropin_loop_matches_oracle(single_box):
    s = single_box
    n = len(s["real"])
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.2), _pdc_cons(s), 60, 3)
    random.seed(3)
    np.random.seed(3)
    E = device_engine(s, [PairDistributionConstraint(s["exp"])])
    dlog = []
    E.run(60, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))


def test_pdc_plus_sfc_binary_alloy():
    s = systems.b2_alloy(cells=(6, 6, 6), a=3.01, sigma=0.08, seed=2)
    pdc_exp = pdc_target(s, 0.08, 3, 0.01, 0.02, 400)
    q = np.linspace(0.5, 25, 120).astype(F32)
    sfc_exp = sfc_target(s, 0.08, 3, q, 0.0, 8.0, 0.02)
    n = len(s["real"])

    def cons(osys, w):
        return [oeng.PDC(osys, pdc_exp, w), oeng.SFC(osys, sfc_exp, w, 0.0, 8.0, 0.02)]

    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.1), cons, 300, 4)
    random.seed(4)
    np.random.seed(4)
    E = device_engine(s, [PairDistributionConstraint(pdc_exp),
                          StructureFactorConstraint(sfc_exp, rmin=0.0, rmax=8.0, dr=0.02)])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    E.run_device(300, k=16, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))


def _target(s, sigma, seed, make):
    """experimentalData of a jittered copy of s: the oracle constraint made
    by make(osys, w) evaluated on it."""
    t = systems.jitter(s, sigma, seed)
    osys = oeng.System(t["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    w = chain.normalized_weighting(osys.counts, {e: WEIGHTS[e] for e in osys.elements})
    c = make(osys, w)
    hi, he = c.hist_full(osys.box, osys.mol, osys.elem)
    exp = np.array(c.exp, dtype=F32)
    exp[:, 1] = c.model(hi, he)
    return exp


def _used_weights(dw):
    """Data weights as the constraints use them: normalised to sum to their count."""
    w = np.array(dw, dtype=F32)
    w /= F32(np.sum(w))
    w *= F32(len(w))
    return w


@pytest.mark.parametrize("runner", ["native", "dropin"])
def test_pcf_plus_reduced_sfc_binary_alloy(runner):
    """PairCorrelationConstraint (g(r), scale factor 1.1, data weights) with a
    ReducedStructureFactorConstraint (scale factor 0.9) on their own grids."""
    s = systems.b2_alloy(cells=(6, 6, 6), a=3.01, sigma=0.08, seed=7)
    x = (0.01 + 0.02 * np.arange(400)).astype(F32)
    pcf_exp = _target(s, 0.08, 8, lambda o, w: oeng.PCF(o, np.stack([x, x], 1), w, 1.1))
    q = np.linspace(0.5, 25, 150).astype(F32)
    rsfc_exp = _target(s, 0.08, 8, lambda o, w: oeng.RSFC(o, np.stack([q, q], 1), w, 0.0, 8.0, 0.02, 0.9))
    dw = (1.0 + np.arange(400) % 3).astype(F32)
    n = len(s["real"])
    nsteps = 300 if runner == "native" else 60

    def cons(osys, w):
        return [oeng.PCF(osys, pcf_exp, w, 1.1, _used_weights(dw)),
                oeng.RSFC(osys, rsfc_exp, w, 0.0, 8.0, 0.02, 0.9)]

    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.1), cons, nsteps, 6)
    random.seed(6)
    np.random.seed(6)
    E = device_engine(s, [PairCorrelationConstraint(pcf_exp, dataWeights=dw, scaleFactor=1.1),
                          ReducedStructureFactorConstraint(rsfc_exp, rmin=0.0, rmax=8.0, dr=0.02,
                                                           scaleFactor=0.9)])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    if runner == "native":
        E.run_device(nsteps, k=16, log=dlog)
    else:
        E.run(nsteps, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    assert 0 < sum(a for _, a, _ in dlog) < nsteps


def test_all_four_constraint_kinds_share_grids():
    """PDC and PCF on one grid, SFC and reduced SFC on another: the device
    engine shares each grid's delta block between its two constraints."""
    s = systems.b2_alloy(cells=(5, 5, 5), a=3.01, sigma=0.08, seed=9)
    x = (0.01 + 0.02 * np.arange(300)).astype(F32)
    q = np.linspace(0.5, 20, 100).astype(F32)
    pdc_exp = _target(s, 0.08, 10, lambda o, w: oeng.PDC(o, np.stack([x, x], 1), w))
    pcf_exp = _target(s, 0.08, 10, lambda o, w: oeng.PCF(o, np.stack([x, x], 1), w))
    sfc_exp = _target(s, 0.08, 10, lambda o, w: oeng.SFC(o, np.stack([q, q], 1), w, 0.0, 6.0, 0.02))
    rsfc_exp = _target(s, 0.08, 10, lambda o, w: oeng.RSFC(o, np.stack([q, q], 1), w, 0.0, 6.0, 0.02))
    n = len(s["real"])

    def cons(osys, w):
        return [oeng.PDC(osys, pdc_exp, w), oeng.PCF(osys, pcf_exp, w),
                oeng.SFC(osys, sfc_exp, w, 0.0, 6.0, 0.02), oeng.RSFC(osys, rsfc_exp, w, 0.0, 6.0, 0.02)]

    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.1), cons, 250, 12)
    random.seed(12)
    np.random.seed(12)
    E = device_engine(s, [PairDistributionConstraint(pdc_exp), PairCorrelationConstraint(pcf_exp),
                          StructureFactorConstraint(sfc_exp, rmin=0.0, rmax=6.0, dr=0.02),
                          ReducedStructureFactorConstraint(rsfc_exp, rmin=0.0, rmax=6.0, dr=0.02)])
    E.set_move_generator(TranslationGenerator(0.1))
    dlog = []
    E.run_device(250, k=32, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))


@pytest.mark.parametrize("flexible", [True, False])
def test_intermolecular_distance_constraint_vetoes_in_dropin_loop(small_triclinic, flexible):
    """InterMolecularDistanceConstraint (rigid) with a PDC, through the
    per-move drop-in loop: the same vetoed, rejected and accepted moves as
    the oracle, identical distance counts and sums, standard errors."""
    s = small_triclinic
    exp = pdc_target(s, 0.1, 31, 0.025, 0.05, 200)
    n = len(s["real"])
    limits = {("C", "H"): 2.2, ("H", "H"): 2.4}
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.4), _pdc_cons_of({"exp": exp}),
                                    120, 17,
                                    rigid_fn=lambda o: [oeng.IMD(o, 2.6, limits, flexible=flexible)])
    random.seed(17)
    np.random.seed(17)
    E = device_engine(s, [PairDistributionConstraint(exp),
                          InterMolecularDistanceConstraint(2.6, pairsDistanceDefinition=[("C", "H", 2.2),
                                                                                         ("H", "H", 2.4)],
                                                           flexible=flexible)])
    E.set_move_generator(TranslationGenerator(0.4))
    dlog = []
    E.run(120, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    c, oc = E.rigidConstraints[0], oe.rigid[0]
    assert c.data["number"].tobytes() == oc.number.tobytes()
    assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert np.float32(c.standardError).tobytes() == np.float32(oc.sd).tobytes()
    assert E.tried < E.generated  # some moves were vetoed before being tried
    assert oc.sd > 0


@pytest.mark.parametrize("kind,flexible,rp", [("inter", True, 1.0), ("inter", True, 0.5), ("inter", False, 1.0),
                                               ("both", True, 0.5), ("both", False, 1.0)])
def test_distance_constraints_veto_on_the_device_engine(small_triclinic, kind, flexible, rp):
    """Rigid distance constraints on the device engine: the veto is decided
    in Python as each window is drawn (a flexible constraint whose error
    grows draws once, and with rejectProbability 0.5 half of those moves go
    on to the accept rule and its own draw; a strict one draws nothing), the
    experimental constraints are evaluated on the device.  The trajectory,
    streams and final state equal the oracle's, and the counters equal the
    per-move drop-in loop's."""
    s = small_triclinic
    exp = pdc_target(s, 0.1, 31, 0.025, 0.05, 200)
    n = len(s["real"])
    inter_lim = {("C", "H"): 2.2, ("H", "H"): 2.4}
    intra_lim = {("H", "H"): 1.7, ("C", "H"): 1.05}

    def rigid_oracle(o):
        r = [oeng.IMD(o, 2.6, inter_lim, flexible=flexible, rejectProbability=rp)]
        if kind == "both":
            r.append(oeng.IMD(o, 1.0, intra_lim, flexible=flexible, rejectProbability=rp, inter=False))
        return r

    def rigid_engine():
        r = [InterMolecularDistanceConstraint(2.6, pairsDistanceDefinition=[(a, b, d) for (a, b), d in inter_lim.items()],
                                              flexible=flexible, rejectProbability=rp)]
        if kind == "both":
            r.append(IntraMolecularDistanceConstraint(1.0, typeDefinition="element",
                                                      pairsDistanceDefinition=[(a, b, d) for (a, b), d in
                                                                               intra_lim.items()],
                                                      flexible=flexible, rejectProbability=rp))
        return r

    nsteps = 300
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.4), _pdc_cons_of({"exp": exp}),
                                    nsteps, 23, rigid_fn=rigid_oracle)
    random.seed(23)
    np.random.seed(23)
    E = device_engine(s, [PairDistributionConstraint(exp)] + rigid_engine())
    E.set_move_generator(TranslationGenerator(0.4))
    dlog = []
    st = E.run_device(nsteps, k=16, log=dlog)
    assert st["windows"] < nsteps
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    for c, oc in zip(E.rigidConstraints, oe.rigid):
        assert c.data["number"].tobytes() == oc.number.tobytes()
        assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
        assert np.float32(c.standardError).tobytes() == np.float32(oc.sd).tobytes()
    # the per-move drop-in loop from the same start: the same counters
    random.seed(23)
    np.random.seed(23)
    P = device_engine(s, [PairDistributionConstraint(exp)] + rigid_engine())
    P.set_move_generator(TranslationGenerator(0.4))
    plog = []
    P.run(nsteps, log=plog)
    assert plog == dlog
    assert (E.generated, E.tried, E.accepted, E.tolerated) == (P.generated, P.tried, P.accepted, P.tolerated)
    assert [(c.tried, c.accepted) for c in E.rigidConstraints] == [(c.tried, c.accepted) for c in P.rigidConstraints]
    assert E.tried < E.generated  # some moves were vetoed before being tried
    assert E.accepted > 10


@pytest.mark.parametrize("flexible", [True, False])
def test_intramolecular_distance_constraint_vetoes_in_dropin_loop(small_triclinic, flexible):
    """IntraMolecularDistanceConstraint (rigid, element types) with a PDC,
    through the per-move drop-in loop: single-atom moves inside the rigid
    rings bring H-H and C-H pairs of one molecule under their limits, and the
    constraint vetoes exactly the moves the oracle vetoes (reference
    DistanceConstraints.py:1093-1210)."""
    s = small_triclinic
    exp = pdc_target(s, 0.1, 33, 0.025, 0.05, 200)
    n = len(s["real"])
    limits = {("H", "H"): 1.7, ("C", "H"): 1.05}
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], oeng.Translation(0.4), _pdc_cons_of({"exp": exp}),
                                    120, 19,
                                    rigid_fn=lambda o: [oeng.IMD(o, 1.0, limits, flexible=flexible, inter=False)])
    random.seed(19)
    np.random.seed(19)
    E = device_engine(s, [PairDistributionConstraint(exp),
                          IntraMolecularDistanceConstraint(1.0, typeDefinition="element",
                                                           pairsDistanceDefinition=[("H", "H", 1.7),
                                                                                    ("C", "H", 1.05)],
                                                           flexible=flexible)])
    E.set_move_generator(TranslationGenerator(0.4))
    dlog = []
    E.run(120, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    c, oc = E.rigidConstraints[0], oe.rigid[0]
    assert c.data["number"].tobytes() == oc.number.tobytes()
    assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert np.float32(c.standardError).tobytes() == np.float32(oc.sd).tobytes()
    assert E.tried < E.generated  # some moves were vetoed before being tried


def test_swap_moves_in_dropin_loop():
    """SwapPositionsGenerator on the Fe atoms of a B2 alloy (partners drawn
    from the Al atoms), translations on the Al atoms: swaps move two atoms
    of different elements, so the pair-type slots change."""
    s = systems.b2_alloy(cells=(5, 5, 5), a=3.01, sigma=0.08, seed=13)
    exp = pdc_target(s, 0.08, 14, 0.01, 0.02, 300)
    n = len(s["real"])
    el = s["elements"]
    al = [[j] for j in range(n) if el[j] == "Al"]
    gp = [SwapPositionsGenerator(swapList=al) if el[i] == "Fe" else TranslationGenerator(0.1) for i in range(n)]
    go = [oeng.Swap(al) if el[i] == "Fe" else oeng.Translation(0.1) for i in range(n)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], go, _pdc_cons_of({"exp": exp}), 200, 21,
                                    tolerance=5.0)
    random.seed(21)
    np.random.seed(21)
    E = device_engine(s, [PairDistributionConstraint(exp)], groups=[[i] for i in range(n)], gen=gp)
    E.set_tolerance(5.0)
    dlog = []
    E.run(200, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    swaps = [a for g, a, _ in dlog if el[g] == "Fe"]
    assert len(swaps) > 20 and any(swaps)


@pytest.mark.parametrize("cells", [0, 1])
def test_swap_moves_on_the_device_engine(cells):
    """The same swap and translation moves in speculative windows on the
    device engine: each proposal names its moved atoms (a swap's group plus
    its drawn partner), evaluated and committed there.  cells=1 keeps k_delta's
    sorted atom copy (kept by the commits; the windows of atom lists take the
    plain k_delta)."""
    from fullrmc_amd import _lib
    s = systems.b2_alloy(cells=(5, 5, 5), a=3.01, sigma=0.08, seed=13)
    exp = pdc_target(s, 0.08, 14, 0.01, 0.02, 300)
    n = len(s["real"])
    el = s["elements"]
    al = [[j] for j in range(n) if el[j] == "Al"]
    gp = [SwapPositionsGenerator(swapList=al) if el[i] == "Fe" else TranslationGenerator(0.1) for i in range(n)]
    go = [oeng.Swap(al) if el[i] == "Fe" else oeng.Translation(0.1) for i in range(n)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], go, _pdc_cons_of({"exp": exp}), 600, 22,
                                    tolerance=5.0)
    L = _lib.lib()
    _lib.check(L.frmc_set_delta_cells(cells))
    try:
        random.seed(22)
        np.random.seed(22)
        E = device_engine(s, [PairDistributionConstraint(exp)], groups=[[i] for i in range(n)], gen=gp)
        E.set_tolerance(5.0)
        dlog = []
        st = E.run_device(600, k=16, log=dlog)
    finally:
        _lib.check(L.frmc_set_delta_cells(-1))
    assert st["windows"] < 600
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    swaps = [a for g, a, _ in dlog if el[g] == "Fe"]
    assert len(swaps) > 50 and sum(swaps) > 3


def test_swap_moves_with_a_distance_constraint_on_the_device_engine():
    """Swaps and translations of the B2 alloy with an inter-molecular
    distance constraint (every atom its own molecule, 2.45 A, strict): the veto
    evaluates the group and its drawn partner, a vetoed swap makes no
    acceptance draw, and the trajectory, streams and final state equal the
    oracle's."""
    s = systems.b2_alloy(cells=(5, 5, 5), a=3.01, sigma=0.08, seed=13)
    exp = pdc_target(s, 0.08, 14, 0.01, 0.02, 300)
    n = len(s["real"])
    el = s["elements"]
    al = [[j] for j in range(n) if el[j] == "Al"]
    gp = [SwapPositionsGenerator(swapList=al) if el[i] == "Fe" else TranslationGenerator(0.3) for i in range(n)]
    go = [oeng.Swap(al) if el[i] == "Fe" else oeng.Translation(0.3) for i in range(n)]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(n)], go, _pdc_cons_of({"exp": exp}), 300, 24,
                                    tolerance=5.0, rigid_fn=lambda o: [oeng.IMD(o, 2.45, flexible=False)])
    random.seed(24)
    np.random.seed(24)
    E = device_engine(s, [PairDistributionConstraint(exp), InterMolecularDistanceConstraint(2.45, flexible=False)],
                      groups=[[i] for i in range(n)], gen=gp)
    E.set_tolerance(5.0)
    dlog = []
    E.run_device(300, k=16, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    c, oc = E.rigidConstraints[0], oe.rigid[0]
    assert c.data["number"].tobytes() == oc.number.tobytes()
    assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert E.tried < E.generated
    swaps = [a for g, a, _ in dlog if el[g] == "Fe"]
    assert len(swaps) > 20 and any(swaps)


def test_swaps_removals_and_a_distance_constraint_together_on_the_device_engine():
    """Swaps, translations and removals of the B2 alloy with a strict
    distance constraint, all on the device engine: atom-list windows, vetoed
    proposals without draws, removals (one rejected, eight accepted, each
    rebuilding the engine and the constraint's resident copies); the same
    trajectory, streams and final state as the oracle."""
    s = systems.b2_alloy(cells=(5, 5, 5), a=3.01, sigma=0.08, seed=13)
    exp = pdc_target(s, 0.08, 14, 0.01, 0.02, 300)
    n = len(s["real"])
    el = s["elements"]
    al = [j for j in range(n) if el[j] == "Al"]
    elw = {e: WEIGHTS[e] for e in sorted(set(el))}
    groups = [[i] for i in range(n)] + [[0]] * 40
    gp = ([SwapPositionsGenerator(swapList=[[j] for j in al]) if el[i] == "Fe" else TranslationGenerator(0.3)
           for i in range(n)] + [AtomsRemoveGenerator(maximumCollected=8, atomsList=al)] * 40)
    go = ([oeng.Swap([[j] for j in al]) if el[i] == "Fe" else oeng.Translation(0.3) for i in range(n)]
          + [oeng.Remove(al, 8)] * 40)
    oe, olog, pys, nps = oracle_run(s, groups, go, lambda osys, w: [oeng.PDC(osys, exp, w, elementWeights=elw)],
                                    300, 25, tolerance=5.0, rigid_fn=lambda o: [oeng.IMD(o, 2.45, flexible=False)])
    random.seed(25)
    np.random.seed(25)
    E = device_engine(s, [PairDistributionConstraint(exp), InterMolecularDistanceConstraint(2.45, flexible=False)],
                      groups=groups, gen=gp)
    E.set_tolerance(5.0)
    dlog = []
    E.run_device(300, k=16, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    c, oc = E.rigidConstraints[0], oe.rigid[0]
    assert c.data["number"].tobytes() == oc.number.tobytes()
    assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert sorted(E.collectedAtoms) == sorted(oe.s.collected) and len(E.collectedAtoms) == 8
    assert E.removed[0] > E.removed[1] == 8
    swaps = [a for g, a, _ in dlog if g < n and el[g] == "Fe"]
    assert sum(swaps) > 5 and E.tried < E.generated


@pytest.mark.parametrize("device", [False, True])
def test_remove_moves_amputation(device):
    """AtomsRemoveGenerator groups beside translations on a B2 alloy with a
    G(r), an S(Q) and a distance constraint: each removal is evaluated as if
    amputated (atom counts, weights, number density without the atom) and,
    accepted, takes the atom out of every array; later moves run on the
    smaller system.  Same trajectory as the oracle, through the per-move
    drop-in loop and on the device engine (a removal ends its window, is
    evaluated on the host, and an accepted one rebuilds the device engine on
    the remaining atoms; the run is split in two calls, so the second starts
    with atoms already removed)."""
    s = systems.b2_alloy(cells=(5, 5, 5), a=3.01, sigma=0.08, seed=17)
    pdc_exp = pdc_target(s, 0.08, 18, 0.01, 0.02, 300)
    q = np.linspace(0.5, 20, 80).astype(F32)
    sfc_exp = sfc_target(s, 0.08, 18, q, 0.0, 6.0, 0.02)
    n = len(s["real"])
    el = s["elements"]
    al = [j for j in range(n) if el[j] == "Al"]
    groups = [[i] for i in range(n)] + [[0]] * 60
    elw = {e: WEIGHTS[e] for e in sorted(set(el))}
    gp = [TranslationGenerator(0.1)] * n + [AtomsRemoveGenerator(maximumCollected=12, atomsList=al)] * 60
    go = [oeng.Translation(0.1)] * n + [oeng.Remove(al, 12)] * 60

    def cons(osys, w):
        return [oeng.PDC(osys, pdc_exp, w, elementWeights=elw),
                oeng.SFC(osys, sfc_exp, w, 0.0, 6.0, 0.02, elementWeights=elw)]

    oe, olog, pys, nps = oracle_run(s, groups, go, cons, 300, 23, tolerance=20.0,
                                    rigid_fn=lambda o: [oeng.IMD(o, 2.3)])
    random.seed(23)
    np.random.seed(23)
    E = device_engine(s, [PairDistributionConstraint(pdc_exp),
                          StructureFactorConstraint(sfc_exp, rmin=0.0, rmax=6.0, dr=0.02),
                          InterMolecularDistanceConstraint(2.3)], groups=groups, gen=gp)
    E.set_tolerance(20.0)
    dlog = []
    if device:
        E.run_device(150, k=16, log=dlog)
        assert 0 < len(E.collectedAtoms)
        E.run_device(150, k=16, log=dlog)
    else:
        E.run(300, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    c, oc = E.rigidConstraints[0], oe.rigid[0]
    assert c.data["number"].tobytes() == oc.number.tobytes()
    assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert 0 < len(E.collectedAtoms) <= 12 and E.numberOfAtoms == n - len(E.collectedAtoms)
    assert sorted(E.collectedAtoms) == sorted(oe.s.collected)
    assert E.removed[0] >= E.removed[1] == len(E.collectedAtoms)


def test_molecular_triclinic_translation_plus_rotation(small_triclinic):
    s = small_triclinic
    exp = pdc_target(s, 0.05, 8, 0.025, 0.05, 140)
    mols = {}
    for i, m in enumerate(s["molecules"]):
        mols.setdefault(int(m), []).append(i)
    groups = list(mols.values())
    ogen = oeng.Combined([oeng.Translation(0.2), oeng.Rotation(10)])
    oe, olog, pys, nps = oracle_run(s, groups, ogen, lambda osys, w: [oeng.PDC(osys, exp, w)], 80, 6)
    random.seed(6)
    np.random.seed(6)
    gen = MoveGeneratorCombinator([TranslationGenerator(0.2), RotationGenerator(10)])
    E = device_engine(s, [PairDistributionConstraint(exp)], groups=groups, gen=gen)
    dlog = []
    E.run_device(80, k=8, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))


@pytest.mark.parametrize("flexible,rp", [(True, 0.5), (False, 1.0)])
def test_molecular_moves_with_a_distance_constraint_on_the_device_engine(small_triclinic, flexible, rp):
    """Whole-molecule translation + rotation windows with an inter-molecular
    distance constraint on the device engine: every veto evaluates the
    13-atom group against all atoms and the group alone on the resident
    copies; trajectory, streams, state and the constraint's data equal the
    oracle's."""
    s = small_triclinic
    exp = pdc_target(s, 0.05, 8, 0.025, 0.05, 140)
    mols = {}
    for i, m in enumerate(s["molecules"]):
        mols.setdefault(int(m), []).append(i)
    groups = list(mols.values())
    lim = {("C", "H"): 2.2, ("H", "H"): 2.4}
    ogen = oeng.Combined([oeng.Translation(0.4), oeng.Rotation(20)])
    oe, olog, pys, nps = oracle_run(s, groups, ogen, lambda osys, w: [oeng.PDC(osys, exp, w)], 150, 9,
                                    rigid_fn=lambda o: [oeng.IMD(o, 2.6, lim, flexible=flexible, rejectProbability=rp)])
    random.seed(9)
    np.random.seed(9)
    gen = MoveGeneratorCombinator([TranslationGenerator(0.4), RotationGenerator(20)])
    imd = InterMolecularDistanceConstraint(2.6, pairsDistanceDefinition=[(a, b, d) for (a, b), d in lim.items()],
                                           flexible=flexible, rejectProbability=rp)
    E = device_engine(s, [PairDistributionConstraint(exp), imd], groups=groups, gen=gen)
    dlog = []
    E.run_device(150, k=8, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
    oc = oe.rigid[0]
    assert imd.data["number"].tobytes() == oc.number.tobytes()
    assert imd.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert np.float32(imd.standardError).tobytes() == np.float32(oc.sd).tobytes()
    assert E.tried < E.generated and E.accepted > 5


def test_b2_13500_size_properties():
    """B2 alloy at full size (13,500 atoms, PDC + SFC): after a device run the
    engine's histograms equal a fresh full recompute in the symmetrised sums
    and its total equals the chain evaluated on that recompute."""
    s = systems.b2_alloy(seed=7)
    pdc_exp = np.stack([(0.01 + 0.02 * np.arange(1000)).astype(F32), np.zeros(1000, F32)], 1)
    q = np.linspace(0.5, 25, 500).astype(F32)
    sfc_exp = np.stack([q, np.ones_like(q)], 1).astype(F32)
    random.seed(1)
    np.random.seed(1)
    E = device_engine(s, [PairDistributionConstraint(pdc_exp),
                          StructureFactorConstraint(sfc_exp, rmin=0.0, rmax=20.0, dr=0.02)])
    E.set_move_generator(TranslationGenerator(0.1))
    st = E.run_device(2000, k=64)
    assert st["generated"] == 2000
    sym = lambda h: h + h.transpose(1, 0, 2)
    chis = []
    for c in E.constraints:
        fi, fe = c._full()
        assert np.array_equal(sym(c.data["intra"] + c.data["inter"]), sym(fi + fe))
        chis.append(c.compute_standard_error(c._model({"intra": fi, "inter": fe})))
        assert np.float32(chis[-1]).tobytes() == np.float32(c.standardError).tobytes()
    assert np.float32(np.sum(chis)).tobytes() == np.float32(E.totalStandardError).tobytes()


@pytest.mark.parametrize("batched", [True, False])
def test_replicas_run_concurrently_and_match_independent_runs(batched):
    """Three replicas of different sizes advanced together (batched: all
    windows of a step in one launch per kernel; else every engine's windows
    in flight on its own stream) end in exactly the state of three
    independent single-engine runs with the same seeds, and replica 0 in the
    oracle's state."""
    from fullrmc_amd.Engine import run_replicas
    seeds, nsteps = [11, 12, 13], 150
    boxes = []
    for i, n in enumerate((300, 500, 420)):
        L = (n / 0.05) ** (1 / 3)
        s = systems.random_box(n, L, 2.0, seed=40 + i)
        s["exp"] = pdc_target(s, 0.1, 50 + i, 0.025, 0.05, 150)
        boxes.append(s)

    def engine(s):
        E = device_engine(s, [PairDistributionConstraint(s["exp"])])
        E.set_move_generator(TranslationGenerator(0.2))
        return E

    reps = [engine(s) for s in boxes]
    stats = run_replicas(reps, nsteps, k=16, seeds=seeds, batched=batched)
    assert all(st["generated"] == nsteps for st in stats)
    for E, s, seed in zip(reps, boxes, seeds):
        E.sync_from_device()
        random.seed(seed)
        np.random.seed(seed)
        S = engine(s)
        S.run_device(nsteps, k=16)
        assert E.realCoordinates.tobytes() == S.realCoordinates.tobytes()
        assert E.boxCoordinates.tobytes() == S.boxCoordinates.tobytes()
        assert np.float32(E.totalStandardError).tobytes() == np.float32(S.totalStandardError).tobytes()
        assert E.accepted == S.accepted
        for c, cs in zip(E.constraints, S.constraints):
            assert c.data["intra"].tobytes() == cs.data["intra"].tobytes()
            assert c.data["inter"].tobytes() == cs.data["inter"].tobytes()
        assert E._replica_rng[0] == random.getstate()
        assert np.array_equal(E._replica_rng[1][1], np.random.get_state()[1])
    s = boxes[0]
    oe, olog, pys, nps = oracle_run(s, [[i] for i in range(len(s["real"]))], oeng.Translation(0.2), _pdc_cons_of(s),
                                    nsteps, seeds[0])
    assert reps[0].realCoordinates.tobytes() == oe.s.real.tobytes()
    assert np.float32(reps[0].totalStandardError).tobytes() == np.float32(oe.tse).tobytes()


def _pdc_cons_of(s):
    return lambda osys, w: [oeng.PDC(osys, s["exp"], w)]


def test_48_replicas_of_1000_to_20000_atoms_match_independent_runs():
    """48 replicas shaped like the 256 of the replica workload (N log-uniform
    in [1000, 20000], density 0.05 per A^3, 2.0 A minimum distance, targets
    from 0.1 A jittered copies on r in [0,15) at 0.05 A, k = 32), batched:
    each ends bit-identical to its own single-engine run with the same seed."""
    from fullrmc_amd.Engine import run_replicas
    sizes = systems.replica_sizes(48, 1000, 20000, 77)
    boxes = []
    for r, n in enumerate(sizes):
        L = float((n / 0.05) ** (1.0 / 3.0))
        b = systems.random_box(int(n), L, 2.0, 300 + r)
        b["exp"] = pdc_target(b, 0.1, 400 + r, 0.025, 0.05, 300)
        boxes.append(b)

    def engine(b):
        E = device_engine(b, [PairDistributionConstraint(b["exp"])])
        E.set_move_generator(TranslationGenerator(0.1))
        return E

    reps = [engine(b) for b in boxes]
    seeds, nsteps = [500 + r for r in range(len(reps))], 150
    stats = run_replicas(reps, nsteps, k=32, seeds=seeds)
    assert all(st["generated"] == nsteps for st in stats)
    for E, b, seed in zip(reps, boxes, seeds):
        E.sync_from_device()
        random.seed(seed)
        np.random.seed(seed)
        S = engine(b)
        S.run_device(nsteps, k=32)
        assert E.realCoordinates.tobytes() == S.realCoordinates.tobytes()
        assert np.float32(E.totalStandardError).tobytes() == np.float32(S.totalStandardError).tobytes()
        assert E.accepted == S.accepted
        c, cs = E.constraints[0], S.constraints[0]
        assert c.data["inter"].tobytes() == cs.data["inter"].tobytes()


def test_batched_replicas_with_tolerance_match_independent_runs():
    """A tolerance above zero makes the accept rule draw for worse moves and
    keep some of them: the batched walk rewinds both streams past those
    draws exactly as the single-engine loop does."""
    from fullrmc_amd.Engine import run_replicas
    boxes = []
    for i, n in enumerate((350, 900, 600)):
        L = (n / 0.05) ** (1 / 3)
        b = systems.random_box(n, L, 2.0, seed=600 + i)
        b["exp"] = pdc_target(b, 0.1, 610 + i, 0.025, 0.05, 150)
        boxes.append(b)

    def engine(b):
        E = device_engine(b, [PairDistributionConstraint(b["exp"])])
        E.set_move_generator(TranslationGenerator(0.2))
        E.set_tolerance(25.0)
        return E

    reps = [engine(b) for b in boxes]
    seeds, nsteps = [620, 621, 622], 300
    stats = run_replicas(reps, nsteps, k=16, seeds=seeds)
    assert all(st["generated"] == nsteps for st in stats)
    tolerated = 0
    for E, b, seed in zip(reps, boxes, seeds):
        E.sync_from_device()
        tolerated += E.tolerated
        random.seed(seed)
        np.random.seed(seed)
        S = engine(b)
        S.run_device(nsteps, k=16)
        assert E.realCoordinates.tobytes() == S.realCoordinates.tobytes()
        assert np.float32(E.totalStandardError).tobytes() == np.float32(S.totalStandardError).tobytes()
        assert (E.accepted, E.tolerated) == (S.accepted, S.tolerated)
        assert E._replica_rng[0] == random.getstate()
        assert np.array_equal(E._replica_rng[1][1], np.random.get_state()[1])
    assert tolerated > 0


def test_batched_replicas_of_mixed_cells_and_constraints(small_triclinic):
    """One batched run over replicas that differ in everything a window
    kernel reads: atom count, cubic or triclinic cell, one or two elements,
    G(r) alone or with S(Q) (two grids), window sizes.  Each ends
    bit-identical to its own single-engine run with the same seed."""
    from fullrmc_amd.Engine import run_replicas
    q = np.linspace(0.5, 12, 90).astype(F32)
    specs = []
    for i, n in enumerate((260, 700, 90, 1500)):
        L = (n / 0.05) ** (1 / 3)
        s = systems.random_box(n, L, 2.0, seed=70 + i, species=("Si", "O") if i % 2 else ("Ar",),
                               fractions=(1 / 3, 2 / 3) if i % 2 else (1.0,))
        s["exp"] = pdc_target(s, 0.1, 80 + i, 0.025, 0.05, 140)
        specs.append((s, i == 1))
    t = dict(small_triclinic)
    t["exp"] = pdc_target(t, 0.1, 90, 0.025, 0.05, 150)
    specs.append((t, True))
    seeds, nsteps = [21, 22, 23, 24, 25], 200

    def engine(s, with_sq):
        cons = [PairDistributionConstraint(s["exp"])]
        if with_sq:
            cons.append(StructureFactorConstraint(np.stack([q, np.ones_like(q)], 1).astype(F32), rmin=0.0, rmax=7.0,
                                                  dr=0.05))
        E = device_engine(s, cons)
        E.set_move_generator(TranslationGenerator(0.2))
        return E

    reps = [engine(s, w) for s, w in specs]
    stats = run_replicas(reps, nsteps, k=24, seeds=seeds)
    assert all(st["generated"] == nsteps for st in stats)
    assert sum(st["accepted"] for st in stats) > 0
    for E, (s, w), seed in zip(reps, specs, seeds):
        E.sync_from_device()
        random.seed(seed)
        np.random.seed(seed)
        S = engine(s, w)
        S.run_device(nsteps, k=24)
        assert E.realCoordinates.tobytes() == S.realCoordinates.tobytes()
        assert np.float32(E.totalStandardError).tobytes() == np.float32(S.totalStandardError).tobytes()
        assert E.accepted == S.accepted
        for c, cs in zip(E.constraints, S.constraints):
            assert c.data["intra"].tobytes() == cs.data["intra"].tobytes()
            assert c.data["inter"].tobytes() == cs.data["inter"].tobytes()


def test_molecular_triclinic_full_size():
    """500 rigid 13-atom rings (6,500 atoms) in the 42/44/46 A, 80/95/105 deg
    cell, molecule translation + rotation through the Python generators:
    the device trajectory equals the oracle's."""
    s = systems.ring_box(nmol=500, cell=(42.0, 44.0, 46.0, 80.0, 95.0, 105.0), seed=61, min_centre=4.5)
    s["elem"] = element_index(s["elements"])
    exp = pdc_target(s, 0.05, 62, 0.025, 0.05, 300)
    mols = {}
    for i, m in enumerate(s["molecules"]):
        mols.setdefault(int(m), []).append(i)
    groups = list(mols.values())
    ogen = oeng.Combined([oeng.Translation(0.2), oeng.Rotation(10)])
    oe, olog, pys, nps = oracle_run(s, groups, ogen, lambda osys, w: [oeng.PDC(osys, exp, w)], 40, 63)
    random.seed(63)
    np.random.seed(63)
    gen = MoveGeneratorCombinator([TranslationGenerator(0.2), RotationGenerator(10)])
    E = device_engine(s, [PairDistributionConstraint(exp)], groups=groups, gen=gen)
    dlog = []
    E.run_device(40, k=8, log=dlog)
    assert_same_state(E, oe, dlog, olog, (pys, nps))
