"""Realistic parity cases from the reference's own example inputs
(Examples/atomicNiTi, Examples/removes, copied to tests/golden/examples):
structures read from PDB, experimental curves read from files, the NiTi
example's constraint set (G(r), reduced S(Q) from rebinned F(Q), an
inter-molecular distance constraint).  The engine's trajectories equal the
oracle engine's move for move, through the per-move drop-in loop and the
device-resident engine."""
import os
import random

import numpy as np
import pytest

from fullrmc_amd.Constraints import (InterMolecularDistanceConstraint, PairDistributionConstraint,
                                     ReducedStructureFactorConstraint)
from fullrmc_amd.Core.Collection import rebin
from fullrmc_amd.Engine import Engine
from fullrmc_amd.Generators import TranslationGenerator
from oracle import chain
from oracle import engine as oeng

pytestmark = pytest.mark.gpu
F32 = np.float32
EX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples")
Z = {"Ni": 28, "Ti": 22, "Co": 27, "Mn": 25, "Li": 3, "O": 8}


def _z(e):
    return Z[e.capitalize()]


def _same_state(E, oe, dlog, olog):
    assert len(dlog) == len(olog)
    first = next((i for i, (a, b) in enumerate(zip(dlog, olog)) if a != b), None)
    assert first is None, "trajectories diverge at step %d: %r vs %r" % (first, dlog[first], olog[first])
    assert np.float32(E.totalStandardError).tobytes() == np.float32(oe.tse).tobytes()
    for c, oc in zip(E.constraints, oe.constraints):
        assert c.data["intra"].tobytes() == oc.intra.tobytes()
        assert c.data["inter"].tobytes() == oc.inter.tobytes()
    for c, oc in zip(E.rigidConstraints, oe.rigid):
        assert c.data["number"].tobytes() == oc.number.tobytes()
        assert c.data["distanceSum"].tobytes() == oc.dsum.tobytes()
    assert E.realCoordinates.tobytes() == oe.s.real.tobytes()


def _oracle(E, cons_fn, rigid_fn, amp, nsteps, seed):
    osys = oeng.System(E.realCoordinates, E.allElements, E.moleculesIndex, E.basisVectors,
                       E.reciprocalBasisVectors)
    w = chain.normalized_weighting(osys.counts, {e: _z(e) for e in osys.elements})
    random.seed(seed)
    np.random.seed(seed)
    oe = oeng.Engine(osys, [[i] for i in range(len(E.realCoordinates))], oeng.Translation(amp),
                     cons_fn(osys, w), 0.0, rigid=rigid_fn(osys))
    oe.initialize()
    log = []
    oe.run(nsteps, log)
    return oe, log


@pytest.fixture(scope="module")
def niti():
    gr = os.path.join(EX, "niti_experimental.gr.gz")
    Sq = np.transpose(rebin(np.loadtxt(os.path.join(EX, "niti_experimental.fq.gz")), bin=0.05)).astype(F32)
    return gr, Sq


def _niti_engine(gr, Sq, with_imd):
    E = Engine()
    E.set_pdb(os.path.join(EX, "niti_system.pdb.gz"))
    cons = [PairDistributionConstraint(experimentalData=gr, weighting="atomicNumber"),
            ReducedStructureFactorConstraint(experimentalData=Sq, weighting="atomicNumber")]
    if with_imd:
        cons.append(InterMolecularDistanceConstraint(defaultDistance=2.2, flexible=True))
    E.add_constraints(cons)
    return E


def _niti_oracle_cons(gr, Sq):
    rmin, rmax, dr = chain.sfc_defaults(Sq[:, 0])
    exp = np.loadtxt(gr, dtype=F32)
    return lambda o, w: [oeng.PDC(o, exp, w), oeng.RSFC(o, Sq, w, rmin, rmax, dr)]


def test_niti_example_dropin_loop(niti):
    gr, Sq = niti
    E = _niti_engine(gr, Sq, with_imd=True)
    oe, olog = _oracle(E, _niti_oracle_cons(gr, Sq), lambda o: [oeng.IMD(o, 2.2)], 0.2, 60, 41)
    random.seed(41)
    np.random.seed(41)
    dlog = []
    E.run(60, log=dlog)
    _same_state(E, oe, dlog, olog)
    assert E.accepted > 0  # the ideal B2 start improves under almost every move


def test_niti_example_device_engine(niti):
    gr, Sq = niti
    E = _niti_engine(gr, Sq, with_imd=False)
    E.set_move_generator(TranslationGenerator(0.1))
    oe, olog = _oracle(E, _niti_oracle_cons(gr, Sq), lambda o: [], 0.1, 400, 43)
    random.seed(43)
    np.random.seed(43)
    dlog = []
    E.run_device(400, k=16, log=dlog)
    _same_state(E, oe, dlog, olog)
    assert E.accepted > 0


@pytest.mark.parametrize("device", [False, True])
def test_removes_triclinic_five_elements(device):
    """Examples/removes: 8,100 atoms of five elements in a triclinic cell,
    G(r) from pdf.exp (first point at r = 0), with the inter-molecular
    distance constraint in the drop-in loop."""
    E = Engine()
    E.set_pdb(os.path.join(EX, "removes_system.pdb.gz"))
    exp_path = os.path.join(EX, "removes_pdf.exp.gz")
    cons = [PairDistributionConstraint(experimentalData=exp_path, weighting="atomicNumber")]
    if not device:
        cons.append(InterMolecularDistanceConstraint())
    E.add_constraints(cons)
    E.set_move_generator(TranslationGenerator(0.05))
    exp = np.loadtxt(exp_path, dtype=F32)
    nsteps = 300 if device else 50
    oe, olog = _oracle(E, lambda o, w: [oeng.PDC(o, exp, w)],
                       (lambda o: []) if device else (lambda o: [oeng.IMD(o, 1.5)]), 0.05, nsteps, 47)
    random.seed(47)
    np.random.seed(47)
    dlog = []
    if device:
        E.run_device(nsteps, k=16, log=dlog)
    else:
        E.run(nsteps, log=dlog)
    _same_state(E, oe, dlog, olog)
