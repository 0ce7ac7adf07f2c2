"""The committed golden vectors (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the oracle): the oracle must still produce
them bit for bit (CPU), and the HIP path must match them (GPU)."""
import os
import random

import numpy as np
import pytest

from oracle import chain, refcpu
from oracle import engine as oeng

F32 = np.float32
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden.npz")


@pytest.fixture(scope="module")
def g():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


# ----------------------------------------------------------------- oracle

def test_oracle_distances(g):
    box, real, basis, recip = g["d_box"], g["d_real"], g["d_basis"], g["d_recip"]
    assert same(refcpu.transform_coordinates(recip, real), g["d_transform"])
    assert same(refcpu.pairs_distances_to_indexcoords(7, box, basis, True), g["d_pbc_atom7"])
    assert same(refcpu.pairs_distances_to_indexcoords(7, box, basis, True, allAtoms=False)[7:],
                g["d_pbc_atom7_upper"])
    assert same(refcpu.pairs_distances_to_indexcoords(7, real, basis, False), g["d_ibc_atom7"])
    assert same(refcpu.pairs_distances_to_multi_indexcoords(np.array([0, 5, 9], np.int32), box, basis, True),
                g["d_pbc_multi"])
    assert same(refcpu.pairs_differences_to_indexcoords(3, box, basis, True), g["d_pbc_diff3"])
    assert same(refcpu.pairs_distances_to_multi_points(g["d_points"], box, basis, True), g["d_pbc_multi_points"])
    rb, vol = refcpu.get_reciprocal_basis(basis)
    assert same(rb, g["d_recip_typo"]) and np.float64(vol) == g["d_recip_vol"]


def _hist_args(g):
    return (g["h_mol"], g["h_elem"], 3, F32(0.0), F32(7.0), F32(0.05), 140)


def test_oracle_histograms(g):
    hi, he = refcpu.full_pairs_histograms_coords(g["h_box"], g["h_basis"], True, *_hist_args(g))
    assert same(hi, g["h_full_intra"]) and same(he, g["h_full_inter"])
    mi, me = refcpu.multiple_pairs_histograms_coords(g["h_multi_idx"], g["h_box"], g["h_basis"], True,
                                                     *_hist_args(g))
    assert same(mi, g["h_multi_intra"]) and same(me, g["h_multi_inter"])
    # every unordered pair once: with 156 atoms and rmax 7 A well below half the
    # cell, the counts are those of pairs closer than 7 A
    assert hi.sum() + he.sum() > 0


def _weights(g):
    return {str(k): F32(v) for k, v in zip(g["r_weight_keys"], g["r_weights"])}


def test_oracle_run(g):
    elements = [str(e) for e in g["r_elements"]]
    random.seed(int(g["r_seed"]))
    np.random.seed(int(g["r_seed"]))
    osys = oeng.System(g["r_real"], elements, g["r_mol"], g["r_basis"], g["r_recip"])
    w = _weights(g)
    assert w == {k: F32(v) for k, v in chain.normalized_weighting(osys.counts, {"O": 8, "Si": 14}).items()}
    oe = oeng.Engine(osys, [[i] for i in range(len(elements))], oeng.Translation(0.2),
                     [oeng.PDC(osys, g["r_pdc_exp"], w), oeng.SFC(osys, g["r_sfc_exp"], w, 0.0, 6.0, 0.05)])
    oe.initialize()
    assert same(F32(oe.tse), g["r_tse0"])
    log = []
    oe.run(int(g["r_nsteps"]), log)
    assert same(np.array([a for a, _, _ in log], np.int32), g["r_log_group"])
    assert same(np.array([b for _, b, _ in log], bool), g["r_log_accepted"])
    assert same(np.array([t for _, _, t in log], F32), g["r_log_tse"])
    assert same(oe.s.real, g["r_final_real"]) and same(oe.s.box, g["r_final_box"])
    for i, c in enumerate(oe.constraints):
        assert same(c.intra, g["r_final_intra%d" % i]) and same(c.inter, g["r_final_inter%d" % i])


# ------------------------------------------------------------------ HIP path

@pytest.mark.gpu
def test_gpu_distances_match_golden(g):
    from fullrmc_amd.Core import boundary_conditions_collection as bc
    from fullrmc_amd.Core import pairs_distances as pd
    box, real, basis, recip = g["d_box"], g["d_real"], g["d_basis"], g["d_recip"]
    assert same(bc.transform_coordinates(recip, real), g["d_transform"])
    assert same(pd.pairs_distances_to_indexcoords(7, box, basis, True), g["d_pbc_atom7"])
    assert same(pd.pairs_distances_to_indexcoords(7, box, basis, True, allAtoms=False)[7:], g["d_pbc_atom7_upper"])
    assert same(pd.pairs_distances_to_indexcoords(7, real, basis, False), g["d_ibc_atom7"])
    assert same(pd.pairs_distances_to_multi_indexcoords(np.array([0, 5, 9], np.int32), box, basis, True),
                g["d_pbc_multi"])
    assert same(pd.pairs_differences_to_indexcoords(3, box, basis, True), g["d_pbc_diff3"])
    assert same(pd.pairs_distances_to_multi_points(g["d_points"], box, basis, True), g["d_pbc_multi_points"])
    rb, vol = bc.get_reciprocal_basis(basis)
    assert same(rb, g["d_recip_typo"]) and np.float64(vol) == g["d_recip_vol"]


@pytest.mark.gpu
def test_gpu_histograms_match_golden(g):
    from fullrmc_amd.Core import pairs_histograms as ph
    hi, he = ph.full_pairs_histograms_coords(g["h_box"], g["h_basis"], True, *_hist_args(g))
    assert same(hi, g["h_full_intra"]) and same(he, g["h_full_inter"])
    mi, me = ph.multiple_pairs_histograms_coords(g["h_multi_idx"], g["h_box"], g["h_basis"], True, *_hist_args(g))
    assert same(mi, g["h_multi_intra"]) and same(me, g["h_multi_inter"])


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 16])
def test_gpu_run_matches_golden(g, k):
    from fullrmc_amd.Constraints import PairDistributionConstraint, StructureFactorConstraint
    from fullrmc_amd.Engine import Engine
    from fullrmc_amd.Generators import TranslationGenerator
    elements = [str(e) for e in g["r_elements"]]
    random.seed(int(g["r_seed"]))
    np.random.seed(int(g["r_seed"]))
    E = Engine()
    E.set_system(g["r_real"], elements, g["r_mol"], g["r_basis"], g["r_recip"])
    E.set_move_generator(TranslationGenerator(0.2))
    E.add_constraints([PairDistributionConstraint(g["r_pdc_exp"]),
                       StructureFactorConstraint(g["r_sfc_exp"], rmin=0.0, rmax=6.0, dr=0.05)])
    log = []
    E.run_device(int(g["r_nsteps"]), k=k, log=log)
    assert same(np.array([a for a, _, _ in log], np.int32), g["r_log_group"])
    assert same(np.array([b for _, b, _ in log], bool), g["r_log_accepted"])
    assert same(np.array([t for _, _, t in log], F32), g["r_log_tse"])
    assert same(E.realCoordinates, g["r_final_real"]) and same(E.boxCoordinates, g["r_final_box"])
    for i, c in enumerate(E.constraints):
        assert same(c.data["intra"], g["r_final_intra%d" % i]) and same(c.data["inter"], g["r_final_inter%d" % i])
        assert same(F32(c.standardError), g["r_final_sd"][i])
