"""Writes tests/golden/golden.npz: input/output vectors of the pair path on
small seeded inputs, computed by the oracle (the C restatement of the
reference's Cython loops and the NumPy float32 constraint chain).

The reference ships no tests, fixtures or golden vectors for this path, and
it cannot be imported or run here, so the vectors come from the oracle,
which the CPU suite pins against NumPy, CPython's random module, brute
force and closed forms.  The fixture freezes those outputs: the CPU suite
checks that the oracle still reproduces them bit for bit, and the GPU suite
checks the HIP path against them.  Inputs are stored with the outputs, so
the check needs neither the generators nor SciPy.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import random
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fullrmc_amd import systems  # noqa: E402
from oracle import chain, refcpu  # noqa: E402
from oracle import engine as oeng  # noqa: E402

F32 = np.float32
WEIGHTS = {"O": 8, "Si": 14, "C": 6, "H": 1}
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def element_index(elements):
    lut = {e: i for i, e in enumerate(sorted(set(elements)))}
    return np.array([lut[e] for e in elements], dtype=np.int32)


def distances_case(g):
    """48 atoms in a triclinic cell: distances, differences, reciprocal
    basis (with the reference's cofactor typo) and the box transform."""
    s = systems.ring_box(nmol=4, cell=(11.0, 12.0, 13.0, 80.0, 95.0, 105.0), seed=31, min_centre=4.5)
    real = s["real"][:48].copy()
    basis, recip = s["basis"], s["recip"]
    box = refcpu.transform_coordinates(recip, real)
    g.update(d_real=real, d_basis=basis, d_recip=recip, d_box=box)
    g["d_pbc_atom7"] = refcpu.pairs_distances_to_indexcoords(7, box, basis, True)
    g["d_pbc_atom7_upper"] = refcpu.pairs_distances_to_indexcoords(7, box, basis, True, allAtoms=False)[7:]
    g["d_ibc_atom7"] = refcpu.pairs_distances_to_indexcoords(7, real, basis, False)
    g["d_pbc_multi"] = refcpu.pairs_distances_to_multi_indexcoords(np.array([0, 5, 9], np.int32), box, basis, True)
    g["d_pbc_diff3"] = refcpu.pairs_differences_to_indexcoords(3, box, basis, True)
    pts = np.array([[0.1, 0.2, 0.3], [0.9, 0.05, 0.5]], dtype=F32).T.copy()  # (3,k) as the reference iterates
    g["d_points"] = pts
    g["d_pbc_multi_points"] = refcpu.pairs_distances_to_multi_points(pts, box, basis, True)
    rb, vol = refcpu.get_reciprocal_basis(basis)
    g["d_recip_typo"], g["d_recip_vol"] = rb, np.float64(vol)
    g["d_transform"] = refcpu.transform_coordinates(recip, real)


def histogram_case(g):
    """12 rigid rings (156 atoms, C/H/O) in a skewed cell: full and
    multiple pair histograms, intra and inter."""
    s = systems.ring_box(nmol=12, cell=(14.0, 15.0, 16.0, 80.0, 95.0, 105.0), seed=32, min_centre=4.0)
    elem = element_index(s["elements"])
    box = refcpu.transform_coordinates(s["recip"], s["real"])
    g.update(h_box=box, h_basis=s["basis"], h_mol=s["molecules"].astype(np.int32), h_elem=elem)
    args = (s["molecules"], elem, 3, F32(0.0), F32(7.0), F32(0.05), 140)
    hi, he = refcpu.full_pairs_histograms_coords(box, s["basis"], True, *args)
    g["h_full_intra"], g["h_full_inter"] = hi, he
    idx = np.array([3, 17, 40, 41], dtype=np.int32)
    mi, me = refcpu.multiple_pairs_histograms_coords(idx, box, s["basis"], True, *args)
    g["h_multi_idx"], g["h_multi_intra"], g["h_multi_inter"] = idx, mi, me


def run_case(g):
    """200 atoms (Si/O) in a 14 A cube, PDC + SFC, 150 single-atom
    translations of 0.2 A with random.seed(13) / np.random.seed(13): the
    per-step (group, accepted, total) log, final coordinates, histograms and
    totals of the sequential engine."""
    s = systems.random_box(200, 14.0, 2.0, seed=33, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    t = systems.jitter(s, 0.1, 34)
    osys_t = oeng.System(t["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    w = chain.normalized_weighting(osys_t.counts, {e: WEIGHTS[e] for e in osys_t.elements})
    x = (0.025 + 0.05 * np.arange(120)).astype(F32)
    pdc_exp = np.stack([x, np.zeros_like(x)], 1).astype(F32)
    p = oeng.PDC(osys_t, pdc_exp, w)
    pdc_exp[:, 1] = p.model(*p.hist_full(osys_t.box, osys_t.mol, osys_t.elem))
    q = np.linspace(0.5, 20.0, 60).astype(F32)
    sfc_exp = np.stack([q, np.zeros_like(q)], 1).astype(F32)
    c = oeng.SFC(osys_t, sfc_exp, w, 0.0, 6.0, 0.05)
    sfc_exp[:, 1] = c.model(*c.hist_full(osys_t.box, osys_t.mol, osys_t.elem))

    g.update(r_real=s["real"], r_basis=s["basis"], r_recip=s["recip"], r_mol=s["molecules"].astype(np.int32),
             r_elem=element_index(s["elements"]), r_pdc_exp=pdc_exp, r_sfc_exp=sfc_exp,
             r_weights=np.array([w[k] for k in sorted(w)], dtype=F32),
             r_weight_keys=np.array(sorted(w)), r_seed=np.int64(13), r_nsteps=np.int64(150))
    g["r_elements"] = np.array(s["elements"])
    random.seed(13)
    np.random.seed(13)
    osys = oeng.System(s["real"], s["elements"], s["molecules"], s["basis"], s["recip"])
    oe = oeng.Engine(osys, [[i] for i in range(200)], oeng.Translation(0.2),
                     [oeng.PDC(osys, pdc_exp, w), oeng.SFC(osys, sfc_exp, w, 0.0, 6.0, 0.05)])
    oe.initialize()
    g["r_tse0"] = F32(oe.tse)
    log = []
    oe.run(150, log)
    g["r_log_group"] = np.array([a for a, _, _ in log], dtype=np.int32)
    g["r_log_accepted"] = np.array([b for _, b, _ in log], dtype=bool)
    g["r_log_tse"] = np.array([t for _, _, t in log], dtype=F32)
    g["r_final_real"] = oe.s.real.copy()
    g["r_final_box"] = oe.s.box.copy()
    g["r_final_sd"] = np.array([c.sd for c in oe.constraints], dtype=F32)
    for i, c in enumerate(oe.constraints):
        g["r_final_intra%d" % i], g["r_final_inter%d" % i] = c.intra.copy(), c.inter.copy()


def main():
    g = {}
    distances_case(g)
    histogram_case(g)
    run_case(g)
    np.savez_compressed(OUT, **g)
    print("wrote %s (%d arrays, %d bytes)" % (OUT, len(g), os.path.getsize(OUT)))


if __name__ == "__main__":
    main()
