"""Input formats: the PDB reader (ATOM/HETATM columns, the molecule rule,
the Boundary Conditions remark and CRYST1), experimental data read from
files, rebin, and the S(Q) grid defaults -- on the CPU, against the
reference's own example files (tests/golden/examples, copied by
tests/golden/make_examples.py) and closed forms."""
import gzip
import math
import os

import numpy as np
import pytest

from fullrmc_amd.Constraints import PairDistributionConstraint, ReducedStructureFactorConstraint
from fullrmc_amd.Core.Collection import rebin
from fullrmc_amd.Core.pdb import cryst1_vectors, read_pdb
from oracle import chain

F32 = np.float32
EX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "examples")


def _atom(serial, name, res, seq, x, y, z, el, seg=""):
    return ("ATOM  %5d %-4s %3s  %4d    %8.3f%8.3f%8.3f  1.00  0.00      %-4s%2s\n"
            % (serial, name, res, seq, x, y, z, seg, el))


def test_niti_pdb_fields():
    p = read_pdb(os.path.join(EX, "niti_system.pdb.gz"))
    n = len(p["names"])
    assert n == 6750 and p["coordinates"].shape == (6750, 3)
    assert np.array_equal(p["coordinates"][:2], np.array([[-22.113, -22.113, -22.113],
                                                          [-19.062, -22.113, -22.113]], F32))
    assert np.array_equal(p["basis"], np.diag([45.75] * 3))
    assert sorted(set(p["elements"])) == ["ni", "ti"]
    # one residue per atom (RMC 1, 2, ...): every atom its own molecule
    assert np.array_equal(p["moleculesIndex"], np.arange(n))
    # the raw text agrees on every coordinate
    with gzip.open(os.path.join(EX, "niti_system.pdb.gz"), "rt") as f:
        rows = [line for line in f if line.startswith("ATOM")]
    raw = np.array([[float(r[30:38]), float(r[38:46]), float(r[46:54])] for r in rows], F32)
    assert np.array_equal(raw, p["coordinates"])


def test_triclinic_remark_and_molecule_rule(tmp_path):
    p = read_pdb(os.path.join(EX, "removes_system.pdb.gz"))
    assert np.allclose(p["basis"], [[42.5, 0, 0], [-21.25, 36.80, 0], [0, 0, 42.28]])
    assert sorted(set(p["elements"])) == ["Co", "Li", "Mn", "Ni", "O"]
    # a new molecule whenever residue name, number or segment changes
    text = "".join([_atom(1, "C1", "THF", 1, 0, 0, 0, "C"), _atom(2, "O1", "THF", 1, 1, 0, 0, "O"),
                    _atom(3, "C1", "THF", 2, 2, 0, 0, "C"), _atom(4, "C1", "THF", 2, 3, 0, 0, "C", seg="B"),
                    _atom(5, "H1", "WAT", 2, 4, 0, 0, ""), _atom(6, "H2", "WAT", 2, 5, 0, 0, "H", seg="")])
    f = tmp_path / "m.pdb"
    f.write_text("REMARK    Boundary Conditions: 12.5\n" + text)
    q = read_pdb(str(f))
    assert list(q["moleculesIndex"]) == [0, 0, 1, 2, 3, 3]
    assert q["elements"][4] == "H"  # blank element column: letters of the atom name
    assert np.array_equal(q["basis"], np.diag([12.5] * 3))


def test_cryst1_cell_closed_forms(tmp_path):
    f = tmp_path / "c.pdb"
    f.write_text("CRYST1   10.000   11.000   12.000  90.00  90.00  90.00 P 1           1\n"
                 + _atom(1, "AR", "ARG", 1, 1, 2, 3, "Ar"))
    assert np.allclose(read_pdb(str(f))["basis"], np.diag([10.0, 11.0, 12.0]))
    B = cryst1_vectors(5.0, 5.0, 8.0, 90.0, 90.0, 120.0)  # hexagonal
    assert np.allclose(B, [[5, 0, 0], [-2.5, 5 * math.sqrt(3) / 2, 0], [0, 0, 8]])
    B = cryst1_vectors(4.0, 5.0, 6.0, 80.0, 95.0, 105.0)
    for (i, j, ang) in ((1, 2, 80.0), (0, 2, 95.0), (0, 1, 105.0)):
        c = B[i] @ B[j] / np.linalg.norm(B[i]) / np.linalg.norm(B[j])
        assert abs(math.degrees(math.acos(c)) - ang) < 1e-9
    assert np.allclose(np.linalg.norm(B, axis=1), [4, 5, 6])


def test_rebin_closed_forms():
    # two points 0.02 either side of each multiple of 0.1: a linear y averages
    # to its value at the bin centre
    c = 1.0 + 0.1 * np.arange(20)
    x = np.sort(np.concatenate([c - 0.02, c + 0.02]))
    X, Y = rebin(np.stack([x, 2 * x + 1], 1), bin=0.1)
    # the first centre is int(x0/bin)*bin = 0.9 (x0 = 0.98 truncates down): an empty bin, 0
    assert np.allclose(X, np.concatenate([[0.9], c]))
    assert np.allclose(Y, np.concatenate([[0.0], 2 * c + 1]))
    X, Y = rebin(np.stack([x, np.full_like(x, 7.0)], 1), bin=0.1)
    assert np.allclose(Y[1:], 7.0)
    X2, Y2 = rebin(np.stack([np.array([1.0, 1.24]), np.array([4.0, 6.0])], 1), bin=0.1)
    assert np.allclose(X2, [1.0, 1.1, 1.2]) and np.allclose(Y2, [4.0, 0.0, 6.0])  # empty bins are 0


def test_experimental_data_from_files_and_sq_defaults():
    gr = os.path.join(EX, "niti_experimental.gr.gz")
    a = PairDistributionConstraint(gr)
    b = PairDistributionConstraint(np.loadtxt(gr, dtype=F32))
    assert a.experimentalData.tobytes() == b.experimentalData.tobytes()
    assert a.histogramSize == 1999 and a.bin == F32(0.01)
    Sq = np.transpose(rebin(np.loadtxt(os.path.join(EX, "niti_experimental.fq.gz")), bin=0.05)).astype(F32)
    c = ReducedStructureFactorConstraint(Sq)
    rmin, rmax, dr = chain.sfc_defaults(Sq[:, 0])
    mn, mx, bn, centers, H, vols = chain.sfc_grid(rmin, rmax, dr)
    assert (c.minimumDistance, c.maximumDistance, c.bin, c.histogramSize) == (mn, mx, bn, H)
    assert c.shellCenters.tobytes() == centers.tobytes()
    assert abs(float(dr) - 0.2) < 1e-6 and abs(float(rmax) - 2 * math.pi / 0.05) < 1e-3
