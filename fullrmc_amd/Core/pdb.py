"""PDB input for the engine: ATOM/HETATM records and the periodic cell.

The reference reads structures through pdbparser (a separate package, not
part of fullrmc); ``Engine.set_pdb`` (Engine.py:2134-2208) then takes from
it the coordinates, atom names, elements, the molecule index (a new
molecule whenever residue name, residue sequence number or segment
identifier changes, Engine.py:2343-2363) and the boundary conditions.  This
module reads the same fields from the fixed PDB columns:

  atom name     13-16   residue name  18-20   residue seq.  23-26
  x, y, z       31-38, 39-46, 47-54           segment id    73-76
  element       77-78 (as written; the atom name's letters if blank)

and the cell from pdbparser's ``REMARK    Boundary Conditions: ...`` line
(9 numbers: the three basis vectors as rows; 3 numbers: an orthorhombic
cell; 1 number: a cube), else from ``CRYST1`` (a, b, c, alpha, beta, gamma,
with a along x and b in the xy plane).  Files ending in ``.gz`` are read
compressed.
"""
import gzip
import math

import numpy as np

from ..Globals import FLOAT_TYPE, INT_TYPE


def _open(path):
    return gzip.open(path, "rt") if str(path).endswith(".gz") else open(path, "r")


def cryst1_vectors(a, b, c, alpha, beta, gamma):
    """Basis vectors (rows) of a cell given by lengths and angles (degrees)."""
    al, be, ga = (math.radians(x) for x in (alpha, beta, gamma))
    ax = np.array([a, 0.0, 0.0])
    bx = np.array([b * math.cos(ga), b * math.sin(ga), 0.0])
    cx = c * math.cos(be)
    cy = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz = math.sqrt(max(c * c - cx * cx - cy * cy, 0.0))
    return np.array([ax, bx, [cx, cy, cz]])


def _bc_vectors(values):
    v = np.asarray(values, dtype=np.float64)
    if v.size == 1:
        return np.diag([v[0]] * 3)
    if v.size == 3:
        return np.diag(v)
    if v.size == 9:
        return v.reshape(3, 3)
    raise ValueError("Boundary Conditions remark needs 1, 3 or 9 numbers, got %d" % v.size)


def read_pdb(path):
    """Parse a PDB file.  Returns a dict with ``coordinates`` (n,3) float32,
    ``names``, ``elements``, ``residues``, ``sequences``, ``segments``
    (lists), ``moleculesIndex`` (int32) and ``basis`` ((3,3) float64 rows,
    or None for a file without a cell)."""
    coords, names, elements, residues, sequences, segments = [], [], [], [], [], []
    bc, cryst = None, None
    with _open(path) as f:
        for line in f:
            line = line.rstrip("\r\n")
            rec = line[:6]
            if rec in ("ATOM  ", "HETATM"):
                line = line.ljust(80)
                name = line[12:16].strip()
                names.append(name)
                residues.append(line[17:20].strip())
                sequences.append(line[22:26].strip())
                segments.append(line[72:76].strip())
                coords.append((float(line[30:38]), float(line[38:46]), float(line[46:54])))
                el = line[76:78].strip()
                if not el:
                    el = "".join(ch for ch in name if ch.isalpha())[:2]
                elements.append(el)
            elif rec == "REMARK" and "Boundary Conditions:" in line:
                bc = _bc_vectors([float(x) for x in line.split("Boundary Conditions:")[1].split()])
            elif rec == "CRYST1":
                p = line[6:54].split()
                cryst = cryst1_vectors(*(float(x) for x in p[:6]))
    mol, cur, m = [], None, -1
    for key in zip(residues, sequences, segments):
        if key != cur:
            m += 1
            cur = key
        mol.append(m)
    return {"coordinates": np.array(coords, dtype=FLOAT_TYPE).reshape(-1, 3), "names": names,
            "elements": elements, "residues": residues, "sequences": sequences, "segments": segments,
            "moleculesIndex": np.array(mol, dtype=INT_TYPE), "basis": bc if bc is not None else cryst}
