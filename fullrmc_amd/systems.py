"""Seeded synthetic configurations of the benchmark workloads.

Every builder returns a dict with ``real`` (n,3) float32 real coordinates,
``elements`` (list of element names per atom), ``molecules`` (n,) int32,
``basis`` (3,3) float32 (rows are the cell vectors), ``recip`` (3,3)
float32 (its inverse) and builder metadata.  Box (fractional) coordinates
follow from :func:`box_coordinates` with the float32 operation order of the
reference's ``transform_coordinates``.

The reference takes the basis, its inverse and the volume from pdbparser's
PeriodicBoundaries (Engine.py:2243-2260); here the inverse is numpy's
float64 inverse rounded to float32 and the volume is |det| rounded to
float32 (parity unpinned for pdbparser, which is not available).
"""
import math

import numpy as np

F32 = np.float32


def cubic_basis(L):
    return np.diag([L, L, L]).astype(F32)


def triclinic_basis(a, b, c, alpha, beta, gamma):
    """Cell vectors as rows: a along x, b in the xy plane (angles in degrees)."""
    al, be, ga = (math.radians(x) for x in (alpha, beta, gamma))
    ax = np.array([a, 0.0, 0.0])
    bx = np.array([b * math.cos(ga), b * math.sin(ga), 0.0])
    cx_ = c * math.cos(be)
    cy_ = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz_ = math.sqrt(max(c * c - cx_ * cx_ - cy_ * cy_, 0.0))
    return np.array([ax, bx, [cx_, cy_, cz_]], dtype=np.float64).astype(F32)


def reciprocal(basis):
    return np.linalg.inv(np.asarray(basis, dtype=np.float64)).astype(F32)


def volume(basis):
    return F32(abs(np.linalg.det(np.asarray(basis, dtype=np.float64))))


def box_coordinates(recip, real):
    """transform_coordinates(recip, real) in float32 with the reference's
    operand order (boundary_conditions_collection.pyx:105-108); NumPy's
    elementwise float32 multiply and add round each operation."""
    M = np.asarray(recip, dtype=F32)
    r = np.asarray(real, dtype=F32)
    x, y, z = r[:, 0:1], r[:, 1:2], r[:, 2:3]
    return ((x * M[0] + y * M[1]) + z * M[2]).astype(F32)


def _rsa(n, basis, rmin, rng, max_tries=200):
    """Random sequential insertion in a periodic cell with a minimum pair
    distance, using a cell list over fractional coordinates."""
    B = np.asarray(basis, dtype=np.float64)
    Binv = np.linalg.inv(B)
    # fractional grid fine enough that neighbours within rmin sit in +-1 cells
    heights = 1.0 / np.linalg.norm(Binv, axis=0)  # plane spacings of the cell
    ncell = np.maximum(1, np.floor(heights / rmin).astype(int))
    cells = {}
    pts = np.empty((n, 3))
    r2 = rmin * rmin
    offs = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]
    placed = 0
    tries = 0
    while placed < n:
        f = rng.random(3)
        c = tuple((f * ncell).astype(int) % ncell)
        ok = True
        for o in offs:
            cc = ((c[0] + o[0]) % ncell[0], (c[1] + o[1]) % ncell[1], (c[2] + o[2]) % ncell[2])
            for q in cells.get(cc, ()):
                d = f - pts[q]
                d -= np.round(d)
                v = d @ B
                if v @ v < r2:
                    ok = False
                    break
            if not ok:
                break
        if ok:
            pts[placed] = f
            cells.setdefault(c, []).append(placed)
            placed += 1
            tries = 0
        else:
            tries += 1
            if tries > max_tries * 1000:
                raise RuntimeError("random insertion jammed at %d of %d atoms" % (placed, n))
    return (pts @ B).astype(F32)


def _rsa_cubic_batched(n, L, rmin, rng, batch=None):
    """Random insertion in a periodic cube, vectorised over batches of
    candidates (periodic k-d trees): a candidate is kept if no kept atom lies
    within rmin and no earlier candidate of its batch does."""
    from scipy.spatial import cKDTree
    if batch is None:
        batch = int(min(40000, max(2000, n // 4)))
    pts = np.empty((0, 3))
    while len(pts) < n:
        cand = rng.random((batch, 3)) * L
        ok = np.ones(batch, dtype=bool)
        if len(pts):
            near = cKDTree(pts, boxsize=L).query_ball_point(cand, rmin, return_length=True)
            ok &= near == 0
        pairs = cKDTree(cand, boxsize=L).query_pairs(rmin, output_type="ndarray")
        if len(pairs):
            ok[pairs.max(axis=1)] = False
        pts = np.concatenate([pts, cand[ok][: n - len(pts)]])
    return pts.astype(F32)


def random_box(n, L, rmin, seed, species=("Ar",), fractions=None):
    """Uniform random atoms in a cube of edge L (Angstrom) with no pair
    closer than rmin; species assigned by fractions (default all the first)."""
    rng = np.random.default_rng(seed)
    basis = cubic_basis(L)
    real = _rsa_cubic_batched(n, L, rmin, rng) if n > 2000 else _rsa(n, basis, rmin, rng)
    if fractions is None:
        elements = [species[0]] * n
    else:
        counts = [int(round(f * n)) for f in fractions]
        counts[-1] = n - sum(counts[:-1])
        elements = []
        for s, c in zip(species, counts):
            elements += [s] * c
        order = rng.permutation(n)
        elements = [elements[i] for i in order]
    return dict(real=real, elements=elements, molecules=np.arange(n, dtype=np.int32), basis=basis,
                recip=reciprocal(basis), name="random_box", seed=seed)


def jitter(system, sigma, seed):
    rng = np.random.default_rng(seed)
    out = dict(system)
    out["real"] = (system["real"].astype(np.float64) + rng.normal(0.0, sigma, system["real"].shape)).astype(F32)
    return out


def b2_alloy(cells=(15, 15, 30), a=3.01, sigma=0.08, seed=0, species=("Fe", "Al")):
    """B2 (CsCl) lattice: species 0 at the cell corner, species 1 at the body
    centre, with a Gaussian jitter sigma (Angstrom)."""
    nx, ny, nz = cells
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    corners = g * a
    centres = (g + 0.5) * a
    real = np.concatenate([corners, centres]).astype(np.float64)
    elements = [species[0]] * len(corners) + [species[1]] * len(centres)
    basis = np.diag([nx * a, ny * a, nz * a]).astype(F32)
    rng = np.random.default_rng(seed)
    real = real + rng.normal(0.0, sigma, real.shape)
    n = len(real)
    return dict(real=real.astype(F32), elements=elements, molecules=np.arange(n, dtype=np.int32), basis=basis,
                recip=reciprocal(basis), name="b2_alloy", seed=seed)


def ring_molecule():
    """13-atom rigid five-membered ring (O + 4 C, two H on each C):
    coordinates in Angstrom around the ring centre, and element names."""
    ring_r = 1.22
    atoms, names = [], []
    for k in range(5):
        t = 2 * math.pi * k / 5
        atoms.append([ring_r * math.cos(t), ring_r * math.sin(t), 0.0])
        names.append("O" if k == 0 else "C")
    for k in range(1, 5):
        t = 2 * math.pi * k / 5
        out = np.array([math.cos(t), math.sin(t), 0.0])
        base = np.array(atoms[k])
        for s in (+1, -1):
            h = base + 0.63 * out + s * 0.89 * np.array([0.0, 0.0, 1.0])
            atoms.append(list(h))
            names.append("H")
    return np.array(atoms), names


def ring_box(nmol=500, cell=(42.0, 44.0, 46.0, 80.0, 95.0, 105.0), seed=0, min_centre=5.0):
    """Rigid 13-atom rings with random orientations and centres in a
    triclinic cell; molecules are contiguous atom ranges."""
    rng = np.random.default_rng(seed)
    basis = triclinic_basis(*cell)
    centres = _rsa(nmol, basis, min_centre, rng).astype(np.float64)
    mol_xyz, names = ring_molecule()
    real, elements, molecules = [], [], []
    for m in range(nmol):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        real.append(mol_xyz @ R.T + centres[m])
        elements += names
        molecules += [m] * len(names)
    real = np.concatenate(real).astype(F32)
    return dict(real=real, elements=elements, molecules=np.array(molecules, dtype=np.int32), basis=basis,
                recip=reciprocal(basis), name="ring_box", seed=seed)


def sio2_box(n=200000, L=145.0, rmin=1.5, seed=0):
    """Amorphous SiO2-like random box: round(n/3) Si (66,667 at n=200,000),
    the rest O, uniform with a minimum pair distance."""
    nsi = int(round(n / 3.0))
    sysm = random_box(n, L, rmin, seed, species=("O", "Si"))
    rng = np.random.default_rng(seed + 1)
    el = np.array(["O"] * n, dtype=object)
    el[rng.choice(n, nsi, replace=False)] = "Si"
    sysm["elements"] = list(el)
    sysm["name"] = "sio2_box"
    return sysm


def replica_sizes(nrep=256, nmin=1000, nmax=20000, seed=0):
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(math.log(nmin), math.log(nmax), nrep))).astype(int)
