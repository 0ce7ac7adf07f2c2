"""The chunk test of k_delta's chunk walk never drops a pair in range.

The engine sorts a periodic box's atoms into chunks of 256 (cells of about
256 atoms, z-major over wrapped fractional coordinates), keeps each chunk's
bounding box in wrapped fractional coordinates, and skips a chunk for a
position whose circular gap to the box along some axis, times that axis'
lattice-plane spacing, exceeds wcut = maxD (1 + 1e-5) + 0.01 + 1e-4 h_max.
Here the same float32 test is applied to random triclinic and cubic boxes
(coordinates unwrapped up to three box lengths away) and checked against the
oracle's distances (pairs_distances_to_point, the reference's minimum-image
float32 arithmetic): every atom of a skipped chunk is at or beyond maxD.  The
test also checks that it skips a good share of the chunks, so that it is not
vacuous."""
import numpy as np
import pytest

from oracle import refcpu

F32 = np.float32
CHUNK = 256


def plane_spacings(basis):
    a = np.asarray(basis, np.float64)
    cx = [np.cross(a[1], a[2]), np.cross(a[2], a[0]), np.cross(a[0], a[1])]
    vol = abs(np.dot(a[0], cx[0]))
    return np.array([vol / np.linalg.norm(c) for c in cx]), vol


def chunk_boxes(box, basis):
    """Sorted order and float32 chunk boxes as the engine builds them."""
    n = len(box)
    h, vol = plane_spacings(basis)
    L = np.cbrt(CHUNK * vol / n)
    nc = np.clip(np.floor(h / L + 0.5), 1, 256).astype(np.int64)
    wf = (box - np.floor(box)).astype(F32)
    ci = np.minimum((wf * nc.astype(F32)).astype(np.int64), nc - 1)
    cell = (ci[:, 2] * nc[1] + ci[:, 1]) * nc[0] + ci[:, 0]
    order = np.argsort(cell, kind="stable")
    nchunk = (n + CHUNK - 1) // CHUNK
    lo = np.full((nchunk, 3), 2.0, F32)
    hi = np.full((nchunk, 3), -1.0, F32)
    for c in range(nchunk):
        sel = wf[order[c * CHUNK:(c + 1) * CHUNK]]
        lo[c] = sel.min(0)
        hi[c] = sel.max(0)
    return order, lo, hi, h


def circ_gap(f, lo, hi):
    f, lo, hi = F32(f), F32(lo), F32(hi)
    if lo <= f <= hi:
        return F32(0)
    direct = lo - f if f < lo else f - hi
    around = (f + F32(1)) - hi if f < lo else (lo + F32(1)) - f
    return min(F32(direct), F32(around))


def far_from(point, lo, hi, hsp, wcut):
    for d in range(3):
        f = F32(point[d]) - F32(np.floor(F32(point[d])))
        if F32(circ_gap(f, lo[d], hi[d]) * hsp[d]) > wcut:
            return True
    return False


def random_case(seed, n, basis, spread):
    rng = np.random.default_rng(seed)
    box = rng.random((n, 3)).astype(F32)
    box += rng.integers(-spread, spread + 1, (n, 3)).astype(F32)  # unwrapped images
    return box.astype(F32), np.asarray(basis, F32)


CASES = {
    "cubic": (np.diag([40.0, 40.0, 40.0]), 8.0),
    "triclinic": (np.array([[38.0, 0.0, 0.0], [7.0, 41.0, 0.0], [-5.0, 6.0, 44.0]]), 9.5),
    "flat": (np.array([[60.0, 0.0, 0.0], [0.0, 55.0, 0.0], [12.0, -9.0, 19.0]]), 6.0),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_skipped_chunks_hold_no_pair_in_range(name):
    basis, maxD = CASES[name]
    box, B = random_case(7, 4000, basis, 3)
    order, lo, hi, h = chunk_boxes(box, B)
    hsp = h.astype(F32)
    wcut = F32(maxD * (1 + 1e-5) + 0.01 + 1e-4 * h.max())
    rng = np.random.default_rng(11)
    # proposal positions: random, on cell edges, and atoms of the box moved a little
    pts = [rng.random(3) * 3 - 1 for _ in range(60)]
    pts += [np.array([0.0, 0.5, 0.999999]), np.array([1.0, -1.0, 0.25])]
    pts += [box[i] + rng.normal(0, 0.01, 3) for i in rng.integers(0, len(box), 40)]
    skipped = total = 0
    for p in pts:
        p = np.asarray(p, F32)
        d = refcpu.pairs_distances_to_point(p, box, B, True)
        for c in range(len(lo)):
            total += 1
            if far_from(p, lo[c], hi[c], hsp, wcut):
                skipped += 1
                idx = order[c * CHUNK:(c + 1) * CHUNK]
                assert d[idx].min() >= F32(maxD), (name, c, float(d[idx].min()))
    assert skipped > 0.15 * total, (skipped, total)


def test_grown_chunk_box_still_covers_moved_atoms():
    """A chunk box widened by moved atoms' new positions (what a commit
    does) keeps the guarantee for the atoms at their new places."""
    basis, maxD = CASES["triclinic"]
    box, B = random_case(3, 3000, basis, 1)
    order, lo, hi, h = chunk_boxes(box, B)
    rng = np.random.default_rng(5)
    moved = box.copy()
    for i in rng.integers(0, len(box), 300):
        moved[i] = (moved[i] + rng.normal(0, 0.05, 3)).astype(F32)
        c = int(np.nonzero(order == i)[0][0]) // CHUNK
        f = (moved[i] - np.floor(moved[i])).astype(F32)
        lo[c] = np.minimum(lo[c], f)
        hi[c] = np.maximum(hi[c], f)
    hsp = h.astype(F32)
    wcut = F32(maxD * (1 + 1e-5) + 0.01 + 1e-4 * h.max())
    for p in [rng.random(3).astype(F32) for _ in range(40)]:
        d = refcpu.pairs_distances_to_point(p, moved, B, True)
        for c in range(len(lo)):
            if far_from(p, lo[c], hi[c], hsp, wcut):
                idx = order[c * CHUNK:(c + 1) * CHUNK]
                assert d[idx].min() >= F32(maxD)


def test_a_too_small_cut_off_is_caught():
    """The check has teeth: with wcut at half of maxD, skipped chunks hold
    pairs in range."""
    basis, maxD = CASES["cubic"]
    box, B = random_case(9, 4000, basis, 0)
    order, lo, hi, h = chunk_boxes(box, B)
    hsp = h.astype(F32)
    wcut = F32(maxD / 2)
    found = False
    for p in [np.random.default_rng(1).random(3).astype(F32) for _ in range(10)]:
        d = refcpu.pairs_distances_to_point(p, box, B, True)
        for c in range(len(lo)):
            if far_from(p, lo[c], hi[c], hsp, wcut) and d[order[c * CHUNK:(c + 1) * CHUNK]].min() < maxD:
                found = True
    assert found
