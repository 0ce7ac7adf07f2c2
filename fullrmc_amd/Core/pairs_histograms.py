"""Drop-in for fullrmc's ``fullrmc.Core.pairs_histograms`` extension
(Extensions/pairs_histograms.pyx), computed by the gfx950 kernels of
``libfrmc_hip.so``.

Same functions, positional order, defaults, dtypes and return values as the
reference.  ``ncores`` is accepted and ignored (the work runs on the GPU).
Counts are exact: the reference's shared-bin ``+=`` inside ``prange`` can
lose increments with ``ncores>1``; these results equal its ``ncores=1``
output.
"""
import numpy as np

from .. import _args as A
from .. import _device as D
from .._lib import check, lib

__all__ = ["single_pairs_histograms", "multiple_pairs_histograms_coords",
           "multiple_pairs_histograms_dists", "full_pairs_histograms_coords",
           "full_pairs_histograms_dists"]


def single_pairs_histograms(atomIndex, distances, moleculeIndex, elementIndex, hintra, hinter,
                            minDistance, maxDistance, bin, allAtoms=True, ncores=1):
    """Add the pair counts of atom ``atomIndex`` given its distances array
    into ``hintra``/``hinter`` in place (pairs_histograms.pyx:77-141).
    Returns None."""
    atomIndex = A.c_int32("atomIndex", atomIndex)
    distances = A.array("distances", distances, np.float32, 1)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    hintra = A.array("hintra", hintra, np.float32, 3)
    hinter = A.array("hinter", hinter, np.float32, 3)
    minDistance, maxDistance, bin = (A.c_float32("minDistance", minDistance),
                                     A.c_float32("maxDistance", maxDistance), A.c_float32("bin", bin))
    A.c_int32("ncores", ncores)
    n = distances.shape[0]
    A.check_index("atomIndex", atomIndex, n)
    if moleculeIndex.shape[0] < n or elementIndex.shape[0] < n:
        raise IndexError("moleculeIndex/elementIndex shorter than distances")
    if hintra.shape != hinter.shape or hintra.shape[0] != hintra.shape[1]:
        raise ValueError("hintra and hinter must both be (nel,nel,H)")
    nel, H = hintra.shape[0], hintra.shape[2]
    A.check_indexes("elementIndex", elementIndex[:n], nel)
    d_dist, d_mol, d_el = D.upload(distances), D.upload(moleculeIndex[:n]), D.upload(elementIndex[:n])
    d_hi, d_he = D.upload(hintra), D.upload(hinter)
    check(lib().frmc_single_pairs_histograms(atomIndex, d_dist.data_ptr(), n, d_mol.data_ptr(), d_el.data_ptr(),
                                             d_hi.data_ptr(), d_he.data_ptr(), nel, H, minDistance, maxDistance, bin,
                                             int(A.bint(allAtoms)), D.stream()))
    np.copyto(hintra, D.download(d_hi))
    np.copyto(hinter, D.download(d_he))


def _coords_common(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements, minDistance,
                   maxDistance, bin, histSize, ncores):
    boxCoords = A.array("boxCoords", boxCoords, np.float32, 2)
    basis = A.array("basis", basis, np.float32, 2)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    nel = A.c_int32("numberOfElements", numberOfElements)
    H = A.c_int32("histSize", histSize)
    A.c_int32("ncores", ncores)
    minDistance, maxDistance, bin = (A.c_float32("minDistance", minDistance),
                                     A.c_float32("maxDistance", maxDistance), A.c_float32("bin", bin))
    if boxCoords.shape[1] != 3 or basis.shape != (3, 3):
        raise ValueError("boxCoords must be (n,3) and basis (3,3)")
    n = boxCoords.shape[0]
    if moleculeIndex.shape[0] < n or elementIndex.shape[0] < n:
        raise IndexError("moleculeIndex/elementIndex shorter than boxCoords")
    A.check_indexes("elementIndex", elementIndex[:n], nel)
    return boxCoords, basis, bool(isPBC), moleculeIndex[:n], elementIndex[:n], nel, H, minDistance, maxDistance, bin


def multiple_pairs_histograms_coords(indexes, boxCoords, basis, isPBC, moleculeIndex, elementIndex,
                                     numberOfElements, minDistance, maxDistance, bin, histSize,
                                     allAtoms=True, ncores=1):
    """Pair histograms of the atoms in ``indexes`` vs all atoms (or vs the
    atoms after each when ``allAtoms`` is False), from box coordinates
    (pairs_histograms.pyx:150-217).  Returns ``(hintra, hinter)``, float32
    (nel, nel, histSize)."""
    indexes = A.array("indexes", indexes, np.int32, 1)
    (boxCoords, basis, isPBC, mol, el, nel, H, mn, mx, bn) = _coords_common(
        boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements, minDistance, maxDistance,
        bin, histSize, ncores)
    n = boxCoords.shape[0]
    A.check_indexes("indexes", indexes, n)
    d_idx, d_c, d_b = D.upload(indexes), D.upload(boxCoords), D.upload(basis)
    d_mol, d_el = D.upload(mol), D.upload(el)
    d_hi, d_he = D.empty((nel, nel, H), np.float32), D.empty((nel, nel, H), np.float32)
    check(lib().frmc_multiple_pairs_histograms_coords(
        d_idx.data_ptr(), indexes.shape[0], d_c.data_ptr(), n, d_b.data_ptr(), int(isPBC), d_mol.data_ptr(),
        d_el.data_ptr(), nel, mn, mx, bn, H, int(A.bint(allAtoms)), d_hi.data_ptr(), d_he.data_ptr(),
        D.stream()))
    return D.download(d_hi), D.download(d_he)


def multiple_pairs_histograms_dists(indexes, distances, moleculeIndex, elementIndex, numberOfElements,
                                    minDistance, maxDistance, bin, histSize, allAtoms=True, ncores=1):
    """Pair histograms from a precomputed (n, k) distance array; column c
    belongs to atom ``indexes[c]`` (pairs_histograms.pyx:225-281)."""
    indexes = A.array("indexes", indexes, np.int32, 1)
    distances = A.array("distances", distances, np.float32, 2)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    nel = A.c_int32("numberOfElements", numberOfElements)
    H = A.c_int32("histSize", histSize)
    A.c_int32("ncores", ncores)
    mn, mx, bn = (A.c_float32("minDistance", minDistance), A.c_float32("maxDistance", maxDistance),
                  A.c_float32("bin", bin))
    n, k = distances.shape
    if indexes.shape[0] > k:
        raise IndexError("more indexes (%d) than distance columns (%d)" % (indexes.shape[0], k))
    if moleculeIndex.shape[0] < n or elementIndex.shape[0] < n:
        raise IndexError("moleculeIndex/elementIndex shorter than distances")
    A.check_indexes("indexes", indexes, n)
    A.check_indexes("elementIndex", elementIndex[:n], nel)
    d_idx, d_d = D.upload(indexes), D.upload(distances)
    d_mol, d_el = D.upload(moleculeIndex[:n]), D.upload(elementIndex[:n])
    d_hi, d_he = D.empty((nel, nel, H), np.float32), D.empty((nel, nel, H), np.float32)
    check(lib().frmc_multiple_pairs_histograms_dists(
        d_idx.data_ptr(), indexes.shape[0], d_d.data_ptr(), n, k, d_mol.data_ptr(), d_el.data_ptr(), nel, mn, mx,
        bn, H, int(A.bint(allAtoms)), d_hi.data_ptr(), d_he.data_ptr(), D.stream()))
    return D.download(d_hi), D.download(d_he)


def full_pairs_histograms_coords(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                 minDistance, maxDistance, bin, histSize, ncores=1):
    """All-pairs histograms (each unordered pair once, slot
    [element(earlier), element(later)]) from box coordinates
    (pairs_histograms.pyx:289-335)."""
    (boxCoords, basis, isPBC, mol, el, nel, H, mn, mx, bn) = _coords_common(
        boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements, minDistance, maxDistance,
        bin, histSize, ncores)
    n = boxCoords.shape[0]
    d_c, d_b, d_mol, d_el = D.upload(boxCoords), D.upload(basis), D.upload(mol), D.upload(el)
    d_hi, d_he = D.empty((nel, nel, H), np.float32), D.empty((nel, nel, H), np.float32)
    check(lib().frmc_full_pairs_histograms_coords(
        d_c.data_ptr(), n, d_b.data_ptr(), int(isPBC), d_mol.data_ptr(), d_el.data_ptr(), nel, mn, mx, bn, H,
        d_hi.data_ptr(), d_he.data_ptr(), D.stream()))
    return D.download(d_hi), D.download(d_he)


def full_pairs_histograms_coords_float64(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                         minDistance, maxDistance, bin, histSize):
    """All-pairs histograms of the binary64 recompute: the same pairs and slots
    as :func:`full_pairs_histograms_coords`, with every distance, range test
    and bin index as if evaluated in double from the float32 coordinates
    (float32 kernel with a double re-check of every pair near a bin edge).
    Not part of the reference's API.  Returns ``(hintra, hinter, rechecked)``
    with float32 histograms and the number of pairs re-checked in double."""
    import ctypes
    (boxCoords, basis, isPBC, mol, el, nel, H, mn, mx, bn) = _coords_common(
        boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements, minDistance, maxDistance,
        bin, histSize, 1)
    n = boxCoords.shape[0]
    size = nel * nel * H
    d_c, d_b, d_mol, d_el = D.upload(boxCoords), D.upload(basis), D.upload(mol), D.upload(el)
    d_cnt = D.empty((2 * size + 4,), np.int32)
    rechecked = ctypes.c_int64(0)
    check(lib().frmc_full_pairs_counts_f64(
        d_c.data_ptr(), n, d_b.data_ptr(), int(isPBC), d_mol.data_ptr(), d_el.data_ptr(), nel, mn, mx, bn, H,
        d_cnt.data_ptr(), ctypes.byref(rechecked), D.stream()))
    c = D.download(d_cnt)
    return (c[:size].astype(np.float32).reshape(nel, nel, H), c[size:2 * size].astype(np.float32).reshape(nel, nel, H),
            int(rechecked.value))


def full_pairs_histograms_dists(distances, moleculeIndex, elementIndex, numberOfElements, minDistance,
                                maxDistance, bin, histSize, ncores=1):
    """All-pairs histograms from an (n, k) distance array with
    ``indexes = arange(k)`` and ``allAtoms=False`` (pairs_histograms.pyx:343-383)."""
    distances = A.array("distances", distances, np.float32, 2)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    nel = A.c_int32("numberOfElements", numberOfElements)
    H = A.c_int32("histSize", histSize)
    A.c_int32("ncores", ncores)
    mn, mx, bn = (A.c_float32("minDistance", minDistance), A.c_float32("maxDistance", maxDistance),
                  A.c_float32("bin", bin))
    n, k = distances.shape
    if k > n:
        raise IndexError("distances has more columns (%d) than rows (%d)" % (k, n))
    if moleculeIndex.shape[0] < n or elementIndex.shape[0] < n:
        raise IndexError("moleculeIndex/elementIndex shorter than distances")
    A.check_indexes("elementIndex", elementIndex[:n], nel)
    d_d, d_mol, d_el = D.upload(distances), D.upload(moleculeIndex[:n]), D.upload(elementIndex[:n])
    d_hi, d_he = D.empty((nel, nel, H), np.float32), D.empty((nel, nel, H), np.float32)
    check(lib().frmc_full_pairs_histograms_dists(
        d_d.data_ptr(), n, k, d_mol.data_ptr(), d_el.data_ptr(), nel, mn, mx, bn, H, d_hi.data_ptr(),
        d_he.data_ptr(), D.stream()))
    return D.download(d_hi), D.download(d_he)
