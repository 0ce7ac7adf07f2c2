"""ORACLE / TEST INFRASTRUCTURE ONLY -- never imported by the product path.

Python face of ``librefcpu.so`` (the plain-C restatement in refcpu.c of the
reference Cython loops), with the reference modules' function names and
signatures so parity tests call the oracle and the GPU drop-in the same way.
Every function cites the reference lines it restates (paths relative to the
fullrmc source tree).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librefcpu.so")

_f = ctypes.c_float
_i = ctypes.c_int32
_int = ctypes.c_int
_p = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("%s not built (make -C oracle)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        protos = {
            "ref_distances_to_indexcoords": ([_i, _p, _i, _p, _int, _int, _p], None),
            "ref_distances_to_point": ([_p, _p, _i, _p, _int, _p], None),
            "ref_point_to_point_distance": ([_p, _p, _p, _int], _f),
            "ref_differences_to_indexcoords": ([_i, _p, _i, _p, _int, _int, _p], None),
            "ref_differences_to_point": ([_p, _p, _i, _p, _int, _p], None),
            "ref_from_to_points_differences": ([_p, _p, _i, _p, _int, _p], None),
            "ref_pair_difference": ([_p, _p, _p, _int, _p], None),
            "ref_single_pairs_histograms": ([_i, _p, _i, _p, _p, _i, _i, _p, _p, _f, _f, _f, _int],
                                            ctypes.c_int64),
            "ref_multiple_pairs_histograms_coords": ([_p, _i, _p, _i, _p, _int, _p, _p, _i, _f, _f, _f, _i, _int,
                                                     _p, _p, _int], ctypes.c_int64),
            "ref_multiple_pairs_histograms_dists": ([_p, _i, _p, _i, _i, _p, _p, _i, _f, _f, _f, _i, _int, _p, _p],
                                                    ctypes.c_int64),
            "ref_full_pairs_counts_f64": ([_p, _i, _p, _int, _p, _p, _i, _f, _f, _f, _i, _p, _int], ctypes.c_int64),
            "ref_full_pairs_counts_cells": ([_p, _i, _p, _p, _p, _i, _f, _f, _f, _i, _int, _p, _int], ctypes.c_int64),
            "ref_get_reciprocal_basis": ([_p, _p], ctypes.c_double),
            "ref_transform_coordinates": ([_p, _p, _i, _p], None),
            "ref_box_coordinates_real_distances": ([_i, _p, _i, _p, _p, _p], None),
            "ref_single_atomic_distances_dists": ([_i, _p, _i, _p, _p, _i, _p, _p, _p, _p, _p, _p, _int, _int],
                                                  None),
            "ref_multiple_atomic_distances_coords": ([_p, _i, _p, _i, _p, _int, _p, _p, _i, _p, _p, _int, _int,
                                                      _p, _p, _p, _p], None),
            "ref_multiple_atomic_distances_dists": ([_p, _i, _p, _i, _i, _p, _p, _i, _p, _p, _int, _int, _p, _p,
                                                     _p, _p], None),
            "ref_pair_elements_stats": ([_p, _p, _i, _i, _p, _p], None),
        }
        for name, (args, res) in protos.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------ pairs_distances

def pairs_distances_to_indexcoords(atomIndex, coords, basis, isPBC, allAtoms=True, ncores=1):
    """pairs_distances.pyx:874-916; unset entries are NaN here."""
    c, b = _c(coords, np.float32), _c(basis, np.float32)
    out = np.full(c.shape[0], np.nan, np.float32)
    lib().ref_distances_to_indexcoords(int(atomIndex), _ptr(c), c.shape[0], _ptr(b), int(bool(isPBC)),
                                       int(bool(allAtoms)), _ptr(out))
    return out


def pairs_distances_to_multi_indexcoords(indexes, coords, basis, isPBC, allAtoms=True, ncores=1):
    """pairs_distances.pyx:975-1021"""
    return np.stack([pairs_distances_to_indexcoords(i, coords, basis, isPBC, allAtoms) for i in indexes], axis=1) \
        if len(indexes) else np.empty((len(coords), 0), np.float32)


def pairs_distances_to_point(point, coords, basis, isPBC, ncores=1):
    """pairs_distances.pyx:827-871"""
    p, c, b = _c(point, np.float32), _c(coords, np.float32), _c(basis, np.float32)
    out = np.empty(c.shape[0], np.float32)
    lib().ref_distances_to_point(_ptr(p), _ptr(c), c.shape[0], _ptr(b), int(bool(isPBC)), _ptr(out))
    return out


def pairs_distances_to_multi_points(points, coords, basis, isPBC, ncores=1):
    """pairs_distances.pyx:924-972 (points is (3,k))"""
    pts = np.asarray(points, np.float32)
    if pts.shape[1] == 0:
        return np.empty((len(coords), 0), np.float32)
    return np.stack([pairs_distances_to_point(pts[:, k], coords, basis, isPBC) for k in range(pts.shape[1])],
                    axis=1)


def point_to_point_distance(point1, point2, basis, isPBC, ncores=1):
    """pairs_distances.pyx:786-824"""
    p1, p2, b = _c(point1, np.float32), _c(point2, np.float32), _c(basis, np.float32)
    return np.float32(lib().ref_point_to_point_distance(_ptr(p1), _ptr(p2), _ptr(b), int(bool(isPBC))))


def pairs_differences_to_indexcoords(atomIndex, coords, basis, isPBC, allAtoms=True, ncores=1):
    """pairs_distances.pyx:626-672; unset rows are NaN here."""
    c, b = _c(coords, np.float32), _c(basis, np.float32)
    out = np.full((c.shape[0], 3), np.nan, np.float32)
    lib().ref_differences_to_indexcoords(int(atomIndex), _ptr(c), c.shape[0], _ptr(b), int(bool(isPBC)),
                                         int(bool(allAtoms)), _ptr(out))
    return out


def pairs_differences_to_multi_indexcoords(indexes, coords, basis, isPBC, allAtoms=True, ncores=1):
    """pairs_distances.pyx:727-783"""
    if not len(indexes):
        return np.empty((len(coords), 3, 0), np.float32)
    return np.stack([pairs_differences_to_indexcoords(i, coords, basis, isPBC, allAtoms) for i in indexes],
                    axis=2)


def pairs_differences_to_point(point, coords, basis, isPBC, ncores=1):
    """pairs_distances.pyx:580-623"""
    p, c, b = _c(point, np.float32), _c(coords, np.float32), _c(basis, np.float32)
    out = np.empty((c.shape[0], 3), np.float32)
    lib().ref_differences_to_point(_ptr(p), _ptr(c), c.shape[0], _ptr(b), int(bool(isPBC)), _ptr(out))
    return out


def pairs_differences_to_multi_points(points, coords, basis, isPBC, ncores=1):
    """pairs_distances.pyx:675-724"""
    pts = np.asarray(points, np.float32)
    if pts.shape[1] == 0:
        return np.empty((len(coords), 3, 0), np.float32)
    return np.stack([pairs_differences_to_point(pts[:, k], coords, basis, isPBC) for k in range(pts.shape[1])],
                    axis=2)


def from_to_points_differences(pointsFrom, pointsTo, basis, isPBC, ncores=1):
    """pairs_distances.pyx:483-527"""
    f, t, b = _c(pointsFrom, np.float32), _c(pointsTo, np.float32), _c(basis, np.float32)
    out = np.empty((f.shape[0], 3), np.float32)
    lib().ref_from_to_points_differences(_ptr(f), _ptr(t), f.shape[0], _ptr(b), int(bool(isPBC)), _ptr(out))
    return out


def pair_difference_to_point(point1, point2, basis, isPBC, ncores=1):
    """pairs_distances.pyx:534-577"""
    p1, p2, b = _c(point1, np.float32), _c(point2, np.float32), _c(basis, np.float32)
    out = np.empty(3, np.float32)
    lib().ref_pair_difference(_ptr(p1), _ptr(p2), _ptr(b), int(bool(isPBC)), _ptr(out))
    return out


# ----------------------------------------------------------- pairs_histograms

def single_pairs_histograms(atomIndex, distances, moleculeIndex, elementIndex, hintra, hinter, minDistance,
                            maxDistance, bin, allAtoms=True, ncores=1):
    """pairs_histograms.pyx:77-141 (in place)"""
    d = _c(distances, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    assert hintra.flags.c_contiguous and hinter.flags.c_contiguous
    nel, H = hintra.shape[0], hintra.shape[2]
    return lib().ref_single_pairs_histograms(int(atomIndex), _ptr(d), d.shape[0], _ptr(m), _ptr(e), nel, H,
                                             _ptr(hintra), _ptr(hinter), np.float32(minDistance),
                                             np.float32(maxDistance), np.float32(bin), int(bool(allAtoms)))


def multiple_pairs_histograms_coords(indexes, boxCoords, basis, isPBC, moleculeIndex, elementIndex,
                                     numberOfElements, minDistance, maxDistance, bin, histSize, allAtoms=True,
                                     ncores=1, return_overflow=False):
    """pairs_histograms.pyx:150-217; ncores>1 uses per-thread private bins."""
    idx = _c(indexes, np.int32)
    c, b = _c(boxCoords, np.float32), _c(basis, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    nel, H = int(numberOfElements), int(histSize)
    hi = np.zeros((nel, nel, H), np.float32)
    he = np.zeros((nel, nel, H), np.float32)
    ov = lib().ref_multiple_pairs_histograms_coords(_ptr(idx), idx.shape[0], _ptr(c), c.shape[0], _ptr(b),
                                                    int(bool(isPBC)), _ptr(m), _ptr(e), nel, np.float32(minDistance),
                                                    np.float32(maxDistance), np.float32(bin), H, int(bool(allAtoms)),
                                                    _ptr(hi), _ptr(he), int(ncores))
    return (hi, he, ov) if return_overflow else (hi, he)


def full_pairs_histograms_coords(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                 minDistance, maxDistance, bin, histSize, ncores=1, return_overflow=False):
    """pairs_histograms.pyx:289-335"""
    n = len(boxCoords)
    return multiple_pairs_histograms_coords(np.arange(n, dtype=np.int32), boxCoords, basis, isPBC,
                                            moleculeIndex, elementIndex, numberOfElements, minDistance,
                                            maxDistance, bin, histSize, allAtoms=False, ncores=ncores,
                                            return_overflow=return_overflow)


def full_pairs_counts_f64(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements, minDistance,
                          maxDistance, bin, histSize, ncores=1):
    """Float64 recompute of the full histogram (the second oracle):
    returns (hintra, hinter) as int64 counts and the past-the-end count."""
    c, b = _c(boxCoords, np.float32), _c(basis, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    nel, H = int(numberOfElements), int(histSize)
    counts = np.zeros(2 * nel * nel * H, np.int32)
    ov = lib().ref_full_pairs_counts_f64(_ptr(c), c.shape[0], _ptr(b), int(bool(isPBC)), _ptr(m), _ptr(e), nel,
                                         np.float32(minDistance), np.float32(maxDistance), np.float32(bin), H,
                                         _ptr(counts), int(ncores))
    counts = counts.astype(np.int64).reshape(2, nel, nel, H)
    return counts[0], counts[1], ov


def full_pairs_counts_cells(boxCoords, basis, moleculeIndex, elementIndex, numberOfElements, minDistance,
                            maxDistance, bin, histSize, float64=False, ncores=None):
    """Periodic full histogram by a cell list (refcpu.c
    ref_full_pairs_counts_cells): the float32 reference binning of
    full_pairs_histograms_coords, or with float64=True the binary64 recompute
    of full_pairs_counts_f64, evaluating only pairs in neighbouring cells
    (the others are provably out of range).  Returns int64 (intra, inter)
    counts and the past-the-end count, or None when the cell is too small
    for three cells per axis.  ncores None: every CPU of the affinity set."""
    c, b = _c(boxCoords, np.float32), _c(basis, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    nel, H = int(numberOfElements), int(histSize)
    if ncores is None:
        ncores = len(os.sched_getaffinity(0))
    counts = np.zeros(2 * nel * nel * H, np.int32)
    ov = lib().ref_full_pairs_counts_cells(_ptr(c), c.shape[0], _ptr(b), _ptr(m), _ptr(e), nel,
                                           np.float32(minDistance), np.float32(maxDistance), np.float32(bin), H,
                                           int(bool(float64)), _ptr(counts), int(ncores))
    if ov < 0:
        return None
    counts = counts.astype(np.int64).reshape(2, nel, nel, H)
    return counts[0], counts[1], ov


def full_pairs_histograms_coords_fast(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                      minDistance, maxDistance, bin, histSize, ncores=None):
    """full_pairs_histograms_coords (pairs_histograms.pyx:289-335) through
    the cell list when the periodic cell holds three cells per axis, else the
    all-pairs loop on ncores threads: the same float32 (hintra, hinter)."""
    if ncores is None:
        ncores = len(os.sched_getaffinity(0))
    if isPBC:
        r = full_pairs_counts_cells(boxCoords, basis, moleculeIndex, elementIndex, numberOfElements, minDistance,
                                    maxDistance, bin, histSize, ncores=ncores)
        if r is not None:
            return r[0].astype(np.float32), r[1].astype(np.float32)
    return full_pairs_histograms_coords(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                        minDistance, maxDistance, bin, histSize, ncores=ncores)


def multiple_pairs_histograms_dists(indexes, distances, moleculeIndex, elementIndex, numberOfElements,
                                    minDistance, maxDistance, bin, histSize, allAtoms=True, ncores=1):
    """pairs_histograms.pyx:225-281"""
    idx = _c(indexes, np.int32)
    d = _c(distances, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    nel, H = int(numberOfElements), int(histSize)
    hi = np.zeros((nel, nel, H), np.float32)
    he = np.zeros((nel, nel, H), np.float32)
    lib().ref_multiple_pairs_histograms_dists(_ptr(idx), idx.shape[0], _ptr(d), d.shape[0], d.shape[1], _ptr(m),
                                              _ptr(e), nel, np.float32(minDistance), np.float32(maxDistance),
                                              np.float32(bin), H, int(bool(allAtoms)), _ptr(hi), _ptr(he))
    return hi, he


def full_pairs_histograms_dists(distances, moleculeIndex, elementIndex, numberOfElements, minDistance,
                                maxDistance, bin, histSize, ncores=1):
    """pairs_histograms.pyx:343-383"""
    k = np.asarray(distances).shape[1]
    return multiple_pairs_histograms_dists(np.arange(k, dtype=np.int32), distances, moleculeIndex, elementIndex,
                                           numberOfElements, minDistance, maxDistance, bin, histSize,
                                           allAtoms=False)


# --------------------------------------------- boundary_conditions_collection

def get_reciprocal_basis(basis):
    """boundary_conditions_collection.pyx:45-80 (with the :59 typo)"""
    b = _c(basis, np.float32)
    r = np.empty((3, 3), np.float32)
    vol = lib().ref_get_reciprocal_basis(_ptr(b), _ptr(r))
    return r, vol


def transform_coordinates(transMatrix, coords):
    """boundary_conditions_collection.pyx:88-110"""
    M, c = _c(transMatrix, np.float32), _c(coords, np.float32)
    out = np.empty((c.shape[0], 3), np.float32)
    lib().ref_transform_coordinates(_ptr(M), _ptr(c), c.shape[0], _ptr(out))
    return out


class MT(ctypes.Structure):
    """refcpu.c's textbook MT19937 (ref_mt_t), loadable from the state of
    CPython's random or NumPy's legacy RandomState."""
    _fields_ = [("mt", ctypes.c_uint32 * 624), ("pos", ctypes.c_int32)]

    @classmethod
    def from_python(cls, state):
        m = cls()
        for q in range(624):
            m.mt[q] = state[1][q]
        m.pos = state[1][624]
        return m

    @classmethod
    def from_numpy(cls, state):
        m = cls()
        for q in range(624):
            m.mt[q] = int(state[1][q])
        m.pos = int(state[2])
        return m

    def next32(self):
        f = lib().ref_mt_next32
        f.argtypes, f.restype = [ctypes.POINTER(MT)], ctypes.c_uint32
        return f(ctypes.byref(self))

    def double(self):
        f = lib().ref_mt_double
        f.argtypes, f.restype = [ctypes.POINTER(MT)], ctypes.c_double
        return f(ctypes.byref(self))

    def randbelow(self, n):
        f = lib().ref_mt_randbelow
        f.argtypes, f.restype = [ctypes.POINTER(MT), ctypes.c_int32], ctypes.c_int32
        return f(ctypes.byref(self), n)


def box_coordinates_real_distances(atomIndex, indexes, boxCoords, basis):
    """boundary_conditions_collection.pyx:119-165"""
    idx, c, b = _c(indexes, np.int32), _c(boxCoords, np.float32), _c(basis, np.float32)
    out = np.empty(idx.shape[0], np.float32)
    lib().ref_box_coordinates_real_distances(int(atomIndex), _ptr(idx), idx.shape[0], _ptr(c), _ptr(b), _ptr(out))
    return out


# ------------------------------------------------------------ atomic_distances

AD_INTER, AD_INTRA, AD_WITHIN, AD_RED_UPPER, AD_RED_LOWER, AD_RED = 1, 2, 4, 8, 16, 32


def ad_flags(interMolecular=True, intraMolecular=True, countWithinLimits=True, reduceDistanceToUpper=False,
             reduceDistanceToLower=False, reduceDistance=False):
    f = 0
    for on, bit in ((interMolecular, AD_INTER), (intraMolecular, AD_INTRA), (countWithinLimits, AD_WITHIN),
                    (reduceDistanceToUpper, AD_RED_UPPER), (reduceDistanceToLower, AD_RED_LOWER),
                    (reduceDistance, AD_RED)):
        if on:
            f |= bit
    return f


def single_atomic_distances_dists(atomIndex, distances, moleculeIndex, elementIndex, dintra, dinter, nintra,
                                  ninter, lowerLimit, upperLimit, flags, allAtoms=True):
    """atomic_distances.pyx:42-118 (in place on the (nel,nel,1) arrays)"""
    d = _c(distances, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    for a, t in ((dintra, np.float32), (dinter, np.float32), (nintra, np.int32), (ninter, np.int32)):
        assert a.flags.c_contiguous and a.dtype == t
    nel = dintra.shape[0]
    lo, up = _c(lowerLimit, np.float32), _c(upperLimit, np.float32)
    lib().ref_single_atomic_distances_dists(int(atomIndex), _ptr(d), d.shape[0], _ptr(m), _ptr(e), nel,
                                            _ptr(dintra), _ptr(dinter), _ptr(nintra), _ptr(ninter), _ptr(lo),
                                            _ptr(up), int(flags), int(bool(allAtoms)))


def multiple_atomic_distances_coords(indexes, boxCoords, basis, isPBC, moleculeIndex, elementIndex,
                                     numberOfElements, lowerLimit, upperLimit, flags, allAtoms=True):
    """atomic_distances.pyx:326-410 -> (nintra, dintra, ninter, dinter)"""
    idx = _c(indexes, np.int32)
    c, b = _c(boxCoords, np.float32), _c(basis, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    nel = int(numberOfElements)
    lo, up = _c(lowerLimit, np.float32), _c(upperLimit, np.float32)
    ni, ne = np.zeros((nel, nel, 1), np.int32), np.zeros((nel, nel, 1), np.int32)
    di, de = np.zeros((nel, nel, 1), np.float32), np.zeros((nel, nel, 1), np.float32)
    lib().ref_multiple_atomic_distances_coords(_ptr(idx), idx.shape[0], _ptr(c), c.shape[0], _ptr(b),
                                               int(bool(isPBC)), _ptr(m), _ptr(e), nel, _ptr(lo), _ptr(up),
                                               int(flags), int(bool(allAtoms)), _ptr(ni), _ptr(di), _ptr(ne),
                                               _ptr(de))
    return ni, di, ne, de


def full_atomic_distances_coords(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                 lowerLimit, upperLimit, flags):
    """atomic_distances.pyx:500-560"""
    n = len(boxCoords)
    return multiple_atomic_distances_coords(np.arange(n, dtype=np.int32), boxCoords, basis, isPBC, moleculeIndex,
                                            elementIndex, numberOfElements, lowerLimit, upperLimit, flags,
                                            allAtoms=False)


def multiple_atomic_distances_dists(indexes, distances, moleculeIndex, elementIndex, numberOfElements, lowerLimit,
                                    upperLimit, flags, allAtoms=True):
    """atomic_distances.pyx:419-491"""
    idx = _c(indexes, np.int32)
    d = _c(distances, np.float32)
    m, e = _c(moleculeIndex, np.int32), _c(elementIndex, np.int32)
    nel = int(numberOfElements)
    lo, up = _c(lowerLimit, np.float32), _c(upperLimit, np.float32)
    ni, ne = np.zeros((nel, nel, 1), np.int32), np.zeros((nel, nel, 1), np.int32)
    di, de = np.zeros((nel, nel, 1), np.float32), np.zeros((nel, nel, 1), np.float32)
    lib().ref_multiple_atomic_distances_dists(_ptr(idx), idx.shape[0], _ptr(d), d.shape[0], d.shape[1], _ptr(m),
                                              _ptr(e), nel, _ptr(lo), _ptr(up), int(flags), int(bool(allAtoms)),
                                              _ptr(ni), _ptr(di), _ptr(ne), _ptr(de))
    return ni, di, ne, de


def pair_elements_stats(elementIndex, moleculeIndex, numberOfElements):
    """atomic_distances.pyx:632-675 -> (nintra, ninter)"""
    e, m = _c(elementIndex, np.int32), _c(moleculeIndex, np.int32)
    nel = int(numberOfElements)
    ni, ne = np.zeros((nel, nel, 1), np.int32), np.zeros((nel, nel, 1), np.int32)
    lib().ref_pair_elements_stats(_ptr(e), _ptr(m), e.shape[0], nel, _ptr(ni), _ptr(ne))
    return ni, ne
