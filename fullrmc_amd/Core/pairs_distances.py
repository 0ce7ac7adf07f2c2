"""Drop-in for fullrmc's ``fullrmc.Core.pairs_distances`` extension
(Extensions/pairs_distances.pyx): all eleven Python-visible functions, same
signatures, computed by ``libfrmc_hip.so`` on gfx950.

Periodic (``isPBC``) inputs are box (fractional) coordinates: the
difference is wrapped by ``d - round(d)`` with the reference's rounding
(``floor(x+0.5)`` for x>0, else ``ceil(x-0.5)``) and mapped through the
basis rows; distances are the correctly rounded float32 square root of
``(rx*rx + ry*ry) + rz*rz``, bit-identical to the reference.  Infinite
boundary inputs are real coordinates.  ``ncores`` is accepted and ignored.
"""
import numpy as np

from .. import _args as A
from .. import _device as D
from .._lib import check, lib

__all__ = ["from_to_points_differences", "pair_difference_to_point", "pairs_differences_to_point",
           "pairs_differences_to_indexcoords", "pairs_differences_to_multi_points",
           "pairs_differences_to_multi_indexcoords", "point_to_point_distance", "pairs_distances_to_point",
           "pairs_distances_to_indexcoords", "pairs_distances_to_multi_points",
           "pairs_distances_to_multi_indexcoords"]


def _coords(name, coords):
    coords = A.array(name, coords, np.float32, 2)
    if coords.shape[1] != 3:
        raise ValueError("%s must be (n,3)" % name)
    return coords


def _basis(basis):
    basis = A.memview("basis", basis, np.float32, 2)
    if basis.shape != (3, 3):
        raise ValueError("basis must be (3,3)")
    return basis


def _point(name, p):
    p = A.array(name, p, np.float32, 1)
    if p.shape[0] < 3:
        raise IndexError("%s must hold 3 coordinates" % name)
    return p[:3]


def from_to_points_differences(pointsFrom, pointsTo, basis, isPBC, ncores=1):
    """differences[i,:] = BC(pointsTo[i,:] - pointsFrom[i,:]), (n,3)
    (pairs_distances.pyx:483-527)."""
    pf, pt, b = _coords("pointsFrom", pointsFrom), _coords("pointsTo", pointsTo), _basis(basis)
    A.c_int32("ncores", ncores)
    if pt.shape[0] < pf.shape[0]:
        raise IndexError("pointsTo shorter than pointsFrom")
    n = pf.shape[0]
    d_f, d_t, d_b = D.upload(pf), D.upload(pt[:n]), D.upload(b)
    out = D.empty((n, 3), np.float32)
    check(lib().frmc_from_to_points_differences(d_f.data_ptr(), d_t.data_ptr(), n, d_b.data_ptr(),
                                                int(A.bint(isPBC)), out.data_ptr(), D.stream()))
    return D.download(out)


def pair_difference_to_point(point1, point2, basis, isPBC, ncores=1):
    """BC(point2 - point1), shape (3,) (pairs_distances.pyx:534-577)."""
    p1, p2, b = _point("point1", point1), _point("point2", point2), _basis(basis)
    A.c_int32("ncores", ncores)
    d1, d2, d_b = D.upload(p1), D.upload(p2), D.upload(b)
    out = D.empty((3,), np.float32)
    check(lib().frmc_pair_difference_to_point(d1.data_ptr(), d2.data_ptr(), d_b.data_ptr(), int(A.bint(isPBC)),
                                              out.data_ptr(), D.stream()))
    return D.download(out)


def pairs_differences_to_point(point, coords, basis, isPBC, ncores=1):
    """differences[i,:] = BC(point - coords[i,:]), (n,3) (pairs_distances.pyx:580-623)."""
    p, c, b = _point("point", point), _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    n = c.shape[0]
    d_p, d_c, d_b = D.upload(p), D.upload(c), D.upload(b)
    out = D.empty((n, 3), np.float32)
    check(lib().frmc_pairs_differences_to_point(d_p.data_ptr(), d_c.data_ptr(), n, d_b.data_ptr(),
                                                int(A.bint(isPBC)), out.data_ptr(), D.stream()))
    return D.download(out)


def pairs_differences_to_indexcoords(atomIndex, coords, basis, isPBC, allAtoms=True, ncores=1):
    """differences[i,:] = BC(coords[atomIndex,:] - coords[i,:]), (n,3); rows
    before atomIndex are left unset when allAtoms is False
    (pairs_distances.pyx:626-672)."""
    atomIndex = A.c_int32("atomIndex", atomIndex)
    c, b = _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    n = c.shape[0]
    A.check_index("atomIndex", atomIndex, n)
    d_c, d_b = D.upload(c), D.upload(b)
    out = D.empty((n, 3), np.float32)
    check(lib().frmc_pairs_differences_to_indexcoords(atomIndex, d_c.data_ptr(), n, d_b.data_ptr(),
                                                      int(A.bint(isPBC)), int(A.bint(allAtoms)), out.data_ptr(),
                                                      D.stream()))
    return D.download(out)


def pairs_differences_to_multi_points(points, coords, basis, isPBC, ncores=1):
    """differences[i,:,k] = BC(points[:,k] - coords[i,:]), (n,3,k).  points
    is (3,k): the reference iterates ``points[:,i]`` (pairs_distances.pyx:675-724)."""
    pts = A.array("points", points, np.float32, 2)
    c, b = _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    if pts.shape[0] < 3:
        raise IndexError("points must be (3,k)")
    k, n = pts.shape[1], c.shape[0]
    d_p, d_c, d_b = D.upload(pts[:3]), D.upload(c), D.upload(b)
    out = D.empty((n, 3, k), np.float32)
    check(lib().frmc_pairs_differences_to_multi_points(d_p.data_ptr(), k, d_c.data_ptr(), n, d_b.data_ptr(),
                                                       int(A.bint(isPBC)), out.data_ptr(), D.stream()))
    return D.download(out)


def pairs_differences_to_multi_indexcoords(indexes, coords, basis, isPBC, allAtoms=True, ncores=1):
    """differences[i,:,k] = BC(coords[indexes[k],:] - coords[i,:]), (n,3,k)
    (pairs_distances.pyx:727-783)."""
    idx = A.array("indexes", indexes, np.int32, 1)
    c, b = _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    n, k = c.shape[0], idx.shape[0]
    A.check_indexes("indexes", idx, n)
    d_i, d_c, d_b = D.upload(idx), D.upload(c), D.upload(b)
    out = D.empty((n, 3, k), np.float32)
    check(lib().frmc_pairs_differences_to_multi_indexcoords(d_i.data_ptr(), k, d_c.data_ptr(), n, d_b.data_ptr(),
                                                            int(A.bint(isPBC)), int(A.bint(allAtoms)),
                                                            out.data_ptr(), D.stream()))
    return D.download(out)


def point_to_point_distance(point1, point2, basis, isPBC, ncores=1):
    """|BC(point1 - point2)| as a float32 scalar (pairs_distances.pyx:786-824)."""
    p1, p2, b = _point("point1", point1), _point("point2", point2), _basis(basis)
    A.c_int32("ncores", ncores)
    d1, d2, d_b = D.upload(p1), D.upload(p2), D.upload(b)
    out = D.empty((1,), np.float32)
    check(lib().frmc_point_to_point_distance(d1.data_ptr(), d2.data_ptr(), d_b.data_ptr(), int(A.bint(isPBC)),
                                             out.data_ptr(), D.stream()))
    return np.float32(D.download(out)[0])


def pairs_distances_to_point(point, coords, basis, isPBC, ncores=1):
    """distances[i] = |BC(point - coords[i,:])|, (n,) (pairs_distances.pyx:827-871)."""
    p, c, b = _point("point", point), _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    n = c.shape[0]
    d_p, d_c, d_b = D.upload(p), D.upload(c), D.upload(b)
    out = D.empty((n,), np.float32)
    check(lib().frmc_pairs_distances_to_point(d_p.data_ptr(), d_c.data_ptr(), n, d_b.data_ptr(),
                                              int(A.bint(isPBC)), out.data_ptr(), D.stream()))
    return D.download(out)


def pairs_distances_to_indexcoords(atomIndex, coords, basis, isPBC, allAtoms=True, ncores=1):
    """distances[i] = |BC(coords[atomIndex,:] - coords[i,:])|, (n,); entries
    before atomIndex are left unset when allAtoms is False
    (pairs_distances.pyx:874-916)."""
    atomIndex = A.c_int32("atomIndex", atomIndex)
    c, b = _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    n = c.shape[0]
    A.check_index("atomIndex", atomIndex, n)
    d_c, d_b = D.upload(c), D.upload(b)
    out = D.empty((n,), np.float32)
    check(lib().frmc_pairs_distances_to_indexcoords(atomIndex, d_c.data_ptr(), n, d_b.data_ptr(),
                                                    int(A.bint(isPBC)), int(A.bint(allAtoms)), out.data_ptr(),
                                                    D.stream()))
    return D.download(out)


def pairs_distances_to_multi_points(points, coords, basis, isPBC, ncores=1):
    """distances[i,k] = |BC(points[:,k] - coords[i,:])|, (n,k); points is (3,k)
    (pairs_distances.pyx:924-972)."""
    pts = A.array("points", points, np.float32, 2)
    c, b = _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    if pts.shape[0] < 3:
        raise IndexError("points must be (3,k)")
    k, n = pts.shape[1], c.shape[0]
    d_p, d_c, d_b = D.upload(pts[:3]), D.upload(c), D.upload(b)
    out = D.empty((n, k), np.float32)
    check(lib().frmc_pairs_distances_to_multi_points(d_p.data_ptr(), k, d_c.data_ptr(), n, d_b.data_ptr(),
                                                     int(A.bint(isPBC)), out.data_ptr(), D.stream()))
    return D.download(out)


def pairs_distances_to_multi_indexcoords(indexes, coords, basis, isPBC, allAtoms=True, ncores=1):
    """distances[i,k] = |BC(coords[indexes[k],:] - coords[i,:])|, (n,k)
    (pairs_distances.pyx:975-1021)."""
    idx = A.array("indexes", indexes, np.int32, 1)
    c, b = _coords("coords", coords), _basis(basis)
    A.c_int32("ncores", ncores)
    n, k = c.shape[0], idx.shape[0]
    A.check_indexes("indexes", idx, n)
    d_i, d_c, d_b = D.upload(idx), D.upload(c), D.upload(b)
    out = D.empty((n, k), np.float32)
    check(lib().frmc_pairs_distances_to_multi_indexcoords(d_i.data_ptr(), k, d_c.data_ptr(), n, d_b.data_ptr(),
                                                          int(A.bint(isPBC)), int(A.bint(allAtoms)),
                                                          out.data_ptr(), D.stream()))
    return D.download(out)
