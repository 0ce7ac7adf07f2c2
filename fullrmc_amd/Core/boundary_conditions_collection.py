"""Drop-in for fullrmc's ``fullrmc.Core.boundary_conditions_collection``
extension (Extensions/boundary_conditions_collection.pyx) on gfx950."""
import ctypes

import numpy as np

from .. import _args as A
from .. import _device as D
from .._lib import check, lib

__all__ = ["get_reciprocal_basis", "transform_coordinates", "box_coordinates_real_distances"]


def get_reciprocal_basis(basis):
    """``(rbasis, vol)`` by cofactors (boundary_conditions_collection.pyx:45-80).

    The reference's cofactor for ``rbasis[0,1]`` reads ``basis[0,1]`` where
    ``basis[0,2]`` belongs (:59); that result is reproduced, so this is not
    a correct reciprocal basis for general cells, exactly like the
    reference.  The engine's setup does not use it."""
    b = A.array("basis", basis, np.float32, 2)
    if b.shape != (3, 3):
        raise IndexError("basis must be (3,3)")
    d_b = D.upload(b)
    out = D.empty((3, 3), np.float32)
    vol = ctypes.c_double(0.0)
    check(lib().frmc_get_reciprocal_basis(d_b.data_ptr(), out.data_ptr(), ctypes.byref(vol), D.stream()))
    return D.download(out), float(vol.value)


def transform_coordinates(transMatrix, coords):
    """out[i,c] = (x*M[0,c] + y*M[1,c]) + z*M[2,c] (boundary_conditions_collection.pyx:88-110)."""
    M = A.array("transMatrix", transMatrix, np.float32, 2)
    c = A.array("coords", coords, np.float32, 2)
    if M.shape != (3, 3) or c.shape[1] != 3:
        raise IndexError("transMatrix must be (3,3) and coords (n,3)")
    n = c.shape[0]
    d_m, d_c = D.upload(M), D.upload(c)
    out = D.empty((n, 3), np.float32)
    check(lib().frmc_transform_coordinates(d_m.data_ptr(), d_c.data_ptr(), n, out.data_ptr(), D.stream()))
    return D.download(out)


def box_coordinates_real_distances(atomIndex, indexes, boxCoords, basis):
    """distances[ii] = |BC(boxCoords[indexes[ii]] - boxCoords[atomIndex])|
    (boundary_conditions_collection.pyx:119-165)."""
    atomIndex = A.c_int32("atomIndex", atomIndex)
    idx = A.array("indexes", indexes, np.int32, 1)
    c = A.array("boxCoords", boxCoords, np.float32, 2)
    b = A.array("basis", basis, np.float32, 2)
    if b.shape != (3, 3) or c.shape[1] != 3:
        raise IndexError("basis must be (3,3) and boxCoords (n,3)")
    n = c.shape[0]
    A.check_index("atomIndex", atomIndex, n)
    A.check_indexes("indexes", idx, n)
    d_i, d_c, d_b = D.upload(idx), D.upload(c), D.upload(b)
    out = D.empty((idx.shape[0],), np.float32)
    check(lib().frmc_box_coordinates_real_distances(atomIndex, d_i.data_ptr(), idx.shape[0], d_c.data_ptr(), n,
                                                    d_b.data_ptr(), out.data_ptr(), D.stream()))
    return D.download(out)
