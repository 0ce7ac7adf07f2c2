"""Drop-in for fullrmc's ``fullrmc.Core.atomic_distances`` extension
(Extensions/atomic_distances.pyx), computed by the gfx950 kernels of
``libfrmc_hip.so``.

Same functions, positional order, defaults, dtypes and return values as the
reference.  Limit arrays and outputs are (nel, nel, 1).  The float32
distance sums are added in the reference's order (atoms in ``indexes``
order, other atoms by index), so they are bit-identical to its ``ncores=1``
result; ``ncores`` is accepted and ignored.
"""
import numpy as np

from .. import _args as A
from .. import _device as D
from .._lib import check, lib

__all__ = ["single_atomic_distances_dists_serial", "single_atomic_distances_dists",
           "multiple_atomic_distances_coords", "multiple_atomic_distances_dists", "full_atomic_distances_coords",
           "full_atomic_distances_dists", "pair_elements_stats"]

INTER, INTRA, WITHIN, RED_UPPER, RED_LOWER, RED = 1, 2, 4, 8, 16, 32


def _flags(interMolecular, intraMolecular, countWithinLimits, reduceDistanceToUpper, reduceDistanceToLower,
           reduceDistance):
    f = 0
    for on, bit in ((interMolecular, INTER), (intraMolecular, INTRA), (countWithinLimits, WITHIN),
                    (reduceDistanceToUpper, RED_UPPER), (reduceDistanceToLower, RED_LOWER), (reduceDistance, RED)):
        if A.bint(on):
            f |= bit
    return f


def _limits(lowerLimit, upperLimit, nel):
    lo = A.array("lowerLimit", lowerLimit, np.float32, 3)
    up = A.array("upperLimit", upperLimit, np.float32, 3)
    shape = lo.shape
    assert shape[0] == shape[1] and shape[0] == nel, \
        "lowerLimit array must have numberOfElements columns and numberOfElements rows"
    assert shape[2] == 1, "lowerLimit array third dimension must have a length of exactly 1"
    if up.shape != lo.shape:
        raise IndexError("upperLimit must have the shape of lowerLimit")
    return lo, up


def _per_atom(name, arr, n):
    if arr.shape[0] < n:
        raise IndexError("%s is shorter than the number of atoms (%d)" % (name, n))
    return arr[:n]


def _outputs(nel):
    return (D.zeros((nel, nel, 1), np.int32), D.zeros((nel, nel, 1), np.float32),
            D.zeros((nel, nel, 1), np.int32), D.zeros((nel, nel, 1), np.float32))


def _single(atomIndex, distances, moleculeIndex, elementIndex, dintra, dinter, nintra, ninter, lowerLimit,
            upperLimit, flags, allAtoms):
    atomIndex = A.c_int32("atomIndex", atomIndex)
    distances = A.array("distances", distances, np.float32, 1)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    dintra = A.array("dintra", dintra, np.float32, 3)
    dinter = A.array("dinter", dinter, np.float32, 3)
    nintra = A.array("nintra", nintra, np.int32, 3)
    ninter = A.array("ninter", ninter, np.int32, 3)
    nel = dintra.shape[0]
    lo = A.array("lowerLimit", lowerLimit, np.float32, 3)
    up = A.array("upperLimit", upperLimit, np.float32, 3)
    for name, arr in (("dinter", dinter), ("nintra", nintra), ("ninter", ninter), ("lowerLimit", lo),
                      ("upperLimit", up)):
        if arr.shape[0] < nel or arr.shape[1] < nel:
            raise IndexError("%s must be (nel,nel,1) with nel = %d" % (name, nel))
    n = distances.shape[0]
    A.check_index("atomIndex", atomIndex, n)
    mol, el = _per_atom("moleculeIndex", moleculeIndex, n), _per_atom("elementIndex", elementIndex, n)
    A.check_indexes("elementIndex", el, nel)
    d = [D.upload(x) for x in (distances, mol, el, dintra, dinter, nintra, ninter, lo[:nel, :nel, :1],
                               up[:nel, :nel, :1])]
    check(lib().frmc_single_atomic_distances_dists(atomIndex, d[0].data_ptr(), n, d[1].data_ptr(), d[2].data_ptr(),
                                                   d[3].data_ptr(), d[4].data_ptr(), d[5].data_ptr(),
                                                   d[6].data_ptr(), nel, d[7].data_ptr(), d[8].data_ptr(), flags,
                                                   int(A.bint(allAtoms)), D.stream()))
    for host, dev in ((dintra, d[3]), (dinter, d[4]), (nintra, d[5]), (ninter, d[6])):
        np.copyto(host, D.download(dev).reshape(host.shape))


def single_atomic_distances_dists_serial(atomIndex, distances, moleculeIndex, elementIndex, dintra, dinter, nintra,
                                         ninter, lowerLimit, upperLimit, interMolecular=True, intraMolecular=True,
                                         countWithinLimits=True, reduceDistanceToUpper=False,
                                         reduceDistanceToLower=False, reduceDistance=False, allAtoms=True):
    """Counted distances of one atom from its distances array, added into
    dintra/dinter/nintra/ninter in place (atomic_distances.pyx:128-232)."""
    _single(atomIndex, distances, moleculeIndex, elementIndex, dintra, dinter, nintra, ninter, lowerLimit,
            upperLimit, _flags(interMolecular, intraMolecular, countWithinLimits, reduceDistanceToUpper,
                               reduceDistanceToLower, reduceDistance), allAtoms)


def single_atomic_distances_dists(atomIndex, distances, moleculeIndex, elementIndex, dintra, dinter, nintra,
                                  ninter, lowerLimit, upperLimit, interMolecular=True, intraMolecular=True,
                                  countWithinLimits=True, reduceDistanceToUpper=False, reduceDistanceToLower=False,
                                  reduceDistance=False, allAtoms=True, ncores=1):
    """As single_atomic_distances_dists_serial (atomic_distances.pyx:241-318)."""
    A.c_int32("ncores", ncores)
    _single(atomIndex, distances, moleculeIndex, elementIndex, dintra, dinter, nintra, ninter, lowerLimit,
            upperLimit, _flags(interMolecular, intraMolecular, countWithinLimits, reduceDistanceToUpper,
                               reduceDistanceToLower, reduceDistance), allAtoms)


def multiple_atomic_distances_coords(indexes, boxCoords, basis, isPBC, moleculeIndex, elementIndex,
                                     numberOfElements, lowerLimit, upperLimit, interMolecular=True,
                                     intraMolecular=True, countWithinLimits=True, reduceDistanceToUpper=False,
                                     reduceDistanceToLower=False, reduceDistance=False, allAtoms=True, ncores=1):
    """``(nintra, dintra, ninter, dinter)`` of the atoms in ``indexes`` vs
    all atoms (atomic_distances.pyx:326-410)."""
    indexes = A.array("indexes", indexes, np.int32, 1)
    boxCoords = A.array("boxCoords", boxCoords, np.float32, 2)
    basis = A.array("basis", basis, np.float32, 2)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    nel = A.c_int32("numberOfElements", numberOfElements)
    lo, up = _limits(lowerLimit, upperLimit, nel)
    A.c_int32("ncores", ncores)
    isPBC = A.bint(isPBC)
    if boxCoords.shape[1] != 3 or basis.shape != (3, 3):
        raise IndexError("boxCoords must be (n,3) and basis (3,3)")
    n = boxCoords.shape[0]
    A.check_indexes("indexes", indexes, n)
    mol, el = _per_atom("moleculeIndex", moleculeIndex, n), _per_atom("elementIndex", elementIndex, n)
    A.check_indexes("elementIndex", el, nel)
    flags = _flags(interMolecular, intraMolecular, countWithinLimits, reduceDistanceToUpper, reduceDistanceToLower,
                   reduceDistance)
    ni, di, ne, de = _outputs(nel)
    d = [D.upload(x) for x in (indexes, boxCoords, basis, mol, el, lo, up)]
    check(lib().frmc_multiple_atomic_distances_coords(d[0].data_ptr(), indexes.shape[0], d[1].data_ptr(), n,
                                                      d[2].data_ptr(), int(isPBC), d[3].data_ptr(), d[4].data_ptr(),
                                                      nel, d[5].data_ptr(), d[6].data_ptr(), flags,
                                                      int(A.bint(allAtoms)), ni.data_ptr(), di.data_ptr(),
                                                      ne.data_ptr(), de.data_ptr(), D.stream()))
    return D.download(ni), D.download(di), D.download(ne), D.download(de)


def multiple_atomic_distances_dists(indexes, distances, moleculeIndex, elementIndex, numberOfElements, lowerLimit,
                                    upperLimit, interMolecular=True, intraMolecular=True, countWithinLimits=True,
                                    reduceDistanceToUpper=False, reduceDistanceToLower=False, reduceDistance=False,
                                    allAtoms=True, ncores=1):
    """The same from an (n,k) distances array, column c for atom
    indexes[c] (atomic_distances.pyx:419-491)."""
    indexes = A.array("indexes", indexes, np.int32, 1)
    distances = A.array("distances", distances, np.float32, 2)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    nel = A.c_int32("numberOfElements", numberOfElements)
    lo, up = _limits(lowerLimit, upperLimit, nel)
    A.c_int32("ncores", ncores)
    n, k = distances.shape
    if indexes.shape[0] > k:
        raise IndexError("indexes has more entries than distances has columns")
    A.check_indexes("indexes", indexes, n)
    mol, el = _per_atom("moleculeIndex", moleculeIndex, n), _per_atom("elementIndex", elementIndex, n)
    A.check_indexes("elementIndex", el, nel)
    flags = _flags(interMolecular, intraMolecular, countWithinLimits, reduceDistanceToUpper, reduceDistanceToLower,
                   reduceDistance)
    ni, di, ne, de = _outputs(nel)
    d = [D.upload(x) for x in (indexes, distances, mol, el, lo, up)]
    check(lib().frmc_multiple_atomic_distances_dists(d[0].data_ptr(), indexes.shape[0], d[1].data_ptr(), n, k,
                                                     d[2].data_ptr(), d[3].data_ptr(), nel, d[4].data_ptr(),
                                                     d[5].data_ptr(), flags, int(A.bint(allAtoms)), ni.data_ptr(),
                                                     di.data_ptr(), ne.data_ptr(), de.data_ptr(), D.stream()))
    return D.download(ni), D.download(di), D.download(ne), D.download(de)


def full_atomic_distances_coords(boxCoords, basis, isPBC, moleculeIndex, elementIndex, numberOfElements,
                                 lowerLimit, upperLimit, interMolecular=True, intraMolecular=True,
                                 reduceDistanceToUpper=False, reduceDistanceToLower=False, reduceDistance=False,
                                 countWithinLimits=True, ncores=1):
    """Every unordered pair once (atomic_distances.pyx:500-560)."""
    boxCoords = A.array("boxCoords", boxCoords, np.float32, 2)
    indexes = np.arange(boxCoords.shape[0], dtype=np.int32)
    return multiple_atomic_distances_coords(indexes, boxCoords, basis, isPBC, moleculeIndex, elementIndex,
                                            numberOfElements, lowerLimit, upperLimit, interMolecular,
                                            intraMolecular, countWithinLimits, reduceDistanceToUpper,
                                            reduceDistanceToLower, reduceDistance, False, ncores)


def full_atomic_distances_dists(distances, moleculeIndex, elementIndex, numberOfElements, lowerLimit, upperLimit,
                                interMolecular=True, intraMolecular=True, reduceDistanceToUpper=False,
                                reduceDistanceToLower=False, reduceDistance=False, countWithinLimits=True):
    """Every pair once from an (n,k) distances array, indexes = arange(k)
    (atomic_distances.pyx:569-623)."""
    distances = A.array("distances", distances, np.float32, 2)
    indexes = np.arange(distances.shape[1], dtype=np.int32)
    return multiple_atomic_distances_dists(indexes, distances, moleculeIndex, elementIndex, numberOfElements,
                                           lowerLimit, upperLimit, interMolecular, intraMolecular,
                                           countWithinLimits, reduceDistanceToUpper, reduceDistanceToLower,
                                           reduceDistance, False)


def pair_elements_stats(elementIndex, moleculeIndex, numberOfElements):
    """``(nintra, ninter)``: per element pair, the number of pairs i<j in the
    same and in different molecules (atomic_distances.pyx:659-675)."""
    elementIndex = A.array("elementIndex", elementIndex, np.int32, 1)
    moleculeIndex = A.array("moleculeIndex", moleculeIndex, np.int32, 1)
    nel = A.c_int32("numberOfElements", numberOfElements)
    n = elementIndex.shape[0]
    mol = _per_atom("moleculeIndex", moleculeIndex, n)
    A.check_indexes("elementIndex", elementIndex, nel)
    ni, ne = D.zeros((nel, nel, 1), np.int32), D.zeros((nel, nel, 1), np.int32)
    d_e, d_m = D.upload(elementIndex), D.upload(mol)
    check(lib().frmc_pair_elements_stats(d_e.data_ptr(), d_m.data_ptr(), n, nel, ni.data_ptr(), ne.data_ptr(),
                                         D.stream()))
    return D.download(ni), D.download(ne)
