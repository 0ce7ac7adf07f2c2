"""Argument checks with the exception types of the reference's Cython
signatures (typed ``ndarray[...]`` buffers, ``not None``, C scalar
conversions), so callers see the same failures for the same misuse:

* ``None`` for a ``not None`` argument -> ``TypeError``
* a non-ndarray where an ndarray is declared -> ``TypeError``
* wrong dtype -> ``ValueError("Buffer dtype mismatch ...")``
* wrong number of dimensions -> ``ValueError("Buffer has wrong number of dimensions ...")``

Out-of-range atom indexes are undefined behaviour in the reference (no
bounds checks); here they raise ``IndexError`` before any kernel runs.
"""
import numpy as np

_CNAME = {np.dtype(np.float32): "C_FLOAT32", np.dtype(np.int32): "C_INT32"}
_FMT = {np.dtype(np.float64): "double", np.dtype(np.float32): "float", np.dtype(np.int64): "long",
        np.dtype(np.int32): "int", np.dtype(np.int16): "short", np.dtype(np.uint8): "unsigned char",
        np.dtype(np.bool_): "'bool'", np.dtype(np.uint32): "unsigned int"}


def array(name, value, dtype, ndim, allow_none=False):
    """Validate a typed buffer argument and return it as a C-contiguous array."""
    if value is None:
        if allow_none:
            return None
        raise TypeError("Argument '%s' must not be None" % name)
    if not isinstance(value, np.ndarray):
        raise TypeError("Argument '%s' has incorrect type (expected numpy.ndarray, got %s)"
                        % (name, type(value).__name__))
    dtype = np.dtype(dtype)
    if value.dtype != dtype:
        raise ValueError("Buffer dtype mismatch, expected '%s' but got %s"
                         % (_CNAME.get(dtype, dtype.name), _FMT.get(value.dtype, "'%s'" % value.dtype.name)))
    if value.ndim != ndim:
        raise ValueError("Buffer has wrong number of dimensions (expected %d, got %d)" % (ndim, value.ndim))
    return value


def memview(name, value, dtype, ndim):
    """A ``C_FLOAT32[:,:] x not None`` memoryview argument."""
    if value is None:
        raise TypeError("Argument '%s' must not be None" % name)
    value = np.asarray(value) if not isinstance(value, np.ndarray) else value
    return array(name, value, dtype, ndim)


def c_int32(name, value):
    """Python object -> C int32 as Cython converts it."""
    if isinstance(value, (float, np.floating)):
        raise TypeError("'%s' object cannot be interpreted as an integer" % type(value).__name__)
    v = int(value)
    if v < -2 ** 31 or v >= 2 ** 31:
        raise OverflowError("value too large to convert to int")
    return v


def c_float32(name, value):
    return np.float32(float(value))


def bint(value):
    return bool(value)


def check_indexes(name, idx, n):
    if idx.size and (int(idx.min()) < 0 or int(idx.max()) >= n):
        raise IndexError("%s holds an index outside [0,%d)" % (name, n))


def check_index(name, i, n):
    if i < 0 or i >= n:
        raise IndexError("%s %d outside [0,%d)" % (name, i, n))
