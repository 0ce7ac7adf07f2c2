"""ctypes binding of ``libfrmc_hip.so`` (the gfx950 C ABI declared in
``include/frmc.h``).

The library is loaded lazily and exactly once.  There is no fallback: if the
shared object is missing, or no ROCm GPU is visible when a kernel is called,
a :class:`FrmcError` is raised.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfrmc_hip.so")

c_f32p = ctypes.POINTER(ctypes.c_float)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_vp = ctypes.c_void_p
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_f32 = ctypes.c_float
c_int = ctypes.c_int

ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ENOMEM", 4: "ESTATE"}


class FrmcError(RuntimeError):
    """An error reported by the HIP library (argument, runtime or memory)."""


class MT19937(ctypes.Structure):
    """``frmc_mt19937``: 624 state words and the read position, the layout of
    ``random.getstate()`` and ``numpy.random.get_state()``."""
    _fields_ = [("mt", ctypes.c_uint32 * 624), ("pos", ctypes.c_int32)]


class RunStats(ctypes.Structure):
    _fields_ = [("generated", c_i64), ("tried", c_i64), ("accepted", c_i64),
                ("tolerated", c_i64), ("windows", c_i64), ("evaluated", c_i64)]


class ConstraintDesc(ctypes.Structure):
    _fields_ = [("kind", c_i32), ("histSize", c_i32), ("minDistance", c_f32),
                ("maxDistance", c_f32), ("bin", c_f32), ("nQ", c_i32),
                ("pairWeights", c_f32p), ("pairDensity", c_f32p),
                ("shellVolumes", c_f32p), ("prefactor", c_f32p),
                ("experimental", c_f32p), ("dataWeights", c_f32p),
                ("sqMatrix", c_f32p), ("scaleFactor", c_f32)]


class EngineDesc(ctypes.Structure):
    _fields_ = [("n", c_i32), ("nel", c_i32), ("isPBC", c_i32),
                ("box", c_f32p), ("real", c_f32p), ("mol", c_i32p), ("elem", c_i32p),
                ("basis", c_f32 * 9), ("recip", c_f32 * 9),
                ("ngroups", c_i32), ("group_offsets", c_i32p), ("group_atoms", c_i32p),
                ("nconstraints", c_i32), ("constraints", ConstraintDesc * 4),
                ("kmax", c_i32), ("stream", c_vp)]


# name -> argtypes (all return int unless listed in _RESTYPES)
_PROTOS = {
    "frmc_last_error": [],
    "frmc_version": [],
    "frmc_last_overflow": [],
    "frmc_set_timing": [c_int],
    "frmc_last_kernel_ms": [],
    "frmc_set_autotune": [c_int],
    "frmc_set_replica_batching": [c_int],
    "frmc_set_replica_threads": [c_int],
    "frmc_replica_batch_times": [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_i64)],
    "frmc_force_full_schedule": [c_i32, c_i32, c_i32],
    "frmc_last_full_schedule": [c_i32p, c_i32p, ctypes.POINTER(c_i64), c_i32p],
    "frmc_last_full_evaluated": [ctypes.POINTER(c_i64)],
    # pairs_distances
    "frmc_pairs_distances_to_indexcoords": [c_i32, c_vp, c_i32, c_vp, c_int, c_int, c_vp, c_vp],
    "frmc_pairs_distances_to_multi_indexcoords": [c_vp, c_i32, c_vp, c_i32, c_vp, c_int, c_int, c_vp, c_vp],
    "frmc_pairs_distances_to_point": [c_vp, c_vp, c_i32, c_vp, c_int, c_vp, c_vp],
    "frmc_pairs_distances_to_multi_points": [c_vp, c_i32, c_vp, c_i32, c_vp, c_int, c_vp, c_vp],
    "frmc_point_to_point_distance": [c_vp, c_vp, c_vp, c_int, c_vp, c_vp],
    "frmc_pairs_differences_to_indexcoords": [c_i32, c_vp, c_i32, c_vp, c_int, c_int, c_vp, c_vp],
    "frmc_pairs_differences_to_multi_indexcoords": [c_vp, c_i32, c_vp, c_i32, c_vp, c_int, c_int, c_vp, c_vp],
    "frmc_pairs_differences_to_point": [c_vp, c_vp, c_i32, c_vp, c_int, c_vp, c_vp],
    "frmc_pairs_differences_to_multi_points": [c_vp, c_i32, c_vp, c_i32, c_vp, c_int, c_vp, c_vp],
    "frmc_from_to_points_differences": [c_vp, c_vp, c_i32, c_vp, c_int, c_vp, c_vp],
    "frmc_pair_difference_to_point": [c_vp, c_vp, c_vp, c_int, c_vp, c_vp],
    # pairs_histograms
    "frmc_single_pairs_histograms": [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_i32, c_i32,
                                     c_f32, c_f32, c_f32, c_int, c_vp],
    "frmc_multiple_pairs_histograms_coords": [c_vp, c_i32, c_vp, c_i32, c_vp, c_int, c_vp, c_vp, c_i32,
                                              c_f32, c_f32, c_f32, c_i32, c_int, c_vp, c_vp, c_vp],
    "frmc_multiple_pairs_histograms_dists": [c_vp, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, c_i32,
                                             c_f32, c_f32, c_f32, c_i32, c_int, c_vp, c_vp, c_vp],
    "frmc_full_pairs_histograms_coords": [c_vp, c_i32, c_vp, c_int, c_vp, c_vp, c_i32,
                                          c_f32, c_f32, c_f32, c_i32, c_vp, c_vp, c_vp],
    "frmc_full_pairs_histograms_dists": [c_vp, c_i32, c_i32, c_vp, c_vp, c_i32,
                                         c_f32, c_f32, c_f32, c_i32, c_vp, c_vp, c_vp],
    "frmc_full_pairs_counts": [c_vp, c_i32, c_vp, c_int, c_vp, c_vp, c_i32,
                               c_f32, c_f32, c_f32, c_i32, c_vp, c_vp],
    "frmc_full_pairs_counts_f64": [c_vp, c_i32, c_vp, c_int, c_vp, c_vp, c_i32,
                                   c_f32, c_f32, c_f32, c_i32, c_vp, ctypes.POINTER(ctypes.c_int64), c_vp],
    # boundary_conditions_collection
    "frmc_get_reciprocal_basis": [c_vp, c_vp, ctypes.POINTER(ctypes.c_double), c_vp],
    "frmc_transform_coordinates": [c_vp, c_vp, c_i32, c_vp, c_vp],
    "frmc_box_coordinates_real_distances": [c_i32, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp, c_vp],
    # atomic_distances
    "frmc_single_atomic_distances_dists": [c_i32, c_vp, c_i32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i32,
                                           c_vp, c_vp, c_int, c_int, c_vp],
    "frmc_multiple_atomic_distances_coords": [c_vp, c_i32, c_vp, c_i32, c_vp, c_int, c_vp, c_vp, c_i32, c_vp, c_vp,
                                              c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "frmc_multiple_atomic_distances_dists": [c_vp, c_i32, c_vp, c_i32, c_i32, c_vp, c_vp, c_i32, c_vp, c_vp,
                                             c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp],
    "frmc_pair_elements_stats": [c_vp, c_vp, c_i32, c_i32, c_vp, c_vp, c_vp],
    # G(r) -> S(Q) transform, reciprocal_space
    "frmc_sq_transform": [c_vp, c_i32, c_i32, c_vp, c_i32, c_int, c_int, c_vp, c_vp],
    "frmc_sq_transform_noalloc": [c_vp, c_i32, c_i32, c_vp, c_i32, c_int, c_int, c_vp, c_vp],
    "frmc_Gr_to_sq": [c_vp, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp],
    "frmc_gr_to_sq": [c_vp, c_vp, c_i32, c_vp, c_i32, c_f32, c_vp, c_vp],
    "frmc_sq_to_Gr": [c_vp, c_vp, c_i32, c_vp, c_i32, c_vp, c_vp],
    # engine
    "frmc_engine_create": [ctypes.POINTER(c_vp), ctypes.POINTER(EngineDesc)],
    "frmc_engine_destroy": [c_vp],
    "frmc_engine_compute_data": [c_vp, c_f32p, c_f32p],
    "frmc_engine_refresh": [c_vp, c_f32p, c_f32p],
    "frmc_engine_eval": [c_vp, c_i32, c_i32p, c_f32p, c_f32p, c_f32p],
    "frmc_engine_eval_atoms": [c_vp, c_i32, c_i32p, c_i32p, c_f32p, c_f32p, c_f32p],
    "frmc_engine_commit": [c_vp, c_i32],
    "frmc_engine_sync": [c_vp],
    "frmc_engine_get_state": [c_vp, c_f32p, c_i32, c_i32p, c_i32p, c_f32p, c_f32p],
    "frmc_engine_run_translation": [c_vp, c_i64, c_i32, c_f32, c_f32, c_f32,
                                    ctypes.POINTER(MT19937), ctypes.POINTER(MT19937),
                                    ctypes.POINTER(RunStats), c_i32p,
                                    ctypes.POINTER(ctypes.c_uint8), c_f32p],
    "frmc_engines_run_translation": [ctypes.POINTER(c_vp), c_i32, c_i64, c_i32, c_f32, c_f32, c_f32,
                                     ctypes.POINTER(MT19937), ctypes.POINTER(MT19937), ctypes.POINTER(RunStats)],
    "frmc_engine_get_real": [c_vp, c_f32p],
    "frmc_engine_set_real": [c_vp, c_f32p],
    "frmc_engine_set_counts": [c_vp, c_i32, c_i32p],
    "frmc_engine_profile": [c_vp, c_int],
    "frmc_engine_time_eval": [c_vp, c_i32, c_i32, c_f32p],
    "frmc_set_eval_shape": [c_i32],
    "frmc_engine_eval_shapes": [c_vp, c_i32p, c_i32],
    "frmc_set_delta_cells": [c_i32],
    "frmc_engine_delta_cells": [c_vp, c_i32p, c_f32p],
    "frmc_set_window_loop": [c_i32],
    "frmc_engine_loop_windows": [c_vp, ctypes.POINTER(c_i64)],
    "frmc_engine_kernel_times": [c_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_i64)],
    "frmc_engine_set_adaptive": [c_vp, c_int],
}
KERNEL_NAMES = ["k_delta", "k_eval_a", "k_commit", "k_eval_b"]
_RESTYPES = {"frmc_last_error": ctypes.c_char_p, "frmc_last_overflow": c_i64, "frmc_set_timing": None,
             "frmc_last_kernel_ms": c_f32, "frmc_set_autotune": None,
             "frmc_set_replica_batching": None, "frmc_set_replica_threads": None}

_lib = None


def exported_symbols():
    """Names of every entry point this binding declares."""
    return sorted(_PROTOS)


def lib():
    """Load ``libfrmc_hip.so`` once and return the ctypes handle."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FrmcError("%s is not built; run __graft_entry__.build() (hipcc, gfx950)" % LIB_PATH)
        # torch first: its HIP runtime is then the one the library binds to,
        # so device pointers and streams from torch are valid in both
        from . import _device
        _device.torch()
        handle = ctypes.CDLL(LIB_PATH)
        for name, args in _PROTOS.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, c_int)
        _lib = handle
    return _lib


def check(rc):
    """Raise :class:`FrmcError` for a non-zero status of a C-ABI call."""
    if rc:
        msg = lib().frmc_last_error().decode(errors="replace")
        if rc == 1:
            raise ValueError(msg)
        raise FrmcError("%s: %s" % (ERRORS.get(rc, rc), msg))


def last_overflow():
    """Pair bins the last histogram call dropped past the end of the arrays."""
    return int(lib().frmc_last_overflow())
