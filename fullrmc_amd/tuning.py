"""Control of the full-histogram schedule autotuner (C ABI
``frmc_set_autotune``, ``frmc_force_full_schedule``,
``frmc_last_full_schedule``).

A schedule of the all-pairs histogram kernel is (threads per workgroup,
which is also the column tile; row atoms per thread; workgroups per CU).
Row atoms -1 selects the cell-list kernel (256 or 512 threads), which
evaluates only pairs in neighbouring cells of a grid at least the
histogram range wide; it exists for periodic cells with three or more such
cells per axis (125 or more in all).
Every schedule gives the same counts, so tuning changes only the time.  With
tuning on (the default), the first all-pairs call of a shape class times each
candidate schedule with HIP events and later calls of the class reuse the
fastest.  The reference has no counterpart: its all-pairs loop
(Extensions/pairs_histograms.pyx:86-176) is one fixed OpenMP loop.
"""
import ctypes

from . import _lib

THREADS = (256, 512, 1024)
ROW_ATOMS = (1, 2, 4)


def set_autotune(enable=True, forget=False):
    """Turn tuning on or off; ``forget`` also drops every tuned class."""
    _lib.lib().frmc_set_autotune(-1 if forget else (1 if enable else 0))
    if forget and enable:
        _lib.lib().frmc_set_autotune(1)


def force_full_schedule(threads=None, rowAtoms=None, workgroupsPerCU=None):
    """Pin one schedule for every all-pairs call; no arguments release it."""
    if rowAtoms is None:
        _lib.check(_lib.lib().frmc_force_full_schedule(0, 0, 0))
        return
    _lib.check(_lib.lib().frmc_force_full_schedule(int(threads), int(rowAtoms), int(workgroupsPerCU)))


def last_full_schedule():
    """The schedule of this thread's last all-pairs call."""
    t, r, tuned = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    w = ctypes.c_int64(0)
    _lib.check(_lib.lib().frmc_last_full_schedule(ctypes.byref(t), ctypes.byref(r), ctypes.byref(w),
                                                  ctypes.byref(tuned)))
    return {"threads": t.value, "rowAtoms": r.value, "workgroups": w.value, "tuned": bool(tuned.value),
            "cells": r.value == -1}
