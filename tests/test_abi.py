"""The C-ABI library loads without a GPU and exports every entry point that
include/frmc.h declares; the ctypes binding declares exactly those."""
import ctypes
import os
import re

import pytest

from fullrmc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "frmc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(frmc_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = header_functions()
    assert "frmc_full_pairs_histograms_coords" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfrmc_hip.so not built")
    handle = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in header_functions() if not hasattr(handle, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert sorted(_lib.exported_symbols()) == header_functions()


def test_binding_loads_and_reports_errors_without_gpu():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfrmc_hip.so not built")
    L = _lib.lib()
    assert L.frmc_version() == 10000
    # argument validation runs on the host before any HIP call
    rc = L.frmc_pairs_distances_to_indexcoords(5, None, 3, None, 1, 1, None, None)
    assert rc == 1
    assert b"out of range" in L.frmc_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)


def test_oracle_symbols():
    from oracle import refcpu
    L = refcpu.lib()
    for name in ("ref_distances_to_indexcoords", "ref_multiple_pairs_histograms_coords", "ref_mt_next32"):
        assert hasattr(L, name)


def test_schedule_and_replica_controls_validate_without_gpu():
    """The tuning and replica setters touch no device: a forced schedule is
    range-checked on the host, and releasing it always succeeds."""
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfrmc_hip.so not built")
    from fullrmc_amd import tuning
    for bad in [(128, 2, 2), (256, 3, 2), (512, 2, 0), (256, 1, 17), (2048, 1, 1)]:
        with pytest.raises(ValueError):
            tuning.force_full_schedule(*bad)
    for good in [(256, 1, 1), (512, 4, 16), (1024, 2, 3)]:
        tuning.force_full_schedule(*good)
    tuning.force_full_schedule()
    tuning.set_autotune(True, forget=True)
    L = _lib.lib()
    L.frmc_set_replica_threads(3)
    L.frmc_set_replica_threads(0)
    L.frmc_set_replica_batching(0)
    L.frmc_set_replica_batching(1)
    ms = (ctypes.c_double * 5)()
    steps = ctypes.c_int64(-1)
    _lib.check(L.frmc_replica_batch_times(ms, ctypes.byref(steps)))
    assert steps.value >= 0
