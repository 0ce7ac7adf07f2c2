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
