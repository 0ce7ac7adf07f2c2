"""The full-histogram autotuner: every schedule it can pick (workgroup size,
row atoms per thread, workgroups per CU) gives counts bit-exact against the
oracle, on the LDS-privatised and the global-atomics kernels, with and
without periodic boundaries, and in the float64-recompute mode; the tuner
caches its choice per shape class."""
import itertools

import numpy as np
import pytest

from fullrmc_amd import systems, tuning
from fullrmc_amd.Core import pairs_histograms as gph
from oracle import refcpu

pytestmark = pytest.mark.gpu
F32 = np.float32

SCHEDULES = [(t, r, w) for t, r, w in itertools.product(tuning.THREADS, tuning.ROW_ATOMS, (1, 2, 16))]


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


@pytest.fixture(scope="module")
def si_o():
    """12,000 atoms (Si:O 1:2), molecules of four atoms, cubic 54 A."""
    s = systems.random_box(12000, 54.0, 1.2, seed=21, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    lut = {e: i for i, e in enumerate(sorted(set(s["elements"])))}
    elem = np.array([lut[e] for e in s["elements"]], dtype=np.int32)
    mol = (np.arange(12000) // 4).astype(np.int32)
    box = systems.box_coordinates(s["recip"], s["real"])
    return {"box": box, "real": s["real"], "basis": s["basis"], "mol": mol, "elem": elem}


@pytest.fixture(autouse=True)
def _release():
    yield
    tuning.force_full_schedule()
    tuning.set_autotune(True)


CASES = {
    # LDS bins: 2*4*240*4 = 7.5 KB
    "lds": lambda s: ((s["box"], s["basis"], True, s["mol"], s["elem"], 2, F32(0.0), F32(6.0), F32(0.025), 240),
                      False),
    # 2*4*16000*4 = 500 KB of bins: global atomics
    "global": lambda s: ((s["box"], s["basis"], True, s["mol"], s["elem"], 2, F32(0.0), F32(8.0), F32(0.0005),
                          16000), False),
    "no_pbc": lambda s: ((s["real"], s["basis"], False, s["mol"], s["elem"], 2, F32(0.5), F32(7.0), F32(0.05), 130),
                         False),
    "float64": lambda s: ((s["box"], s["basis"], True, s["mol"], s["elem"], 2, F32(0.0), F32(6.0), F32(0.004),
                           1500), True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_every_schedule_bit_exact(case, si_o):
    args, f64 = CASES[case](si_o)
    if f64:
        ci, ce, _ = refcpu.full_pairs_counts_f64(*args, ncores=8)
        ci, ce = ci.astype(F32), ce.astype(F32)
    else:
        ci, ce = refcpu.full_pairs_histograms_coords(*args, ncores=8)
    assert ci.sum() > 0 and ce.sum() > 0
    for t, r, w in SCHEDULES:
        print("schedule", case, t, r, w, flush=True)
        tuning.force_full_schedule(t, r, w)
        if f64:
            gi, ge, _ = gph.full_pairs_histograms_coords_float64(*args)
        else:
            gi, ge = gph.full_pairs_histograms_coords(*args)
        got = tuning.last_full_schedule()
        assert (got["threads"], got["rowAtoms"], got["tuned"]) == (t, r, False)
        assert got["workgroups"] >= 1
        assert same(gi, ci) and same(ge, ce), (t, r, w)


def test_tuner_picks_caches_and_matches(si_o):
    args, _ = CASES["lds"](si_o)
    ci, ce = refcpu.full_pairs_histograms_coords(*args, ncores=8)
    tuning.set_autotune(True, forget=True)
    gi, ge = gph.full_pairs_histograms_coords(*args)
    first = tuning.last_full_schedule()
    assert first["tuned"] and first["threads"] in tuning.THREADS and first["rowAtoms"] in tuning.ROW_ATOMS
    assert same(gi, ci) and same(ge, ce)
    gi, ge = gph.full_pairs_histograms_coords(*args)
    assert tuning.last_full_schedule() == first
    assert same(gi, ci) and same(ge, ce)
    tuning.set_autotune(False)
    gi, ge = gph.full_pairs_histograms_coords(*args)
    fixed = tuning.last_full_schedule()
    assert not fixed["tuned"] and fixed["threads"] == 256
    assert same(gi, ci) and same(ge, ce)


def test_force_schedule_rejects_unknown_shapes():
    for bad in [(128, 2, 2), (256, 3, 2), (512, 2, 0), (256, 1, 17), (2048, 1, 1)]:
        with pytest.raises(ValueError):
            tuning.force_full_schedule(*bad)


@pytest.fixture(scope="module")
def big_cells():
    """40,000 atoms (Si:O 1:2, molecules of five atoms) in a 90 A cube, the
    box coordinates shifted by whole cells (-2..2) per atom, and the same
    fractional coordinates in a triclinic cell of 95/100/105 A, 80/95/105 deg:
    large enough against a 12 A range for the cell-list kernel."""
    s = systems.random_box(40000, 90.0, 1.2, seed=31, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    lut = {e: i for i, e in enumerate(sorted(set(s["elements"])))}
    elem = np.array([lut[e] for e in s["elements"]], dtype=np.int32)
    mol = (np.arange(40000) // 5).astype(np.int32)
    box = systems.box_coordinates(s["recip"], s["real"])
    rng = np.random.default_rng(32)
    box = (box + rng.integers(-2, 3, box.shape)).astype(F32)
    tb = systems.triclinic_basis(95.0, 100.0, 105.0, 80.0, 95.0, 105.0)
    return {"box": box, "basis": s["basis"], "tbasis": tb, "mol": mol, "elem": elem}


CELL_CASES = {
    "cubic_lds": lambda s: ((s["box"], s["basis"], True, s["mol"], s["elem"], 2, F32(0.0), F32(12.0), F32(0.01),
                             1200), False),
    "cubic_global": lambda s: ((s["box"], s["basis"], True, s["mol"], s["elem"], 2, F32(0.5), F32(12.0),
                                F32(0.0005), 23000), False),
    "triclinic": lambda s: ((s["box"], s["tbasis"], True, s["mol"], s["elem"], 2, F32(0.0), F32(12.0), F32(0.02),
                             600), False),
    "float64": lambda s: ((s["box"], s["basis"], True, s["mol"], s["elem"], 2, F32(0.0), F32(12.0), F32(0.004),
                           3000), True),
}


@pytest.mark.parametrize("case", sorted(CELL_CASES))
def test_cell_list_schedules_bit_exact(case, big_cells):
    """The cell-list kernel (rowAtoms -1) at 256 and 512 threads and several
    workgroups per CU: slot for slot the oracle's all-pairs counts, with
    LDS-private and global bins, in a triclinic cell and in the float64
    mode; the tuner, free to choose, also matches."""
    args, f64 = CELL_CASES[case](big_cells)
    if f64:
        ci, ce, _ = refcpu.full_pairs_counts_f64(*args, ncores=16)
        ci, ce = ci.astype(F32), ce.astype(F32)
    else:
        ci, ce = refcpu.full_pairs_histograms_coords(*args, ncores=16)
    assert ci.sum() > 0 and ce.sum() > 1e6
    for t, w in ((256, 1), (256, 2), (512, 1), (512, 4)):
        print("cell schedule", case, t, w, flush=True)
        tuning.force_full_schedule(t, -1, w)
        if f64:
            gi, ge, _ = gph.full_pairs_histograms_coords_float64(*args)
        else:
            gi, ge = gph.full_pairs_histograms_coords(*args)
        got = tuning.last_full_schedule()
        assert (got["threads"], got["rowAtoms"], got["cells"]) == (t, -1, True)
        assert same(gi, ci) and same(ge, ce), (case, t, w)
    tuning.force_full_schedule()
    tuning.set_autotune(True, forget=True)
    gi, ge = (gph.full_pairs_histograms_coords_float64(*args)[:2] if f64 else gph.full_pairs_histograms_coords(*args))
    print("tuned", case, tuning.last_full_schedule(), flush=True)
    assert same(gi, ci) and same(ge, ce)


def test_cell_list_schedule_refused_on_a_small_cell(si_o):
    """54 A against a 20 A range: fewer than three cells per axis."""
    tuning.force_full_schedule(256, -1, 1)
    with pytest.raises(Exception):
        gph.full_pairs_histograms_coords(si_o["box"], si_o["basis"], True, si_o["mol"], si_o["elem"], 2, F32(0.0),
                                         F32(20.0), F32(0.02), 1000)
