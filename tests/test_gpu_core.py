"""GPU parity of the drop-in extension functions (fullrmc_amd.Core, HIP via
the C ABI) against the CPU oracle: bit-exact on every output."""
import numpy as np
import pytest

from fullrmc_amd import systems
from fullrmc_amd.Core import boundary_conditions_collection as gbc
from fullrmc_amd.Core import pairs_distances as gpd
from fullrmc_amd.Core import pairs_histograms as gph
from oracle import refcpu

pytestmark = pytest.mark.gpu
F32 = np.float32


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()


def _systems(small_random, small_triclinic):
    return [("cubic", small_random), ("triclinic", small_triclinic)]


@pytest.fixture(params=["cubic", "triclinic"])
def sysm(request, small_random, small_triclinic):
    s = small_random if request.param == "cubic" else small_triclinic
    return s


def test_cr_sqrt_and_div_known_answers():
    """The device sqrt / divide of the bin index are correctly rounded: compare
    distances of many exactly representable pairs with the host float64 path."""
    rng = np.random.default_rng(0)
    pts = rng.random((20000, 3)).astype(F32)
    basis = np.diag([31.7, 29.3, 33.1]).astype(F32)
    g = gpd.pairs_distances_to_indexcoords(0, pts, basis, True)
    c = refcpu.pairs_distances_to_indexcoords(0, pts, basis, True)
    assert same(g, c)


@pytest.mark.parametrize("isPBC", [True, False])
def test_distances_to_indexcoords(sysm, isPBC):
    coords = sysm["box"] if isPBC else sysm["real"]
    for atom in (0, 7, len(coords) - 1):
        for allAtoms in (True, False):
            g = gpd.pairs_distances_to_indexcoords(atom, coords, sysm["basis"], isPBC, allAtoms)
            c = refcpu.pairs_distances_to_indexcoords(atom, coords, sysm["basis"], isPBC, allAtoms)
            start = 0 if allAtoms else atom
            assert same(g[start:], c[start:])


@pytest.mark.parametrize("isPBC", [True, False])
def test_distances_and_differences_all_forms(sysm, isPBC):
    coords = sysm["box"] if isPBC else sysm["real"]
    B = sysm["basis"]
    idx = np.array([3, 0, 11, 3], dtype=np.int32)  # includes a duplicate
    pts = coords[[5, 9, 2]].T.copy()  # (3,k)
    assert same(gpd.pairs_distances_to_multi_indexcoords(idx, coords, B, isPBC),
                refcpu.pairs_distances_to_multi_indexcoords(idx, coords, B, isPBC))
    assert same(gpd.pairs_distances_to_point(coords[4], coords, B, isPBC),
                refcpu.pairs_distances_to_point(coords[4], coords, B, isPBC))
    assert same(gpd.pairs_distances_to_multi_points(pts, coords, B, isPBC),
                refcpu.pairs_distances_to_multi_points(pts, coords, B, isPBC))
    assert same(gpd.point_to_point_distance(coords[1], coords[8], B, isPBC),
                refcpu.point_to_point_distance(coords[1], coords[8], B, isPBC))
    assert same(gpd.pairs_differences_to_indexcoords(6, coords, B, isPBC),
                refcpu.pairs_differences_to_indexcoords(6, coords, B, isPBC))
    assert same(gpd.pairs_differences_to_multi_indexcoords(idx, coords, B, isPBC),
                refcpu.pairs_differences_to_multi_indexcoords(idx, coords, B, isPBC))
    assert same(gpd.pairs_differences_to_point(coords[4], coords, B, isPBC),
                refcpu.pairs_differences_to_point(coords[4], coords, B, isPBC))
    assert same(gpd.pairs_differences_to_multi_points(pts, coords, B, isPBC),
                refcpu.pairs_differences_to_multi_points(pts, coords, B, isPBC))
    other = coords[::-1].copy()
    assert same(gpd.from_to_points_differences(coords, other, B, isPBC),
                refcpu.from_to_points_differences(coords, other, B, isPBC))
    assert same(gpd.pair_difference_to_point(coords[1], coords[8], B, isPBC),
                refcpu.pair_difference_to_point(coords[1], coords[8], B, isPBC))


def test_boundary_conditions(sysm):
    assert same(gbc.transform_coordinates(sysm["recip"], sysm["real"]),
                refcpu.transform_coordinates(sysm["recip"], sysm["real"]))
    r, v = gbc.get_reciprocal_basis(sysm["basis"])
    rr, vv = refcpu.get_reciprocal_basis(sysm["basis"])
    assert same(r, rr) and v == vv
    idx = np.arange(0, len(sysm["box"]), 3, dtype=np.int32)
    assert same(gbc.box_coordinates_real_distances(2, idx, sysm["box"], sysm["basis"]),
                refcpu.box_coordinates_real_distances(2, idx, sysm["box"], sysm["basis"]))


def _hargs(s, nel, mn=0.0, mx=9.0, bn=0.05, H=None):
    if H is None:
        H = int(round((mx - mn) / bn))
    return (s["box"], s["basis"], True, s["molecules"], s["elem"], nel, F32(mn), F32(mx), F32(bn), H)


@pytest.mark.parametrize("grid", [(0.0, 9.0, 0.05, 180), (1.0, 6.0, 0.02, 250), (0.0, 5.0, 0.01, 500)])
def test_full_pairs_histograms(sysm, grid):
    nel = int(sysm["elem"].max()) + 1
    args = _hargs(sysm, nel, *grid)
    g = gph.full_pairs_histograms_coords(*args)
    c = refcpu.full_pairs_histograms_coords(*args)
    assert same(g[0], c[0]) and same(g[1], c[1])
    assert g[0].sum() + g[1].sum() > 0


def test_full_pairs_histograms_global_atomics_path(small_triclinic):
    """2*nel^2*H*4 bytes above the LDS budget -> the global-atomics kernel."""
    args = _hargs(small_triclinic, 3, 0.0, 12.0, 0.002, 6000)
    g = gph.full_pairs_histograms_coords(*args)
    c = refcpu.full_pairs_histograms_coords(*args)
    assert same(g[0], c[0]) and same(g[1], c[1])


def test_full_pairs_infinite_boundaries(small_random):
    s = small_random
    args = (s["real"], s["basis"], False, s["molecules"], s["elem"], 2, F32(0.0), F32(8.0), F32(0.05), 160)
    g = gph.full_pairs_histograms_coords(*args)
    c = refcpu.full_pairs_histograms_coords(*args)
    assert same(g[0], c[0]) and same(g[1], c[1])


@pytest.mark.parametrize("allAtoms", [True, False])
def test_multiple_pairs_histograms_coords(sysm, allAtoms):
    nel = int(sysm["elem"].max()) + 1
    args = _hargs(sysm, nel)
    for idx in ([0], [13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25], [5, 5, 400], list(range(0, 500, 9)), []):
        idx = np.array(idx, dtype=np.int32)
        g = gph.multiple_pairs_histograms_coords(idx, *args, allAtoms=allAtoms)
        c = refcpu.multiple_pairs_histograms_coords(idx, *args, allAtoms=allAtoms)
        assert same(g[0], c[0]) and same(g[1], c[1])


def test_dists_variants(small_triclinic):
    s = small_triclinic
    n = len(s["box"])
    D = refcpu.pairs_distances_to_multi_indexcoords(np.arange(n, dtype=np.int32), s["box"], s["basis"], True)
    D = np.ascontiguousarray(D)
    idx = np.array([4, 100, 7], dtype=np.int32)
    sub = np.ascontiguousarray(D[:, idx])
    for allAtoms in (True, False):
        g = gph.multiple_pairs_histograms_dists(idx, sub, s["molecules"], s["elem"], 3, F32(0), F32(9), F32(0.05),
                                                180, allAtoms)
        c = refcpu.multiple_pairs_histograms_dists(idx, sub, s["molecules"], s["elem"], 3, F32(0), F32(9),
                                                   F32(0.05), 180, allAtoms)
        assert same(g[0], c[0]) and same(g[1], c[1])
    g = gph.full_pairs_histograms_dists(D, s["molecules"], s["elem"], 3, F32(0), F32(9), F32(0.05), 180)
    c = refcpu.full_pairs_histograms_dists(D, s["molecules"], s["elem"], 3, F32(0), F32(9), F32(0.05), 180)
    assert same(g[0], c[0]) and same(g[1], c[1])
    # single_pairs_histograms accumulates in place
    hi = np.ones((3, 3, 180), F32)
    he = np.ones((3, 3, 180), F32)
    hi2, he2 = hi.copy(), he.copy()
    gph.single_pairs_histograms(4, np.ascontiguousarray(D[:, 4]), s["molecules"], s["elem"], hi, he, F32(0), F32(9),
                                F32(0.05))
    refcpu.single_pairs_histograms(4, np.ascontiguousarray(D[:, 4]), s["molecules"], s["elem"], hi2, he2, F32(0),
                                   F32(9), F32(0.05))
    assert same(hi, hi2) and same(he, he2)


def test_bin_index_H_quirk_on_gpu():
    mn, bn, H = F32(0.0), F32(0.1), 10
    mx = F32(1.0000001)
    d = F32(0.99999994)
    while not (int(F32(F32(d - mn) / bn)) == H and d < mx):
        d = np.nextafter(d, F32(2))
    dist = np.array([0.0, d], dtype=F32)
    for elem in ([0, 0], [1, 1]):
        hi, he = np.zeros((2, 2, H), F32), np.zeros((2, 2, H), F32)
        hi2, he2 = hi.copy(), he.copy()
        e = np.array(elem, np.int32)
        m = np.array([0, 1], np.int32)
        gph.single_pairs_histograms(0, dist, m, e, hi, he, mn, mx, bn)
        refcpu.single_pairs_histograms(0, dist, m, e, hi2, he2, mn, mx, bn)
        assert same(hi, hi2) and same(he, he2)


def test_argument_errors_match_reference_types(small_random):
    s = small_random
    args = list(_hargs(s, 2))
    with pytest.raises(ValueError, match="Buffer dtype mismatch"):
        gph.full_pairs_histograms_coords(s["box"].astype(np.float64), *args[1:])
    with pytest.raises(TypeError):
        gph.full_pairs_histograms_coords(None, *args[1:])
    with pytest.raises(ValueError, match="wrong number of dimensions"):
        gph.full_pairs_histograms_coords(s["box"].reshape(-1), *args[1:])
    with pytest.raises(IndexError):
        gph.multiple_pairs_histograms_coords(np.array([10 ** 6], np.int32), *args)


def test_full_histogram_b2_13500_atoms():
    """B2 alloy, 13,500 atoms, PDC grid r in [0,20] @0.02: full histogram
    bit-exact against the oracle (8 threads)."""
    s = systems.b2_alloy(seed=1)
    lut = {e: i for i, e in enumerate(sorted(set(s["elements"])))}
    elem = np.array([lut[e] for e in s["elements"]], dtype=np.int32)
    box = systems.box_coordinates(s["recip"], s["real"])
    args = (box, s["basis"], True, s["molecules"], elem, 2, F32(0.0), F32(20.0), F32(0.02), 1000)
    g = gph.full_pairs_histograms_coords(*args)
    c = refcpu.full_pairs_histograms_coords(*args, ncores=8)
    assert same(g[0], c[0]) and same(g[1], c[1])
    n = len(box)
    assert g[1].sum() > 0.1 * n * (n - 1) / 2


def test_full_histogram_two_rows_per_thread_and_partials():
    """40,000 atoms (two row atoms per thread, thousands of tiles, LDS
    partials summed per workgroup) on a small grid, with molecules of four
    atoms: bit-exact against the oracle."""
    s = systems.random_box(40000, 80.0, 1.2, seed=12, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    lut = {e: i for i, e in enumerate(sorted(set(s["elements"])))}
    elem = np.array([lut[e] for e in s["elements"]], dtype=np.int32)
    mol = (np.arange(40000) // 4).astype(np.int32)
    box = systems.box_coordinates(s["recip"], s["real"])
    args = (box, s["basis"], True, mol, elem, 2, F32(0.0), F32(6.0), F32(0.05), 120)
    g = gph.full_pairs_histograms_coords(*args)
    c = refcpu.full_pairs_histograms_coords(*args, ncores=8)
    assert same(g[0], c[0]) and same(g[1], c[1])
    assert g[0].sum() > 0 and g[1].sum() > 0


@pytest.mark.parametrize("case", ["cubic", "triclinic", "small_cell"])
def test_full_histogram_float64_recompute(case, small_random, small_triclinic):
    """The float64 mode equals the binary64 recompute oracle exactly, with a
    non-zero number of double re-checks (pairs near edges exist).  small_cell:
    a 10 A cube binned to 9 A, where pairs sit at fractional differences near
    +-1/2 (minimum image not unique)."""
    if case == "small_cell":
        s = systems.random_box(150, 10.0, 1.0, seed=4)
        nel, grid = 1, (0.0, 9.0, 0.005, 1800)
        mol, el = s["molecules"], np.zeros(150, np.int32)
        box = systems.box_coordinates(s["recip"], s["real"])
    else:
        s = small_random if case == "cubic" else small_triclinic
        nel, grid = int(s["elem"].max()) + 1, (0.5, 8.0, 0.004, 1875)
        mol, el, box = s["molecules"], s["elem"], s["box"]
    args = (box, s["basis"], True, mol, el, nel, F32(grid[0]), F32(grid[1]), F32(grid[2]), grid[3])
    gi, ge, rechecked = gph.full_pairs_histograms_coords_float64(*args)
    ci, ce, ov = refcpu.full_pairs_counts_f64(*args, ncores=8)
    assert same(gi, ci.astype(F32)) and same(ge, ce.astype(F32))
    assert rechecked > 0
    # the float32 (reference) histogram is a different oracle: same total, a few pairs moved
    fi, fe = gph.full_pairs_histograms_coords(*args)
    assert fi.sum() + fe.sum() == gi.sum() + ge.sum()


def test_full_histogram_float64_recompute_40000_atoms():
    s = systems.random_box(40000, 80.0, 1.2, seed=13, species=("Si", "O"), fractions=(1 / 3, 2 / 3))
    lut = {e: i for i, e in enumerate(sorted(set(s["elements"])))}
    elem = np.array([lut[e] for e in s["elements"]], dtype=np.int32)
    box = systems.box_coordinates(s["recip"], s["real"])
    args = (box, s["basis"], True, s["molecules"], elem, 2, F32(0.0), F32(6.0), F32(0.01), 600)
    gi, ge, rechecked = gph.full_pairs_histograms_coords_float64(*args)
    ci, ce, ov = refcpu.full_pairs_counts_f64(*args, ncores=8)
    assert same(gi, ci.astype(F32)) and same(ge, ce.astype(F32))
    assert 0 < rechecked < 0.05 * (gi.sum() + ge.sum())
