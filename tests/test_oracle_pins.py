"""Pins of the CPU oracle (oracle/refcpu.c, oracle/chain.py, oracle/engine.py)
against what fixes it independently of the restatement: float64 brute force,
closed forms, invariants, NumPy / CPython themselves as semantics oracles,
and known answers chosen so that a plausible restatement mistake (FMA
contraction, rint instead of the reference's round, wrong operand order,
dropped term, transposed basis, wrong slot) fails.

The reference ships no tests, golden vectors or fixtures for this path
(its tests/ directory is git-ignored); see tests/golden/ for the vectors
committed here.
"""
import itertools
import math
import random

import numpy as np
import pytest

from fullrmc_amd import systems
from oracle import chain, refcpu

F32 = np.float32


# ------------------------------------------------------------------ MT19937

@pytest.mark.parametrize("seed", [0, 1, 12345, 2 ** 32 - 1, 2 ** 40 + 7])
def test_mt19937_matches_cpython_random(seed):
    random.seed(seed)
    m = refcpu.MT.from_python(random.getstate())
    for n in (1, 2, 3, 7, 600, 2000, 13500, 200000):
        assert m.randbelow(n) == random.randint(0, n - 1)
        assert m.double() == random.random()
    for _ in range(3000):  # crosses several twists
        assert m.double() == random.random()


@pytest.mark.parametrize("seed", [0, 1, 12345, 2 ** 32 - 1])
def test_mt19937_matches_numpy_legacy(seed):
    np.random.seed(seed)
    m = refcpu.MT.from_numpy(np.random.get_state())
    for _ in range(2000):
        v = np.random.random(3)
        assert [m.double(), m.double(), m.double()] == list(v)


def test_linalg_norm_rule():
    """np.linalg.norm of a float32 3-vector == sqrt(float32(double sum of
    float32 products)) -- the rule the native generator uses."""
    rng = np.random.default_rng(0)
    bad = 0
    for _ in range(20000):
        v = np.array(1 - 2 * rng.random(3), dtype=F32)
        p = [float(F32(x * x)) for x in v]
        acc = 0.0
        for q in p:
            acc += q
        rule = np.sqrt(F32(acc))
        bad += rule.tobytes() != np.linalg.norm(v).tobytes()
    assert bad == 0


# ------------------------------------------------------- NumPy semantics

def _pairwise(a):
    """NumPy's pairwise float32 summation, restated from its algorithm."""
    n = len(a)
    if n < 8:
        r = F32(0)
        for x in a:
            r = F32(r + x)
        return r
    if n <= 128:
        r = [F32(a[j]) for j in range(8)]
        i = 8
        while i < n - (n % 8):
            for j in range(8):
                r[j] = F32(r[j] + a[i + j])
            i += 8
        res = F32(F32(F32(r[0] + r[1]) + F32(r[2] + r[3])) + F32(F32(r[4] + r[5]) + F32(r[6] + r[7])))
        while i < n:
            res = F32(res + a[i])
            i += 1
        return res
    n2 = n // 2
    n2 -= n2 % 8
    return F32(_pairwise(a[:n2]) + _pairwise(a[n2:]))


@pytest.mark.parametrize("n", [1, 5, 8, 9, 127, 128, 129, 300, 500, 999, 1000, 1001, 2000, 2047])
def test_pairwise_sum_structure(n):
    rng = np.random.default_rng(n)
    for _ in range(3):
        a = (rng.standard_normal(n) ** 2 * rng.uniform(1e-3, 1e3)).astype(F32)
        assert _pairwise(a).tobytes() == np.add.reduce(a).tobytes()


@pytest.mark.parametrize("H,nQ", [(300, 7), (1000, 500), (2000, 1000)])
def test_axis0_sum_is_sequential(H, nQ):
    rng = np.random.default_rng(H)
    P = rng.standard_normal((H, nQ)).astype(F32)
    s = np.sum(P, axis=0)
    seq = P[0].copy()
    for r in range(1, H):
        seq = (seq + P[r]).astype(F32)
    assert s.tobytes() == seq.tobytes()


def test_nep50_scalars():
    x = F32(0.3)
    assert type(0.123 * x) is np.float32 and (0.123 * x) == F32(F32(0.123) * x)
    # random() > tolerance compares in float32
    t = F32(0.1)
    assert not (0.10000000149011613 > t)


# ------------------------------------------------------------- distances

def _min_image_f64(a, c, basis):
    """float64 brute force over the 27 neighbouring images."""
    B = basis.astype(np.float64)
    d = (a.astype(np.float64) - c.astype(np.float64))
    best = np.inf
    for s in itertools.product((-1, 0, 1), repeat=3):
        v = (d + np.array(s) - np.round(d)) @ B
        best = min(best, float(np.sqrt(v @ v)))
    return best


def test_pbc_distance_vs_float64_brute_force_cubic(small_random):
    s = small_random
    d = refcpu.pairs_distances_to_indexcoords(17, s["box"], s["basis"], True)
    for j in range(0, 600, 7):
        ref = _min_image_f64(s["box"][17], s["box"][j], s["basis"])
        assert abs(float(d[j]) - ref) <= 2e-6 * max(ref, 1.0)


def test_pbc_distance_is_reference_fractional_rounding_in_triclinic(small_triclinic):
    """In a skewed cell the reference wraps each fractional component
    independently (not the true minimum image); the float64 evaluation of
    that same rule agrees to float32 precision."""
    s = small_triclinic
    B = s["basis"].astype(np.float64)
    d = refcpu.pairs_distances_to_indexcoords(3, s["box"], s["basis"], True)
    a = s["box"][3].astype(np.float64)
    for j in range(0, len(s["box"]), 5):
        df = a - s["box"][j].astype(np.float64)
        b = df - np.where(df > 0, np.floor(df + 0.5), np.ceil(df - 0.5))
        v = b @ B
        assert abs(float(d[j]) - math.sqrt(v @ v)) <= 3e-6 * max(math.sqrt(v @ v), 1.0)


def test_distance_symmetry_and_self(small_triclinic):
    s = small_triclinic
    n = len(s["box"])
    D = refcpu.pairs_distances_to_multi_indexcoords(np.arange(0, n, 13, dtype=np.int32), s["box"], s["basis"], True)
    for c, i in enumerate(range(0, n, 13)):
        assert D[i, c] == 0.0
        for c2, j in enumerate(range(0, n, 13)):
            assert D[j, c].tobytes() == D[i, c2].tobytes()  # bitwise symmetric


def test_round_quirk_known_answer():
    """round(0.49999997f) = floor(0.49999997f + 0.5f) = floor(1.0f) = 1, so a
    fractional difference of 0.49999997 maps to -0.50000003: the distance is
    0.50000003*L, not 0.49999997*L as rint/round-half-even would give."""
    L = F32(10.0)
    basis = np.diag([L, L, L]).astype(F32)
    x = F32(0.49999997)
    coords = np.array([[x, 0, 0], [0, 0, 0]], dtype=F32)
    d = refcpu.pairs_distances_to_indexcoords(0, coords, basis, True)[1]
    expect = F32(F32(x - F32(1.0)) * L)
    assert d == abs(expect)
    assert d != F32(x * L)


def _min_image_inputs():
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-3, 3, 200000), rng.uniform(-1e4, 1e4, 20000)]).astype(F32)
    halves = (np.arange(-40, 41) * 0.5).astype(F32)
    near = np.concatenate([halves, np.nextafter(halves, F32(np.inf)), np.nextafter(halves, F32(-np.inf))])
    special = np.array([0.0, -0.0, 0.49999997, -0.49999997, 1.0, -1.0, 1e-45, -1e-45, 8388607.5, -8388607.5],
                       dtype=F32)
    return np.concatenate([x, near, special]).astype(F32)


def test_min_image_four_op_form_is_the_reference_round():
    """The kernels compute b = x - round(x) as x - copysign(floor(|x|+0.5f), x)
    and, where only b*b is used, as |x| - floor(|x|+0.5f).  Both equal the
    reference's x - (x>0 ? floor(x+0.5f) : ceil(x-0.5f)) (pairs_distances.pyx:31-32,
    :373-375): the first bit for bit, the second up to sign, on random values and
    on every half-integer and its float neighbours."""
    x = _min_image_inputs()
    half = F32(0.5)
    with np.errstate(all="ignore"):
        ref = (x - np.where(x > 0, np.floor(x + half), np.ceil(x - half))).astype(F32)
        four = (x - np.copysign(np.floor(np.abs(x) + half), x)).astype(F32)
        pm = (np.abs(x) - np.floor(np.abs(x) + half)).astype(F32)
    assert ref.tobytes() == four.tobytes()
    assert np.abs(ref).tobytes() == np.abs(pm).tobytes()
    assert (ref * ref).tobytes() == (pm * pm).tobytes()
    # a plausible slip, rounding half to even, is told apart
    assert (x - np.rint(x)).astype(F32).tobytes() != ref.tobytes()


def test_no_fma_known_answer():
    """A squared norm whose FMA-contracted value differs from the separately
    rounded one: the oracle must return the non-contracted result."""
    rng = np.random.default_rng(7)
    basis = np.diag([1.0, 1.0, 1.0]).astype(F32)
    found = 0
    for _ in range(2000):
        p = rng.random((2, 3)).astype(F32) * F32(0.4)
        d = p[0] - p[1]
        sep = F32(F32(F32(d[0] * d[0]) + F32(d[1] * d[1])) + F32(d[2] * d[2]))
        fma = F32(np.float64(F32(F32(d[0] * d[0]) + F32(d[1] * d[1]))) + np.float64(d[2]) * np.float64(d[2]))
        got = refcpu.point_to_point_distance(p[0], p[1], basis, True)
        assert got == F32(math.sqrt(float(sep)))
        found += F32(math.sqrt(float(sep))) != F32(math.sqrt(float(fma)))
    assert found > 0  # the test can tell the two apart


def test_transform_coordinates_order_and_float64(small_triclinic):
    s = small_triclinic
    M = s["recip"]
    out = refcpu.transform_coordinates(M, s["real"])
    assert out.tobytes() == systems.box_coordinates(M, s["real"]).tobytes()
    ref64 = s["real"].astype(np.float64) @ M.astype(np.float64)
    assert np.max(np.abs(out - ref64)) < 1e-6


def test_reciprocal_basis_closed_form():
    b = np.diag([30.0, 40.0, 50.0]).astype(F32)
    r, vol = refcpu.get_reciprocal_basis(b)
    assert np.allclose(r, np.linalg.inv(b.astype(np.float64)), rtol=1e-6)
    assert vol == 60000.0
    # with basis[0,1] != basis[0,2] the :59 cofactor typo shows
    t = systems.triclinic_basis(42, 44, 46, 80, 95, 105)
    t[0, 1], t[0, 2] = F32(0.5), F32(1.5)
    r, vol = refcpu.get_reciprocal_basis(t)
    B = t.astype(np.float64)
    cof01 = -(B[0, 1] * B[2, 2] - B[2, 1] * B[0, 1])
    assert abs(r[0, 1] - cof01 / vol) < 1e-6 * abs(cof01 / vol) + 1e-12
    assert not np.allclose(r, np.linalg.inv(B), rtol=1e-4)


# ------------------------------------------------------------ histograms

def _bins_f64(d, mn, mx, bn):
    d = np.asarray(d, dtype=np.float64)
    ok = (d >= mn) & (d < mx)
    return np.floor((d[ok] - mn) / bn).astype(int), ok


def test_full_histogram_vs_float64_pair_enumeration(small_random):
    """Every in-range unordered pair counted once in slot [e(earlier), e(later)];
    away from bin edges the float64 bin equals the float32 bin."""
    s = small_random
    mn, mx, bn, H = F32(0.0), F32(8.0), F32(0.05), 160
    hi, he = refcpu.full_pairs_histograms_coords(s["box"], s["basis"], True, s["molecules"], s["elem"], 2, mn, mx,
                                                 bn, H)
    B = s["basis"].astype(np.float64)
    f = s["box"].astype(np.float64)
    n = len(f)
    expect = np.zeros((2, 2, H))
    edge = 0
    for i in range(n - 1):
        d = f[i] - f[i + 1:]
        d = d - np.where(d > 0, np.floor(d + 0.5), np.ceil(d - 0.5))
        r = np.sqrt(np.sum((d @ B) ** 2, axis=1))
        ok = (r >= mn) & (r < mx)
        x = (r[ok] - float(mn)) / float(bn)
        near = np.abs(x - np.rint(x)) < 1e-4
        edge += int(near.sum())
        b = np.floor(x).astype(int)
        js = np.arange(i + 1, n)[ok]
        np.add.at(expect, (s["elem"][i], s["elem"][js][~near], b[~near]), 1)
    assert hi.sum() == 0  # one atom per molecule: no intra pairs
    assert expect.sum() > 1000
    # every non-edge pair agrees; each edge pair is at most one extra count
    assert expect.sum() <= he.sum() <= expect.sum() + edge
    assert np.abs(he - expect).sum() <= edge


def test_multiple_all_atoms_is_twice_full_symmetrised(small_triclinic):
    s = small_triclinic
    n = len(s["box"])
    args = (s["box"], s["basis"], True, s["molecules"], s["elem"], 3, F32(0.0), F32(7.0), F32(0.05), 140)
    fi, fe = refcpu.full_pairs_histograms_coords(*args)
    mi, me = refcpu.multiple_pairs_histograms_coords(np.arange(n, dtype=np.int32), *args, allAtoms=True)
    sym = lambda h: h + h.transpose(1, 0, 2)
    assert np.array_equal(sym(mi), 2 * sym(fi)) and np.array_equal(sym(me), 2 * sym(fe))
    assert fi.sum() > 0 and fe.sum() > 0  # rings give intra-molecular pairs


def test_move_delta_equals_recompute(small_triclinic):
    """data - (M-F)before + (M-F)after == full recompute after the move, in the
    symmetrised sums G(r) consumes (slot-level equality does not hold)."""
    s = small_triclinic
    box = s["box"].copy()
    args = (s["basis"], True, s["molecules"], s["elem"], 3, F32(0.0), F32(7.0), F32(0.05), 140)
    fi, fe = refcpu.full_pairs_histograms_coords(box, *args)
    idx = np.arange(13, 26, dtype=np.int32)  # molecule 1

    def mf(b):
        mi, me = refcpu.multiple_pairs_histograms_coords(idx, b, *args)
        gi, ge = refcpu.full_pairs_histograms_coords(b[idx], args[0], True, s["molecules"][idx], s["elem"][idx],
                                                     *args[4:])
        return mi - gi, me - ge

    bi, be = mf(box)
    box2 = box.copy()
    box2[idx] += F32(0.03)
    ai, ae = mf(box2)
    ri, re_ = refcpu.full_pairs_histograms_coords(box2, *args)
    sym = lambda h: h + h.transpose(1, 0, 2)
    assert np.array_equal(sym(fi - bi + ai), sym(ri)) and np.array_equal(sym(fe - be + ae), sym(re_))


def test_bin_index_H_aliases_next_slot():
    """A distance just below maxDistance whose (d-min)/bin rounds to H lands in
    slot [a, b+1, 0] of the C-ordered array, as the reference's unchecked write."""
    mn, bn, H = F32(0.0), F32(0.1), 10
    mx = F32(1.0000001)
    # find a float d < mx with int((d-mn)/bn) == H
    d = F32(0.99999994)
    while not (int(F32(F32(d - mn) / bn)) == H and d < mx):
        d = np.nextafter(d, F32(2))
    assert d < mx
    dist = np.array([0.0, d], dtype=F32)
    hi = np.zeros((2, 2, H), F32)
    he = np.zeros((2, 2, H), F32)
    ov = refcpu.single_pairs_histograms(0, dist, np.array([0, 1], np.int32), np.array([0, 0], np.int32), hi, he, mn,
                                        mx, bn)
    assert he[0, 1, 0] == 1 and he.sum() == 1 and ov == 0
    # in the last slot the write would leave the array: counted, not stored
    hi[:] = 0
    he[:] = 0
    ov = refcpu.single_pairs_histograms(0, dist, np.array([0, 1], np.int32), np.array([1, 1], np.int32), hi, he, mn,
                                        mx, bn)
    assert he.sum() == 0 and ov == 1


# ----------------------------------------------------------------- chain

def test_chain_ideal_counts_give_zero_gr():
    """With pair counts equal to N_ij/V * shell volume, g(r)=1 and G(r)=0 up to
    float32 rounding (closed form)."""
    x = (0.025 + 0.05 * np.arange(300)).astype(F32)
    exp = np.stack([x, np.zeros_like(x)], 1)
    bn, mn, mx, centers, H, vols = chain.pdc_grid(exp)
    counts = {"A": 500, "B": 1000}
    V = F32(27000.0)
    w = chain.normalized_weighting(counts, {"A": 8, "B": 14})
    intra = np.zeros((2, 2, H), F32)
    inter = np.zeros((2, 2, H), F32)
    inter[0, 0] = 500 * 499 / 2 / 27000.0 * vols
    inter[1, 1] = 1000 * 999 / 2 / 27000.0 * vols
    inter[0, 1] = 500 * 1000 / 27000.0 * vols
    G = chain.total_gr(intra, inter, ["A", "B"], counts, w, V, F32(1500) / V, vols, centers)
    assert np.max(np.abs(G)) < 2e-4 * np.max(4 * np.pi * centers * 1500 / 27000.0)
    assert sum(w.values()) == pytest.approx(1.0, rel=1e-6)


def test_chain_float32_close_to_float64(small_random):
    s = small_random
    x = (0.025 + 0.05 * np.arange(160)).astype(F32)
    exp = np.stack([x, np.zeros_like(x)], 1)
    bn, mn, mx, centers, H, vols = chain.pdc_grid(exp)
    hi, he = refcpu.full_pairs_histograms_coords(s["box"], s["basis"], True, s["molecules"], s["elem"], 2, mn, mx,
                                                 bn, H)
    counts = {"O": 400, "Si": 200}
    w = chain.normalized_weighting(counts, {"O": 8, "Si": 14})
    V = systems.volume(s["basis"])
    G32 = chain.total_gr(hi, he, ["O", "Si"], counts, w, V, F32(600) / V, vols, centers)
    # float64 evaluation of the same formula
    g = np.zeros(H)
    for a, b in [("O", "O"), ("O", "Si"), ("Si", "Si")]:
        i, j = ["O", "Si"].index(a), ["O", "Si"].index(b)
        if i == j:
            nij = hi[i, i] + he[i, i]
            D = counts[a] * (counts[a] - 1) / 2 / float(V)
        else:
            nij = hi[i, j] + hi[j, i] + he[i, j] + he[j, i]
            D = counts[a] * counts[b] / float(V)
        g += float(w[a + "-" + b]) * nij.astype(np.float64) / D
    e = np.arange(161) * 0.05
    vol64 = 4 / 3 * np.pi * (e[1:] ** 3 - e[:-1] ** 3)
    G64 = 4 * np.pi * centers.astype(np.float64) * (600 / float(V)) * (g / vol64 - 1)
    assert np.max(np.abs(G32 - G64)) < 1e-4 * (1 + np.max(np.abs(G64)))


def test_sq_matrix_vs_float64():
    centers = (0.01 + 0.02 * np.arange(1000)).astype(F32)
    q = np.linspace(0.5, 25, 500).astype(F32)
    M = chain.sq_matrix(centers, q)
    ref = 0.02 * np.sin(np.outer(centers.astype(np.float64), q.astype(np.float64))) / q.astype(np.float64)
    assert np.max(np.abs(M - ref)) < 5e-6


def _ideal_gas_histograms(n_a, n_b, L, mn, bn, H, seed):
    """Uniform random atoms (one per molecule) in a cubic periodic box."""
    rng = np.random.default_rng(seed)
    n = n_a + n_b
    box = rng.random((n, 3)).astype(F32)
    basis = np.diag([L, L, L]).astype(F32)
    elem = np.array([0] * n_a + [1] * n_b, np.int32)
    mol = np.arange(n, dtype=np.int32)
    hi, he = refcpu.full_pairs_histograms_coords(box, basis, True, mol, elem, 2, mn, F32(mn + bn * H), bn, H)
    return hi, he, systems.volume(basis)


def test_chain_gofr_is_one_for_an_ideal_gas():
    """g(r) of uniformly random atoms is 1 at every r (pins the pair counts
    N_ij, the volume, the shell volumes and the weights' normalisation
    independently of G(r)); statistical bound from the expected pair count."""
    counts = {"A": 600, "B": 1200}
    x = (3.05 + 0.1 * np.arange(80)).astype(F32)
    exp = np.stack([x, np.zeros_like(x)], 1)
    bn, mn, mx, centers, H, vols = chain.pdc_grid(exp)
    hi, he, V = _ideal_gas_histograms(600, 1200, F32(24.0), mn, bn, H, 11)
    w = chain.normalized_weighting(counts, {"A": 8, "B": 14})
    g = chain.total_gofr(hi, he, ["A", "B"], counts, w, V, F32(1800) / V, vols, centers)
    # expected pairs per bin ~ N^2/2 * v/V; the weighted mix has at most ~2x the Poisson variance
    npairs = 1800 * 1799 / 2 * vols.astype(np.float64) / float(V)
    sigma = np.sqrt(2.0 / npairs)
    assert np.all(np.abs(g - 1) < 6 * sigma)
    assert abs(np.mean(g) - 1) < 6 * np.sqrt(np.mean(sigma ** 2) / H)


def test_chain_gofr_and_gr_are_one_transform(small_random):
    """G(r) = 4*PI*r*rho0*(g(r)-1) exactly in float32 without a scale factor,
    and a scale factor multiplies g(r)-1 (PairCorrelationConstraints.py:
    26-40, 160-163), against float64."""
    s = small_random
    x = (0.025 + 0.05 * np.arange(160)).astype(F32)
    exp = np.stack([x, np.zeros_like(x)], 1)
    bn, mn, mx, centers, H, vols = chain.pdc_grid(exp)
    hi, he = refcpu.full_pairs_histograms_coords(s["box"], s["basis"], True, s["molecules"], s["elem"], 2, mn, mx,
                                                 bn, H)
    counts = {"O": 400, "Si": 200}
    w = chain.normalized_weighting(counts, {"O": 8, "Si": 14})
    V = systems.volume(s["basis"])
    rho0 = F32(600) / V
    args = (hi, he, ["O", "Si"], counts, w, V, rho0, vols, centers)
    g = chain.total_gofr(*args)
    G = chain.total_gr(*args)
    pf = 4. * chain.PI * centers * rho0
    assert G.tobytes() == (pf * (g - 1)).astype(F32).tobytes()
    gs = chain.total_gofr(*args, scaleFactor=1.25)
    ref = 1 + 1.25 * (g.astype(np.float64) - 1)
    assert np.max(np.abs(gs - ref) / (1 + np.abs(ref))) < 4e-6
    assert not np.allclose(gs, 1.25 * g, atol=1e-3)  # not a plain multiply of g


def test_chain_reduced_sq_is_sq_minus_one_and_vs_float64(small_random):
    """Reduced S(Q) is the S(Q) sum without the +1, scaled without the shift
    (StructureFactorConstraints.py:1255-1261); the transform against a float64
    evaluation of sum_r G(r) dr sin(Qr)/Q from the same histograms."""
    s = small_random
    mn, mx, bn, centers, H, vols = chain.sfc_grid(0.0, 8.0, 0.05)
    hi, he = refcpu.full_pairs_histograms_coords(s["box"], s["basis"], True, s["molecules"], s["elem"], 2, mn, mx,
                                                 bn, H)
    counts = {"O": 400, "Si": 200}
    w = chain.normalized_weighting(counts, {"O": 8, "Si": 14})
    V = systems.volume(s["basis"])
    rho0 = F32(600) / V
    q = np.linspace(0.5, 20, 90).astype(F32)
    M = chain.sq_matrix(centers, q)
    args = (hi, he, ["O", "Si"], counts, w, V, rho0, vols, centers, M)
    S = chain.total_sq(*args)
    R = chain.total_sq(*args, reduced=True)
    assert (R + 1).astype(F32).tobytes() == S.tobytes()
    Rs = chain.total_sq(*args, scaleFactor=0.8, reduced=True)
    assert Rs.tobytes() == (F32(0.8) * R).astype(F32).tobytes()
    Ss = chain.total_sq(*args, scaleFactor=0.8)
    assert np.max(np.abs((Ss - 1) - Rs)) < 1e-5 * (1 + np.max(np.abs(Rs)))
    # float64: G(r) from the histograms, then the sine transform
    g = np.zeros(H)
    for a, b in [("O", "O"), ("O", "Si"), ("Si", "Si")]:
        i, j = ["O", "Si"].index(a), ["O", "Si"].index(b)
        if i == j:
            nij = hi[i, i] + he[i, i]
            D = counts[a] * (counts[a] - 1) / 2 / float(V)
        else:
            nij = hi[i, j] + hi[j, i] + he[i, j] + he[j, i]
            D = counts[a] * counts[b] / float(V)
        g += float(w[a + "-" + b]) * nij.astype(np.float64) / D
    e = np.arange(H + 1) * 0.05
    r = (e[1:] + e[:-1]) / 2
    G64 = 4 * np.pi * r * (600 / float(V)) * (g / (4 / 3 * np.pi * (e[1:] ** 3 - e[:-1] ** 3)) - 1)
    q64 = q.astype(np.float64)
    R64 = (G64[:, None] * 0.05 * np.sin(np.outer(r, q64)) / q64).sum(0)
    assert np.max(np.abs(R - R64)) < 2e-4 * (1 + np.max(np.abs(R64)))


def test_chi_square_is_pairwise_and_close_to_fsum():
    rng = np.random.default_rng(1)
    e = rng.random(1000).astype(F32)
    m = rng.random(1000).astype(F32)
    c = chain.chi_square(e, m)
    assert c.tobytes() == _pairwise(((e - m) ** 2).astype(F32)).tobytes()
    assert c == pytest.approx(math.fsum(((e.astype(np.float64) - m) ** 2)), rel=1e-5)


def _f64_brute_force(box, basis, mol, elem, nel, mn, mx, bn, H):
    """Vectorised binary64 enumeration of all pairs i<j (numpy elementwise
    float64 operations: separately rounded, no fused multiply-add)."""
    c = box.astype(np.float64)
    B = basis.astype(np.float64)
    i, j = np.triu_indices(len(c), 1)
    d = c[i] - c[j]
    b = d - np.where(d > 0, np.floor(d + 0.5), np.ceil(d - 0.5))
    r = (b[:, 0:1] * B[0] + b[:, 1:2] * B[1]) + b[:, 2:3] * B[2]
    dist = np.sqrt((r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1]) + r[:, 2] * r[:, 2])
    mn, mx, bn = float(F32(mn)), float(F32(mx)), float(F32(bn))
    ok = (dist >= mn) & (dist < mx)
    k = ((dist[ok] - mn) / bn).astype(np.int64)
    flat = (elem[i[ok]].astype(np.int64) * nel + elem[j[ok]]) * H + k
    intra = mol[i[ok]] == mol[j[ok]]
    size = nel * nel * H
    keep = flat < size
    out = np.zeros(2 * size, np.int64)
    np.add.at(out, flat[keep] + np.where(intra[keep], 0, size), 1)
    return out[:size].reshape(nel, nel, H), out[size:].reshape(nel, nel, H), int((~keep).sum())


def test_float64_recompute_oracle_vs_brute_force(small_triclinic):
    """The float64 full-histogram oracle equals an independent vectorised
    binary64 enumeration, and differs from the float32 (reference) histogram
    only by pairs moved across a bin edge (same total)."""
    s = small_triclinic
    args = (s["box"], s["basis"], True, s["molecules"], s["elem"], 3, 0.0, 9.0, 0.01, 900)
    hi, he, ov = refcpu.full_pairs_counts_f64(*args, ncores=4)
    bi, be, bov = _f64_brute_force(s["box"], s["basis"], s["molecules"], s["elem"], 3, 0.0, 9.0, 0.01, 900)
    assert np.array_equal(hi, bi) and np.array_equal(he, be) and ov == bov
    fi, fe = refcpu.full_pairs_histograms_coords(s["box"], s["basis"], True, s["molecules"], s["elem"], 3,
                                                 F32(0.0), F32(9.0), F32(0.01), 900)
    assert hi.sum() + he.sum() == fi.sum() + fe.sum()
    moved = np.abs(hi - fi).sum() + np.abs(he - fe).sum()
    assert moved <= 0.002 * (hi.sum() + he.sum())


@pytest.mark.parametrize("cell", ["cubic", "triclinic"])
def test_cell_list_full_histogram_equals_all_pairs(cell):
    """The cell-list form of the full histogram (the checker used at
    200,000 atoms) counts exactly the pairs of the all-pairs loops: float32
    reference binning and the binary64 recompute, box coordinates outside
    [0, 1) included (cells are assigned after wrapping), three or more cells
    per axis; with fewer it declines (None)."""
    rng = np.random.default_rng(17 if cell == "cubic" else 18)
    if cell == "cubic":
        B = systems.cubic_basis(38.0)
    else:
        B = systems.triclinic_basis(40.0, 42.0, 44.0, 80.0, 95.0, 105.0)
    R = systems.reciprocal(B)
    n = 4000
    frac = (rng.random((n, 3)) * 3.0 - 1.0).astype(F32)  # several images of the cell
    real = (frac.astype(np.float64) @ B.astype(np.float64)).astype(F32)
    box = systems.box_coordinates(R, real)
    el = rng.integers(0, 3, n).astype(np.int32)
    mol = (np.arange(n) // 5).astype(np.int32)
    for maxD, H in ((9.5, 950), (12.0, 1200)):
        c32 = refcpu.full_pairs_counts_cells(box, B, mol, el, 3, 0.0, maxD, 0.01, H)
        assert c32 is not None
        hi, he = refcpu.full_pairs_histograms_coords(box, B, True, mol, el, 3, 0.0, maxD, 0.01, H, ncores=4)
        assert np.array_equal(c32[0], hi) and np.array_equal(c32[1], he)
        assert c32[1].sum() > 100000
        c64 = refcpu.full_pairs_counts_cells(box, B, mol, el, 3, 0.0, maxD, 0.01, H, float64=True)
        r64 = refcpu.full_pairs_counts_f64(box, B, True, mol, el, 3, 0.0, maxD, 0.01, H, ncores=4)
        assert np.array_equal(c64[0], r64[0]) and np.array_equal(c64[1], r64[1]) and c64[2] == r64[2]
    assert refcpu.full_pairs_counts_cells(box, B, mol, el, 3, 0.0, 14.0, 0.01, 1400) is None
