"""The G(r) -> S(Q) transform entry and the reciprocal_space drop-in:
oracle pins on the CPU (closed forms, identities, float64 round trips) and
GPU parity of the HIP kernels against the oracle and NumPy."""
import numpy as np
import pytest

from fullrmc_amd.Core import reciprocal_space as grs
from oracle import chain
from oracle import reciprocal as orc

F32 = np.float32


def _ulps(a, b):
    a = np.asarray(a, F32).view(np.int32).astype(np.int64)
    b = np.asarray(b, F32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


# ------------------------------------------------------------ oracle pins

def test_oracle_Gr_to_sq_single_shell_closed_form():
    """One non-zero G(r) bin at r0: S(q) = 1 + dr*sin(q*r0)/q*G0."""
    r = (0.05 + 0.1 * np.arange(60)).astype(F32)
    G = np.zeros(60, F32)
    G[23] = F32(2.5)
    q = np.linspace(0.3, 20, 40).astype(F32)
    S = orc.Gr_to_sq(r, G, q)
    q64 = q.astype(np.float64)
    ref = 1 + 0.1 * np.sin(q64 * float(r[23])) / q64 * 2.5
    assert np.max(np.abs(S - ref)) < 1e-6


def test_oracle_Gr_to_sq_matches_the_constraint_transform():
    """Gr_to_sq and the S(Q) constraints' matrix transform are the same sum
    up to float32 rounding (each pinned against float64)."""
    rng = np.random.default_rng(3)
    r = (0.01 + 0.02 * np.arange(250)).astype(F32)
    G = (rng.standard_normal(250) * np.exp(-r / 3)).astype(F32)
    q = np.linspace(0.5, 25, 60).astype(F32)
    S = orc.Gr_to_sq(r, G, q)
    S_m = np.sum(G.reshape((-1, 1)) * chain.sq_matrix(r, q), axis=0) + 1
    q64, r64 = q.astype(np.float64), r.astype(np.float64)
    S64 = 1 + (G.astype(np.float64)[:, None] * 0.02 * np.sin(np.outer(r64, q64)) / q64).sum(0)
    assert np.max(np.abs(S - S64)) < 2e-6 * (1 + np.max(np.abs(S64)))
    assert np.max(np.abs(S - S_m)) < 4e-6 * (1 + np.max(np.abs(S64)))


def test_oracle_gr_to_sq_is_Gr_to_sq_of_the_matching_G():
    """g(r) = 1 + G(r)/(4*PI*rho*r) gives the same S(q) through gr_to_sq as
    G(r) through Gr_to_sq (pins 4*PI*rho, the factor r and the -1)."""
    rng = np.random.default_rng(4)
    r = (0.05 + 0.1 * np.arange(120)).astype(F32)
    G = (rng.standard_normal(120) * 0.5).astype(F32)
    rho = F32(0.07)
    g = (1 + G.astype(np.float64) / (4 * np.pi * 0.07 * r.astype(np.float64))).astype(F32)
    q = np.linspace(0.4, 18, 50).astype(F32)
    a = orc.gr_to_sq(r, g, q, rho)
    b = orc.Gr_to_sq(r, G, q)
    assert np.max(np.abs(a - b)) < 2e-5 * (1 + np.max(np.abs(b)))


def test_oracle_sq_to_Gr_inverts_the_forward_transform():
    """A smooth G(r) through the forward sine transform (float64) and back
    through sq_to_Gr: G(r) = (2/PI) int Q (S-1) sin(Qr) dQ."""
    r = np.arange(0.0, 12.0, 0.01)
    G = np.exp(-((r - 3.0) ** 2) / (2 * 0.25 ** 2)) - 0.5 * np.exp(-((r - 5.0) ** 2) / (2 * 0.4 ** 2))
    q = np.arange(0.005, 40.0, 0.01)
    S = 1 + (G[:, None] * 0.01 * np.sin(np.outer(r, q)) / q).sum(0)
    rv = np.linspace(1.0, 8.0, 71).astype(F32)
    back = orc.sq_to_Gr(q.astype(F32), rv, S.astype(F32))
    ref = np.exp(-((rv - 3.0) ** 2) / (2 * 0.25 ** 2)) - 0.5 * np.exp(-((rv - 5.0) ** 2) / (2 * 0.4 ** 2))
    assert np.max(np.abs(back - ref)) < 2e-3


def test_argument_errors_match_reference_types():
    r = np.arange(5, dtype=F32)
    with pytest.raises(ValueError, match="Buffer dtype mismatch"):
        grs.Gr_to_sq(r.astype(np.float64), r, r)
    with pytest.raises(TypeError):
        grs.gr_to_sq(None, r, r, 1.0)
    with pytest.raises(ValueError, match="wrong number of dimensions"):
        grs.sq_to_Gr(r, r.reshape(5, 1), r)
    with pytest.raises(ValueError):
        grs.sq_transform(r, np.zeros((4, 3), F32))


# ------------------------------------------------------------- GPU parity

@pytest.mark.gpu
@pytest.mark.parametrize("k,H,nQ", [(1, 300, 7), (5, 1000, 500), (3, 4500, 300), (130, 700, 65)])
@pytest.mark.parametrize("addOne", [True, False])
def test_sq_transform_exact_is_numpy_axis0_sum(k, H, nQ, addOne):
    rng = np.random.default_rng(k * 7 + H)
    G = rng.standard_normal((k, H)).astype(F32)
    M = (rng.standard_normal((H, nQ)) * 0.02).astype(F32)
    S = grs.sq_transform(G, M, addOne=addOne)
    for p in range(k):
        ref = np.sum(G[p].reshape((-1, 1)) * M, axis=0)
        if addOne:
            ref = ref + 1
        assert S[p].tobytes() == ref.astype(F32).tobytes(), "curve %d" % p


@pytest.mark.gpu
def test_sq_transform_exact_zero_sign_and_the_constraint_S():
    # NumPy's reduction starts from +0: a sum of -0 products is +0
    H, nQ = 50, 9
    M = np.abs(np.random.default_rng(1).standard_normal((H, nQ))).astype(F32)
    G = np.full(H, -0.0, F32)
    S = grs.sq_transform(G, M, addOne=False)
    assert S.tobytes() == np.sum(G.reshape((-1, 1)) * M, axis=0).tobytes()
    assert not np.any(np.signbit(S))
    # the S(Q) constraint's own matrix and G(r)
    centers = (0.01 + 0.02 * np.arange(1000)).astype(F32)
    q = np.linspace(0.5, 25, 500).astype(F32)
    Mc = chain.sq_matrix(centers, q)
    Gc = (np.sin(centers) * np.exp(-centers / 5)).astype(F32)
    ref = np.sum(Gc.reshape((-1, 1)) * Mc, axis=0) + 1
    assert grs.sq_transform(Gc, Mc).tobytes() == ref.tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("k,H,nQ", [(1, 5, 3), (16, 64, 16), (37, 1003, 501), (128, 2000, 1000)])
def test_sq_transform_mfma_vs_float64(k, H, nQ):
    """Matrix-core contraction: within float32 accumulation error of the
    float64 product (bound: 1e-6 of the sum of |terms| per Q)."""
    rng = np.random.default_rng(H + nQ)
    G = rng.standard_normal((k, H)).astype(F32)
    M = (rng.standard_normal((H, nQ)) * 0.02).astype(F32)
    S = grs.sq_transform(G, M, addOne=True, mode="mfma")
    S64 = G.astype(np.float64) @ M.astype(np.float64) + 1
    scale = np.abs(G.astype(np.float64)) @ np.abs(M.astype(np.float64)) + 1
    assert np.max(np.abs(S - S64) / scale) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("n,m", [(2, 1), (120, 40), (2300, 37)])
def test_Gr_to_sq_and_gr_to_sq_vs_oracle(n, m):
    """Terms in double, float32 accumulator: the device and host double sin
    may differ in the last bit, which reaches the float32 result only at a
    rounding boundary; every value within 1 ulp, nearly all identical."""
    rng = np.random.default_rng(n + m)
    r = (0.05 + 0.1 * np.arange(n)).astype(F32)
    G = rng.standard_normal(n).astype(F32)
    q = np.linspace(0.3, 20, m).astype(F32)
    for got, ref in [(grs.Gr_to_sq(r, G, q), orc.Gr_to_sq(r, G, q)),
                     (grs.gr_to_sq(r, 1 + G, q, 0.05), orc.gr_to_sq(r, 1 + G, q, F32(0.05)))]:
        u = _ulps(got, ref)
        assert u.max() <= 1
        assert np.mean(u == 0) >= 0.99


@pytest.mark.gpu
def test_sq_to_Gr_vs_oracle():
    q = np.arange(0.01, 30.0, 0.02).astype(F32)
    S = (1 + np.sin(q) / q).astype(F32)
    rv = np.linspace(0.5, 10, 200).astype(F32)
    got = grs.sq_to_Gr(q, rv, S)
    ref = orc.sq_to_Gr(q, rv, S)
    assert _ulps(got, ref).max() <= 1 or np.max(np.abs(got - ref)) < 1e-6 * np.max(np.abs(ref))
