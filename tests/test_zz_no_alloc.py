"""The no-allocation S(Q) transform entry (frmc_sq_transform_noalloc),
called through the C ABI: bit-identical to frmc_sq_transform in both modes
on two calls in a row (the second smaller, so a stale scratch tail would
show), and a clear refusal of a shape whose split partials exceed its fixed
scratch array."""
import ctypes

import numpy as np
import pytest

from fullrmc_amd import _device as D
from fullrmc_amd import _lib

pytestmark = pytest.mark.gpu


def _curves(k, H, nQ, seed):
    rng = np.random.default_rng(seed)
    r = (0.005 + 0.01 * np.arange(H)).astype(np.float32)
    q = np.linspace(0.3, 30.0, nQ).astype(np.float32)
    M = ((r[1] - r[0]) * (np.sin(r[:, None] * q[None, :]) / q)).astype(np.float32)
    G = (rng.standard_normal((k, H)) * np.exp(-r / 6)).astype(np.float32)
    return G, M


def _both(L, G, M, mode, add_one):
    k, H = G.shape
    nQ = M.shape[1]
    d_g, d_m = D.upload(G), D.upload(M)
    d_a, d_b = D.empty((k, nQ), np.float32), D.empty((k, nQ), np.float32)
    d_b.fill_(np.nan)
    _lib.check(L.frmc_sq_transform(d_g.data_ptr(), k, H, d_m.data_ptr(), nQ, add_one, mode, d_a.data_ptr(),
                                   D.stream()))
    _lib.check(L.frmc_sq_transform_noalloc(d_g.data_ptr(), k, H, d_m.data_ptr(), nQ, add_one, mode,
                                           d_b.data_ptr(), D.stream()))
    return D.download(d_a), D.download(d_b)


def test_noalloc_transform_equals_the_allocating_entry():
    L = _lib.lib()
    D.require_gpu()
    # 128 curves x 2,000 r x 1,000 Q: 32 r splits, 4,096,000 partial floats;
    # then a smaller call (16 x 500 x 300, 8 splits) over the same scratch
    for k, H, nQ, seed in ((128, 2000, 1000, 1), (16, 500, 300, 2), (3, 40, 70, 3)):
        G, M = _curves(k, H, nQ, seed)
        for mode in (1, 0):
            for add_one in (1, 0):
                a, b = _both(L, G, M, mode, add_one)
                assert np.isfinite(a).all()
                assert a.tobytes() == b.tobytes(), (k, H, nQ, mode, add_one)
        # the matrix-core result is the transform (float64 check, loose)
        S64 = G.astype(np.float64) @ M.astype(np.float64)
        a, b = _both(L, G, M, 1, 0)
        assert np.max(np.abs(b - S64)) <= 1e-5 * max(1.0, np.max(np.abs(S64)))


def test_noalloc_transform_refuses_a_shape_beyond_its_scratch():
    L = _lib.lib()
    D.require_gpu()
    # 1,600 curves x 2,560 Q: 25 x 40 output tiles, 2 r splits,
    # 8,192,000 partial floats > 2^22
    k, H, nQ = 1600, 128, 2560
    d_g, d_m = D.empty((k, H), np.float32), D.empty((H, nQ), np.float32)
    d_s = D.empty((k, nQ), np.float32)
    rc = L.frmc_sq_transform_noalloc(d_g.data_ptr(), k, H, d_m.data_ptr(), nQ, 1, 1, d_s.data_ptr(), D.stream())
    assert rc != 0
    msg = L.frmc_last_error().decode()
    assert "8192000" in msg and "4194304" in msg, msg
    # sizes that need no partials still run
    assert L.frmc_sq_transform_noalloc(d_g.data_ptr(), 0, H, d_m.data_ptr(), nQ, 1, 1, d_s.data_ptr(),
                                       D.stream()) == 0
