"""Drop-in for fullrmc's ``fullrmc.Core.reciprocal_space`` extension
(Extensions/reciprocal_space.pyx) on gfx950, plus ``sq_transform``: the
G(r) -> S(Q) transform of the S(Q) constraints for many curves in one call.

Same functions, argument order, dtypes and return values as the reference.
The reference sums each term in double precision into a float32
accumulator (Python-float arithmetic inside its Cython loop); the kernels do
the same, with the device's double ``sin``.
"""
import numpy as np

from .. import _args as A
from .. import _device as D
from .._lib import check, lib

__all__ = ["gr_to_sq", "Gr_to_sq", "sq_to_Gr", "sq_transform"]

SQ_EXACT, SQ_MFMA = 0, 1


def _vec(name, value):
    return A.array(name, value, np.float32, 1)


def gr_to_sq(distances, gr, qrange, rho):
    """S(q) = 1 + 4*PI*rho * sum_r dr*r*sin(q*r)/q*(g(r)-1) from the pair
    correlation function g(r) (reciprocal_space.pyx:42-80)."""
    distances, gr, qrange = _vec("distances", distances), _vec("gr", gr), _vec("qrange", qrange)
    rho = A.c_float32("rho", rho)
    n, m = distances.shape[0], qrange.shape[0]
    if gr.shape[0] < n:
        raise IndexError("gr is shorter than distances")
    if n < 2:
        raise IndexError("distances needs at least two values (dr = distances[1]-distances[0])")
    out = D.empty((m,), np.float32)
    d_r, d_g, d_q = D.upload(distances), D.upload(gr[:n]), D.upload(qrange)
    check(lib().frmc_gr_to_sq(d_r.data_ptr(), d_g.data_ptr(), n, d_q.data_ptr(), m, rho, out.data_ptr(),
                              D.stream()))
    return D.download(out)


def Gr_to_sq(distances, Gr, qrange):
    """S(q) = 1 + sum_r dr*sin(q*r)/q*G(r) from the pair distribution
    function G(r) (reciprocal_space.pyx:89-122)."""
    distances, Gr, qrange = _vec("distances", distances), _vec("Gr", Gr), _vec("qrange", qrange)
    n, m = distances.shape[0], qrange.shape[0]
    if Gr.shape[0] < n:
        raise IndexError("Gr is shorter than distances")
    if n < 2:
        raise IndexError("distances needs at least two values (dr = distances[1]-distances[0])")
    out = D.empty((m,), np.float32)
    d_r, d_g, d_q = D.upload(distances), D.upload(Gr[:n]), D.upload(qrange)
    check(lib().frmc_Gr_to_sq(d_r.data_ptr(), d_g.data_ptr(), n, d_q.data_ptr(), m, out.data_ptr(), D.stream()))
    return D.download(out)


def sq_to_Gr(qValues, rValues, sq):
    """G(r) = (2/PI) * sum_q dq*q*(S(q)-1)*sin(q*r) (reciprocal_space.pyx:
    131-162).  The reference's loop hands an array to the C ``sin`` and
    raises TypeError for every input; this evaluates the transform it
    documents, in double precision rounded once to float32."""
    qValues, rValues, sq = _vec("qValues", qValues), _vec("rValues", rValues), _vec("sq", sq)
    m, nr = qValues.shape[0], rValues.shape[0]
    if sq.shape[0] < m:
        raise IndexError("sq is shorter than qValues")
    if m < 2:
        raise IndexError("qValues needs at least two values (dq = qValues[1]-qValues[0])")
    out = D.empty((nr,), np.float32)
    d_q, d_r, d_s = D.upload(qValues), D.upload(rValues), D.upload(sq[:m])
    check(lib().frmc_sq_to_Gr(d_q.data_ptr(), d_r.data_ptr(), nr, d_s.data_ptr(), m, out.data_ptr(), D.stream()))
    return D.download(out)


def sq_transform(Gr, Gr2SqMatrix, addOne=True, mode="exact"):
    """S(Q) of one G(r) curve (H,) or of k curves (k,H) through the S(Q)
    constraints' matrix (H,nQ): ``np.sum(Gr.reshape((-1,1))*M, axis=0) (+1)``
    per curve (StructureFactorConstraints.py:772-773; addOne=False is the
    reduced form, :1255-1256).  mode "exact" reproduces that float32 sum
    bit for bit; "mfma" runs the contraction on the matrix cores with float32
    accumulation (not bit-identical)."""
    G = np.asarray(Gr)
    if G.dtype != np.float32:
        raise ValueError("Gr must be float32")
    M = A.array("Gr2SqMatrix", Gr2SqMatrix, np.float32, 2)
    single = G.ndim == 1
    G2 = G.reshape(1, -1) if single else G
    if G2.ndim != 2 or G2.shape[1] != M.shape[0]:
        raise ValueError("Gr must be (H,) or (k,H) with H = Gr2SqMatrix.shape[0]")
    modes = {"exact": SQ_EXACT, "mfma": SQ_MFMA}
    if mode not in modes:
        raise ValueError("mode must be 'exact' or 'mfma'")
    k, H, nQ = G2.shape[0], M.shape[0], M.shape[1]
    out = D.empty((k, nQ), np.float32)
    d_g, d_m = D.upload(G2), D.upload(M)
    check(lib().frmc_sq_transform(d_g.data_ptr(), k, H, d_m.data_ptr(), nQ, int(bool(addOne)), modes[mode],
                                  out.data_ptr(), D.stream()))
    S = D.download(out)
    return S[0] if single else S
