"""ORACLE / TEST INFRASTRUCTURE ONLY -- never imported by the product path.

Restatement of Extensions/reciprocal_space.pyx.  The Cython loops mix C
float32 arithmetic (typed variables) with Python-object arithmetic (every
expression that touches ``np.sin``), so each term is a Python float (double)
added to the float32 accumulator in double and rounded back to float32.
Pure-Python loops, one ``np.sin`` call per term as in the reference: small
cases only.

  gr_to_sq   reciprocal_space.pyx:42-80
  Gr_to_sq   reciprocal_space.pyx:89-122
  sq_to_Gr   reciprocal_space.pyx:131-162 raises TypeError there for every
             input (an array handed to libc sin); this is the float64
             transform it documents
"""
import numpy as np

F32 = np.float32


def Gr_to_sq(distances, Gr, qrange):
    r = np.asarray(distances, F32)
    G = np.asarray(Gr, F32)
    qs = np.asarray(qrange, F32)
    sq = np.ones(len(qs), dtype=F32)
    dr = F32(r[1] - r[0])                          # C float subtraction
    for qi in range(len(qs)):
        q = qs[qi]
        acc = sq[qi]
        for ri in range(len(r)):
            x = float(F32(q * r[ri]))              # C float product, to a Python float
            term = float(dr) * (float(np.sin(x)) / float(q)) * float(G[ri])
            acc = F32(float(acc) + term)           # Python-float +=, stored as C float
        sq[qi] = acc
    return sq


def gr_to_sq(distances, gr, qrange, rho):
    r = np.asarray(distances, F32)
    g = np.asarray(gr, F32)
    qs = np.asarray(qrange, F32)
    sq = np.ones(len(qs), dtype=F32)
    dr = F32(r[1] - r[0])
    fact = F32(F32(F32(4.0) * F32(3.1415927)) * F32(rho))
    for qi in range(len(qs)):
        q = qs[qi]
        acc = sq[qi]
        for ri in range(len(r)):
            drr = float(F32(dr * r[ri]))           # dr*r: C float product
            x = float(F32(q * r[ri]))
            g1 = float(F32(g[ri] - F32(1)))        # gr[r]-1: C float
            term = float(fact) * ((drr * (float(np.sin(x)) / float(q))) * g1)
            acc = F32(float(acc) + term)
        sq[qi] = acc
    return sq


def sq_to_Gr(qValues, rValues, sq):
    q = np.asarray(qValues, F32).astype(np.float64)
    r = np.asarray(rValues, F32).astype(np.float64)
    s = np.asarray(sq, F32).astype(np.float64)
    dq = q[1] - q[0]
    return (2.0 / np.pi * dq * (np.sin(np.outer(r, q)) * (q * (s - 1.0))).sum(1)).astype(F32)
