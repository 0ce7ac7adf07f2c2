"""ORACLE / TEST INFRASTRUCTURE ONLY -- never imported by the product path.

NumPy 2.x (NEP 50) float32 restatement of the constraint chain that turns
pair histograms into G(r), S(Q), chi-square and the engine total.  NumPy is
the semantics oracle here: the reference computes exactly these NumPy
expressions, so the restatement keeps their operand order and dtypes.

  pdc_grid          PairDistributionConstraints.py:722, 748-758
  sfc_grid          StructureFactorConstraints.py:330-360
  sfc_defaults      StructureFactorConstraints.py:502-575 (rmin, rmax, dr
                    when not given)
  sq_matrix         StructureFactorConstraints.py:302-312
  total_gr          PairDistributionConstraints.py:847-895 (no shape function,
                    no window, constant scale factor)
  total_gofr        PairCorrelationConstraints.py:126-169 (no shape function,
                    no window, constant scale factor)
  total_sq          StructureFactorConstraints.py:780-822, 772-778;
                    reduced=True: ReducedStructureFactorConstraint,
                    StructureFactorConstraints.py:1255-1261
  chi_square        PairDistributionConstraints.py:810-838,
                    StructureFactorConstraints.py:742-770
  total_error       Engine.py:3000-3029 (varianceSquared = 1, Core/Constraint.py:39)
"""
import itertools

import numpy as np

F32 = np.float32
PI = F32(np.pi)


def pdc_grid(experimentalData):
    """bin, minDistance, maxDistance, shellCenters, histSize, shellVolumes."""
    x = np.asarray(experimentalData, dtype=F32)
    bin_ = F32(x[1, 0] - x[0, 0])
    mn = F32(x[0, 0] - bin_ / 2.)
    mx = F32(x[-1, 0] + bin_ / 2.)
    centers = np.array([x[i, 0] for i in range(x.shape[0])], dtype=F32)
    edges = [x[i, 0] - bin_ / 2. for i in range(x.shape[0])]
    edges.append(x[-1, 0] + bin_ / 2.)
    edges = np.array(edges, dtype=F32)
    H = np.int32(len(edges) - 1)
    vols = F32(4.0 / 3.) * PI * ((edges[1:]) ** 3 - edges[0:-1] ** 3)
    return bin_, mn, mx, centers, H, vols


def sfc_grid(rmin, rmax, dr):
    """minDistance, maxDistance, bin, shellCenters, histSize, shellVolumes."""
    mn, mx, bin_ = F32(rmin), F32(rmax), F32(dr)
    edges = np.arange(mn, mx + bin_, bin_).astype(F32)
    mn, mx = edges[0], edges[-1]
    centers = (edges[0:-1] + edges[1:]) / F32(2.)
    H = np.int32(len(edges) - 1)
    vols = F32(4.0 / 3.) * PI * ((edges[1:]) ** 3 - edges[0:-1] ** 3)
    return mn, mx, bin_, centers, H, vols


def sfc_defaults(qvalues):
    """rmin = 2*PI/qmax, rmax = 2*PI/dq, dr = 2*PI/qmax rounded down to a
    multiple of 0.1, all in the float32 the reference computes them in."""
    q = np.asarray(qvalues, F32)
    qmax = q[-1]
    rmin = F32(2. * PI / qmax)
    rmax = F32(2. * PI / (q[1] - q[0]))
    b = 2. * PI / qmax
    rb = round(b, 1)
    if rb > b:
        rb -= 0.1
    return rmin, rmax, F32(rb)


def sq_matrix(centers, qvalues):
    Rs, Qs = np.asarray(centers, F32), np.asarray(qvalues, F32)
    dr = Rs[1] - Rs[0]
    qr = Rs.reshape((-1, 1)) * (np.ones((len(Rs), 1), dtype=F32) * Qs)
    return dr * (np.sin(qr) / Qs)


def _gr_sum(intra, inter, elements, counts, weights, volume, cast_density):
    """sum over sorted element pairs of w_ij * n_ij / D_ij (float32)."""
    H = intra.shape[2]
    Gr = np.zeros(H, dtype=F32)
    for a, b in sorted(itertools.combinations_with_replacement(elements, 2)):
        w = weights.get(a + "-" + b, weights.get(b + "-" + a))
        ni, nj = counts[a], counts[b]
        i, j = elements.index(a), elements.index(b)
        if i == j:
            Nij = ni * (ni - 1) / 2.0
            nij = intra[i, j, :] + inter[i, j, :]
        else:
            Nij = ni * nj
            nij = intra[i, j, :] + intra[j, i, :] + inter[i, j, :] + inter[j, i, :]
        Dij = Nij / volume
        if cast_density:
            Dij = F32(Dij)
        Gr += w * nij / Dij
    return Gr


def total_gr(intra, inter, elements, counts, weights, volume, rho0, shellVolumes, centers, scaleFactor=1.0):
    """PDC model G(r)."""
    Gr = _gr_sum(intra, inter, elements, counts, weights, volume, cast_density=True)
    Gr /= shellVolumes
    Gr = (4. * PI * centers * rho0) * (Gr - 1)
    if scaleFactor != 1:
        Gr *= F32(scaleFactor)
    return Gr


def total_gofr(intra, inter, elements, counts, weights, volume, rho0, shellVolumes, centers, scaleFactor=1.0):
    """PCF model g(r): the G(r) sum divided by the shell volumes; a scale
    factor acts on 4*PI*r*rho0*(g-1) and is mapped back to g."""
    gr = _gr_sum(intra, inter, elements, counts, weights, volume, cast_density=False)
    gr /= shellVolumes
    if scaleFactor != 1:
        Gr = (4. * PI * centers * rho0) * (gr - 1)
        Gr *= F32(scaleFactor)
        gr = 1. + Gr / (4. * PI * centers * rho0)
    return gr


def total_sq(intra, inter, elements, counts, weights, volume, rho0, shellVolumes, centers, Gr2Sq,
             scaleFactor=1.0, reduced=False):
    """SFC model S(Q); reduced=True: the reduced S(Q) (no +1, plain scaling)."""
    Gr = _gr_sum(intra, inter, elements, counts, weights, volume, cast_density=False)
    Gr /= shellVolumes
    Gr = (F32(4.) * PI * centers * rho0) * (Gr - 1)
    if reduced:
        Sq = np.sum(Gr.reshape((-1, 1)) * Gr2Sq, axis=0)
        if scaleFactor != 1:
            Sq = F32(scaleFactor) * Sq
        return Sq
    Sq = np.sum(Gr.reshape((-1, 1)) * Gr2Sq, axis=0) + 1
    if scaleFactor != 1:
        Sq = scaleFactor * (Sq - 1) + 1
    return Sq


def chi_square(experimental, model, dataWeights=None):
    diff = experimental - model
    if dataWeights is None:
        return np.add.reduce((diff) ** 2)
    return np.add.reduce(dataWeights * ((diff) ** 2))


def total_error(chi2s):
    return np.sum([c / 1 for c in chi2s])


def normalized_weighting(counts, weights):
    """Faber-Ziman pair weights c_i c_j b_i b_j / (sum_k c_k b_k)^2, doubled for
    unlike pairs, as float32.  The reference takes these from pdbparser
    (absent here), so this formula is an input convention, not pinned."""
    els = sorted(counts)
    total = float(sum(counts.values()))
    c = {e: counts[e] / total for e in els}
    mean = sum(c[e] * float(weights[e]) for e in els)
    out = {}
    for a, b in itertools.combinations_with_replacement(els, 2):
        w = c[a] * c[b] * float(weights[a]) * float(weights[b]) / mean ** 2
        if a != b:
            w *= 2.0
        out[a + "-" + b] = F32(w)
    return out
