"""Helpers of fullrmc's Core/Collection.py that the pair path's inputs use."""
import numpy as np


def rebin(data, bin=0.05, check=False):
    """Re-bin (N,2) data onto equally spaced bins of width ``bin``
    (Core/Collection.py:274-314): bins centred on multiples of ``bin`` from
    the first to the last abscissa; each new value is the mean of the
    ordinates falling in its bin (0 for an empty bin).  Returns (X, Y)."""
    if check:
        data = np.asarray(data)
        if data.ndim != 2 or data.shape[1] != 2:
            raise ValueError("data must be an (N,2) array")
        bin = float(bin)
        if not bin > 0:
            raise ValueError("bin must be positive")
    x = data[:, 0].astype(float)
    y = data[:, 1].astype(float)
    lo = int(x[0] / bin) * bin - bin / 2.
    hi = int(x[-1] / bin) * bin + bin / 2.
    edges = np.arange(lo, hi, bin)
    if edges[-1] != hi:
        edges = np.append(edges, hi)
    counts, _ = np.histogram(x, bins=edges)
    counts[counts == 0] = 1
    sums, edges = np.histogram(x, bins=edges, weights=y)
    return (edges[1:] + edges[:-1]) / 2., sums / counts
