"""Flock statistics: polarisation and mean nearest-neighbour distance.

Polarisation is the length of the mean unit heading, 1 for a perfectly
aligned flock and about 1/sqrt(n) for random headings.  The nearest-neighbour
distance uses the torus metric of the field (toroidal_distance per axis,
field_2d.rs:988-1002), by brute force in float64 blocks.
"""
import numpy as np


def polarisation(hx, hy):
    hx = np.asarray(hx, np.float64)
    hy = np.asarray(hy, np.float64)
    n = np.hypot(hx, hy)
    ok = n > 0
    if not ok.any():
        return 0.0
    return float(np.hypot((hx[ok] / n[ok]).mean(), (hy[ok] / n[ok]).mean()))


def mean_nn_distance(x, y, width, height, block=2048):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    n = len(x)
    best = np.full(n, np.inf)
    for s in range(0, n, block):
        dx = np.abs(x[s:s + block, None] - x[None, :])
        dy = np.abs(y[s:s + block, None] - y[None, :])
        dx = np.minimum(dx, width - dx)
        dy = np.minimum(dy, height - dy)
        d = np.hypot(dx, dy)
        idx = np.arange(s, min(n, s + block))
        d[idx - s, idx] = np.inf
        best[s:s + block] = d.min(axis=1)
    return float(best.mean())
