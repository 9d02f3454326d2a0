"""Shared test helpers: seeded fields and comparisons."""
import numpy as np


def rand_storage(shape, seed, lo=-1.0, hi=1.0):
    return np.asfortranarray(np.random.default_rng(seed).uniform(lo, hi, size=shape))


def rel_err(a, b):
    """max |a - b| / max |b| (the reference's agreement measure, test/support/agreement.jl:60-70)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0))


def interior(a, H):
    return a[H:-H, H:-H, H:-H]


def cell_avgs_sin(x0, h, off0, n):
    """Cell averages of sin on the n-cell stencil whose cell 0 is left of x0
    (test/advection.jl:21-22 with F = -cos)."""
    out = []
    for j in range(n):
        xc = x0 - h / 2 + (off0 + j) * h
        out.append((-np.cos(xc + h / 2) + np.cos(xc - h / 2)) / h)
    return out
