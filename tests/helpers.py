"""Shared helpers of the test suite (no GPU work here)."""
import numpy as np


def interior_mask(n):
    """Boolean [3][n3+1][n2+1][n1+1] mask of the edges that carry an equation."""
    n1, n2, n3 = n
    k, j, i = np.meshgrid(np.arange(n3 + 1), np.arange(n2 + 1), np.arange(n1 + 1), indexing="ij")
    p = (i, j, k)
    m = np.zeros((3, n3 + 1, n2 + 1, n1 + 1), bool)
    nn = (n1, n2, n3)
    for a in range(3):
        b, c = (a + 1) % 3, (a + 2) % 3
        m[a] = (p[a] < nn[a]) & (p[b] >= 1) & (p[b] <= nn[b] - 1) & (p[c] >= 1) & (p[c] <= nn[c] - 1)
    return m


def random_field(n, seed, interior_only=True):
    rng = np.random.default_rng(seed)
    n1, n2, n3 = n
    v = rng.standard_normal((3, n3 + 1, n2 + 1, n1 + 1)) + 1j * rng.standard_normal((3, n3 + 1, n2 + 1, n1 + 1))
    if interior_only:
        v = np.where(interior_mask(n), v, 0)
    return np.ascontiguousarray(v.reshape(-1))


def rel(a, b):
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
