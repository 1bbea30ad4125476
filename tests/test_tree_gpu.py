"""Morton keys, stable-sort permutation and leaf offsets built on the GPU
match the CPU oracle bit for bit."""
import math

import numpy as np
import pytest

from phyflow_amd import tree

pytestmark = pytest.mark.gpu


def _domain(x, y, L):
    x0, y0 = x.min(), y.min()
    ext = max(x.max() - x0, y.max() - y0)
    if not ext > 0:
        ext = 1.0
    _, e = math.frexp(ext)
    unit = math.ldexp(1.0, e - 5)
    D = math.ceil(ext / unit) * unit
    return x0, y0, math.ldexp(1.0, L) / D


@pytest.mark.parametrize("n,L,seed", [(1, 0, 0), (2, 1, 1), (1000, 3, 2), (4097, 5, 3),
                                      (100000, 7, 4), (262144, 8, 5), (50000, 10, 6)])
def test_tree_bitexact(n, L, seed, O):
    rng = np.random.default_rng(seed)
    x, y = rng.random((2, n))
    if n > 100:
        x[: n // 10] = x[n // 10: 2 * (n // 10)]  # exact duplicates: ties keep input order
        y[: n // 10] = y[n // 10: 2 * (n // 10)]
        x[-1], y[-1] = x.max(), y.max()          # a point on the far corner (clamped)
    keys, perm, start, x0, y0, inv_s = tree.build(x, y, L)
    ex0, ey0, einv = _domain(x, y, L)
    assert (x0, y0, inv_s) == (ex0, ey0, einv)
    ok = O.morton_keys(x, y, x0, y0, inv_s, L)
    assert np.array_equal(keys, ok)
    sk, operm = O.stable_sort_keys(ok)
    assert np.array_equal(perm, operm)
    assert np.array_equal(start, O.leaf_offsets(sk, 4 ** L))


def test_tree_lattice_ties():
    # lattice points sit exactly on box edges: binning must still agree
    g = np.arange(64) / 64.0
    X, Y = np.meshgrid(g, g, indexing="ij")
    x, y = X.ravel(), Y.ravel()
    from oracle import oracle as O
    keys, perm, start, x0, y0, inv_s = tree.build(x, y, 6)
    assert np.array_equal(keys, O.morton_keys(x, y, x0, y0, inv_s, 6))
    assert np.array_equal(perm, O.stable_sort_keys(keys)[1])
