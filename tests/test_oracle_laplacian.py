"""Pins the numpy restatement of the reference's kNN graph weights (oracle/laplacian.py)
against a pure-Python computation that shares nothing with it: distances as sums of
squared coordinate differences (not norms minus twice the dot product), selection by
sorting (value, column) pairs, the lower median by statistics.median_low, and the
exponential weights row by row."""
import math
import statistics

import numpy as np
import pytest

from oracle import laplacian as L


def brute(x, k):
    N = len(x)
    rows = []
    for i in range(N):
        d = [(sum((a - b) ** 2 for a, b in zip(x[i], x[j])), j) for j in range(N)]
        rows.append(sorted(d)[:k])
    med = statistics.median_low([v for r in rows for v, _ in r])
    cols, vals = [], []
    for r in rows:
        e = [math.exp(-(v / med)) for v, _ in r]
        s = sum(e)
        cols.append([j for _, j in r])
        vals.append([w / s for w in e])
    return np.array(cols), np.array(vals)


@pytest.mark.parametrize("N, d, k, seed", [(12, 3, 4, 0), (17, 5, 6, 1), (9, 1, 9, 2), (20, 8, 3, 3)])
def test_matches_pure_python(N, d, k, seed):
    x = np.random.default_rng(seed).random((N, d))
    cols, vals = L.laplacian(x, k)
    bc, bv = brute(x.tolist(), k)
    np.testing.assert_array_equal(cols, bc)
    np.testing.assert_allclose(vals, bv, rtol=2e-7)   # float32 values


def test_duplicates_break_ties_by_column_and_self_first():
    x = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 0.0], [5.0, 5.0], [1.0, 0.0]])
    cols, vals = L.laplacian(x, 3)
    assert cols.tolist() == [[0, 2, 1], [1, 4, 0], [0, 2, 1], [3, 1, 4], [1, 4, 0]]
    # most selected distances are exact zeros, so the median is 0 and the reference's
    # exp(-d / median) / sum is 0/0: NaN in every row
    assert np.isnan(vals).all()


def test_invariants():
    x = np.random.default_rng(5).random((40, 6))
    cols, vals = L.laplacian(x, 7)
    assert (cols[:, 0] == np.arange(40)).all()          # every point is its own nearest neighbour
    np.testing.assert_allclose(vals.sum(1), 1.0, rtol=1e-6)
    assert (np.diff(vals, axis=1) <= 1e-7).all()         # weights fall with distance
    d = L.distances(x)
    np.testing.assert_allclose(d, d.T, atol=1e-12)
    # permuting the points permutes the graph
    p = np.random.default_rng(6).permutation(40)
    c2, v2 = L.laplacian(x[p], 7)
    inv = np.argsort(p)
    np.testing.assert_allclose(v2[inv], vals, rtol=1e-6)
    assert (p[c2[inv]] == cols).all()
