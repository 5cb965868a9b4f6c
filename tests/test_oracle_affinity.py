"""Pins the oracle's affinity construction (kNN, bandwidth search, symmetrisation, normalisation)
against the reference's own known answers, closed forms and brute force."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle
from tests._util import f32_seq_dist, lcg_samples


# --- kNN -------------------------------------------------------------------------------------

@pytest.mark.parametrize("n,dim,k", [(80, 4, 30), (200, 7, 12), (64, 3, 63)])
def test_knn_matches_bruteforce_with_index_tiebreak(n, dim, k):
    # brute_force_neighbors (src/test.rs:866-881) with the (distance, index) order
    x = lcg_samples(n, dim, 11)
    dist = f32_seq_dist(x)
    np.fill_diagonal(dist, np.inf)
    order = np.lexsort((np.broadcast_to(np.arange(n), (n, n)), dist), axis=1)[:, :k]
    idx, d = oracle.knn(x, k)
    assert np.array_equal(idx, order.astype(np.uint32))
    assert np.array_equal(d, np.take_along_axis(dist, order, 1))


def test_knn_ties_break_by_index():
    # four coincident points and a far one: every row's ties resolve by ascending index
    x = np.zeros((5, 2), np.float32)
    x[4] = [10.0, 0.0]
    idx, d = oracle.knn(x, 3)
    assert idx[0].tolist() == [1, 2, 3]
    assert idx[2].tolist() == [0, 1, 3]
    assert idx[4].tolist() == [0, 1, 2]
    assert np.all(d[:4] == 0.0)


# --- bandwidth search ------------------------------------------------------------------------

def _kat_row():
    # src/test.rs:1023-1049: 90 distances sqrt(20 + 100 (i+1)/90)
    return np.sqrt(20.0 + 100.0 * (np.arange(90) + 1.0) / 90.0)


@pytest.mark.parametrize("f64", [True, False])
def test_search_beta_known_answer(f64):
    p, _ = oracle.search_beta_row(_kat_row(), 30.0, f64=f64)
    p = p.astype(np.float64)
    h = -(p[p > 0] * np.log(p[p > 0])).sum()
    assert abs(math.exp(h) - 30.0) < 0.1


def test_search_beta_matches_fp64_resolve():
    # Independent solve: find beta with H(beta) = ln(perp) by Brent's method in f64, then compare
    # the resulting row against the f32 bisection (catches a flipped bracket or a missing square).
    rng = np.random.default_rng(3)
    for _ in range(20):
        d = np.sort(rng.uniform(0.5, 6.0, 90)).astype(np.float32)
        perp = 30.0

        d2 = d.astype(np.float64) ** 2

        def h_minus(b):
            # entropy with the exponent shifted by the smallest distance (no underflow)
            w = np.exp(-b * (d2 - d2.min()))
            s = w.sum()
            return b * (w * d2).sum() / s - b * d2.min() + np.log(s) - np.log(perp)

        beta = brentq(h_minus, 1e-8, 1e4, xtol=1e-14)
        w = np.exp(-beta * (d2 - d2.min()))
        want = w / w.sum()
        got, _ = oracle.search_beta_row(d, perp)
        assert np.max(np.abs(got - want) / want.max()) < 2e-3


def test_search_beta_equal_distances_is_uniform():
    # equal distances: H = ln k for every beta, so perplexity k is met on the first step and P
    # is uniform 1/k
    d = np.full(10, 2.0, np.float32)
    p, it = oracle.search_beta_row(d, 10.0)
    assert it == 1
    assert np.allclose(p, 0.1, rtol=1e-6)


# --- symmetrisation --------------------------------------------------------------------------

def _random_knn(n, k, seed):
    rng = np.random.default_rng(seed)
    cols = np.zeros((n, k), np.uint32)
    for i in range(n):
        cand = np.delete(np.arange(n), i)
        cols[i] = rng.choice(cand, k, replace=False)
    vals = rng.uniform(0.01, 1.0, (n, k)).astype(np.float32)
    return cols, vals


def _q6_rows(cols, n):
    """Row order derived from src/tsne/mod.rs:343-397 (one ordered list per row)."""
    lists = [sorted(int(c) for c in cols[i]) for i in range(len(cols))]
    sets = [set(l) for l in lists]
    out = []
    for r in range(n):
        first = [m for m in range(r) if r in sets[m]]
        second = [c for c in lists[r] if not (c < r and r in sets[c])]
        third = [m for m in range(r + 1, n) if r in sets[m] and m not in sets[r]]
        out.append(first + second + third)
    return out


@pytest.mark.parametrize("n,k,seed", [(12, 3, 0), (40, 6, 1), (60, 10, 2), (25, 24, 3)])
def test_symmetrize_structure_and_values(n, k, seed):
    cols, vals = _random_knn(n, k, seed)
    rows, sc, sv = oracle.symmetrize(cols, vals)
    # row structure: the insertion-order rule
    want = _q6_rows(cols, n)
    for r in range(n):
        got = sc[rows[r]:rows[r + 1]].tolist()
        assert got == want[r], r
    # values: dense (P + P^T)/2 restricted to the union pattern
    dense = np.zeros((n, n), np.float64)
    for i in range(n):
        dense[i, cols[i]] = vals[i]
    sym = 0.5 * (dense + dense.T)
    for r in range(n):
        for e in range(int(rows[r]), int(rows[r + 1])):
            assert abs(sv[e] - sym[r, sc[e]]) <= 1e-7 * max(1.0, sym[r, sc[e]])
    # every nonzero of the dense symmetric matrix appears exactly once per row
    assert int(rows[-1]) == int((sym > 0).sum())


def test_symmetrize_self_loop_counted_once():
    # caller-supplied neighbours may name the sample itself; the reference keeps one entry p
    cols = np.array([[0, 1], [0, 2], [1, 0]], np.uint32)
    vals = np.array([[0.5, 0.25], [0.125, 0.5], [0.75, 0.0625]], np.float32)
    rows, sc, sv = oracle.symmetrize(cols, vals)
    r0 = dict(zip(sc[rows[0]:rows[1]].tolist(), sv[rows[0]:rows[1]].tolist()))
    assert r0[0] == pytest.approx(0.5)


def test_normalize_and_stop_lying():
    rng = np.random.default_rng(0)
    v = rng.uniform(0, 1, 1000).astype(np.float32)
    oracle.normalize_p(v, 12.0)
    assert abs(float(v.astype(np.float64).sum()) - 12.0) < 1e-3
    oracle.stop_lying(v, 12.0)
    assert abs(float(v.astype(np.float64).sum()) - 1.0) < 1e-4


def test_auto_learning_rate_known_answers():
    # src/test.rs:26-69
    assert oracle.auto_learning_rate(100, 12.0) == 50.0
    assert oracle.auto_learning_rate(120_000, 12.0) == np.float32(120_000.0 / 12.0 / 4.0)
    assert oracle.auto_learning_rate(120_000, 6.0) == 2 * oracle.auto_learning_rate(120_000, 12.0)
    assert 205.0 <= oracle.auto_learning_rate(10_000, 12.0) <= 212.0


def test_build_affinities_row_sums():
    # after normalisation with exaggeration 1 the symmetric P sums to 1 (src/test.rs:1344-1394)
    x = lcg_samples(150, 5, 7)
    rows, cols, vals = oracle.build_affinities(x, 10.0)
    v = vals.copy()
    oracle.normalize_p(v, 1.0)
    assert abs(float(v.astype(np.float64).sum()) - 1.0) < 1e-4
    assert int(rows[-1]) >= 150 * 30
