"""Pins of the oracle's KL divergence and run-loop schedule that do not retype its formulas:

* the KL is recomputed in FP64 numpy from the definition of the reference's formula
  (src/tsne/mod.rs:738-775: every edge of row i weighted by p_values[i], the sample index) and
  of the true KL, on a non-uniform P where the two differ, with planted mistakes (a dropped 1/Z,
  the edge value in place of p_values[i]) shown to fail the same comparison;
* kl_divergence() is checked against the FP64 exact Z (all pairs) within the approximation's
  own accuracy;
* the schedule (src/lib.rs:1162-1179: stop-lying and the momentum switch fire after the update of
  epoch == k, so epochs 0..=k run exaggerated with the first momentum) is pinned by composing a run
  from pieces with no schedule at all: k + 1 epochs, then an explicit stop_lying and the final
  momentum, then the rest.  The composed run must equal the scheduled one bit for bit, and the
  compositions one epoch early or late must not.
"""
import numpy as np
import pytest

import oracle
from tests._util import lcg_samples

FLT_MIN = np.float64(np.finfo(np.float32).tiny)  # T::min_positive_value() for f32


def _kl_numpy(y, rows, cols, vals, inv_z, weight="sample"):
    y = np.asarray(y, np.float64)
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64)
    n = y.shape[0]
    row_of = np.repeat(np.arange(n), np.diff(rows))
    d2 = ((y[row_of] - y[cols]) ** 2).sum(1)
    q = inv_z / (1.0 + d2)
    p_ref = vals[row_of] if weight == "sample" else vals
    ref = float((p_ref * np.log((p_ref + FLT_MIN) / (q + FLT_MIN))).sum())
    true = float((vals * np.log(vals / q)).sum())
    return ref, true


@pytest.fixture(scope="module")
def graph():
    x = lcg_samples(200, 6, 21)
    rows, cols, vals = oracle.build_affinities(x, 8.0)
    oracle.normalize_p(vals, 1.0)
    y = (lcg_samples(200, 2, 22) * np.float32(6.0)).astype(np.float32)
    return x, rows, cols, vals, y


def test_sparse_kl_against_numpy_on_non_uniform_p(graph):
    _, rows, cols, vals, y = graph
    assert vals.std() > 0.3 * vals.mean()  # P is far from uniform: the two formulas differ
    inv_z = np.float32(1.0 / 1234.5)
    r, t = oracle.sparse_kl(y, rows, cols, vals, inv_z)
    nr, nt = _kl_numpy(y, rows, cols, vals, float(inv_z))
    assert abs(r - nr) / abs(nr) < 1e-5
    assert abs(t - nt) / abs(nt) < 1e-6  # q is formed in f32 as the reference does
    assert abs(nr - nt) / abs(nt) > 1e-2  # the quirk matters on this P
    # planted mistakes move the value far beyond the tolerance above
    for bad in (_kl_numpy(y, rows, cols, vals, 1.0),                # 1/Z dropped
                _kl_numpy(y, rows, cols, vals, float(inv_z), "edge")):  # edge value as weight
        assert abs(r - bad[0]) / abs(nr) > 1e-3


@pytest.mark.parametrize("method,theta,tol", [("fitsne", 0.5, 2e-3), ("bh", 0.2, 2e-3)])
def test_kl_divergence_uses_the_fit_z(graph, method, theta, tol):
    _, rows, cols, vals, y = graph
    _, z = oracle.exact_repulsion_f64(y)
    r, t = oracle.kl_divergence(y, rows, cols, vals, method=method, theta=theta)
    nr, nt = _kl_numpy(y, rows, cols, vals, 1.0 / z)
    # the fit's Z approximates the exact one to ~1e-3 (interpolation.rs:520-581); the KL moves by
    # sum(P) * log(Z ratio) at most
    assert abs(r - nr) / abs(nr) < tol
    assert abs(t - nt) / abs(nt) < tol


def _run(state, rows, cols, vals, e0, e1, momentum, sl=10**9, ms=10**9):
    y, uy, g = state
    return oracle.run_epochs(y, uy, g, rows, cols, vals, method="fitsne",
                             learning_rate=oracle.auto_learning_rate(y.shape[0], 12.0),
                             momentum=momentum, final_momentum=0.8, early_exaggeration=12.0,
                             momentum_switch_epoch=ms, stop_lying_epoch=sl, e0=e0, e1=e1)


def _fresh(graph):
    _, rows, cols, vals, y = graph
    v = vals.copy()
    oracle.normalize_p(v, 12.0)
    y0 = (y * np.float32(1e-3)).astype(np.float32)
    return [y0.copy(), np.zeros_like(y0), np.ones_like(y0)], v


def _composed(graph, k, total):
    """k + 1 plain epochs, then stop_lying and the final momentum by hand, then the rest."""
    _, rows, cols, _, _ = graph
    st, v = _fresh(graph)
    _run(st, rows, cols, v, 0, k + 1, 0.5)
    oracle.stop_lying(v, 12.0)
    _run(st, rows, cols, v, k + 1, total, 0.8)
    return st, v


def test_schedule_fires_after_epoch_k(graph):
    _, rows, cols, _, _ = graph
    k, total = 5, 12
    st, v = _fresh(graph)
    m = _run(st, rows, cols, v, 0, total, 0.5, sl=k, ms=k)
    assert m == np.float32(0.8)
    ref_st, ref_v = _composed(graph, k, total)
    for a, b in zip(st, ref_st):
        assert np.array_equal(a, b)
    assert np.array_equal(v, ref_v)
    # one epoch early or late is a different trajectory
    for kk in (k - 1, k + 1):
        other, _ = _composed(graph, kk, total)
        assert not np.array_equal(other[0], st[0])
        d = np.abs(other[0] - st[0]).max() / np.abs(st[0]).max()
        assert d > 1e-3


def test_stop_lying_epoch_zero_never_rescales_in_the_loop(graph):
    # lib.rs:667-669 applies it up front; epoch_tail skips epoch 0 (lib.rs:1168)
    _, rows, cols, _, _ = graph
    a, va = _fresh(graph)
    _run(a, rows, cols, va, 0, 4, 0.5, sl=0, ms=10**9)
    b, vb = _fresh(graph)
    _run(b, rows, cols, vb, 0, 4, 0.5)
    assert np.array_equal(va, vb)
    assert np.array_equal(a[0], b[0])
