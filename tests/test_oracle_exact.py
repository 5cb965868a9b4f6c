"""Pins of the oracle's restatement of the exact fit (src/lib.rs:489-646) by checks independent of
its code: numpy distances with the metric's f32 order, the perplexity the bandwidth search is
built to hit, symmetry and mass of P, the gradient rebuilt from the FP64 yardstick (a different
code path, itself pinned), and the KL from its definition."""
import numpy as np
import pytest

import oracle
from tests._util import f32_seq_dist, lcg_samples


@pytest.fixture(scope="module")
def x60():
    return lcg_samples(60, 4, 7)  # the reference's exact tests (src/test.rs:527-546)


def _sq_seq(x):
    x = np.asarray(x, np.float32)
    n, d = x.shape
    s = np.zeros((n, n), np.float32)
    for t in range(d):
        diff = (x[:, None, t] - x[None, :, t]).astype(np.float32)
        s = (s + (diff * diff).astype(np.float32)).astype(np.float32)
    return s


def test_dense_distances_follow_the_metric(x60):
    e = oracle.exact_distances(x60, "euclidean")
    assert np.array_equal(e, f32_seq_dist(x60))
    assert np.all(np.diag(e) == 0)
    s = oracle.exact_distances(x60, "sqeuclidean")
    expect = _sq_seq(x60)
    np.fill_diagonal(expect, 0.0)
    assert np.array_equal(s, expect)


@pytest.mark.parametrize("metric", ["euclidean", "sqeuclidean"])
def test_dense_rows_hit_the_perplexity(x60, metric):
    # the conditional rows (self entry included, d = 0) have entropy ln(perplexity) to the
    # bisection's tolerance
    d = oracle.exact_distances(x60, metric)
    cond = oracle.search_beta(d, 10.0)
    assert np.allclose(cond.sum(1), 1.0, atol=1e-5)
    h = -(cond.astype(np.float64) * np.log(np.maximum(cond.astype(np.float64), 1e-300))).sum(1)
    assert np.max(np.abs(h - np.log(10.0))) < 2e-4


def test_dense_p_is_the_symmetrised_normalised_conditional(x60):
    d = oracle.exact_distances(x60, "sqeuclidean")
    cond = oracle.search_beta(d, 10.0).astype(np.float64)
    P = oracle.exact_affinities(x60, 10.0, "sqeuclidean", exag=12.0).astype(np.float64)
    assert np.array_equal(P, P.T)
    sym = cond + cond.T
    np.fill_diagonal(sym, np.diag(cond))  # the diagonal is not doubled
    expect = sym * 12.0 / sym.sum()
    assert np.max(np.abs(P - expect)) / expect.max() < 1e-5
    assert np.all(np.diag(P) > 0)  # the reference keeps each point's own affinity
    P0 = oracle.exact_affinities(x60, 10.0, "sqeuclidean", exag=12.0, stop_lying_epoch=0)
    assert abs(float(P0.astype(np.float64).sum()) - 1.0) < 1e-5


def test_exact_gradient_from_the_fp64_yardstick(x60):
    P = oracle.exact_affinities(x60, 10.0, "sqeuclidean").astype(np.float64)
    y = (lcg_samples(60, 2, 99) * np.float32(3.0)).astype(np.float32)
    dy, inv = oracle.exact_gradient(y, P)
    # quirk-free FP64 repulsion (self excluded) + the n diagonal q = 1 terms of the exact path's Z
    neg, z = oracle.exact_repulsion_f64(y)
    z_exact = z + 60.0
    yd = y.astype(np.float64)
    diff = yd[:, None, :] - yd[None, :, :]
    q = 1.0 / (1.0 + (diff ** 2).sum(-1))
    attr = ((P * q)[:, :, None] * diff).sum(1)
    expect = attr - neg / z_exact
    assert abs(inv * z_exact - 1.0) < 1e-6
    assert np.sqrt(((dy - expect) ** 2).sum() / (expect ** 2).sum()) < 1e-5


def test_exact_kl_is_its_definition(x60):
    P = oracle.exact_affinities(x60, 10.0, "sqeuclidean", stop_lying_epoch=0)
    y = lcg_samples(60, 2, 99)
    r, t = oracle.exact_kl(y, P)
    yd, Pd = y.astype(np.float64), P.astype(np.float64)
    q = 1.0 / (1.0 + ((yd[:, None, :] - yd[None, :, :]) ** 2).sum(-1))
    qn = q / q.sum()
    m = np.finfo(np.float32).tiny
    expect = (Pd * np.log((Pd + m) / (qn + m))).sum()
    assert abs(t - expect) / abs(expect) < 1e-5 and abs(r - expect) / abs(expect) < 1e-4
    assert t >= 0.0  # P and Q both sum to one


def test_exact_epochs_compose_the_pinned_steps(x60):
    P0 = oracle.exact_affinities(x60, 10.0, "sqeuclidean")
    y0 = lcg_samples(60, 2, 99)
    lr = oracle.auto_learning_rate(60, 12.0)
    # three epochs with stop-lying after epoch 1: P is divided once, before epoch 2
    y, uy, g, P = y0.copy(), np.zeros_like(y0), np.ones_like(y0), P0.copy()
    oracle.exact_epochs(y, uy, g, P, learning_rate=lr, stop_lying_epoch=1,
                        momentum_switch_epoch=1, e0=0, e1=3)
    ym, uym, gm, Pm = y0.copy(), np.zeros_like(y0), np.ones_like(y0), P0.copy()
    mom = 0.5
    for e in range(3):
        dy, _ = oracle.exact_gradient(ym, Pm)
        oracle.gradient_step(ym, dy, np.zeros_like(dy), uym, gm, lr, mom, 0.0)
        oracle.zero_mean(ym)
        if e == 1:
            oracle.stop_lying(Pm, 12.0)
            mom = 0.8
    assert np.array_equal(y, ym) and np.array_equal(uy, uym) and np.array_equal(P, Pm)
