"""Pins the oracle's Barnes-Hut arena, the attraction/update step and the run loop."""
import numpy as np
import pytest

import oracle
from tests._util import lcg_cloud, lcg_samples, same_cluster_fraction, two_clusters


def _decode(code):
    def compact(x):
        x &= 0x5555555555555555
        x = (x | (x >> 1)) & 0x3333333333333333
        x = (x | (x >> 2)) & 0x0F0F0F0F0F0F0F0F
        x = (x | (x >> 4)) & 0x00FF00FF00FF00FF
        x = (x | (x >> 8)) & 0x0000FFFF0000FFFF
        return (x | (x >> 16)) & 0xFFFFFFFF
    return compact(code), compact(code >> 1)


# --- Morton codec (morton/d2.rs:56-122) ------------------------------------------------------

def test_morton_known_z_order():
    assert [oracle.morton_encode2(*c) for c in [(0, 0), (1, 0), (0, 1), (1, 1)]] == [0, 1, 2, 3]


def test_morton_round_trip():
    rng = np.random.default_rng(0x12345678)
    for _ in range(2000):
        a, b = (int(v) for v in rng.integers(0, 2**32, 2))
        assert _decode(oracle.morton_encode2(a, b)) == (a, b)


def test_quantize_corners_and_zero_width():
    mx = 2**32 - 1
    inv0 = np.float32(2**32) / np.float32(4.0)
    inv1 = np.float32(2**32) / np.float32(8.0)
    assert oracle.quantize(0.0, 0.0, inv0) == 0 and oracle.quantize(-2.0, -2.0, inv1) == 0
    assert oracle.quantize(4.0, 0.0, inv0) == mx and oracle.quantize(6.0, -2.0, inv1) == mx
    assert 0 < oracle.quantize(2.0, 0.0, inv0) < mx
    assert oracle.quantize(1.0, 1.0, 0.0) == 0


# --- arena invariants (arena.rs:507-574, test.rs:1164-1175) ----------------------------------

def test_arena_conserves_points_and_coms_in_cells():
    y = lcg_cloud(2000, 2, 17) + np.float32(100.0)
    a = oracle.bh_build(y)
    leaves = a["first_child"] == 0xFFFFFFFF
    assert int(a["count"][leaves].sum()) == 2000
    assert int(a["count"][0]) == 2000
    # every internal node has 2..4 children whose counts add up
    for i in np.nonzero(~leaves)[0]:
        fc, cc = int(a["first_child"][i]), int(a["child_count"][i])
        assert 2 <= cc <= 4
        assert int(a["count"][fc:fc + cc].sum()) == int(a["count"][i])
        assert fc > i  # breadth-first: children after parents
    # every centre of mass lies inside the bounding box
    assert np.all(a["com"] >= y.min(0) - 1e-3) and np.all(a["com"] <= y.max(0) + 1e-3)


def test_arena_root_com_is_mean():
    y = lcg_cloud(1000, 2, 5)
    a = oracle.bh_build(y)
    assert np.all(np.abs(a["com"][0] - y.mean(0)) < 1e-3)


def test_arena_duplicates_single_and_empty():
    a = oracle.bh_build(np.full((500, 2), 3.5, np.float32))
    assert len(a["count"]) == 1 and a["count"][0] == 500 and a["first_child"][0] == 0xFFFFFFFF
    a = oracle.bh_build(np.array([[2.0, -1.0]], np.float32))
    assert len(a["count"]) == 1 and a["com"][0].tolist() == [2.0, -1.0]
    neg, q, s = oracle.bh_repulsion(np.zeros((0, 2), np.float32), 0.5)
    assert s == 0.0


def test_bh_converges_to_exact_as_theta_shrinks():
    rng = np.random.default_rng(4)
    y = (rng.standard_normal((700, 2)) * 3).astype(np.float32)
    ref, z = oracle.exact_repulsion_f64(y)
    errs = []
    for theta in (1.0, 0.5, 0.1, 0.01):
        neg, q, s = oracle.bh_repulsion(y, theta)
        errs.append(float(np.sqrt(((neg - ref) ** 2).sum() / (ref ** 2).sum())))
        if theta == 0.01:
            assert abs(s - z) / z < 1e-5
    assert errs[-1] < 1e-5 and all(a > b for a, b in zip(errs, errs[1:])) and errs[1] < 0.1


# --- attraction, update, recentring ----------------------------------------------------------

def test_edge_forces_match_fp64():
    x = lcg_samples(120, 6, 3)
    rows, cols, vals = oracle.build_affinities(x, 10.0)
    y = (lcg_samples(120, 2, 99) * 4).astype(np.float32)
    pos = oracle.edge_forces(y, rows, cols, vals)
    ref = oracle.exact_attraction_f64(y, rows, cols, vals)
    assert np.max(np.abs(pos - ref)) / np.abs(ref).max() < 1e-5


def test_gradient_step_signum_of_positive_zero():
    # Rust signum(+0.0) = +1 (mod.rs:503): at epoch 0 (uy = +0.0) a positive gradient decays the
    # gain (0.8) and a negative one increments it (1.2)
    y = np.zeros(4, np.float32)
    uy = np.zeros(4, np.float32)
    gains = np.ones(4, np.float32)
    pos = np.array([1.0, -1.0, 0.0, 2.0], np.float32)
    neg = np.zeros(4, np.float32)
    oracle.gradient_step(y, pos, neg, uy, gains, 10.0, 0.5, 1.0)
    assert gains.tolist() == [np.float32(0.8), np.float32(1.2), np.float32(0.8), np.float32(0.8)]
    assert uy[0] == np.float32(-10.0 * np.float32(0.8) * 1.0)
    assert np.array_equal(y, uy)


def test_gradient_step_min_gain_and_momentum():
    y = np.zeros(1, np.float32)
    uy = np.array([1.0], np.float32)
    gains = np.array([0.011], np.float32)
    oracle.gradient_step(y, np.array([1.0], np.float32), np.zeros(1, np.float32), uy, gains,
                         2.0, 0.5, 1.0)
    # same sign -> decay -> clamped to 0.01; uy = 0.5*1 - 2*0.01*1
    assert gains[0] == np.float32(0.01)
    assert uy[0] == np.float32(np.float32(0.5) - np.float32(2.0) * np.float32(0.01))


def test_zero_mean():
    y = (lcg_samples(333, 2, 8) + np.float32(5)).astype(np.float32)
    oracle.zero_mean(y)
    assert np.all(np.abs(y.astype(np.float64).mean(0)) < 2e-5)


# --- run loop (lib.rs:961-1014) --------------------------------------------------------------

def _prepared(x, perplexity, exag=12.0):
    rows, cols, vals = oracle.build_affinities(x, perplexity)
    oracle.normalize_p(vals, exag)
    return rows, cols, vals


@pytest.mark.parametrize("method", ["fitsne", "bh"])
def test_separates_two_clusters(method):
    # src/test.rs:1055-1120 and 1733-1790 (500 epochs, perplexity 30, >95% same-cluster NN)
    x = two_clusters(150, 10)
    n = x.shape[0]
    rows, cols, vals = _prepared(x, 30.0)
    y = (lcg_samples(n, 2, 99) * np.float32(1e-2)).astype(np.float32)
    uy = np.zeros_like(y)
    gains = np.ones_like(y)
    oracle.run_epochs(y, uy, gains, rows, cols, vals, method=method,
                      learning_rate=oracle.auto_learning_rate(n, 12.0), e0=0, e1=500)
    assert same_cluster_fraction(y, 150) > 0.95
    # stop-lying fired: P sums to ~1 again
    assert abs(float(vals.astype(np.float64).sum()) - 1.0) < 1e-3
    r, t = oracle.kl_divergence(y, rows, cols, vals, method=method)
    assert np.isfinite(r) and r >= 0 and np.isfinite(t) and t >= 0


def test_kl_quirk_vanishes_for_uniform_p():
    # with equal P values p_values[index] equals the edge value, so both formulas agree
    x = lcg_samples(60, 3, 4)
    rows, cols, vals = oracle.build_affinities(x, 5.0)
    vals[:] = np.float32(1.0 / len(vals))
    y = (lcg_samples(60, 2, 5) * 3).astype(np.float32)
    r, t = oracle.kl_divergence(y, rows, cols, vals)
    assert abs(r - t) / abs(t) < 1e-4
