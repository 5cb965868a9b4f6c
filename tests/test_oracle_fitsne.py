"""Pins the oracle's FIt-SNE repulsion and its FFT (src/tsne/interpolation.rs, src/tsne/fft.rs)."""
import numpy as np
import pytest

import oracle


def _brute(y):
    neg, z = oracle.exact_repulsion_f64(y)
    return neg, z


# --- FFT (fft.rs:255-333) --------------------------------------------------------------------

def test_fft_round_trip_2d():
    rng = np.random.default_rng(0xABCD)
    m = 12
    g = rng.uniform(-0.5, 0.5, (m, m))
    back = oracle.fft2_c2r_f64(oracle.fft2_r2c_f64(g)) / (m * m)
    assert np.max(np.abs(back - g)) < 1e-12


@pytest.mark.parametrize("m", [8, 12, 20, 60, 14])
def test_fft_matches_numpy_rfft2(m):
    # pins the sign convention and the half-spectrum layout against numpy
    rng = np.random.default_rng(m)
    g = rng.uniform(-0.5, 0.5, (m, m))
    assert np.max(np.abs(oracle.fft2_r2c_f64(g) - np.fft.rfft2(g))) < 1e-11


def test_fft_circular_convolution_matches_direct():
    rng = np.random.default_rng(0x5151)
    m = 8
    a = rng.uniform(-0.5, 0.5, (m, m))
    b = rng.uniform(-0.5, 0.5, (m, m))
    direct = np.zeros((m, m))
    for ki in range(m):
        for kj in range(m):
            acc = 0.0
            for ji in range(m):
                for jj in range(m):
                    acc += a[(ki - ji) % m, (kj - jj) % m] * b[ji, jj]
            direct[ki, kj] = acc
    out = oracle.fft2_c2r_f64(oracle.fft2_r2c_f64(a) * oracle.fft2_r2c_f64(b)) / (m * m)
    assert np.max(np.abs(out - direct)) < 1e-10


# --- grid sizing and Lagrange basis ----------------------------------------------------------

def test_next_smooth_brute_force():
    def smooth(v):
        for p in (2, 3, 5, 7):
            while v % p == 0:
                v //= p
        return v == 1
    for n in range(1, 700):
        want = next(c for c in range(n, 10 * n + 10) if smooth(c))
        assert oracle.next_smooth(n) == want
    # grid lengths quoted for 10, 50, 200 and 1000 boxes
    assert [oracle.next_smooth(6 * b - 1) for b in (10, 50, 200, 1000)] == [60, 300, 1200, 6000]


def test_box_count_clamps():
    assert oracle.box_count(0.001) == 10
    assert oracle.box_count(37.2) == 38
    assert oracle.box_count(1e7) == 1000
    assert oracle.box_count(37.2, 50) == 50


def test_lagrange_partition_of_unity_and_cardinal():
    # interpolation.rs:585-606 (f64 there, 1e-12); f32 here
    for step in range(11):
        w = oracle.lagrange_weights(step / 10.0)
        assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    for j in range(3):
        w = oracle.lagrange_weights((j + 0.5) / 3.0)
        for k in range(3):
            assert abs(w[k] - (1.0 if k == j else 0.0)) < 1e-6


# --- repulsion against brute force (interpolation.rs:520-581) --------------------------------

def _rel_l2(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def test_repulsion_matches_brute_force_2d():
    rng = np.random.default_rng(0xBADC0DE)
    y = rng.uniform(-5.0, 5.0, (800, 2)).astype(np.float32)
    neg, z, info = oracle.fitsne_repulsion(y)
    ref_neg, ref_z = _brute(y)
    assert info["boxes"] == 10 and info["fft_len"] == 60
    assert abs(z - ref_z) / ref_z < 1e-3
    assert _rel_l2(neg, ref_neg) < 3e-2


def test_repulsion_matches_brute_force_1d():
    rng = np.random.default_rng(0xFEEDF00D)
    y = rng.uniform(-4.0, 4.0, (500, 1)).astype(np.float32)
    neg, z, _ = oracle.fitsne_repulsion(y)
    ref_neg, ref_z = _brute(y)
    assert abs(z - ref_z) / ref_z < 1e-3
    assert _rel_l2(neg, ref_neg) < 3e-2


def test_finer_grid_is_more_accurate():
    # min_boxes raises the resolution; the error against brute force must shrink
    rng = np.random.default_rng(5)
    y = (rng.standard_normal((1500, 2)) * 6).astype(np.float32)
    ref_neg, ref_z = _brute(y)
    coarse, zc, _ = oracle.fitsne_repulsion(y, 10)
    fine, zf, info = oracle.fitsne_repulsion(y, 50)
    assert info["boxes"] == 50
    assert _rel_l2(fine, ref_neg) < _rel_l2(coarse, ref_neg)
    assert abs(zf - ref_z) / ref_z < 1e-3


@pytest.mark.parametrize("D,n,lo,hi", [(2, 800, -5.0, 5.0), (1, 500, -4.0, 4.0)])
def test_f64_repulsion_matches_brute_force_and_f32(D, n, lo, hi):
    # the f64 instantiation meets the same brute-force bounds, and the f32 path stays within
    # f32 rounding of it
    rng = np.random.default_rng(0xBADC0DE + D)
    y = rng.uniform(lo, hi, (n, D)).astype(np.float32)
    ref_neg, ref_z = _brute(y)
    neg64, z64 = oracle.fitsne_repulsion_f64(y)
    assert abs(z64 - ref_z) / ref_z < 1e-3
    assert _rel_l2(neg64, ref_neg) < 3e-2
    neg32, z32, _ = oracle.fitsne_repulsion(y)
    assert _rel_l2(neg32, neg64) < 1e-5
    assert abs(z32 - z64) / z64 < 1e-5


def test_repulsion_is_translation_invariant():
    rng = np.random.default_rng(9)
    y = rng.uniform(-3, 3, (300, 2)).astype(np.float32)
    n1, z1, _ = oracle.fitsne_repulsion(y)
    n2, z2, _ = oracle.fitsne_repulsion(y + np.float32(2.0))
    assert abs(z1 - z2) / z1 < 1e-4
    assert _rel_l2(n2, n1) < 1e-3
