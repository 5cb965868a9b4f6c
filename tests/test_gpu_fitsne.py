"""GPU parity of the FIt-SNE epoch: forces and 1/Z from an identical state, the fused update, the
grid plan, the point renumbering, and end-to-end quality, against the oracle.

Tolerances: the attraction from an identical (y, P) state agrees with the oracle to 1e-5
relative L2, and so does the whole gradient wherever the f32 rounding of the FFT grids allows it;
on large grids (hundreds of boxes) the gradient is required to be no further from the
interpolation scheme evaluated in FP64 than the oracle's own f32 arithmetic is (1.5x)."""
import numpy as np
import pytest

import oracle
from bhtsne_amd import datagen
from tests._util import same_cluster_fraction, two_clusters

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


@pytest.fixture(scope="module")
def problem():
    x, labels = datagen.gmm(5000, 20, 10, 16.0, 1)
    rows, cols, vals = oracle.build_affinities(x, 30.0)
    return x, rows, cols, vals


def _job(bt, x, rows, cols, vals, y0, **kw):
    t = bt.TSNE(x)
    for k, v in kw.items():
        getattr(t, k)(v)
    t.with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(y0)
    t.prepare("fitsne")
    return t


@pytest.mark.parametrize("scale,min_boxes", [(1e-2, 10), (8.0, 10), (40.0, 10), (8.0, 50)])
def test_forces_from_identical_state(bt, problem, scale, min_boxes):
    x, rows, cols, vals = problem
    n = x.shape[0]
    rng = np.random.default_rng(5)
    y = (rng.standard_normal((n, 2)) * scale).astype(np.float32)
    t = _job(bt, x, rows, cols, vals, y, min_boxes=min_boxes)
    pos, neg, inv = t.forces()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    rpos, rneg, rinv = oracle.forces(y, rows, cols, pv, min_boxes=min_boxes)
    _, _, info = oracle.fitsne_repulsion(y, min_boxes)
    g = t.grid_info()
    assert (g["boxes"], g["fft_len"]) == (info["boxes"], info["fft_len"])
    e_att = _rel(pos, rpos)
    e_rep = _rel(neg, rneg)
    e_z = abs(inv - rinv) / abs(rinv)
    grad, rgrad = pos - neg * inv, rpos - rneg * rinv
    e_grad = _rel(grad, rgrad)
    # the interpolation scheme without f32 rounding (T = f64): the yardstick for the FP32 noise of
    # both the oracle's f32 path (which stands for the reference's) and the GPU's
    neg64, z64 = oracle.fitsne_repulsion_f64(y, min_boxes)
    grad64 = rpos.astype(np.float64) - neg64 / z64
    with oracle.f32_fft():
        fpos, fneg, finv = oracle.forces(y, rows, cols, pv, min_boxes=min_boxes)
    noise_ref = _rel(rgrad, grad64)                 # f32 grids, exact transforms
    noise_ref32 = _rel(fpos - fneg * finv, grad64)  # f32 grids, f32 transforms (as rustfft)
    noise_gpu = _rel(grad, grad64)
    print(f"scale={scale} boxes={g['boxes']} L={g['fft_len']} att={e_att:.2e} rep={e_rep:.2e} "
          f"invZ={e_z:.2e} grad={e_grad:.2e} | vs f64 scheme: oracle {noise_ref:.2e} "
          f"oracle-f32-fft {noise_ref32:.2e} gpu {noise_gpu:.2e}", flush=True)
    assert e_att < 1e-5
    # gradient parity: 1e-5 relative, or, where the f32 rounding of the grids alone exceeds that
    # (large grids: y * phi0 - phi1 cancels), within twice the distance of an f32-FFT
    # restatement of the reference from the exact scheme
    assert e_grad < 1e-5 or noise_gpu <= max(1e-5, 2.0 * noise_ref32)
    assert e_z < 1e-4 or abs(1.0 / inv - z64) / z64 <= 2.0 * abs(1.0 / finv - z64) / z64


def test_one_epoch_update_matches(bt, problem):
    x, rows, cols, vals = problem
    n = x.shape[0]
    y0 = datagen.seed_embedding(n, 2, 3)
    t = _job(bt, x, rows, cols, vals, y0)
    t.step(1)
    gy, guy, gg = t.get_state()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    y, uy, gains = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.run_epochs(y, uy, gains, rows, cols, pv, learning_rate=oracle.auto_learning_rate(n, 12),
                      e0=0, e1=1)
    print("1 epoch: y", _rel(gy, y), "uy", _rel(guy, uy), flush=True)
    assert _rel(guy, uy) < 1e-4
    assert _rel(gy, y) < 1e-4
    assert np.array_equal(gg, gains) or np.mean(gg == gains) > 0.999


def test_schedule_and_stop_lying(bt, problem):
    # 10 epochs with stop-lying and the momentum switch at epoch 4: P must be rescaled exactly once
    x, rows, cols, vals = problem
    n = x.shape[0]
    y0 = datagen.seed_embedding(n, 2, 4)
    t = _job(bt, x, rows, cols, vals, y0, stop_lying_epoch=4, momentum_switch_epoch=4)
    t.step(10).finish()
    aff = t.affinities()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    oracle.stop_lying(pv, 12.0)
    assert np.max(np.abs(aff.values - pv) / pv.max()) < 1e-6
    assert t.params()["momentum"] == pytest.approx(0.8)


def test_fit_sne_separates_clusters(bt):
    # src/test.rs:1733-1790
    x = two_clusters(150, 10)
    k = 90
    idx, dist = oracle.knn(x, k)
    t = bt.TSNE(x).perplexity(30.0).epochs(500)
    t.initial_embedding(datagen.seed_embedding(300, 2, 9))
    t.fit_sne_with_neighbors((idx.astype(np.uint64), dist))
    emb = t.embedding().reshape(-1, 2)
    assert np.all(np.isfinite(emb))
    assert same_cluster_fraction(emb, 150) > 0.95
    # no collapse (src/test.rs:1794-1825)
    assert len({(round(a * 100), round(b * 100)) for a, b in emb}) > 150


def test_run_tracks_oracle_for_50_epochs(bt, problem):
    x, rows, cols, vals = problem
    n = x.shape[0]
    y0 = datagen.seed_embedding(n, 2, 5)
    t = _job(bt, x, rows, cols, vals, y0)
    t.step(50)
    gy, _, _ = t.get_state()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    y, uy, gains = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.run_epochs(y, uy, gains, rows, cols, pv, learning_rate=oracle.auto_learning_rate(n, 12),
                      e0=0, e1=50)
    e = _rel(gy, y)
    print("50 epochs rel", e, flush=True)
    # t-SNE trajectories amplify last-bit differences; after 50 epochs they stay close
    assert e < 3e-2


def test_callback_reports_each_epoch(bt, problem):
    x, rows, cols, vals = problem
    seen = []
    t = bt.TSNE(x).perplexity(30.0).epochs(7)
    t.with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.epoch_callback(lambda e, y: seen.append((e, float(np.abs(y.reshape(-1, 2).mean(0)).max()))))
    t.fit_sne()
    assert [e for e, _ in seen] == list(range(7))
    assert all(m < 1e-5 for _, m in seen)


@pytest.fixture(scope="module")
def big_problem(bt):
    # above the renumbering threshold (32768 points)
    x, _ = datagen.gmm(40000, 20, 10, 16.0, 12)
    t = bt.TSNE(x).perplexity(30.0).epochs(0)
    t.fit_sne()
    return x, t.affinities()


def test_renumbering_keeps_results(bt, big_problem):
    x, aff = big_problem
    n = x.shape[0]
    y0 = datagen.seed_embedding(n, 2, 13)
    runs = {}
    for flags in (0, 1):
        t = bt.TSNE(x).perplexity(30.0)
        t._set(flags=flags)
        t.with_affinities(aff).initial_embedding(y0).prepare("fitsne")
        t.step(20)
        y, uy, g = t.get_state()
        rng = np.random.default_rng(1)
        ys = (rng.standard_normal((n, 2)) * 6).astype(np.float32)
        t.set_state(ys, uy, g)
        pos, neg, inv = t.forces()
        t.finish()
        runs[flags] = (y, pos, neg, inv, t.affinities(), t.kl_divergence_pair())
    (ya, pa, na, ia, aa, ka), (yb, pb, nb, ib, ab, kb) = runs[0], runs[1]
    print("renumbered vs input order: y", _rel(ya, yb), "pos", _rel(pa, pb), "neg",
          _rel(na, nb), flush=True)
    # forces of the same state: the same sums per row, summation order of the spread aside
    assert _rel(pa, pb) < 1e-6
    assert _rel(na, nb) < 1e-5 and abs(ia - ib) / ib < 1e-6
    # 20 epochs of trajectory: equal up to summation order
    assert _rel(ya, yb) < 1e-3
    # the affinity export is in the reference's layout either way
    assert np.array_equal(aa.rows, ab.rows) and np.array_equal(aa.columns, ab.columns)
    assert np.array_equal(aa.values, ab.values)
    assert abs(ka[0] - kb[0]) / abs(kb[0]) < 1e-5


def test_one_dimensional_embedding(bt):
    rng = np.random.default_rng(0xFEEDF00D)
    x, _ = datagen.gmm(800, 10, 5, 16.0, 4)
    rows, cols, vals = oracle.build_affinities(x, 20.0)
    y = rng.uniform(-4, 4, (800, 1)).astype(np.float32)
    t = bt.TSNE(x, d_out=1)
    t.with_affinities(bt.SparseAffinities(rows, cols, vals, 20.0)).initial_embedding(y)
    t.prepare("fitsne")
    pos, neg, inv = t.forces()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    rpos, rneg, rinv = oracle.forces(y, rows, cols, pv)
    assert _rel(pos, rpos) < 1e-5
    assert _rel(neg, rneg) < 1e-4
    assert abs(inv - rinv) / abs(rinv) < 1e-5


def _oracle_run(x_vals, rows, cols, y0, epochs, sl, ms):
    pv = x_vals.copy()
    oracle.normalize_p(pv, 12.0)
    y, uy, g = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.run_epochs(y, uy, g, rows, cols, pv, learning_rate=oracle.auto_learning_rate(
        y0.shape[0], 12), stop_lying_epoch=sl, momentum_switch_epoch=ms, e0=0, e1=epochs)
    return y


def test_schedule_epoch_tracks_oracle(bt, problem):
    # stop-lying and the momentum switch at epoch 5 (they fire after epoch 5's update,
    # src/lib.rs:1162-1174): the GPU's trajectory over 12 epochs follows the oracle's with the
    # same schedule, and is far from the oracle's with the schedule one epoch early or late
    x, rows, cols, vals = problem
    n = x.shape[0]
    y0 = (np.random.default_rng(7).standard_normal((n, 2)) * 1e-2).astype(np.float32)
    t = _job(bt, x, rows, cols, vals, y0, stop_lying_epoch=5, momentum_switch_epoch=5)
    t.step(12)
    gy, _, _ = t.get_state()
    e_same = _rel(gy, _oracle_run(vals, rows, cols, y0, 12, 5, 5))
    e_early = _rel(gy, _oracle_run(vals, rows, cols, y0, 12, 4, 4))
    e_late = _rel(gy, _oracle_run(vals, rows, cols, y0, 12, 6, 6))
    print(f"12 epochs, schedule at 5: same {e_same:.2e}, one early {e_early:.2e}, one late "
          f"{e_late:.2e}", flush=True)
    assert e_same < 1e-4
    assert e_early > 100 * e_same and e_late > 100 * e_same
