"""GPU parity of the Barnes-Hut path (Morton arena + traversal), the KL divergence, and the
builder's rejection messages, against the oracle and the reference's tests."""
import numpy as np
import pytest

import oracle
from bhtsne_amd import datagen
from tests._util import lcg_samples, same_cluster_fraction, two_clusters

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
    x, _ = datagen.gmm(3000, 50, 10, 25.0, 0)
    rows, cols, vals = oracle.build_affinities(x, 30.0)
    return x, rows, cols, vals


@pytest.mark.parametrize("theta,scale", [(0.5, 1e-2), (0.5, 10.0), (0.2, 10.0), (1.0, 30.0)])
def test_bh_forces_from_identical_state(bt, problem, theta, scale):
    x, rows, cols, vals = problem
    n = x.shape[0]
    y = (np.random.default_rng(2).standard_normal((n, 2)) * scale).astype(np.float32)
    t = bt.TSNE(x).with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(y).prepare("bh", theta)
    pos, neg, inv = t.forces()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    rpos, rneg, rinv = oracle.forces(y, rows, cols, pv, method="bh", theta=theta)
    e_rep, e_z = _rel(neg, rneg), abs(inv - rinv) / abs(rinv)
    e_grad = _rel(pos - neg * inv, rpos - rneg * rinv)
    print(f"theta={theta} scale={scale} rep={e_rep:.2e} invZ={e_z:.2e} grad={e_grad:.2e}",
          flush=True)
    assert e_rep < 1e-5 and e_z < 1e-5 and e_grad < 1e-5


def test_bh_duplicates_and_tiny_n(bt):
    # coincident points collapse to one leaf that skips itself (arena.rs:541-551, 369-373)
    x = np.zeros((40, 3), np.float32)
    x[20:] = 1.0
    rows, cols, vals = oracle.build_affinities(x + lcg_samples(40, 3, 1) * 1e-3, 3.0)
    y = np.zeros((40, 2), np.float32)
    y[20:] = 2.0
    t = bt.TSNE(x).with_affinities(bt.SparseAffinities(rows, cols, vals, 3.0))
    t.initial_embedding(y).prepare("bh", 0.5)
    pos, neg, inv = t.forces()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    rpos, rneg, rinv = oracle.forces(y, rows, cols, pv, method="bh", theta=0.5)
    assert np.allclose(neg, rneg, rtol=1e-6, atol=1e-7)
    assert abs(inv - rinv) / rinv < 1e-6


def test_barnes_hut_separates_clusters(bt):
    # src/test.rs:1055-1120
    x = two_clusters(150, 10)
    t = bt.TSNE(x).perplexity(30.0).epochs(500)
    t.initial_embedding(datagen.seed_embedding(300, 2, 1))
    t.barnes_hut(0.5)
    emb = t.embedding().reshape(-1, 2)
    assert same_cluster_fraction(emb, 150) > 0.95
    kl = t.kl_divergence()
    assert np.isfinite(kl) and kl >= 0


def test_bh_tracks_oracle_for_30_epochs(bt, problem):
    x, rows, cols, vals = problem
    n = x.shape[0]
    y0 = datagen.seed_embedding(n, 2, 6)
    t = bt.TSNE(x).with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(y0).prepare("bh", 0.5)
    t.step(30)
    gy, _, _ = t.get_state()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    y, uy, g = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.run_epochs(y, uy, g, rows, cols, pv, method="bh", theta=0.5,
                      learning_rate=oracle.auto_learning_rate(n, 12), e0=0, e1=30)
    e = _rel(gy, y)
    print("bh 30 epochs rel", e, flush=True)
    assert e < 1e-2


@pytest.mark.parametrize("method", ["fitsne", "bh"])
def test_kl_divergence_matches_oracle(bt, problem, method):
    x, rows, cols, vals = problem
    n = x.shape[0]
    t = bt.TSNE(x).perplexity(30.0).epochs(60)
    t.with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(datagen.seed_embedding(n, 2, 8))
    if method == "fitsne":
        t.fit_sne()
    else:
        t.barnes_hut(0.5)
    r, tr = t.kl_divergence_pair()
    emb = t.embedding().reshape(-1, 2)
    aff = t.affinities()
    pv = aff.values * np.float32(12.0)  # stop-lying has not fired: stored P is exaggerated
    rr, rt = oracle.kl_divergence(emb, aff.rows, aff.columns, pv, method=method)
    print(method, "kl gpu", r, tr, "oracle", rr, rt, flush=True)
    assert abs(r - rr) / abs(rr) < 1e-4
    assert abs(tr - rt) / abs(rt) < 1e-4


def test_kl_is_none_before_fitting(bt):
    t = bt.TSNE(lcg_samples(50, 4, 1))
    assert t.kl_divergence() is None


# ---- rejection messages (src/test.rs:622-671, 930-970, 1869-1907) -------------------------------

def test_rejects_zero_theta(bt):
    t = bt.TSNE(lcg_samples(80, 4, 11)).perplexity(10.0)
    with pytest.raises(bt.InvalidInput, match="theta value must be greater than 0.0"):
        t.barnes_hut(0.0)


def test_rejects_large_perplexity(bt):
    t = bt.TSNE(lcg_samples(20, 4, 11)).perplexity(10.0)
    with pytest.raises(bt.InvalidInput, match="perplexity is too large"):
        t.fit_sne()


@pytest.mark.parametrize("method", ["bh", "fitsne"])
def test_rejects_ragged_and_out_of_range_neighbors(bt, method):
    x = lcg_samples(80, 4, 11)
    idx, dist = oracle.knn(x, 30)
    nb = [[(int(i), float(d)) for i, d in zip(r, dr)] for r, dr in zip(idx, dist)]
    ragged = [list(r) for r in nb]
    ragged[0].pop()
    t = bt.TSNE(x).perplexity(10.0).epochs(1)
    fit = (lambda n: t.barnes_hut_with_neighbors(0.5, n)) if method == "bh" else \
        t.fit_sne_with_neighbors
    with pytest.raises(bt.InvalidInput, match="same length"):
        fit(ragged)
    bad = [list(r) for r in nb]
    bad[0][0] = (80, bad[0][0][1])
    with pytest.raises(bt.InvalidInput, match="out of range"):
        fit(bad)


def test_rejects_wrong_initial_embedding(bt):
    t = bt.TSNE(lcg_samples(80, 4, 11)).perplexity(10.0).epochs(1)
    t.initial_embedding(np.zeros(7, np.float32))
    with pytest.raises(bt.InvalidInput, match="initial embedding has"):
        t.barnes_hut(0.5)


def test_with_neighbors_matches_data_path(bt):
    # src/test.rs:888-928: fed the exact neighbours, the *_with_neighbors entry reproduces the
    # data entry bit for bit (both run the same device kernels on the same inputs)
    x = lcg_samples(300, 4, 11)
    seed = lcg_samples(300, 2, 99)
    a = bt.TSNE(x).perplexity(10.0).epochs(100).initial_embedding(seed).barnes_hut(0.5)
    idx, dist = oracle.knn(x, 30)
    b = bt.TSNE(x).perplexity(10.0).epochs(100).initial_embedding(seed)
    b.barnes_hut_with_neighbors(0.5, (idx.astype(np.uint64), dist))
    assert np.array_equal(a.embedding(), b.embedding())
