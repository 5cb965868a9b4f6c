"""The affinity cache on the device: the reference's own cache tests (src/test.rs:1229-1712) run
through the GPU path -- export / import round trips, the export's un-exaggeration, the
neighbour-match rule of the *_with_neighbors fits, the key (n + 1 rows and the perplexity), the
reset of stop-lying on a cached re-fit, and with_affinities adopting the perplexity.
(src/lib.rs:1019-1051, 1192-1248)."""
import numpy as np
import pytest

from tests._util import f32_seq_dist, lcg_samples

pytestmark = pytest.mark.gpu

D = 4
PERPLEXITY = 10.0
THETA = 0.5
EPOCHS = 2000  # src/test.rs:7-10


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


def _bh(bt, x, epochs, perplexity=PERPLEXITY, seed=None, aff=None, **kw):
    t = bt.TSNE(x).perplexity(perplexity).epochs(epochs)
    for k, v in kw.items():
        getattr(t, k)(v)
    if seed is not None:
        t.initial_embedding(seed)
    if aff is not None:
        t.with_affinities(aff)
    return t


def _mean_point_distance(a, b, dim):
    a = np.asarray(a, np.float64).reshape(-1, dim)
    b = np.asarray(b, np.float64).reshape(-1, dim)
    return float(np.linalg.norm(a - b, axis=1).mean())


def _neighbours(x, k):
    # brute force with the metric's f32 arithmetic, (distance, index) order (src/test.rs:866-881)
    d = f32_seq_dist(x)
    np.fill_diagonal(d, np.inf)
    idx = np.lexsort((np.broadcast_to(np.arange(len(x)), d.shape), d), axis=1)[:, :k]
    return idx.astype(np.uint64), np.take_along_axis(d, idx, 1).astype(np.float32)


def test_affinities_round_trip_barnes_hut(bt):
    x = lcg_samples(200, D, 42)
    t1 = _bh(bt, x, EPOCHS)
    t1.barnes_hut(THETA)
    e1 = t1.embedding()
    aff = t1.affinities()
    t2 = _bh(bt, x, 500, seed=e1, aff=aff)
    t2.barnes_hut(THETA)
    e2 = t2.embedding()
    tr = _bh(bt, x, 500)
    tr.barnes_hut(THETA)
    assert _mean_point_distance(e1, e2, 2) < _mean_point_distance(e1, tr.embedding(), 2)
    # as the reference measures it: the first 40 values of both as ten 4-vectors
    d1 = _mean_point_distance(x.ravel()[:40], e1[:40], D)
    d2 = _mean_point_distance(x.ravel()[:40], e2[:40], D)
    assert abs(d1 - d2) < d1 * 0.5


def test_affinities_equivalence_first_step(bt):
    x = lcg_samples(100, D, 99)
    ref = _bh(bt, x, EPOCHS)
    ref.barnes_hut(THETA)
    aff, seed = ref.affinities(), ref.embedding()
    a = _bh(bt, x, 2, seed=seed)
    a.barnes_hut(THETA)
    b = _bh(bt, x, 2, seed=seed, aff=aff)
    b.barnes_hut(THETA)
    ra, rb = a.embedding(), b.embedding()
    assert np.max(np.abs(ra - rb)) / np.max(np.abs(ra)) < 1e-3


def test_affinities_pristine_independent_of_run_length(bt):
    x = lcg_samples(100, D, 7)
    short = _bh(bt, x, 50)
    short.barnes_hut(THETA)
    long = _bh(bt, x, 1000)
    long.barnes_hut(THETA)
    a, b = short.affinities(), long.affinities()
    assert abs(float(a.values.sum()) - 1.0) < 0.05 and abs(float(b.values.sum()) - 1.0) < 0.05
    assert np.array_equal(a.rows, b.rows) and np.array_equal(a.columns, b.columns)
    assert np.max(np.abs(a.values - b.values)) < 1e-6


def test_affinities_work_with_custom_neighbors(bt):
    x = lcg_samples(100, D, 42)
    ref = _bh(bt, x, EPOCHS)
    ref.barnes_hut(THETA)
    aff, seed = ref.affinities(), ref.embedding()
    rng = np.random.default_rng(5)
    rows = []
    for i in range(100):
        cand = [j for j in range(100) if j != i][:15]
        rows.append([(j, float(rng.uniform(0.0, 100.0))) for j in cand])
    b = _bh(bt, x, 50, seed=seed, aff=aff)
    b.barnes_hut_with_neighbors(THETA, rows)  # differs from the cached neighbours: rebuilt
    c = _bh(bt, x, 50, seed=seed, aff=aff)
    c.barnes_hut(THETA)  # the cache is used
    rb, rc = b.embedding(), c.embedding()
    assert np.max(np.abs(rb - rc)) / np.max(np.abs(rb)) > 0.05


def test_cached_affinities_random_seed(bt):
    x = lcg_samples(50, D, 7)
    ref = _bh(bt, x, EPOCHS)
    ref.barnes_hut(THETA)
    t = _bh(bt, x, 50, aff=ref.affinities())
    t.barnes_hut(THETA)
    e = t.embedding()
    assert e.size == 100 and np.all(np.isfinite(e)) and np.any(e != 0.0)


def test_cached_affinities_discarded_on_dataset_mismatch(bt):
    small = _bh(bt, lcg_samples(50, D, 7), EPOCHS)
    small.barnes_hut(THETA)
    large = bt.TSNE(lcg_samples(100, D, 7)).perplexity(PERPLEXITY)
    large.with_affinities(small.affinities())  # another n: discarded, rebuilt from the data
    large.barnes_hut(THETA)
    e = large.embedding()
    assert e.size == 200 and np.all(np.isfinite(e))


def test_cached_affinities_invalidated_on_perplexity_change(bt):
    x = lcg_samples(80, D, 42)
    new_perp = 2.0
    nb = _neighbours(x, int(3 * new_perp))
    seed_run = _bh(bt, x, 20)
    seed_run.barnes_hut(THETA)
    seed = seed_run.embedding()
    ref = bt.TSNE(x).perplexity(new_perp)
    ref.barnes_hut_with_neighbors(THETA, nb)
    aff = ref.affinities()
    a = _bh(bt, x, 50, perplexity=new_perp, seed=seed, aff=aff)
    a.perplexity(PERPLEXITY)  # after caching: the key no longer matches
    a.barnes_hut_with_neighbors(THETA, nb)
    b = _bh(bt, x, 50, perplexity=new_perp, seed=seed, aff=aff)
    b.barnes_hut_with_neighbors(THETA, nb)
    assert np.any(np.abs(a.embedding() - b.embedding()) > 1e-6)


def test_cached_affinities_reset_stop_lying_flag(bt):
    x = lcg_samples(80, D, 42)
    nb = _neighbours(x, int(3 * PERPLEXITY))
    ref = bt.TSNE(x).perplexity(PERPLEXITY)
    ref.barnes_hut_with_neighbors(THETA, nb)
    aff = ref.affinities()
    seed_run = _bh(bt, x, 20)
    seed_run.barnes_hut(THETA)
    seed = seed_run.embedding()
    a = _bh(bt, x, 20, seed=seed, aff=aff, stop_lying_epoch=5)
    a.barnes_hut_with_neighbors(THETA, nb)
    b = _bh(bt, x, 20, stop_lying_epoch=5)
    b.barnes_hut_with_neighbors(THETA, nb)
    b.epochs(20).initial_embedding(seed)
    b.barnes_hut_with_neighbors(THETA, nb)  # the cached re-fit fires stop-lying again
    assert np.array_equal(a.embedding(), b.embedding())


def test_with_affinities_adopts_perplexity_from_affinities(bt):
    x = np.arange(800, dtype=np.float32).reshape(200, 4)
    t = bt.TSNE(x).perplexity(30.0).epochs(20)
    t.barnes_hut(THETA)
    aff = t.affinities()
    t2 = bt.TSNE(x).perplexity(5.0)
    t2.with_affinities(aff)
    assert t2.params()["perplexity"] == 30.0
