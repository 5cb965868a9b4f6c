"""The builder contract around the affinity cache and the limits of the engine, on the device:
malformed injected graphs are rejected before any upload, a renumbered job keeps its embedding in
sample order across with_affinities, the export follows affinities()'s stop_lying_fired rule
(src/lib.rs:1019-1037) including stop_lying_epoch == 0, and the neighbour-count ceiling is an
error with a message rather than a fault."""
import numpy as np
import pytest

import oracle
from bhtsne_amd import datagen
from tests._util import lcg_samples

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


def test_with_affinities_rejects_malformed_graph(bt):
    x = lcg_samples(50, 4, 3)
    rows, cols, vals = oracle.build_affinities(x, 5.0)
    t = bt.TSNE(x).perplexity(5.0).epochs(3)
    bad = cols.copy()
    bad[7] = 50
    with pytest.raises(bt.InvalidInput, match="out of range"):
        t.with_affinities(bt.SparseAffinities(rows, bad, vals, 5.0))
    r2 = rows.copy()
    r2[3], r2[4] = r2[4], r2[3]
    with pytest.raises(bt.InvalidInput, match="non-decreasing"):
        t.with_affinities(bt.SparseAffinities(r2, cols, vals, 5.0))
    with pytest.raises(bt.InvalidInput):
        t.with_affinities(bt.SparseAffinities(rows, cols, vals[:-1], 5.0))
    # the job is untouched and still fits from its data
    t.fit_sne()
    assert np.all(np.isfinite(t.embedding()))


def test_renumbered_job_keeps_sample_order_across_with_affinities(bt):
    # n >= 32768: the fit renumbers the points; with_affinities must leave y as it was
    x, _ = datagen.gmm(33000, 16, 8, 16.0, 31)
    n = x.shape[0]
    t = bt.TSNE(x).perplexity(30.0).epochs(30)
    t.initial_embedding(datagen.seed_embedding(n, 2, 31)).fit_sne()
    before = t.embedding()
    aff = t.affinities()
    t.with_affinities(aff)
    after = t.embedding()
    assert np.array_equal(before, after)
    y, uy, g = t.get_state()
    assert np.array_equal(y.ravel(), before)


def test_export_after_stop_lying_epoch_zero_divides_once_more(bt):
    # stop_lying_epoch == 0: finalize_p_and_seed rescales P up front without setting
    # stop_lying_fired, so affinities() divides by the exaggeration again (lib.rs:667-669,
    # 1024-1029)
    x, _ = datagen.gmm(600, 10, 4, 16.0, 5)
    rows, cols, vals = oracle.build_affinities(x, 10.0)
    t = bt.TSNE(x).perplexity(10.0).epochs(5).stop_lying_epoch(0)
    t.with_affinities(bt.SparseAffinities(rows, cols, vals, 10.0))
    t.initial_embedding(datagen.seed_embedding(600, 2, 5)).fit_sne()
    ref = vals.copy()
    oracle.normalize_p(ref, 12.0)
    oracle.stop_lying(ref, 12.0)
    ref = (ref * np.float32(1.0 / 12.0)).astype(np.float32)
    got = t.affinities().values
    assert np.max(np.abs(got - ref) / ref.max()) < 1e-6
    # and with the default schedule (fired at epoch 250 of 300) the export is P itself
    t2 = bt.TSNE(x).perplexity(10.0).epochs(300)
    t2.with_affinities(bt.SparseAffinities(rows, cols, vals, 10.0))
    t2.initial_embedding(datagen.seed_embedding(600, 2, 6)).fit_sne()
    ref2 = vals.copy()
    oracle.normalize_p(ref2, 12.0)
    oracle.stop_lying(ref2, 12.0)
    assert np.max(np.abs(t2.affinities().values - ref2) / ref2.max()) < 1e-6


def test_perplexity_above_the_neighbour_ceiling_is_rejected(bt):
    # k = (int)(3 * 65) = 195 > 192 neighbours per row from sample vectors
    x = lcg_samples(300, 4, 9)
    t = bt.TSNE(x).perplexity(65.0).epochs(1)
    with pytest.raises(bt.InvalidInput, match="at most 192 neighbours per row"):
        t.fit_sne()
    # 64 (k = 192) is accepted
    t = bt.TSNE(x).perplexity(64.0).epochs(1)
    t.fit_sne()
    assert np.all(np.isfinite(t.embedding()))
