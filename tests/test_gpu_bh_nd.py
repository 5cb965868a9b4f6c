"""GPU parity of Barnes-Hut with D = 3..7 output dimensions (the generic arena, src/tsne/arena.rs;
morton/d3.rs .. d7.rs; the reference's D = 3..7 tests, src/test.rs:1914-2016) against the oracle.

The tree, its walk and the attraction follow the reference's operation order, so the repulsion and
the attraction sums equal the oracle's up to the last bits of P's normalisation (FP64 here) and of
Z (an FP64 sum here, sequential f32 there)."""
import numpy as np
import pytest

import oracle
from bhtsne_amd import datagen
from tests._util import lcg_samples, same_cluster_fraction, two_clusters

pytestmark = pytest.mark.gpu

DIMS = [3, 4, 5, 6, 7]


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
    x, _ = datagen.gmm(1500, 20, 8, 25.0, 3)
    rows, cols, vals = oracle.build_affinities(x, 30.0)
    return x, rows, cols, vals


@pytest.mark.parametrize("D", DIMS)
@pytest.mark.parametrize("theta,scale", [(0.5, 1e-2), (0.5, 10.0), (0.2, 5.0)])
def test_nd_forces_from_identical_state(bt, problem, D, theta, scale):
    x, rows, cols, vals = problem
    n = x.shape[0]
    print(f"D={D} theta={theta} scale={scale}", flush=True)
    y = (np.random.default_rng(10 + D).standard_normal((n, D)) * scale).astype(np.float32)
    t = bt.TSNE(x, d_out=D).with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(y).prepare("bh", theta)
    pos, neg, inv = t.forces()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    rpos, rneg, rinv = oracle.forces(y, rows, cols, pv, method="bh", theta=theta)
    e_att, e_rep, e_z = _rel(pos, rpos), _rel(neg, rneg), abs(inv - rinv) / abs(rinv)
    e_grad = _rel(pos - neg * inv, rpos - rneg * rinv)
    print(f"  att={e_att:.2e} rep={e_rep:.2e} invZ={e_z:.2e} grad={e_grad:.2e}", flush=True)
    assert e_att < 1e-6 and e_rep < 1e-6 and e_z < 1e-5 and e_grad < 1e-5


@pytest.mark.parametrize("D", [3, 5])
def test_nd_duplicates_collapse(bt, D):
    # coincident points collapse to one leaf that skips itself (arena.rs:541-551, 369-373)
    x = np.zeros((60, 4), np.float32)
    x[30:] = 1.0
    rows, cols, vals = oracle.build_affinities(x + lcg_samples(60, 4, 2) * 1e-3, 3.0)
    y = np.zeros((60, D), np.float32)
    y[30:] = 2.0
    t = bt.TSNE(x, d_out=D).with_affinities(bt.SparseAffinities(rows, cols, vals, 3.0))
    t.initial_embedding(y).prepare("bh", 0.5)
    pos, neg, inv = t.forces()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    rpos, rneg, rinv = oracle.forces(y, rows, cols, pv, method="bh", theta=0.5)
    assert np.allclose(neg, rneg, rtol=1e-6, atol=1e-7)
    assert abs(inv - rinv) / rinv < 1e-6


@pytest.mark.parametrize("D", [3, 5, 7])
def test_nd_tracks_oracle_for_30_epochs(bt, problem, D):
    x, rows, cols, vals = problem
    n = x.shape[0]
    y0 = datagen.seed_embedding(n, D, 20 + D)
    t = bt.TSNE(x, d_out=D).with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(y0).prepare("bh", 0.5)
    t.step(30)
    gy, _, _ = t.get_state()
    pv = vals.copy()
    oracle.normalize_p(pv, 12.0)
    y, uy, g = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.run_epochs(y, uy, g, rows, cols, pv, method="bh", theta=0.5,
                      learning_rate=oracle.auto_learning_rate(n, 12), e0=0, e1=30)
    e = _rel(gy, y)
    print(f"D={D} 30 epochs rel {e:.2e}", flush=True)
    assert e < 1e-2


@pytest.mark.parametrize("D", [3, 6])
def test_nd_kl_divergence_matches_oracle(bt, problem, D):
    x, rows, cols, vals = problem
    n = x.shape[0]
    t = bt.TSNE(x, d_out=D).perplexity(30.0).epochs(60)
    t.with_affinities(bt.SparseAffinities(rows, cols, vals, 30.0))
    t.initial_embedding(datagen.seed_embedding(n, D, 30 + D))
    t.barnes_hut(0.5)
    r, tr = t.kl_divergence_pair()
    emb = t.embedding().reshape(-1, D)
    aff = t.affinities()
    pv = aff.values * np.float32(12.0)  # stop-lying has not fired: stored P is exaggerated
    rr, rt = oracle.kl_divergence(emb, aff.rows, aff.columns, pv, method="bh")
    print(f"D={D} kl gpu {r} {tr} oracle {rr} {rt}", flush=True)
    assert abs(r - rr) / abs(rr) < 1e-4
    assert abs(tr - rt) / abs(rt) < 1e-4


@pytest.mark.parametrize("D,din,seed", [(3, 5, 43), (4, 5, 42), (5, 6, 44), (6, 7, 45),
                                         (7, 8, 46)])
def test_reference_runs_in_d_dimensions(bt, D, din, seed):
    # src/test.rs:1914-2004: N = 200, perplexity 10, 50 epochs of Barnes-Hut theta 0.5; the
    # embedding holds N * D finite values
    x = lcg_samples(200, din, seed)
    t = bt.TSNE(x, d_out=D).perplexity(10.0).epochs(50)
    t.barnes_hut(0.5)
    emb = t.embedding()
    assert emb.size == 200 * D and np.all(np.isfinite(emb))


def test_nd_separates_clusters(bt):
    # src/test.rs:1055-1120 in three output dimensions
    x = two_clusters(150, 10)
    t = bt.TSNE(x, d_out=3).perplexity(30.0).epochs(500)
    t.initial_embedding(datagen.seed_embedding(300, 3, 1))
    t.barnes_hut(0.5)
    emb = t.embedding().reshape(-1, 3)
    assert same_cluster_fraction(emb, 150) > 0.95
    kl = t.kl_divergence()
    assert np.isfinite(kl) and kl >= 0


def test_nd_large_job_forces(bt):
    # 40,000 points in 3-D (a size at which 2-D jobs are renumbered and forked; D = 3 runs in
    # order on one stream over the CSR in input order), affinities built on the device
    x, _ = datagen.gmm(40000, 16, 10, 25.0, 12)
    y = (np.random.default_rng(5).standard_normal((40000, 3)) * 8).astype(np.float32)
    t = bt.TSNE(x, d_out=3).perplexity(30.0)
    t.initial_embedding(y).prepare("bh", 0.5)
    pos, neg, inv = t.forces()
    aff = t.affinities()
    pv = aff.values * np.float32(12.0)
    rpos, rneg, rinv = oracle.forces(y, aff.rows, aff.columns, pv, method="bh", theta=0.5)
    e_att, e_rep, e_z = _rel(pos, rpos), _rel(neg, rneg), abs(inv - rinv) / abs(rinv)
    print(f"40k 3-D att={e_att:.2e} rep={e_rep:.2e} invZ={e_z:.2e}", flush=True)
    assert e_att < 1e-6 and e_rep < 1e-6 and e_z < 1e-5


def test_nd_batch_members_equal_solo_fits(bt):
    # a batch holding 2-, 3- and 5-dimensional Barnes-Hut jobs: one launch group per dimension,
    # each member equal to its fit alone bit for bit
    specs = [(900, 16, 3, 40, 1), (1300, 20, 5, 35, 2), (700, 16, 2, 45, 3),
             (1100, 12, 3, 30, 4), (600, 8, 7, 40, 5)]

    def job(s):
        n, d_in, d_out, epochs, seed = s
        x, _ = datagen.gmm(n, d_in, 6, 25.0, seed)
        t = bt.TSNE(x, d_out=d_out).perplexity(30.0).epochs(epochs)
        t.stop_lying_epoch(20).momentum_switch_epoch(20)
        return t.initial_embedding(datagen.seed_embedding(n, d_out, seed))

    alone = []
    for s in specs:
        t = job(s)
        t.barnes_hut(0.5)
        alone.append((t.embedding(), t.kl_divergence_pair()))
    for graph in (True, False):
        b = bt.Batch(graph=graph)
        jobs = []
        for s in specs:
            t = job(s)
            b.add(t, "bh", 0.5)
            jobs.append(t)
        b.run()
        for s, t, (emb, kl) in zip(specs, jobs, alone):
            e = t.embedding()
            print(graph, s, "max |diff|", float(np.max(np.abs(e - emb))), flush=True)
            assert np.array_equal(e, emb), s
            assert t.kl_divergence_pair() == kl
        b.close()


def test_rejects_unsupported_dimensions(bt):
    x = lcg_samples(80, 4, 11)
    with pytest.raises(bt.InvalidInput, match="Barnes-Hut path supports d_out in 2..7"):
        bt.TSNE(x, d_out=8).perplexity(10.0).barnes_hut(0.5)
    with pytest.raises(bt.InvalidInput, match="FIt-SNE path supports d_out in"):
        bt.TSNE(x, d_out=3).perplexity(10.0).fit_sne()
