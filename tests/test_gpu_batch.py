"""Batched and served jobs: every member of a batch computes what its fit alone computes.

The batch runs the single-job kernels over a device job table (one launch per epoch kernel for
all members of a method / d_out, a segmented sort, graph replay), with the same block partition
of every reduction, so the bar is bit-for-bit equality with the job run alone (fit_sne /
barnes_hut on the same data, parameters and initial embedding)."""
import numpy as np
import pytest

from bhtsne_amd import datagen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


# (n, d_in, d_out, method, epochs, seed): uneven sizes, both methods, both FIt-SNE dimensions
SPECS = [
    (700, 16, 2, "fitsne", 60, 1),
    (2500, 32, 2, "fitsne", 45, 2),
    (1200, 20, 1, "fitsne", 50, 3),
    (1500, 64, 2, "bh", 40, 4),
    (400, 16, 2, "fitsne", 70, 5),
    (900, 32, 2, "bh", 55, 6),
    (3000, 20, 1, "fitsne", 35, 7),
    # above 32768 points: renumbered before the first epoch, alone and in the batch alike
    (36000, 16, 2, "fitsne", 30, 8),
    (33000, 16, 2, "bh", 25, 9),
]


def _data(n, d_in, seed):
    x, _ = datagen.gmm(n, d_in, 6, 25.0, seed)
    return x


def _job(bt, spec):
    n, d_in, d_out, method, epochs, seed = spec
    t = bt.TSNE(_data(n, d_in, seed), d_out=d_out).perplexity(30.0).epochs(epochs)
    t.stop_lying_epoch(20).momentum_switch_epoch(20)
    t.initial_embedding(datagen.seed_embedding(n, d_out, seed))
    return t


# Solo fits of 32768 points and more run each epoch's repulsion on a second stream beside the
# attraction, and such runs have differed from one another in the last bits now and then; the
# solo references here run the same kernels in order on one stream (BHT_FLAG_SERIAL_EPOCH), which
# is how a batch runs them.
SERIAL_EPOCH = 2


def _alone(bt, spec):
    t = _job(bt, spec)
    t._set(flags=SERIAL_EPOCH)
    if spec[3] == "fitsne":
        t.fit_sne()
    else:
        t.barnes_hut(0.5)
    return t.embedding(), t.kl_divergence_pair()


@pytest.fixture(scope="module")
def alone(bt):
    return [_alone(bt, s) for s in SPECS]


@pytest.mark.parametrize("graph,concurrent", [(True, False), (False, False), (True, True)])
def test_batch_equals_alone(bt, alone, graph, concurrent):
    # with the repulsions beside the attractions on a second stream or every kernel in order,
    # replayed from a graph or launched plainly: each member is its fit alone, bit for bit
    b = bt.Batch(graph=graph, concurrent=concurrent)
    jobs = []
    for s in SPECS:
        t = _job(bt, s)
        b.add(t, s[3], 0.5)
        jobs.append(t)
    assert len(b) == len(SPECS)
    done = b.run(chunk=1 << 20)
    assert len(done) == len(SPECS) and len(b) == 0
    for s, t, (emb, kl) in zip(SPECS, jobs, alone):
        e = t.embedding()
        print(s, "max |diff|", float(np.max(np.abs(e - emb))), flush=True)
        assert np.array_equal(e, emb), s
        assert t.kl_divergence_pair() == kl
    b.close()


def test_continuous_admission_equals_alone(bt, alone):
    # jobs join while others run (the serving loop's admission) and leave as they finish
    b = bt.Batch()
    order = [0, 3, 1, 7, 2, 4, 8, 5, 6]
    jobs = {}
    pending = list(order)
    finished = []
    while pending or len(b):
        if pending:
            i = pending.pop(0)
            jobs[i] = _job(bt, SPECS[i])
            b.add(jobs[i], SPECS[i][3], 0.5)
        b.step(7)
        finished += b.take_finished()
    assert len(finished) == len(SPECS)
    for i, t in jobs.items():
        assert np.array_equal(t.embedding(), alone[i][0]), SPECS[i]


def test_graph_replays_match_eager_each_step(bt):
    # graph replay checked against plain launches after every step, with host work between
    specs = SPECS[:4]
    runs = {}
    for graph in (True, False):
        b = bt.Batch(graph=graph)
        js = [_job(bt, s) for s in specs]
        for s, t in zip(specs, js):
            b.add(t, s[3], 0.5)
        snaps = []
        for _ in range(4):
            b.step(5)
            b.synchronize()
            snaps.append([j.get_state()[0].copy() for j in js])
        runs[graph] = snaps
        b.close()
    for a, c in zip(runs[True], runs[False]):
        for ya, yc in zip(a, c):
            assert np.array_equal(ya, yc)


def test_batch_rejections(bt):
    b = bt.Batch()
    n = 60
    t = bt.TSNE(_data(n, 8, 9)).perplexity(30.0)
    with pytest.raises(bt.InvalidInput, match="perplexity is too large"):
        b.add(t, "fitsne")
    t2 = _job(bt, SPECS[0])
    with pytest.raises(bt.InvalidInput, match="theta value must be greater than 0.0"):
        b.add(t2, "bh", 0.0)
    b.add(t2, "fitsne")
    with pytest.raises(bt.BhtError, match="already belongs to a batch"):
        b.add(t2, "fitsne")
    b.run()
    assert len(b) == 0


def test_serving_stream_equals_plain(bt):
    # continuous batching of a request stream: each result equals the same request served alone
    from bhtsne_amd import serving
    reqs = []
    for r, (n, d_in, d_out) in enumerate([(900, 20, 2), (2600, 50, 1), (600, 20, 2),
                                          (1800, 50, 2), (1200, 20, 1)]):
        x = _data(n, d_in, 100 + r)
        reqs.append(serving.Request(r, x, d_out=d_out, epochs=40 + 5 * r, arrival=0.05 * r,
                                    y0=datagen.seed_embedding(n, d_out, r)))
    ctx = bt.Context(0)
    got = serving.serve(reqs, ctx, step_epochs=8)
    ref = serving.serve_plain(reqs, ctx)
    for a, c in zip(got, ref):
        assert a.rid == c.rid
        assert np.array_equal(a.embedding, c.embedding), a.rid
        assert a.kl == c.kl
        assert a.finished >= a.admitted >= a.arrival - 1e-3
    s = serving.latency_summary(got)
    assert s["jobs"] == len(reqs) and s["p99_s"] >= s["p50_s"] > 0


def test_serving_respects_hbm_budget(bt):
    # a budget that holds about two jobs at a time: admission waits for room, every request is
    # still served, and the later requests are admitted only after earlier ones finished
    from bhtsne_amd import serving
    n, d_in = 1500, 20
    need = bt.job_footprint(n, d_in, 2, "fitsne", max_boxes=100)
    ctx = bt.Context(0)
    reqs = [serving.Request(r, _data(n, d_in, 200 + r), epochs=30, arrival=0.0,
                            y0=datagen.seed_embedding(n, 2, r), max_boxes=100)
            for r in range(4)]
    got = serving.serve(reqs, ctx, step_epochs=10, hbm_budget=int(2.5 * need), want_kl=False)
    assert all(g.embedding is not None for g in got)
    first_done = min(g.finished for g in got)
    assert sorted(g.admitted for g in got)[2] >= first_done - 1e-3


def test_footprint_covers_allocation(bt):
    ctx = bt.Context(0)
    for spec in (SPECS[1], SPECS[3], SPECS[2], SPECS[7]):
        n, d_in, d_out, method, _, _ = spec
        before = ctx.bytes_in_use()
        t = bt.TSNE(_data(n, d_in, 11), d_out=d_out, ctx=ctx).perplexity(30.0).epochs(3)
        b = bt.Batch(ctx=ctx)
        b.add(t, method, 0.5)
        used = ctx.bytes_in_use() - before
        bound = bt.job_footprint(n, d_in, d_out, method)
        print(spec, "allocated", used, "bound", bound, flush=True)
        assert used <= bound
        b.run()
        b.close()


def test_mixed_batch_of_varied_jobs_equals_alone(bt):
    # Many jobs of both methods (two launch groups) with uneven sizes, some above the renumbering
    # threshold, stepped with plain launches and with graph replay; sampled members are compared
    # bit for bit with their fits alone.
    specs = [datagen.config_c4_job(j) for j in range(24)]
    ns = [x.shape[0] for x, _ in specs]
    sample = sorted({j for j in range(24) if ns[j] >= 32768} | {0, 1, 5, 6})
    ctx = bt.Context(0)

    def make(j):
        x, _ = specs[j]
        t = bt.TSNE(x, ctx=ctx).perplexity(30.0).epochs(300)
        return t.initial_embedding(datagen.seed_embedding(x.shape[0], 2, j))

    alone = {}
    for j in sample:
        t = make(j)
        t._set(flags=SERIAL_EPOCH)
        if specs[j][1] == "fitsne":
            t.fit_sne()
        else:
            t.barnes_hut(0.5)
        alone[j] = t.embedding()
    for graph in (False, True):
        b = bt.Batch(ctx, graph=graph)
        jobs = [make(j) for j in range(24)]
        for j, t in enumerate(jobs):
            b.add(t, specs[j][1], 0.5)
        b.run()
        for j in sample:
            assert np.array_equal(jobs[j].embedding(), alone[j]), (graph, j, ns[j])
        b.close()
