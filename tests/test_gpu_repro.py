"""Run-to-run reproducibility of jobs large enough to run each epoch's repulsion on a second
stream beside the attraction (32768 points or more).  The two streams compute exactly what the
serial epoch computes (the attraction's work queue does not change any sum), so a forked fit must
repeat bit for bit and equal the serial fit bit for bit.  Each stream that consumes what the other
produced first writes back and invalidates every XCD's L2 (the per-XCD L2s are not coherent with
each other for ordinary device memory)."""
import numpy as np
import pytest

from bhtsne_amd import datagen

pytestmark = pytest.mark.gpu

SERIAL = 2  # BHT_FLAG_SERIAL_EPOCH
NO_AUTOTUNE = 4  # BHT_FLAG_NO_AUTOTUNE
GLOBAL_LINES = 8  # BHT_FLAG_FFT_GLOBAL_LINES


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


def _fit(bt, x, y0, aff, flags, epochs, d_out=2):
    t = bt.TSNE(x, d_out=d_out).perplexity(30.0).epochs(epochs)
    t._set(flags=flags)
    t.with_affinities(aff).initial_embedding(y0)
    t.fit_sne()
    return t.embedding(), t.kl_divergence_pair()


@pytest.mark.parametrize("request_id,epochs", [(8, 1000), (16, 400)])
def test_forked_fit_repeats_and_equals_serial(bt, request_id, epochs):
    # request 8 of the serving stream (111,573 points, D_in 20) differed from run to run in 6 of 6
    # runs before the cross-stream cache maintenance
    x, d_out = datagen.config_c5_request(request_id)
    n = x.shape[0]
    print(f"request {request_id}: n={n} d_out={d_out} epochs={epochs}", flush=True)
    t = bt.TSNE(x, d_out=d_out).perplexity(30.0).epochs(0)
    t.fit_sne()
    aff = t.affinities()
    y0 = datagen.seed_embedding(n, d_out, request_id)
    a, ka = _fit(bt, x, y0, aff, 0, epochs, d_out)
    b, kb = _fit(bt, x, y0, aff, 0, epochs, d_out)
    c, kc = _fit(bt, x, y0, aff, SERIAL, epochs, d_out)
    d, kd = _fit(bt, x, y0, aff, NO_AUTOTUNE, epochs, d_out)
    # every FFT line transformed in global memory instead of LDS: the same arithmetic
    e, ke = _fit(bt, x, y0, aff, GLOBAL_LINES, epochs, d_out)
    diff_ab = int((a != b).sum())
    diff_ac = int((a != c).sum())
    diff_ad = int((a != d).sum())
    diff_ae = int((a != e).sum())
    print(f"forked vs forked: {diff_ab} differing values; forked vs serial: {diff_ac}; "
          f"autotuned vs built-in form: {diff_ad}; LDS vs global FFT lines: {diff_ae}; "
          f"KL {ka} {kb} {kc} {kd} {ke}", flush=True)
    assert np.all(np.isfinite(a))
    assert diff_ab == 0 and ka == kb
    assert diff_ac == 0 and ka == kc
    assert diff_ad == 0 and ka == kd
    assert diff_ae == 0 and ka == ke


def test_autotuner_choice_is_reported_and_cached(bt):
    x, _ = datagen.gmm(40000, 16, 8, 16.0, 5)
    ctx = bt.Context(0)
    forms = []
    for _ in range(2):
        t = bt.TSNE(x, ctx=ctx).perplexity(30.0).epochs(5)
        t.initial_embedding(datagen.seed_embedding(40000, 2, 5)).fit_sne()
        forms.append(t.epoch_form())
    print(f"autotuned epoch form at n=40000: {forms}", flush=True)
    assert forms[0] == forms[1] and forms[0] >= -1
    t = bt.TSNE(x[:1000], ctx=ctx).perplexity(30.0).epochs(3)
    t.fit_sne()
    assert t.epoch_form() == -1  # small jobs run in order on one stream
