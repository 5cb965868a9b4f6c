"""GPU parity of the exact fit (src/lib.rs:489-646) against the oracle's restatement, and the
reference's own exact-path tests (src/test.rs:190-204, 520-617) on the device.

Tolerances: with the f32-sequential bandwidth search and correctly rounded exp / log on both
sides, P matches to summation order (the normalising sum is FP64 here, a pairwise f32 tree in
the oracle), and the gradient is summed in FP64 here against the reference's left-to-right f32.
Under early exaggeration a 60-point fit is chaotic: a relative change of 3e-7 in P moves the
oracle's own embedding by 12% after 20 epochs, so, as the reference's own exact tests do, it is
compared after one epoch (1e-5); without exaggeration (stop_lying_epoch 0) the same change stays
at 6e-7 after 50 epochs, and 50 epochs are compared at 1e-5.  The final KL of a 1000-point,
300-epoch fit (through stop-lying; the oracle's own KL moves 4e-7 under that change) agrees to
1%."""
import numpy as np
import pytest

import oracle
from tests._util import lcg_samples

pytestmark = pytest.mark.gpu

SEQ = 1  # BHT_BANDWIDTH_F32_SEQ


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


def _fit(bt, x, y0, epochs, metric="sqeuclidean", mode=SEQ, **kw):
    t = bt.TSNE(x).perplexity(10.0).epochs(epochs).bandwidth_mode(mode)
    for k, v in kw.items():
        getattr(t, k)(v)
    if y0 is not None:
        t.initial_embedding(y0)
    t.exact(metric)
    return t


def _oracle(x, y0, epochs, metric="sqeuclidean", exag=12.0, stop_lying_epoch=250):
    with oracle.cr_explog():
        P = oracle.exact_affinities(x, 10.0, metric, exag=exag, stop_lying_epoch=stop_lying_epoch)
    y, uy, g = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.exact_epochs(y, uy, g, P, learning_rate=oracle.auto_learning_rate(len(x), exag),
                        early_exaggeration=exag, stop_lying_epoch=stop_lying_epoch,
                        e0=0, e1=epochs)
    return y, P


@pytest.mark.parametrize("metric", ["sqeuclidean", "euclidean"])
@pytest.mark.parametrize("epochs,stop_lying", [(1, 250), (50, 0)])
def test_exact_epochs_match_oracle(bt, metric, epochs, stop_lying):
    x = lcg_samples(60, 4, 7)
    y0 = lcg_samples(60, 2, 99)
    t = _fit(bt, x, y0, epochs, metric, stop_lying_epoch=stop_lying)
    y = t.embedding().reshape(-1, 2)
    yr, _ = _oracle(x, y0, epochs, metric, stop_lying_epoch=stop_lying)
    print(f"{metric} {epochs} epochs: rel {_rel(y, yr):.2e}", flush=True)
    assert _rel(y, yr) < 1e-5


@pytest.mark.parametrize("D,epochs", [(1, 1), (3, 50), (5, 50), (7, 50)])
def test_exact_epochs_match_oracle_in_d_dimensions(bt, D, epochs):
    # the reference's exact fit takes any output dimension (src/lib.rs:489-646).  In one
    # dimension the points' order along the line amplifies last-bit differences (50 epochs:
    # 9.3e-5 against the oracle), so D = 1 is compared after one epoch.
    x = lcg_samples(60, 4, 7)
    y0 = lcg_samples(60, D, 99 + D)
    t = bt.TSNE(x, d_out=D).perplexity(10.0).epochs(epochs).bandwidth_mode(SEQ)
    t.stop_lying_epoch(0).initial_embedding(y0).exact("sqeuclidean")
    y = t.embedding().reshape(-1, D)
    yr, _ = _oracle(x, y0, epochs, "sqeuclidean", stop_lying_epoch=0)
    print(f"exact D={D} {epochs} epochs: rel {_rel(y, yr):.2e}", flush=True)
    assert _rel(y, yr) < 1e-5


def test_exact_final_kl_within_one_percent(bt):
    # crosses stop-lying (250) and the momentum switch; KL against the reference's formula
    x = lcg_samples(1000, 10, 3)
    y0 = (lcg_samples(1000, 2, 99) * np.float32(1e-2)).astype(np.float32)
    t = _fit(bt, x, y0, 300)
    kr, kt = t.kl_divergence_pair()
    y = t.embedding().reshape(-1, 2)
    yr, P = _oracle(x, y0, 300)
    rr, rt = oracle.exact_kl(yr, P)
    # the same embedding through the oracle's KL: the device KL itself
    sr, st = oracle.exact_kl(y, P)
    print(f"KL device {kr:.6f} {kt:.6f}; oracle run {rr:.6f} {rt:.6f}; device state {sr:.6f}",
          flush=True)
    assert abs(kt - st) / abs(st) < 1e-5
    assert abs(kr - rr) / abs(rr) < 1e-2 and abs(kt - rt) / abs(rt) < 1e-2


def test_exact_repeats_and_has_no_affinities(bt):
    x = lcg_samples(60, 4, 7)
    y0 = lcg_samples(60, 2, 99)
    a = _fit(bt, x, y0, 50, mode=0)
    b = _fit(bt, x, y0, 50, mode=0)
    assert np.array_equal(a.embedding(), b.embedding())
    assert a.affinities() is None


def _snapshot0(bt, **kw):
    x = lcg_samples(60, 4, 7)
    seed = lcg_samples(60, 2, 99)
    got = []
    t = bt.TSNE(x).perplexity(10.0).epochs(1).initial_embedding(seed)
    for k, v in kw.items():
        getattr(t, k)(v)
    t.epoch_callback(lambda e, y: got.append(y.copy()) if e == 0 else None)
    t.exact("sqeuclidean")
    return got[0].reshape(-1, 2)


def _diag(y):
    return float(np.linalg.norm(y.max(0) - y.min(0)))


def test_reference_exaggeration_tests(bt):
    # src/test.rs:560-617
    default = _snapshot0(bt)
    assert np.array_equal(default, _snapshot0(bt, early_exaggeration=12.0))
    weak = _snapshot0(bt, early_exaggeration=4.0)
    assert np.linalg.norm(default - weak, axis=1).mean() > 0.05 * _diag(default)
    one = _snapshot0(bt, early_exaggeration=1.0)
    zero = _snapshot0(bt, stop_lying_epoch=0)
    assert np.linalg.norm(one - zero, axis=1).mean() <= 1e-4 * _diag(one) + 1e-6


def test_reference_exact_kl_is_finite_and_nonnegative(bt):
    # src/test.rs:190-204
    x = lcg_samples(60, 4, 7)
    t = bt.TSNE(x).perplexity(10.0).epochs(100)
    t.exact("sqeuclidean")
    kl = t.kl_divergence()
    assert np.isfinite(kl) and kl >= 0.0


def test_exact_rejects_wrong_seed_length(bt):
    # src/test.rs:648-669
    x = lcg_samples(60, 4, 7)
    t = bt.TSNE(x).perplexity(10.0).epochs(1).initial_embedding(np.zeros(7, np.float32))
    with pytest.raises(bt.InvalidInput, match="initial embedding has"):
        t.exact("sqeuclidean")
