"""Golden vectors (tests/golden/golden_vectors.npz, written by tests/golden/make_golden.py from the
oracle on the reference tests' LCG data): the oracle must keep reproducing them bit for bit (a
regression pin of the restatement, with an independent numpy check of the kNN), and the device
path must match them to the parity bars without needing the oracle at all."""
import os

import numpy as np
import pytest

from tests._util import f32_seq_dist

HERE = os.path.dirname(os.path.abspath(__file__))
G = dict(np.load(os.path.join(HERE, "golden", "golden_vectors.npz")))


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(((a - b) ** 2).sum() / max((b ** 2).sum(), 1e-300)))


# ---- the oracle against its own fixtures (CPU) ----------------------------------------------

def test_fixture_knn_is_brute_force():
    # independent of the oracle: sequential f32 distances, (distance, index) order
    d = f32_seq_dist(G["knn_x"])
    np.fill_diagonal(d, np.inf)
    order = np.lexsort((np.broadcast_to(np.arange(200), d.shape), d), axis=1)[:, :12]
    assert np.array_equal(order, G["knn_idx"])
    assert np.array_equal(np.take_along_axis(d, order, 1), G["knn_dist"])


def test_oracle_reproduces_fixtures():
    import oracle
    idx, dist = oracle.knn(G["knn_x"], 12)
    assert np.array_equal(idx, G["knn_idx"]) and np.array_equal(dist, G["knn_dist"])
    assert np.array_equal(oracle.search_beta(G["beta_dist"], 15.0), G["beta_p_libm"])
    with oracle.cr_explog():
        assert np.array_equal(oracle.search_beta(G["beta_dist"], 15.0), G["beta_p_cr"])
    rows, cols, vals = oracle.symmetrize(idx, oracle.search_beta(dist, 4.0))
    assert np.array_equal(rows, G["sym_rows"]) and np.array_equal(cols, G["sym_cols"])
    assert np.array_equal(vals, G["sym_vals"])
    pn = vals.copy()
    oracle.normalize_p(pn, 12.0)
    assert np.array_equal(pn, G["p_norm"])
    for name, method in (("fit", "fitsne"), ("bh", "bh")):
        pos, neg, inv = oracle.forces(G["y0"], rows, cols, pn, method=method, theta=0.5)
        assert np.array_equal(pos, G[f"{name}_pos"]) and np.array_equal(neg, G[f"{name}_neg"])
        assert np.float32(inv) == G[f"{name}_invz"]
    for D in range(3, 8):
        pos, neg, inv = oracle.forces(G[f"y0_d{D}"], rows, cols, pn, method="bh", theta=0.5)
        assert np.array_equal(pos, G[f"bh_pos_d{D}"]) and np.array_equal(neg, G[f"bh_neg_d{D}"])
        assert np.float32(inv) == G[f"bh_invz_d{D}"]


# ---- the device against the fixtures --------------------------------------------------------

@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


@pytest.mark.gpu
def test_device_affinities_match_fixtures(bt):
    idx, dist = bt.knn(G["knn_x"], 12)
    assert np.array_equal(idx, G["knn_idx"]) and np.array_equal(dist, G["knn_dist"])
    p_cr = bt.search_beta(G["beta_dist"], 15.0, mode=1)
    assert np.array_equal(p_cr, G["beta_p_cr"])
    rel = (np.abs(p_cr - G["beta_p_libm"]) / G["beta_p_libm"].max(1, keepdims=True)).max(1)
    assert np.mean(rel < 1e-5) >= 0.99 and rel.max() < 2e-4
    p = bt.search_beta(G["knn_dist"], 4.0, mode=1)
    rows, cols, vals = bt.symmetrize(G["knn_idx"], p)
    assert np.array_equal(rows, G["sym_rows"]) and np.array_equal(cols, G["sym_cols"])
    assert np.max(np.abs(vals - G["sym_vals"]) / G["sym_vals"].max()) < 1e-5


def _job(bt, method):
    t = bt.TSNE(G["knn_x"]).perplexity(4.0)
    t.with_affinities(bt.SparseAffinities(G["sym_rows"], G["sym_cols"], G["sym_vals"], 4.0))
    t.initial_embedding(G["y0"]).prepare(method, 0.5)
    return t


@pytest.mark.gpu
@pytest.mark.parametrize("name,method", [("fit", "fitsne"), ("bh", "bh")])
def test_device_forces_match_fixtures(bt, name, method):
    t = _job(bt, method)
    pos, neg, inv = t.forces()
    rp, rn, ri = G[f"{name}_pos"], G[f"{name}_neg"], float(G[f"{name}_invz"])
    e_grad = _rel(pos - neg * inv, rp - rn * ri)
    print(f"{method}: attraction {_rel(pos, rp):.2e} repulsion {_rel(neg, rn):.2e} "
          f"gradient {e_grad:.2e}", flush=True)
    assert _rel(pos, rp) < 1e-5 and e_grad < 1e-5 and abs(inv - ri) / ri < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("D", [3, 4, 5, 6, 7])
def test_device_bh_forces_in_d_dimensions_match_fixtures(bt, D):
    t = bt.TSNE(G["knn_x"], d_out=D).perplexity(4.0)
    t.with_affinities(bt.SparseAffinities(G["sym_rows"], G["sym_cols"], G["sym_vals"], 4.0))
    t.initial_embedding(G[f"y0_d{D}"]).prepare("bh", 0.5)
    pos, neg, inv = t.forces()
    rp, rn, ri = G[f"bh_pos_d{D}"], G[f"bh_neg_d{D}"], float(G[f"bh_invz_d{D}"])
    print(f"D={D}: attraction {_rel(pos, rp):.2e} repulsion {_rel(neg, rn):.2e}", flush=True)
    assert _rel(pos, rp) < 1e-6 and _rel(neg, rn) < 1e-6 and abs(inv - ri) / ri < 1e-5


@pytest.mark.gpu
def test_device_epoch_and_kl_match_fixtures(bt):
    t = _job(bt, "fitsne")
    t.step(1)
    y, uy, g = t.get_state()
    assert _rel(y, G["epoch1_y"]) < 1e-4 and _rel(uy, G["epoch1_uy"]) < 1e-4
    assert np.mean(g == G["epoch1_gains"]) > 0.99
    for name, method in (("fit", "fitsne"), ("bh", "bh")):
        t = _job(bt, method)
        t.finish()  # a run of no epochs: KL of the seed state with the fit's own Z
        r, tr = t.kl_divergence_pair()
        rr, rt = G[f"kl_{name}"]
        assert abs(r - rr) / abs(rr) < 1e-4 and abs(tr - rt) / abs(rt) < 1e-4
