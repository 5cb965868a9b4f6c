"""The FP64 brute-force repulsion on the device against the oracle's, and the error of the
approximate gradients (FIt-SNE interpolation, Barnes-Hut) against the exact FP64 gradient at the
sizes of the C1 (Barnes-Hut, N = 5,000) and C2 (FIt-SNE, N = 70,000) workloads, measured on a
5,000-row sample of a mid-run state.  The exact gradient is grad_i = sum_j p_ij q_ij (y_i - y_j) -
sum_j q_ij^2 (y_i - y_j) / Z (src/lib.rs:570-622)."""
import numpy as np
import pytest

import oracle
from bhtsne_amd import datagen
from bhtsne_amd.metrics import gradient_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import bhtsne_amd
    return bhtsne_amd


@pytest.mark.parametrize("n,d,seed", [(3000, 2, 1), (1500, 1, 2), (257, 2, 3)])
def test_exact_repulsion_matches_oracle(bt, n, d, seed):
    rng = np.random.default_rng(seed)
    y = (rng.standard_normal((n, d)) * 7).astype(np.float32)
    neg, z = bt.exact_repulsion_f64(y)
    rneg, rz = oracle.exact_repulsion_f64(y)
    assert abs(z - rz) / rz < 1e-12
    assert np.sqrt(((neg - rneg) ** 2).sum() / (rneg ** 2).sum()) < 1e-12


def test_gradient_error_c1_barnes_hut(bt):
    x, _ = datagen.config_c1()
    n = x.shape[0]
    t = bt.TSNE(x).perplexity(30.0).epochs(1000)
    t.initial_embedding(datagen.seed_embedding(n, 2, 0)).prepare("bh", 0.5)
    t.step(300)
    err = gradient_error(t)
    print(f"C1 Barnes-Hut theta 0.5, N={n}, epoch 300: gradient rel L2 error vs FP64 exact "
          f"{err:.3e}", flush=True)
    assert err < 1e-1


def test_gradient_error_c2_fitsne(bt):
    x, _ = datagen.config_c2()
    n = x.shape[0]
    t = bt.TSNE(x).perplexity(30.0).epochs(750)
    t.initial_embedding(datagen.seed_embedding(n, 2, 1)).prepare("fitsne")
    t.step(300)
    err = gradient_error(t)
    print(f"C2 FIt-SNE, N={n}, epoch 300: gradient rel L2 error vs FP64 exact (5k rows) "
          f"{err:.3e}", flush=True)
    assert err < 5e-2
