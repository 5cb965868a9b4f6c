"""Writes the golden vectors under tests/golden/ from the oracle (the C restatement of the
reference's arithmetic).  Run from the repository root: ``python tests/golden/make_golden.py``.

Inputs are the reference tests' deterministic LCG data (src/test.rs:973-982), so the fixtures do
not depend on any random generator.  Every array is stored in one numpy archive (no pickles):

* knn_x (200 x 8), knn_idx / knn_dist (k = 12): brute-force kNN, (distance, index) order;
* beta_dist (64 x 45), beta_p_libm / beta_p_cr: the bandwidth search (perplexity 15) with the host
  libm's expf / logf and with correctly rounded exp / ln;
* sym_rows / sym_cols / sym_vals: symmetrisation of (knn_idx, conditional P of knn_dist);
* p_norm: P normalised with early exaggeration 12;
* y0 (200 x 2): a state; fit_pos / fit_neg / fit_invz and bh_pos / bh_neg / bh_invz: forces of
  that state (FIt-SNE, 10-box floor; Barnes-Hut theta 0.5);
* epoch1_y / epoch1_uy / epoch1_gains: one FIt-SNE epoch from (y0, 0, 1);
* kl_fit / kl_bh: (reference formula, true KL) at y0;
* y0_d3 .. y0_d7 (200 x D) and bh_pos_dD / bh_neg_dD / bh_invz_dD: Barnes-Hut forces (theta 0.5)
  of a D-dimensional state over the same P, for D = 3..7.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from tests._util import lcg_samples  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_vectors.npz")


def main():
    g = {}
    x = lcg_samples(200, 8, 1234)
    idx, dist = oracle.knn(x, 12)
    g.update(knn_x=x, knn_idx=idx, knn_dist=dist)
    bd = lcg_samples(64, 45, 77)
    bd = np.sort(np.abs(bd) * np.float32(10.0), axis=1).astype(np.float32)
    g["beta_dist"] = bd
    g["beta_p_libm"] = oracle.search_beta(bd, 15.0)
    with oracle.cr_explog():
        g["beta_p_cr"] = oracle.search_beta(bd, 15.0)
    p = oracle.search_beta(dist, 4.0)
    rows, cols, vals = oracle.symmetrize(idx, p)
    g.update(sym_rows=rows, sym_cols=cols, sym_vals=vals)
    pn = vals.copy()
    oracle.normalize_p(pn, 12.0)
    g["p_norm"] = pn
    y0 = (lcg_samples(200, 2, 99) * np.float32(4.0)).astype(np.float32)
    g["y0"] = y0
    for name, method in (("fit", "fitsne"), ("bh", "bh")):
        pos, neg, inv = oracle.forces(y0, rows, cols, pn, method=method, theta=0.5)
        g[f"{name}_pos"], g[f"{name}_neg"], g[f"{name}_invz"] = pos, neg, np.float32(inv)
        g[f"kl_{name}"] = np.array(oracle.kl_divergence(y0, rows, cols, pn, method=method),
                                   np.float64)
    for D in range(3, 8):
        yd = (lcg_samples(200, D, 99 + D) * np.float32(4.0)).astype(np.float32)
        pos, neg, inv = oracle.forces(yd, rows, cols, pn, method="bh", theta=0.5)
        g[f"y0_d{D}"] = yd
        g[f"bh_pos_d{D}"], g[f"bh_neg_d{D}"], g[f"bh_invz_d{D}"] = pos, neg, np.float32(inv)
    y, uy, gains = y0.copy(), np.zeros_like(y0), np.ones_like(y0)
    oracle.run_epochs(y, uy, gains, rows, cols, pn.copy(),
                      learning_rate=oracle.auto_learning_rate(200, 12.0), e0=0, e1=1)
    g.update(epoch1_y=y, epoch1_uy=uy, epoch1_gains=gains)
    np.savez(OUT, **g)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(g)} arrays)")


if __name__ == "__main__":
    main()
