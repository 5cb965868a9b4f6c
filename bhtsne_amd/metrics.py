"""Error of a job's approximate gradient against the exact FP64 gradient.

The exact gradient of the reference's exact path (src/lib.rs:570-622) is
grad_i = sum_j p_ij q_ij (y_i - y_j) - sum_j q_ij^2 (y_i - y_j) / Z, q_ij = 1 / (1 + |y_i - y_j|^2),
Z = sum of q over ordered pairs.  The non-edge terms and Z come from the device's FP64
brute-force kernel (all pairs); the edge terms are summed here in FP64 over a row sample."""
from __future__ import annotations

import numpy as np


def exact_attraction_rows(y, aff, sel):
    """sum_j p_ij q_ij (y_i - y_j) in FP64 for the rows in sel (CSR in sample order)."""
    rows = np.asarray(aff.rows, np.int64)
    sel = np.asarray(sel, np.int64)
    lens = rows[sel + 1] - rows[sel]
    rid = np.repeat(np.arange(len(sel)), lens)
    starts = np.repeat(rows[sel], lens)
    off = np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens)
    e = starts + off
    yi = y[sel][rid].astype(np.float64)
    yj = y[np.asarray(aff.columns, np.int64)[e]].astype(np.float64)
    dl = yi - yj
    w = np.asarray(aff.values, np.float64)[e] / (1.0 + (dl ** 2).sum(1))
    out = np.zeros((len(sel), y.shape[1]))
    np.add.at(out, rid, w[:, None] * dl)
    return out


def exact_gradient_rows(y, aff, sel, ctx=None):
    """The exact FP64 gradient of the rows in sel at embedding y (n x d, sample order)."""
    from . import exact_repulsion_f64
    eneg, ez = exact_repulsion_f64(y, ctx)
    return exact_attraction_rows(y, aff, sel) - eneg[np.asarray(sel)] / ez


def rel_l2(g, ge) -> float:
    g = np.asarray(g, np.float64)
    return float(np.sqrt(((g - ge) ** 2).sum() / (ge ** 2).sum()))


def gradient_sample(job, sample: int = 5000, seed: int = 0):
    """(the prepared job's gradient on a row sample, the exact FP64 one, the sample, y, the
    affinities): the job's gradient is attraction - repulsion / Z as its epoch computes it."""
    pos, neg, inv = job.forces()
    y, _, _ = job.get_state()
    n, d = job._n, job._d
    y = y.reshape(n, d)
    pos, neg = pos.reshape(n, d), neg.reshape(n, d)
    aff = job.affinities()
    if job.run_state()[1]:  # the epoch's P: exported values times the exaggeration
        aff.values = aff.values * np.float32(job.params()["early_exaggeration"])
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(n, size=min(sample, n), replace=False))
    g = pos[sel].astype(np.float64) - neg[sel].astype(np.float64) * inv
    return g, exact_gradient_rows(y, aff, sel, job._ctx), sel, y, aff


def gradient_error(job, sample: int = 5000, seed: int = 0) -> float:
    """Relative L2 error, over `sample` rows, of the prepared job's gradient at its current state
    against the exact FP64 gradient."""
    g, ge, _, _, _ = gradient_sample(job, sample, seed)
    return rel_l2(g, ge)


def error_report(job, sample: int = 5000, seed: int = 0) -> dict:
    """The gradient's error against the exact FP64 one, and the error of its repulsive term
    (F- / Z) alone.  Near equilibrium attraction and repulsion nearly cancel, so the gradient's
    relative error grows while the approximation's own accuracy is the repulsion's."""
    from . import exact_repulsion_f64
    pos, neg, inv = job.forces()
    y, _, _ = job.get_state()
    n, d = job._n, job._d
    y = y.reshape(n, d)
    pos, neg = pos.reshape(n, d), neg.reshape(n, d)
    aff = job.affinities()
    if job.run_state()[1]:
        aff.values = aff.values * np.float32(job.params()["early_exaggeration"])
    rng = np.random.default_rng(seed)
    sel = np.sort(rng.choice(n, size=min(sample, n), replace=False))
    eneg, ez = exact_repulsion_f64(y, job._ctx)
    rep = neg[sel].astype(np.float64) * inv
    erep = eneg[sel] / ez
    g = pos[sel].astype(np.float64) - rep
    ge = exact_attraction_rows(y, aff, sel) - erep
    return {"gradient": rel_l2(g, ge), "repulsion": rel_l2(rep, erep),
            "z": abs(1.0 / inv - ez) / ez, "rows": int(len(sel))}
