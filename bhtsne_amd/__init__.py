"""bhtsne_amd -- the bhtsne hot path (sparse affinities + per-epoch gradient) on MI355X.

``TSNE`` mirrors the reference builder ``tSNE<'data, f32, &[f32], D>`` (src/lib.rs:149-1268):
the same setters and defaults, ``barnes_hut`` / ``fit_sne`` and their ``*_with_neighbors``
variants, ``embedding()``, ``kl_divergence()``, ``affinities()`` / ``with_affinities()``, and the
same rejection messages (raised as ``InvalidInput``).  Every numeric step runs in the HIP library
(libbhtsne_hip.so, C ABI in include/bhtsne_hip.h); the metric is Euclidean.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
from dataclasses import dataclass
from typing import Callable, NamedTuple, Optional, Sequence

import numpy as np

from . import _lib
from ._lib import BhtError, InvalidInput, check, ptr

__all__ = ["Context", "Neighbor", "SparseAffinities", "TSNE", "Batch", "BhtError",
           "InvalidInput", "knn", "search_beta", "symmetrize", "check_affinities",
           "default_context", "job_footprint"]


# Live handles, released in dependency order (batches, jobs, contexts) when the interpreter exits,
# while the library and the HIP runtime are still fully up: objects left to the interpreter's
# final teardown would otherwise be released in no particular order.
_live = weakref.WeakSet()


def _release_all():
    objs = list(_live)
    for kind in ("Batch", "TSNE", "Context"):
        for o in objs:
            if type(o).__name__ == kind:
                try:
                    o.close()
                except Exception:
                    pass


atexit.register(_release_all)


class Context:
    """One device context (twiddle tables, HBM budget accounting)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = _lib.load()
        h = C.c_void_p()
        check(L.bht_ctx_create(int(device), int(hbm_budget), C.byref(h)))
        self._h = h
        self._L = L
        _live.add(self)

    @property
    def handle(self):
        return self._h

    def bytes_in_use(self) -> int:
        return int(self._L.bht_ctx_bytes_in_use(self._h))

    def close(self):
        if self._h:
            self._L.bht_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class Neighbor(NamedTuple):
    """A sample's nearest neighbour: index and distance (src/lib.rs:107-111)."""
    index: int
    distance: float


@dataclass
class SparseAffinities:
    """CSR affinities (src/lib.rs:116-144): usize rows, u32 columns, f32 values, perplexity."""
    rows: np.ndarray
    columns: np.ndarray
    values: np.ndarray
    perplexity: float


def _neighbors_arrays(neighbors):
    """Flatten &[Vec<Neighbor<f32>>]-like input into (index u64, distance f32, row_lens u64)."""
    if isinstance(neighbors, tuple) and len(neighbors) == 2:
        idx = np.ascontiguousarray(neighbors[0], np.uint64)
        dist = np.ascontiguousarray(neighbors[1], np.float32)
        n, k = idx.shape
        return idx.ravel(), dist.ravel(), np.full(n, k, np.uint64)
    lens = np.array([len(r) for r in neighbors], np.uint64)
    idx = np.array([nb[0] for r in neighbors for nb in r], np.uint64)
    dist = np.array([nb[1] for r in neighbors for nb in r], np.float32)
    return idx, dist, lens


class TSNE:
    """tSNE builder (src/lib.rs:241-435) over the MI355X engine; Euclidean metric only."""

    def __init__(self, data, d_out: int = 2, ctx: Optional[Context] = None):
        self._ctx = ctx or default_context()
        self._L = _lib.load()
        x = np.ascontiguousarray(data, np.float32)
        if x.ndim == 1:
            x = x.reshape(-1, 1)
        self._x = x
        self._n = x.shape[0]
        self._d = int(d_out)
        p = _lib.Params()
        self._L.bht_params_default(C.byref(p))
        p.d_out = self._d
        h = C.c_void_p()
        check(self._L.bht_job_create(self._ctx.handle, self._n, C.byref(p), C.byref(h)))
        self._h = h
        if x.shape[1] > 0 and self._n > 0:
            check(self._L.bht_job_set_data_f32(self._h, ptr(x), self._n, x.shape[1]))
        self._callback = None
        self._fitted = False
        _live.add(self)

    def close(self):
        """Releases the job's device memory (the object is unusable afterwards)."""
        if getattr(self, "_h", None):
            self._L.bht_job_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters (src/lib.rs:277-371) ----------------------------------------------------
    def _params(self):
        p = _lib.Params()
        check(self._L.bht_job_get_params(self._h, C.byref(p)))
        return p

    def _set(self, **kv):
        p = self._params()
        for k, v in kv.items():
            setattr(p, k, v)
        check(self._L.bht_job_set_params(self._h, C.byref(p)))
        return self

    def learning_rate(self, v: float):
        return self._set(learning_rate=float(v))

    def epochs(self, v: int):
        return self._set(epochs=int(v))

    def momentum(self, v: float):
        return self._set(momentum=float(v))

    def final_momentum(self, v: float):
        return self._set(final_momentum=float(v))

    def momentum_switch_epoch(self, v: int):
        return self._set(momentum_switch_epoch=int(v))

    def stop_lying_epoch(self, v: int):
        return self._set(stop_lying_epoch=int(v))

    def early_exaggeration(self, v: float):
        return self._set(early_exaggeration=float(v))

    def perplexity(self, v: float):
        return self._set(perplexity=float(v))

    def min_boxes(self, v: int):
        """FIt-SNE grid lower bound per axis (10 reproduces the reference)."""
        return self._set(min_boxes=int(v))

    def max_boxes(self, v: int):
        return self._set(max_boxes=int(v))

    def bandwidth_mode(self, v: int):
        return self._set(bandwidth_mode=int(v))

    def epoch_callback(self, cb: Callable[[int, np.ndarray], None]):
        self._callback = cb
        return self

    def initial_embedding(self, y):
        y = np.ascontiguousarray(y, np.float32).ravel()
        check(self._L.bht_job_set_initial_embedding(self._h, ptr(y), y.size))
        return self

    def params(self) -> dict:
        p = self._params()
        return {name: getattr(p, name) for name, _ in _lib.Params._fields_}

    # ---- fits --------------------------------------------------------------------------------
    def _cb(self):
        if self._callback is None:
            return _lib.NULL_CB
        user = self._callback

        def tramp(epoch, data, length, _u):
            user(int(epoch), np.ctypeslib.as_array(data, (int(length),)).copy())
        self._tramp = _lib.EPOCH_CB(tramp)
        return self._tramp

    @staticmethod
    def _euclid(metric):
        if metric not in (None, "euclidean"):
            raise NotImplementedError("the MI355X engine implements the Euclidean metric only")

    def exact(self, metric="euclidean"):
        """The exact O(n^2) fit (src/lib.rs:489-646) with the Euclidean distance or, as the
        reference's exact tests use, the squared Euclidean distance ("sqeuclidean")."""
        m = {"euclidean": _lib.METRIC_EUCLIDEAN, "sqeuclidean": _lib.METRIC_SQEUCLIDEAN}.get(metric)
        if m is None:
            raise NotImplementedError("the exact path takes metric 'euclidean' or 'sqeuclidean'")
        check(self._L.bht_job_exact(self._h, m, self._cb(), None))
        self._fitted = True
        return self

    def barnes_hut(self, theta: float, metric="euclidean"):
        self._euclid(metric)
        check(self._L.bht_job_barnes_hut(self._h, float(theta), self._cb(), None))
        self._fitted = True
        return self

    def barnes_hut_with_neighbors(self, theta: float, neighbors):
        idx, dist, lens = _neighbors_arrays(neighbors)
        check(self._L.bht_job_barnes_hut_with_neighbors(self._h, float(theta), ptr(idx),
                                                        ptr(dist), ptr(lens), lens.size,
                                                        self._cb(), None))
        self._fitted = True
        return self

    def fit_sne(self, metric="euclidean"):
        self._euclid(metric)
        check(self._L.bht_job_fit_sne(self._h, self._cb(), None))
        self._fitted = True
        return self

    def fit_sne_with_neighbors(self, neighbors):
        idx, dist, lens = _neighbors_arrays(neighbors)
        check(self._L.bht_job_fit_sne_with_neighbors(self._h, ptr(idx), ptr(dist), ptr(lens),
                                                     lens.size, self._cb(), None))
        self._fitted = True
        return self

    # ---- outputs -----------------------------------------------------------------------------
    def embedding(self) -> np.ndarray:
        out = np.zeros(self._n * self._d, np.float32)
        check(self._L.bht_job_get_embedding(self._h, ptr(out), out.size))
        return out

    def kl_divergence_pair(self):
        """(reference-compatible KL, true KL) or None before any fit."""
        r = C.c_double()
        t = C.c_double()
        st = self._L.bht_job_kl_divergence(self._h, C.byref(r), C.byref(t))
        if st == _lib.BHT_ERR_STATE and not self._fitted:
            return None
        check(st)
        return r.value, t.value

    def kl_divergence(self):
        pair = self.kl_divergence_pair()
        return None if pair is None else pair[0]

    def affinities(self) -> Optional[SparseAffinities]:
        nnz = C.c_uint64(0)
        perp = C.c_float(0)
        st = self._L.bht_job_get_affinities(self._h, None, None, None, C.byref(nnz),
                                            C.byref(perp))
        if st == _lib.BHT_ERR_STATE:
            return None
        check(st)
        rows = np.zeros(self._n + 1, np.uint64)
        cols = np.zeros(nnz.value, np.uint32)
        vals = np.zeros(nnz.value, np.float32)
        check(self._L.bht_job_get_affinities(self._h, ptr(rows), ptr(cols), ptr(vals),
                                             C.byref(nnz), C.byref(perp)))
        return SparseAffinities(rows, cols, vals, perp.value)

    def with_affinities(self, aff: SparseAffinities):
        rows = np.ascontiguousarray(aff.rows, np.uint64)
        cols = np.ascontiguousarray(aff.columns, np.uint32)
        vals = np.ascontiguousarray(aff.values, np.float32)
        if vals.size != cols.size:
            raise InvalidInput(_lib.BHT_ERR_INVALID, f"affinities have {cols.size} columns but "
                                                     f"{vals.size} values")
        check(self._L.bht_job_set_affinities(self._h, ptr(rows), ptr(cols), ptr(vals), rows.size,
                                             cols.size, float(aff.perplexity)))
        return self

    # ---- stepping interface (bench, parity checks) --------------------------------------------
    def prepare(self, method: str = "fitsne", theta: float = 0.5):
        m = _lib.METHOD_FITSNE if method == "fitsne" else _lib.METHOD_BARNES_HUT
        check(self._L.bht_job_prepare(self._h, m, float(theta)))
        return self

    def step(self, n_epochs: int):
        check(self._L.bht_job_step(self._h, int(n_epochs), self._cb(), None))
        return self

    def finish(self):
        check(self._L.bht_job_finish(self._h))
        self._fitted = True
        return self

    def synchronize(self):
        check(self._L.bht_job_synchronize(self._h))

    def forces(self):
        len_ = self._n * self._d
        pos = np.zeros(len_, np.float32)
        neg = np.zeros(len_, np.float32)
        inv = C.c_float(0)
        check(self._L.bht_job_forces(self._h, ptr(pos), ptr(neg), C.byref(inv)))
        return pos.reshape(self._n, self._d), neg.reshape(self._n, self._d), inv.value

    def get_state(self):
        len_ = self._n * self._d
        y, uy, g = (np.zeros(len_, np.float32) for _ in range(3))
        check(self._L.bht_job_get_state(self._h, ptr(y), ptr(uy), ptr(g)))
        return y.reshape(self._n, self._d), uy.reshape(self._n, self._d), g.reshape(self._n,
                                                                                    self._d)

    def set_state(self, y, uy, gains):
        y, uy, gains = (np.ascontiguousarray(a, np.float32).ravel() for a in (y, uy, gains))
        check(self._L.bht_job_set_state(self._h, ptr(y), ptr(uy), ptr(gains)))
        return self

    def set_timing(self, enable: bool):
        check(self._L.bht_job_set_timing(self._h, 1 if enable else 0))
        return self

    def timing(self):
        """(repulsion ms, attraction ms, epochs) measured with HIP events while timing was
        enabled; the two phases run concurrently after each epoch's recentring."""
        out = (C.c_double * 3)()
        check(self._L.bht_job_get_timing(self._h, out))
        return float(out[0]), float(out[1]), int(out[2])

    def run_state(self):
        """(next epoch index of the prepared run, whether its P is still exaggerated)."""
        e = C.c_uint64(0)
        x = C.c_int(0)
        check(self._L.bht_job_run_state(self._h, C.byref(e), C.byref(x)))
        return int(e.value), bool(x.value)

    def epoch_form(self) -> int:
        """How the last run's epochs ran (the autotuner's choice): -1 in order on one stream, 0
        the repulsion beside an attraction with a block per row block, k > 0 beside an
        attraction on k blocks."""
        v = C.c_int32(0)
        check(self._L.bht_job_epoch_form(self._h, C.byref(v)))
        return int(v.value)

    def grid_info(self):
        b, L_ = C.c_uint32(), C.c_uint32()
        lo, hi = C.c_float(), C.c_float()
        check(self._L.bht_job_grid_info(self._h, C.byref(b), C.byref(L_), C.byref(lo),
                                        C.byref(hi)))
        return dict(boxes=b.value, fft_len=L_.value, min=lo.value, max=hi.value)


# ---- batched jobs ---------------------------------------------------------------------------------

class Batch:
    """Many independent TSNE jobs advanced together: every epoch kernel is one launch for all
    members of a method / d_out, replayed from a captured graph.  Each member's result equals its
    fit alone (``TSNE.fit_sne`` / ``barnes_hut``) bit for bit.  Jobs join with ``add``
    (which builds their affinities unless cached) and leave when their ``epochs`` are done."""

    def __init__(self, ctx: Optional[Context] = None, graph: bool = True,
                 concurrent: bool = False):
        self._ctx = ctx or default_context()
        self._L = _lib.load()
        h = C.c_void_p()
        check(self._L.bht_batch_create(self._ctx.handle, C.byref(h)))
        self._h = h
        self._members = {}  # handle value -> TSNE (kept alive while a member)
        _live.add(self)
        if not graph:
            check(self._L.bht_batch_set_graph(self._h, 0))
        if concurrent:
            check(self._L.bht_batch_set_concurrent(self._h, 1))

    def add(self, job: TSNE, method: str = "fitsne", theta: float = 0.5):
        m = _lib.METHOD_FITSNE if method == "fitsne" else _lib.METHOD_BARNES_HUT
        check(self._L.bht_batch_add(self._h, job._h, m, float(theta)))
        self._members[job._h.value] = job
        return self

    def __len__(self):
        return int(self._L.bht_batch_size(self._h))

    def step(self, max_epochs: int) -> int:
        k = C.c_uint64(0)
        check(self._L.bht_batch_step(self._h, int(max_epochs), C.byref(k)))
        return int(k.value)

    def take_finished(self) -> list:
        out = []
        buf = (C.c_void_p * 64)()
        while True:
            cnt = C.c_uint64(0)
            check(self._L.bht_batch_take_finished(self._h, buf, 64, C.byref(cnt)))
            for i in range(cnt.value):
                job = self._members.pop(buf[i])
                job._fitted = True
                out.append(job)
            if cnt.value < 64:
                return out

    def run(self, chunk: int = 1 << 30) -> list:
        """Steps until every member has finished; returns the jobs in completion order."""
        done = []
        while len(self):
            self.step(chunk)
            done += self.take_finished()
        return done + self.take_finished()

    def synchronize(self):
        check(self._L.bht_batch_synchronize(self._h))

    def close(self):
        if self._h:
            self._L.bht_batch_destroy(self._h)
            self._h = None
            self._members.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_footprint(n: int, d_in: int, d_out: int = 2, method: str = "fitsne", **params) -> int:
    """Upper bound of the device bytes of one job (HBM admission of a request loop)."""
    L = _lib.load()
    p = _lib.Params()
    L.bht_params_default(C.byref(p))
    p.d_out = int(d_out)
    for k, v in params.items():
        setattr(p, k, v)
    m = _lib.METHOD_FITSNE if method == "fitsne" else _lib.METHOD_BARNES_HUT
    out = C.c_uint64(0)
    check(L.bht_job_footprint(int(n), int(d_in), C.byref(p), m, C.byref(out)))
    return int(out.value)


# ---- operators (the reference's internal steps) -------------------------------------------------

def check_affinities(aff: SparseAffinities) -> None:
    """The host-side structure check ``with_affinities`` makes (rows[0] == 0, non-decreasing
    offsets, rows[n] == nnz, every column < n); raises ``InvalidInput``.  Needs no device."""
    L = _lib.load()
    rows = np.ascontiguousarray(aff.rows, np.uint64)
    cols = np.ascontiguousarray(aff.columns, np.uint32)
    vals = np.asarray(aff.values)
    if vals.size != cols.size:
        raise InvalidInput(_lib.BHT_ERR_INVALID, f"affinities have {cols.size} columns but "
                                                 f"{vals.size} values")
    check(L.bht_check_affinities(ptr(rows), ptr(cols), rows.size, cols.size))


def knn(x, k: int, ctx: Optional[Context] = None, mode: str = "auto", stats: bool = False):
    """Exact kNN, self excluded, ascending (distance, index) (src/tsne/vptree.rs:274-319).
    mode "auto": matrix-core candidates, exact re-rank, per-row certificate (uncertified rows
    recomputed exactly); "exact": the exact kernel for every row.  Both give the same result;
    with stats=True the number of rows the exact kernel computed is returned as well."""
    ctx = ctx or default_context()
    L = _lib.load()
    x = np.ascontiguousarray(x, np.float32)
    n, d = x.shape
    idx = np.zeros((n, k), np.uint32)
    dist = np.zeros((n, k), np.float32)
    rows = C.c_uint64(0)
    m = _lib.KNN_EXACT if mode == "exact" else _lib.KNN_AUTO
    check(L.bht_knn_f32_ex(ctx.handle, ptr(x), n, d, k, m, ptr(idx), ptr(dist), C.byref(rows)))
    return (idx, dist, int(rows.value)) if stats else (idx, dist)


def exact_repulsion_f64(y, ctx: Optional[Context] = None):
    """FP64 brute-force non-edge gradient terms and Z (the exact path, src/lib.rs:570-622)."""
    ctx = ctx or default_context()
    L = _lib.load()
    y = np.ascontiguousarray(y, np.float32)
    if y.ndim == 1:
        y = y.reshape(-1, 1)
    n, d = y.shape
    neg = np.zeros((n, d), np.float64)
    z = C.c_double(0.0)
    check(L.bht_exact_repulsion_f64(ctx.handle, ptr(y), n, d, ptr(neg), C.byref(z)))
    return neg, z.value


def search_beta(dist, perplexity: float, mode: int = _lib.BANDWIDTH_FP64,
                ctx: Optional[Context] = None):
    """Per-row bandwidth search (src/tsne/mod.rs:201-275)."""
    ctx = ctx or default_context()
    L = _lib.load()
    dist = np.ascontiguousarray(dist, np.float32)
    n, k = dist.shape
    p = np.zeros_like(dist)
    check(L.bht_search_beta_f32(ctx.handle, ptr(dist), n, k, float(perplexity), int(mode),
                                ptr(p)))
    return p


def symmetrize(cols, vals, ctx: Optional[Context] = None):
    """Symmetrisation into CSR in the reference's row order (src/tsne/mod.rs:306-406)."""
    ctx = ctx or default_context()
    L = _lib.load()
    cols = np.ascontiguousarray(cols, np.uint32)
    vals = np.ascontiguousarray(vals, np.float32)
    n, k = cols.shape
    nnz = C.c_uint64(0)
    check(L.bht_symmetrize(ctx.handle, ptr(cols), ptr(vals), n, k, None, None, None,
                           C.byref(nnz)))
    rows = np.zeros(n + 1, np.uint64)
    oc = np.zeros(nnz.value, np.uint32)
    ov = np.zeros(nnz.value, np.float32)
    check(L.bht_symmetrize(ctx.handle, ptr(cols), ptr(vals), n, k, ptr(rows), ptr(oc), ptr(ov),
                           C.byref(nnz)))
    return rows, oc, ov
