"""CPU oracle for the bhtsne hot path -- TEST INFRASTRUCTURE ONLY.

A ctypes front-end over ``oracle/libbhtsne_oracle.so`` (built from ``oracle/bhtsne_oracle.c`` by
``oracle/Makefile``), a restatement of the reference crate's f32 arithmetic.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import this package,
and only as the checker or the timed CPU baseline, never as the product path.

Each wrapper names the reference lines it restates (paths relative to the reference crate root).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libbhtsne_oracle.so")
_lib = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    sig = {
        "oracle_set_threads": (C.c_int, [C.c_int]),
        "oracle_knn": (None, [f32p, C.c_int64, C.c_int, C.c_int, u32p, f32p]),
        "oracle_knn_rows": (None, [f32p, C.c_int64, C.c_int, C.c_int, i64p, C.c_int64, u32p,
                                   f32p]),
        "oracle_search_beta_row": (C.c_int, [f32p, f32p, C.c_int, C.c_float]),
        "oracle_search_beta_row_f64": (C.c_int, [f64p, f64p, C.c_int, C.c_double]),
        "oracle_search_beta": (None, [f32p, f32p, C.c_int64, C.c_int, C.c_float]),
        "oracle_symmetrize": (C.c_int64, [u32p, f32p, C.c_int64, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
        "oracle_normalize_p": (C.c_float, [f32p, C.c_int64, C.c_float]),
        "oracle_stop_lying": (None, [f32p, C.c_int64, C.c_float]),
        "oracle_auto_learning_rate": (C.c_float, [C.c_int64, C.c_float]),
        "oracle_edge_forces": (None, [f32p, C.c_int64, C.c_int, u64p, u32p, f32p, f32p]),
        "oracle_gradient_step": (None, [f32p, f32p, f32p, f32p, f32p, C.c_int64, C.c_float,
                                        C.c_float, C.c_float]),
        "oracle_zero_mean": (None, [f32p, C.c_int64, C.c_int, f32p]),
        "oracle_fft2_r2c_f64": (None, [f64p, f64p, C.c_int]),
        "oracle_fft2_c2r_f64": (None, [f64p, f64p, C.c_int]),
        "oracle_next_smooth": (C.c_int64, [C.c_int64]),
        "oracle_box_count": (C.c_int, [C.c_float, C.c_int]),
        "oracle_lagrange_weights": (None, [C.c_float, f32p]),
        "oracle_fitsne_repulsion": (None, [f32p, C.c_int64, C.c_int, C.c_int, f32p,
                                           C.POINTER(C.c_float), f64p]),
        "oracle_set_fft_f32": (None, [C.c_int]),
        "oracle_set_cr_explog": (None, [C.c_int]),
        "oracle_fitsne_repulsion_f64": (None, [f32p, C.c_int64, C.c_int, C.c_int, f64p,
                                               C.POINTER(C.c_double)]),
        "oracle_morton_encode2": (C.c_uint64, [C.c_uint32, C.c_uint32]),
        "oracle_quantize": (C.c_uint32, [C.c_float, C.c_float, C.c_float, C.c_uint32]),
        "oracle_morton_encode": (None, [u32p, C.c_int, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
        "oracle_morton_bits": (C.c_int, [C.c_int]),
        "oracle_bh_build": (C.c_int64, [f32p, C.c_int64, f32p, u32p, u32p, u8p, u8p, f32p]),
        "oracle_bh_build_d": (C.c_int64, [f32p, C.c_int64, C.c_int, f32p, u32p, u32p, u8p, u8p,
                                          f32p]),
        "oracle_bh_build_ranges": (C.c_int64, [f32p, C.c_int64, C.c_int, f32p, u32p, u32p, u8p,
                                               u8p, f32p, u32p, u32p]),
        "oracle_bh_repulsion": (C.c_float, [f32p, C.c_int64, C.c_float, f32p, f32p]),
        "oracle_bh_repulsion_d": (C.c_float, [f32p, C.c_int64, C.c_int, C.c_float, f32p, f32p]),
        "oracle_forces": (C.c_float, [f32p, C.c_int64, C.c_int, u64p, u32p, f32p, C.c_int,
                                      C.c_float, C.c_int, f32p, f32p]),
        "oracle_run_epochs": (None, [f32p, f32p, f32p, C.c_int64, C.c_int, u64p, u32p, f32p,
                                     C.c_int, C.c_float, C.c_int, f32p, i64p, C.c_int64,
                                     C.c_int64, C.POINTER(C.c_float)]),
        "oracle_sparse_kl": (None, [f32p, C.c_int64, C.c_int, u64p, u32p, f32p, C.c_float,
                                    C.POINTER(C.c_float), C.POINTER(C.c_double)]),
        "oracle_kl_divergence": (None, [f32p, C.c_int64, C.c_int, u64p, u32p, f32p, C.c_int,
                                        C.c_float, C.c_int, C.POINTER(C.c_float),
                                        C.POINTER(C.c_double)]),
        "oracle_exact_repulsion_f64": (C.c_double, [f32p, C.c_int64, C.c_int, f64p]),
        "oracle_exact_repulsion_rows_f64": (C.c_double, [f32p, C.c_int64, C.c_int, i64p,
                                                         C.c_int64, f64p]),
        "oracle_exact_attraction_f64": (None, [f32p, C.c_int64, C.c_int, u64p, u32p, f32p,
                                               f64p]),
        "oracle_exact_distances": (None, [f32p, C.c_int64, C.c_int, C.c_int, f32p]),
        "oracle_exact_affinities": (None, [f32p, C.c_int64, C.c_int, C.c_int, C.c_float,
                                           C.c_float, C.c_int, f32p]),
        "oracle_exact_gradient": (C.c_float, [f32p, C.c_int64, C.c_int, f32p, f32p]),
        "oracle_exact_epochs": (None, [f32p, f32p, f32p, C.c_int64, C.c_int, f32p, f32p, i64p,
                                       C.c_int64, C.c_int64, C.POINTER(C.c_float)]),
        "oracle_exact_kl": (C.c_float, [f32p, C.c_int64, C.c_int, f32p, C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


# --- affinities ------------------------------------------------------------------------------

def knn(x: np.ndarray, k: int):
    """Brute-force kNN with the reference metric (src/test.rs:856-881), (dist, idx) order."""
    x = _f32(x)
    n, d = x.shape
    idx = np.zeros((n, k), np.uint32)
    dist = np.zeros((n, k), np.float32)
    lib().oracle_knn(x, n, d, k, idx, dist)
    return idx, dist


def knn_rows(x: np.ndarray, k: int, sel):
    """Brute-force kNN of the rows `sel` only (same order and arithmetic as knn)."""
    x = _f32(x)
    n, d = x.shape
    sel = np.ascontiguousarray(sel, np.int64)
    idx = np.zeros((len(sel), k), np.uint32)
    dist = np.zeros((len(sel), k), np.float32)
    lib().oracle_knn_rows(x, n, d, k, sel, len(sel), idx, dist)
    return idx, dist


def search_beta(dist: np.ndarray, perplexity: float) -> np.ndarray:
    """Per-row bandwidth search in f32 (src/tsne/mod.rs:201-275)."""
    dist = _f32(dist)
    n, k = dist.shape
    p = np.zeros((n, k), np.float32)
    lib().oracle_search_beta(p, dist, n, k, float(perplexity))
    return p


def search_beta_row(dist_row, perplexity, f64=False):
    if f64:
        d = np.ascontiguousarray(dist_row, np.float64)
        p = np.zeros_like(d)
        it = lib().oracle_search_beta_row_f64(p, d, len(d), float(perplexity))
    else:
        d = _f32(dist_row)
        p = np.zeros_like(d)
        it = lib().oracle_search_beta_row(p, d, len(d), float(perplexity))
    return p, it


def symmetrize(cols: np.ndarray, vals: np.ndarray):
    """Sparse symmetrisation into CSR (src/tsne/mod.rs:306-406). Returns rows(u64), cols, vals."""
    cols = np.ascontiguousarray(cols, np.uint32)
    vals = _f32(vals)
    n, k = cols.shape
    L = lib()
    nnz = L.oracle_symmetrize(cols, vals, n, k, None, None, None)
    rows = np.zeros(n + 1, np.uint64)
    oc = np.zeros(nnz, np.uint32)
    ov = np.zeros(nnz, np.float32)
    L.oracle_symmetrize(cols, vals, n, k, rows.ctypes.data, oc.ctypes.data, ov.ctypes.data)
    return rows, oc, ov


def normalize_p(vals: np.ndarray, exag: float):
    """P <- P * exag / (sum P + eps) in place (src/tsne/mod.rs:284-293); returns the scale."""
    return lib().oracle_normalize_p(vals, vals.size, float(exag))


def stop_lying(vals: np.ndarray, exag: float):
    lib().oracle_stop_lying(vals, vals.size, float(exag))


def auto_learning_rate(n: int, exag: float) -> float:
    return lib().oracle_auto_learning_rate(int(n), float(exag))


# --- per-epoch -----------------------------------------------------------------------------

def edge_forces(y, rows, cols, vals):
    y = _f32(y)
    n, D = y.shape
    pos = np.zeros_like(y)
    lib().oracle_edge_forces(y, n, D, rows, cols, vals, pos)
    return pos


def gradient_step(y, pos, neg, uy, gains, lr, momentum, inverse_norm):
    """Fused update (src/tsne/mod.rs:470-516), in place on y, uy, gains."""
    lib().oracle_gradient_step(y, _f32(pos), _f32(neg), uy, gains, y.size, float(lr),
                               float(momentum), float(inverse_norm))


def zero_mean(y):
    n, D = y.shape
    m = np.zeros(8, np.float32)
    lib().oracle_zero_mean(y, n, D, m)
    return m[:D].copy()


def fitsne_repulsion(y, min_boxes=10):
    """Interpolated repulsion and Z (src/tsne/interpolation.rs:100-134)."""
    y = _f32(y)
    n, D = y.shape
    neg = np.zeros_like(y)
    z = C.c_float(0.0)
    info = np.zeros(4, np.float64)
    lib().oracle_fitsne_repulsion(y, n, D, int(min_boxes), neg, C.byref(z), info)
    return neg, z.value, dict(boxes=int(info[0]), fft_len=int(info[1]), min=info[2], max=info[3])


class cr_explog:
    """Context manager: the bandwidth search's exp / ln correctly rounded (as the device computes
    them) instead of the host libm's expf / logf (as the reference calls them)."""

    def __enter__(self):
        lib().oracle_set_cr_explog(1)

    def __exit__(self, *exc):
        lib().oracle_set_cr_explog(0)


class f32_fft:
    """Context manager: the f32 pipeline's transforms run in f32 arithmetic (a representative
    single-precision FFT, as the reference's rustfft computes) instead of exact double."""

    def __enter__(self):
        lib().oracle_set_fft_f32(1)

    def __exit__(self, *exc):
        lib().oracle_set_fft_f32(0)


def fitsne_repulsion_f64(y, min_boxes=10):
    """The interpolated repulsion and Z in T = f64 throughout (same grid plan as the f32 path)."""
    y = _f32(y)
    n, D = y.shape
    neg = np.zeros((n, D), np.float64)
    z = C.c_double(0.0)
    lib().oracle_fitsne_repulsion_f64(y, n, D, int(min_boxes), neg, C.byref(z))
    return neg, z.value


def next_smooth(n):
    return lib().oracle_next_smooth(int(n))


def box_count(extent, min_boxes=10):
    return lib().oracle_box_count(float(extent), int(min_boxes))


def lagrange_weights(local):
    out = np.zeros(3, np.float32)
    lib().oracle_lagrange_weights(float(local), out)
    return out


def fft2_r2c_f64(grid):
    grid = np.ascontiguousarray(grid, np.float64)
    m = grid.shape[0]
    spec = np.zeros((m, m // 2 + 1, 2), np.float64)
    lib().oracle_fft2_r2c_f64(grid, spec, m)
    return spec[..., 0] + 1j * spec[..., 1]


def fft2_c2r_f64(spec):
    m = spec.shape[0]
    s = np.ascontiguousarray(np.stack([spec.real, spec.imag], -1), np.float64)
    grid = np.zeros((m, m), np.float64)
    lib().oracle_fft2_c2r_f64(s, grid, m)
    return grid


def morton_encode2(c0, c1):
    return lib().oracle_morton_encode2(int(c0), int(c1))


def quantize(p, mn, inv_scale, max_bucket=0xFFFFFFFF):
    return lib().oracle_quantize(float(p), float(mn), float(inv_scale), int(max_bucket))


def morton_encode(coords):
    """Z-order code of D quantised coordinates (D = 2..7, morton/d2.rs .. d7.rs) as an int."""
    c = np.ascontiguousarray(coords, np.uint32)
    lo, hi = C.c_uint64(0), C.c_uint64(0)
    lib().oracle_morton_encode(c, int(c.size), C.byref(lo), C.byref(hi))
    return (hi.value << 64) | lo.value


def morton_bits(D):
    """Bits per axis of the D-dimensional Morton code (0 for an unsupported D)."""
    return lib().oracle_morton_bits(int(D))


def bh_build(y):
    """Morton arena for D = y.shape[1] in 2..7 (src/tsne/arena.rs:137-328). Returns a dict of
    node arrays."""
    y = _f32(y)
    n, D = y.shape
    cap = max(2 * n - 1, 1)
    com = np.zeros((cap, D), np.float32)
    cnt = np.zeros(cap, np.uint32)
    fc = np.zeros(cap, np.uint32)
    cc = np.zeros(cap, np.uint8)
    lv = np.zeros(cap, np.uint8)
    hw2 = np.zeros(morton_bits(D) + 1, np.float32)
    ranges = np.zeros((cap, 2), np.uint32)
    order = np.zeros(max(n, 1), np.uint32)
    nn = lib().oracle_bh_build_ranges(y, n, D, com, cnt, fc, cc, lv, hw2, ranges, order)
    if nn < 0:
        raise ValueError(f"Barnes-Hut arena: D = {D} is not in 2..7")
    return dict(com=com[:nn], count=cnt[:nn], first_child=fc[:nn], child_count=cc[:nn],
                level=lv[:nn], hw2=hw2, ranges=ranges[:nn], order=order[:n])


def bh_repulsion(y, theta):
    """Barnes-Hut non-edge forces and q sums (arena.rs:338-392, repulsion.rs:95-158), D = 2..7."""
    y = _f32(y)
    n, D = y.shape
    if morton_bits(D) == 0:
        raise ValueError(f"Barnes-Hut repulsion: D = {D} is not in 2..7")
    neg = np.zeros_like(y)
    q = np.zeros(n, np.float32)
    s = lib().oracle_bh_repulsion_d(y, n, D, float(theta), neg, q)
    return neg, q, s


def forces(y, rows, cols, vals, method="fitsne", theta=0.5, min_boxes=10):
    """Attractive + repulsive forces and 1/Z for one state (repulsion.rs:95-158, 206-234)."""
    y = _f32(y)
    n, D = y.shape
    pos = np.zeros_like(y)
    neg = np.zeros_like(y)
    inv = lib().oracle_forces(y, n, D, rows, cols, _f32(vals), 0 if method == "fitsne" else 1,
                              float(theta), int(min_boxes), pos, neg)
    return pos, neg, inv


def run_epochs(y, uy, gains, rows, cols, vals, *, method="fitsne", theta=0.5, min_boxes=10,
               learning_rate, momentum=0.5, final_momentum=0.8, early_exaggeration=12.0,
               momentum_switch_epoch=250, stop_lying_epoch=250, e0=0, e1=1):
    """Epochs [e0, e1) of the approximate run loop (src/lib.rs:961-1014, 1162-1179), in place.
    Returns the momentum after the last epoch."""
    n, D = y.shape
    sched = np.array([learning_rate, momentum, final_momentum, early_exaggeration], np.float32)
    isched = np.array([momentum_switch_epoch, stop_lying_epoch], np.int64)
    mom = C.c_float(momentum)
    lib().oracle_run_epochs(y, uy, gains, n, D, rows, cols, vals, 0 if method == "fitsne" else 1,
                            float(theta), int(min_boxes), sched, isched, int(e0), int(e1),
                            C.byref(mom))
    return mom.value


def sparse_kl(y, rows, cols, vals, inverse_z):
    """sparse_kl_divergence (mod.rs:738-775) with a given 1/Z: (reference-compatible KL in f32,
    true KL in double)."""
    y = _f32(y)
    n, D = y.shape
    r = C.c_float(0.0)
    t = C.c_double(0.0)
    lib().oracle_sparse_kl(y, n, D, np.ascontiguousarray(rows, np.uint64),
                           np.ascontiguousarray(cols, np.uint32), _f32(vals), float(inverse_z),
                           C.byref(r), C.byref(t))
    return r.value, t.value


def kl_divergence(y, rows, cols, vals, method="fitsne", theta=0.5, min_boxes=10):
    """(reference-compatible KL, true KL) as kl_divergence() (lib.rs:448-477, mod.rs:635-775)."""
    y = _f32(y)
    n, D = y.shape
    r = C.c_float(0.0)
    t = C.c_double(0.0)
    lib().oracle_kl_divergence(y, n, D, rows, cols, _f32(vals), 0 if method == "fitsne" else 1,
                               float(theta), int(min_boxes), C.byref(r), C.byref(t))
    return r.value, t.value


def exact_repulsion_f64(y):
    y = _f32(y)
    n, D = y.shape
    neg = np.zeros((n, D), np.float64)
    z = lib().oracle_exact_repulsion_f64(y, n, D, neg)
    return neg, z


def exact_repulsion_rows_f64(y, sel):
    y = _f32(y)
    n, D = y.shape
    sel = np.ascontiguousarray(sel, np.int64)
    neg = np.zeros((len(sel), D), np.float64)
    z = lib().oracle_exact_repulsion_rows_f64(y, n, D, sel, len(sel), neg)
    return neg, z


def exact_attraction_f64(y, rows, cols, vals):
    y = _f32(y)
    n, D = y.shape
    pos = np.zeros((n, D), np.float64)
    lib().oracle_exact_attraction_f64(y, n, D, rows, cols, _f32(vals), pos)
    return pos


def build_affinities(x, perplexity):
    """kNN -> bandwidth search -> symmetrisation (src/lib.rs:1061-1127)."""
    k = int(np.float32(3.0) * np.float32(perplexity))
    idx, dist = knn(x, k)
    p = search_beta(dist, perplexity)
    return symmetrize(idx, p)


# --- the exact fit (src/lib.rs:489-646), with the reference's own semantics ----------------------

def exact_distances(x, metric="euclidean"):
    x = _f32(x)
    n, d = x.shape
    out = np.zeros((n, n), np.float32)
    lib().oracle_exact_distances(x, n, d, 0 if metric == "euclidean" else 1, out)
    return out


def exact_affinities(x, perplexity, metric="euclidean", exag=12.0, stop_lying_epoch=250):
    """Dense P as the exact fit builds it: normalised and exaggerated (divided back at once when
    stop_lying_epoch is 0)."""
    x = _f32(x)
    n, d = x.shape
    P = np.zeros((n, n), np.float32)
    lib().oracle_exact_affinities(x, n, d, 0 if metric == "euclidean" else 1, float(perplexity),
                                  float(exag), 1 if stop_lying_epoch == 0 else 0, P)
    return P


def exact_gradient(y, P):
    """(dy, 1/Z) of the exact path for one state."""
    y = _f32(y)
    n, D = y.shape
    dy = np.zeros_like(y)
    inv = lib().oracle_exact_gradient(y, n, D, _f32(P), dy)
    return dy, inv


def exact_epochs(y, uy, gains, P, *, learning_rate, momentum=0.5, final_momentum=0.8,
                 early_exaggeration=12.0, momentum_switch_epoch=250, stop_lying_epoch=250, e0=0,
                 e1=1):
    """Epochs [e0, e1) of the exact loop, in place (P is rescaled at stop-lying)."""
    n, D = y.shape
    sched = np.array([learning_rate, momentum, final_momentum, early_exaggeration], np.float32)
    isched = np.array([momentum_switch_epoch, stop_lying_epoch], np.int64)
    mom = C.c_float(momentum)
    lib().oracle_exact_epochs(y, uy, gains, n, D, P, sched, isched, int(e0), int(e1),
                              C.byref(mom))
    return mom.value


def exact_kl(y, P):
    """(reference formula in f32, the same sum in double) of evaluate_error (mod.rs:579-618)."""
    y = _f32(y)
    n, D = y.shape
    t = C.c_double(0.0)
    r = lib().oracle_exact_kl(y, n, D, _f32(P), C.byref(t))
    return float(r), t.value
