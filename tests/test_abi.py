"""CPU-side checks of the C-ABI library: it loads and exports every function the header declares
(no compute call is made without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bhtsne_hip.h")
LIB = os.path.join(ROOT, "bhtsne_amd", "libbhtsne_hip.so")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bht_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        import bhtsne_amd.build as b
        b.build()
    return ctypes.CDLL(LIB)


def test_header_declares_the_builder_surface():
    names = _declared()
    for required in ["bht_job_barnes_hut", "bht_job_barnes_hut_with_neighbors",
                     "bht_job_fit_sne", "bht_job_fit_sne_with_neighbors", "bht_job_get_embedding",
                     "bht_job_kl_divergence", "bht_job_get_affinities", "bht_job_set_affinities",
                     "bht_job_set_initial_embedding", "bht_knn_f32", "bht_search_beta_f32",
                     "bht_symmetrize"]:
        assert required in names


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_signatures_cover_the_header():
    from bhtsne_amd import _lib
    assert set(_declared()) == set(_lib.SIGNATURES)


def _header_prototypes():
    """name -> number of parameters of every function the header declares."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(bht_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def _header_constants():
    text = open(HEADER).read()
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define\s+(BHT_[A-Z0-9_]+)\s+(\d+)u?\b", text)}


def test_rust_sys_binding_covers_the_header():
    """The Rust -sys crate (not compiled here: no cargo) declares every function of the header
    with the same number of parameters, and every constant with the header's value."""
    src = open(os.path.join(ROOT, "rust", "bhtsne-hip-sys", "src", "lib.rs")).read()
    src = re.sub(r"//[^\n]*", "", src)
    fns = {}
    for m in re.finditer(r"pub fn\s+(bht_[a-z0-9_]+)\s*\(([^)]*)\)", src):
        args = m.group(2).strip().rstrip(",")
        fns[m.group(1)] = 0 if not args else args.count(",") + 1
    assert fns == _header_prototypes()
    consts = {m.group(1): int(m.group(2))
              for m in re.finditer(r"pub const\s+(BHT_[A-Z0-9_]+)\s*:\s*\w+\s*=\s*(\d+)", src)}
    assert consts == _header_constants()
    # bht_params: the same fields, types and order
    ctypes_of = {"uint32_t": "u32", "uint64_t": "u64", "float": "f32"}
    hdr = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = re.search(r"typedef struct bht_params \{(.*?)\} bht_params;", hdr, re.S).group(1)
    c_fields = [(n, ctypes_of[t]) for t, n in re.findall(r"(\w+)\s+(\w+);", body)]
    rs_body = re.search(r"pub struct bht_params \{(.*?)\}", src, re.S).group(1)
    rs_fields = re.findall(r"pub (\w+):\s*(\w+),", rs_body)
    assert rs_fields == c_fields


def test_params_default_is_the_reference_builder_default(lib):
    from bhtsne_amd import _lib
    L = _lib.load()
    p = _lib.Params()
    L.bht_params_default(ctypes.byref(p))
    # src/lib.rs:241-266
    assert (p.epochs, p.momentum, p.final_momentum, p.momentum_switch_epoch,
            p.stop_lying_epoch, p.early_exaggeration, p.perplexity, p.learning_rate) == \
        (1000, 0.5, pytest.approx(0.8), 250, 250, 12.0, 20.0, 0.0)
    assert p.d_out == 2 and p.min_boxes == 10 and p.max_boxes == 1000


def test_ctx_create_without_gpu_reports_an_error(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from bhtsne_amd import _lib
    L = _lib.load()
    h = ctypes.c_void_p()
    st = L.bht_ctx_create(0, 0, ctypes.byref(h))
    assert st != 0 and L.bht_last_error()


def _csr():
    import numpy as np
    rows = np.array([0, 2, 3, 5], np.uint64)
    cols = np.array([1, 2, 0, 0, 1], np.uint32)
    return rows, cols


def test_check_affinities_accepts_a_valid_graph(lib):
    import bhtsne_amd as bt
    import numpy as np
    rows, cols = _csr()
    bt.check_affinities(bt.SparseAffinities(rows, cols, np.ones(5, np.float32), 3.0))


@pytest.mark.parametrize("case", ["empty_rows", "first_offset", "decreasing", "last_offset",
                                  "column_range", "value_count"])
def test_check_affinities_rejects_malformed_graphs(lib, case):
    # the checks with_affinities makes on the host before any upload (no device involved)
    import bhtsne_amd as bt
    import numpy as np
    rows, cols = _csr()
    vals = np.ones(5, np.float32)
    if case == "empty_rows":
        rows = np.zeros(0, np.uint64)
        cols = cols[:0]
        vals = vals[:0]
    elif case == "first_offset":
        rows = rows.copy()
        rows[0] = 1
    elif case == "decreasing":
        rows = np.array([0, 3, 2, 5], np.uint64)
    elif case == "last_offset":
        rows = np.array([0, 2, 3, 4], np.uint64)
    elif case == "column_range":
        cols = cols.copy()
        cols[3] = 3
    else:
        vals = vals[:4]
    with pytest.raises(bt.InvalidInput):
        bt.check_affinities(bt.SparseAffinities(rows, cols, vals, 3.0))
