"""ctypes binding of libbhtsne_hip.so (the C ABI declared in include/bhtsne_hip.h).

Loading fails loudly when the HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbhtsne_hip.so")

BHT_OK = 0
BHT_ERR_INVALID = 1
BHT_ERR_HIP = 2
BHT_ERR_OOM = 3
BHT_ERR_STATE = 4

METHOD_BARNES_HUT = 0
METHOD_FITSNE = 1
KNN_AUTO = 0
KNN_EXACT = 1
METRIC_EUCLIDEAN = 0
METRIC_SQEUCLIDEAN = 1
BANDWIDTH_FP64 = 0
BANDWIDTH_F32_SEQ = 1


class Params(C.Structure):
    _fields_ = [
        ("d_out", C.c_uint32),
        ("bandwidth_mode", C.c_uint32),
        ("epochs", C.c_uint64),
        ("momentum_switch_epoch", C.c_uint64),
        ("stop_lying_epoch", C.c_uint64),
        ("learning_rate", C.c_float),
        ("momentum", C.c_float),
        ("final_momentum", C.c_float),
        ("early_exaggeration", C.c_float),
        ("perplexity", C.c_float),
        ("min_boxes", C.c_uint32),
        ("max_boxes", C.c_uint32),
        ("flags", C.c_uint32),
    ]


EPOCH_CB = C.CFUNCTYPE(None, C.c_uint64, C.POINTER(C.c_float), C.c_uint64, C.c_void_p)

# exported symbol -> (restype, argtypes); the set include/bhtsne_hip.h declares
SIGNATURES = {
    "bht_params_default": (None, [C.POINTER(Params)]),
    "bht_ctx_create": (C.c_int32, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "bht_ctx_destroy": (None, [C.c_void_p]),
    "bht_last_error": (C.c_char_p, []),
    "bht_ctx_bytes_in_use": (C.c_uint64, [C.c_void_p]),
    "bht_job_create": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(Params),
                                   C.POINTER(C.c_void_p)]),
    "bht_job_destroy": (None, [C.c_void_p]),
    "bht_job_set_params": (C.c_int32, [C.c_void_p, C.POINTER(Params)]),
    "bht_job_get_params": (C.c_int32, [C.c_void_p, C.POINTER(Params)]),
    "bht_job_set_data_f32": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32]),
    "bht_job_set_initial_embedding": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "bht_job_barnes_hut": (C.c_int32, [C.c_void_p, C.c_float, EPOCH_CB, C.c_void_p]),
    "bht_job_barnes_hut_with_neighbors": (C.c_int32, [C.c_void_p, C.c_float, C.c_void_p,
                                                      C.c_void_p, C.c_void_p, C.c_uint64,
                                                      EPOCH_CB, C.c_void_p]),
    "bht_job_fit_sne": (C.c_int32, [C.c_void_p, EPOCH_CB, C.c_void_p]),
    "bht_job_exact": (C.c_int32, [C.c_void_p, C.c_uint32, EPOCH_CB, C.c_void_p]),
    "bht_job_fit_sne_with_neighbors": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p,
                                                   C.c_void_p, C.c_uint64, EPOCH_CB,
                                                   C.c_void_p]),
    "bht_job_get_embedding": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "bht_job_kl_divergence": (C.c_int32, [C.c_void_p, C.POINTER(C.c_double),
                                          C.POINTER(C.c_double)]),
    "bht_job_get_affinities": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "bht_job_set_affinities": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_uint64, C.c_uint64, C.c_float]),
    "bht_check_affinities": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64]),
    "bht_job_prepare": (C.c_int32, [C.c_void_p, C.c_uint32, C.c_float]),
    "bht_job_step": (C.c_int32, [C.c_void_p, C.c_uint64, EPOCH_CB, C.c_void_p]),
    "bht_job_finish": (C.c_int32, [C.c_void_p]),
    "bht_job_synchronize": (C.c_int32, [C.c_void_p]),
    "bht_job_forces": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "bht_job_get_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bht_job_set_state": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "bht_job_grid_info": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "bht_job_set_timing": (C.c_int32, [C.c_void_p, C.c_int]),
    "bht_job_get_timing": (C.c_int32, [C.c_void_p, C.POINTER(C.c_double)]),
    "bht_knn_f32": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                C.c_void_p, C.c_void_p]),
    "bht_search_beta_f32": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                        C.c_float, C.c_uint32, C.c_void_p]),
    "bht_symmetrize": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "bht_knn_f32_ex": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                   C.c_uint32, C.c_void_p, C.c_void_p, C.POINTER(C.c_uint64)]),
    "bht_job_run_state": (C.c_int32, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    "bht_job_epoch_form": (C.c_int32, [C.c_void_p, C.POINTER(C.c_int32)]),
    "bht_exact_repulsion_f64": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32,
                                            C.c_void_p, C.POINTER(C.c_double)]),
    "bht_exact_repulsion_f64_noalloc": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint64,
                                                    C.c_uint32, C.c_void_p, C.c_void_p,
                                                    C.c_void_p]),
    "bht_batch_create": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "bht_batch_destroy": (None, [C.c_void_p]),
    "bht_batch_add": (C.c_int32, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_float]),
    "bht_batch_step": (C.c_int32, [C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]),
    "bht_batch_take_finished": (C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint64,
                                            C.POINTER(C.c_uint64)]),
    "bht_batch_size": (C.c_uint64, [C.c_void_p]),
    "bht_batch_set_graph": (C.c_int32, [C.c_void_p, C.c_int]),
    "bht_batch_set_concurrent": (C.c_int32, [C.c_void_p, C.c_int]),
    "bht_batch_synchronize": (C.c_int32, [C.c_void_p]),
    "bht_job_footprint": (C.c_int32, [C.c_uint64, C.c_uint32, C.POINTER(Params), C.c_uint32,
                                      C.POINTER(C.c_uint64)]),
}

_lib = None


class BhtError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


class InvalidInput(BhtError, ValueError):
    """The reference would panic with this message."""


def load():
    """Load the HIP library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m bhtsne_amd.build` "
            "(the engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status == BHT_OK:
        return
    msg = load().bht_last_error().decode(errors="replace")
    if status == BHT_ERR_INVALID:
        raise InvalidInput(status, msg)
    raise BhtError(status, msg)


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


NULL_CB = EPOCH_CB()
