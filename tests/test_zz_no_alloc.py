"""The allocation-free variant of the FP64 exact repulsion (bht_exact_repulsion_f64_noalloc),
called through the C ABI on device buffers the caller owns, against the allocating entry point
(bht_exact_repulsion_f64): equal bit for bit on two calls in a row, the second smaller (the shared
device scratch keeps the first call's partials, which the second must not pick up), and the size
limit rejected with a message."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


class Hip:
    """The caller's side: device buffers and a stream from the HIP runtime the library itself
    links (a second HIP runtime in the process, such as the one torch bundles, would not see the
    device once the library's is up)."""

    def __init__(self):
        import bhtsne_amd  # noqa: F401  (loads libbhtsne_hip.so and its HIP runtime)
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.h.hipStreamSynchronize.argtypes = [C.c_void_p]
        self.h.hipStreamDestroy.argtypes = [C.c_void_p]

    def ok(self, e):
        assert e == 0, f"HIP error {e}"

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.h.hipMalloc(C.byref(p), nbytes))
        return p

    def upload(self, a):
        p = self.malloc(a.nbytes)
        self.ok(self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D))
        return p

    def download(self, p, out):
        self.ok(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H))
        return out


@pytest.fixture(scope="module")
def env():
    import bhtsne_amd as bt
    from bhtsne_amd import _lib
    return Hip(), bt, _lib.load()


def _noalloc(hip, L, ctx, y, stream):
    from bhtsne_amd._lib import check
    n, d = y.shape
    y_dev = hip.upload(y)
    neg_dev = hip.malloc(n * d * 8)
    z_dev = hip.upload(np.array([-1.0]))
    check(L.bht_exact_repulsion_f64_noalloc(ctx.handle, y_dev, n, d, neg_dev, z_dev, stream))
    hip.ok(hip.h.hipStreamSynchronize(stream))
    neg = hip.download(neg_dev, np.zeros((n, d), np.float64))
    z = hip.download(z_dev, np.zeros(1, np.float64))[0]
    for p in (y_dev, neg_dev, z_dev):
        hip.ok(hip.h.hipFree(p))
    return neg, float(z)


def test_noalloc_equals_allocating_entry_point(env):
    hip, bt, L = env
    ctx = bt.Context(0)
    stream = C.c_void_p()
    hip.ok(hip.h.hipStreamCreate(C.byref(stream)))
    rng = np.random.default_rng(11)
    for n, d in ((5000, 2), (1234, 1), (700, 2)):  # each call smaller than the one before
        y = (rng.standard_normal((n, d)) * 5).astype(np.float32)
        print(f"no-alloc exact repulsion n={n} d={d}", flush=True)
        neg, z = _noalloc(hip, L, ctx, y, stream)
        rneg, rz = bt.exact_repulsion_f64(y, ctx)
        assert z == rz
        assert np.array_equal(neg, rneg)
    hip.ok(hip.h.hipStreamDestroy(stream))
    ctx.close()


def test_noalloc_rejects_sizes_beyond_its_scratch(env):
    hip, bt, L = env
    ctx = bt.Context(0)
    p = hip.malloc(64)
    st = L.bht_exact_repulsion_f64_noalloc(ctx.handle, p, (1 << 24) + 1, 2, p, p, None)
    assert st == 1 and b"16777216" in L.bht_last_error()
    st = L.bht_exact_repulsion_f64_noalloc(ctx.handle, p, 10, 3, p, p, None)
    assert st == 1
    hip.ok(hip.h.hipFree(p))
    ctx.close()
