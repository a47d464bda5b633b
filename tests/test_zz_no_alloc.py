"""The allocation-free UDT entry point (dqmc_la_udt_noalloc) against dqmc_la_udt.

Both run the same kernels; the no-allocation form keeps its working storage in the
library's static device array, so its outputs must equal the allocating entry point's bit
for bit.  Two calls in a row, the second smaller, show that the shared array's leftover
contents do not leak into a later call.  Device memory and the stream come from the HIP
runtime the library is linked against; the library is called through its C ABI."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """The HIP runtime the library itself is linked against (device memory and a stream for
    the test; one runtime per process)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        for name, args in (("hipMalloc", [C.POINTER(C.c_void_p), C.c_size_t]),
                           ("hipFree", [C.c_void_p]),
                           ("hipMemcpy", [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
                           ("hipMemset", [C.c_void_p, C.c_int, C.c_size_t]),
                           ("hipStreamCreate", [C.POINTER(C.c_void_p)]),
                           ("hipStreamSynchronize", [C.c_void_p]),
                           ("hipStreamDestroy", [C.c_void_p])):
            f = getattr(self.L, name)
            f.argtypes = args
            f.restype = C.c_int
        self.bufs = []

    def ok(self, rc, what):
        assert rc == 0, f"{what}: {rc}"

    def alloc(self, nbytes):
        p = C.c_void_p()
        self.ok(self.L.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        self.bufs.append(p)
        return p

    def put(self, arr):
        arr = np.ascontiguousarray(arr) if arr.ndim < 2 else np.asfortranarray(arr)
        p = self.alloc(arr.nbytes)
        self.ok(self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1), "hipMemcpy H2D")
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype=dtype, order="F")
        self.ok(self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def _noalloc(hip, H, stream, X: np.ndarray, apply_pivot: bool):
    n = X.shape[0]
    Xd = H.put(X)
    Ud, Td = H.alloc(8 * n * n), H.alloc(8 * n * n)
    Dd, Pd = H.alloc(8 * n), H.alloc(4 * n)
    Ed = H.put(np.zeros(1, dtype=np.uint64))
    f = hip.load().dqmc_la_udt_noalloc
    rc = f(n, Xd, int(apply_pivot), Ud, Dd, Td, Pd, Ed, stream)
    assert rc == 0, rc
    H.ok(H.L.hipStreamSynchronize(stream), "hipStreamSynchronize")
    out = (H.get(Ud, (n, n), np.float64), H.get(Dd, (n,), np.float64), H.get(Td, (n, n), np.float64),
           H.get(Pd, (n,), np.int32), int(H.get(Ed, (1,), np.uint64)[0]))
    H.free()
    return out


def test_noalloc_udt_equals_allocating_udt(hip):
    hip.load()
    H = _Hip()
    stream = C.c_void_p()
    H.ok(H.L.hipStreamCreate(C.byref(stream)), "hipStreamCreate")
    try:
        rng = np.random.default_rng(12)
        # two calls in a row with the second smaller, then a larger one and a small one again
        for n, apply_pivot in ((300, True), (100, False), (576, False), (64, True)):
            # graded columns, as the stack's products have
            X = np.asfortranarray(rng.standard_normal((n, n)) * np.logspace(0, -12, n)[None, :])
            U0, D0, T0, P0 = hip.la_udt(X, apply_pivot)
            U1, D1, T1, P1, errs = _noalloc(hip, H, stream, X, apply_pivot)
            assert errs == 0
            assert np.array_equal(P1, P0), n
            assert np.array_equal(D1, D0), n
            assert np.array_equal(U1, U0), n
            assert np.array_equal(T1, T0), n
    finally:
        H.L.hipStreamDestroy(stream)


def test_noalloc_refuses_sizes_beyond_its_array(hip):
    f = hip.load().dqmc_la_udt_noalloc
    dummy = C.c_void_p(16)   # never dereferenced: the size check comes first
    assert f(1025, dummy, 1, C.c_void_p(32), dummy, C.c_void_p(48), dummy, None, None) == -1
    assert f(0, dummy, 1, C.c_void_p(32), dummy, C.c_void_p(48), dummy, None, None) == -1
