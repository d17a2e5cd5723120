"""mps2_tridiag_lowest_noalloc (the allocation-free entry point) against
mps2_tridiag_lowest on the same matrices: two calls in a row, the second
smaller, results identical (the same kernel on the same input), with its
work space reused between the calls; sizes beyond the static array are
rejected.  Device buffers come from the HIP runtime the engine links."""
import ctypes as C

import numpy as np
import pytest

from mps2_amd import engine as E

pytestmark = pytest.mark.gpu


def _hip():
    E.lib()   # loads libamdhip64.so.7 as the engine's dependency; this is the same instance
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0

    def put(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size <= self.n
        assert self.hip.hipMemcpy(self.p, x.ctypes.data, 8 * x.size, H2D) == 0

    def get(self, n):
        out = np.zeros(n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * n, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_tridiag_noalloc_matches_tridiag_lowest():
    L = E.lib()
    hip = _hip()
    ctx = E.Context(0)
    rng = np.random.default_rng(21)
    nmax = 200
    da, db, dout = DevBuf(hip, nmax), DevBuf(hip, nmax), DevBuf(hip, nmax + 1)
    try:
        for n in (200, 37):   # the second call is smaller: stale work space must not matter
            a = rng.uniform(-2, 2, n)
            b = rng.uniform(0.1, 1.0, n - 1)
            ev_ref, vec_ref = ctx.tridiag_lowest(a, b, want_vec=True)
            da.put(a)
            db.put(b)
            E.check(L.mps2_tridiag_lowest_noalloc(ctx.h, n, da.p, db.p, 1, dout.p))
            ctx.sync()
            o = dout.get(n + 1)
            assert o[0] == ev_ref, (n, o[0], ev_ref)
            assert np.array_equal(o[1:], vec_ref), (n, np.abs(o[1:] - vec_ref).max())
            # and against the dense lowest eigenpair
            T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
            w, V = np.linalg.eigh(T)
            assert abs(o[0] - w[0]) < 1e-12 * max(1.0, abs(w[0]))
            assert abs(abs(float(o[1:] @ V[:, 0])) - 1.0) < 1e-10
        # eigenvalue only
        E.check(L.mps2_tridiag_lowest_noalloc(ctx.h, 37, da.p, db.p, 0, dout.p))
        ctx.sync()
        assert dout.get(1)[0] == ev_ref
        # beyond the static work space, and bad arguments: error statuses, nothing launched
        assert L.mps2_tridiag_lowest_noalloc(ctx.h, 4097, da.p, db.p, 1, dout.p) != 0
        assert "4096" in E.lib().mps2_last_error().decode()
        assert L.mps2_tridiag_lowest_noalloc(ctx.h, 0, da.p, db.p, 1, dout.p) != 0
        assert L.mps2_tridiag_lowest_noalloc(ctx.h, 5, None, db.p, 1, dout.p) != 0
    finally:
        for x in (da, db, dout):
            x.free()
        ctx.close()
