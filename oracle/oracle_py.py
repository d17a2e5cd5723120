"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, where it is the checker.  The product package
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from mps2_amd.flat import BTensor

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.oracle_ten_import.restype = vp
        L.oracle_ten_import.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.oracle_ten_import_complex.restype = vp
        L.oracle_ten_import_complex.argtypes = [C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.oracle_ten_is_complex.restype = C.c_int
        L.oracle_ten_is_complex.argtypes = [vp]
        L.oracle_ten_export_sizes.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.oracle_ten_export.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        L.oracle_ten_free.argtypes = [vp]
        L.oracle_contract.restype = vp
        L.oracle_contract.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.oracle_dag.restype = vp
        L.oracle_dag.argtypes = [vp]
        L.oracle_normalize.restype = C.c_double
        L.oracle_normalize.argtypes = [vp]
        L.oracle_dot.restype = C.c_double
        L.oracle_dot.argtypes = [vp, vp]
        L.oracle_svd.argtypes = [vp, C.c_int, C.POINTER(C.c_int32), C.c_double, C.c_int64, C.c_int64,
                                 C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int64)]
        L.oracle_tridiag_lowest.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.oracle_lanczos.argtypes = [vp, vp, vp, vp, vp, C.c_char_p, C.c_double, C.c_int64,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_double), C.POINTER(vp)]
        L.oracle_heff_mul.restype = vp
        L.oracle_heff_mul.argtypes = [vp, vp, vp, vp, vp, C.c_char_p]
        L.oracle_centralize0.argtypes = [C.c_int, C.POINTER(vp)]
        L.oracle_dmrg.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp), C.c_int64, C.c_int64, C.c_int64,
                                  C.c_double, C.c_double, C.c_int64, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double)]
        L.oracle_expect.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(C.c_double)]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_set_threads.restype = C.c_int
        _lib = L
    return _lib


def _err():
    return lib().oracle_last_error().decode()


class OTen:
    """Owning handle of an oracle tensor."""

    def __init__(self, h):
        if not h:
            raise RuntimeError("oracle: " + _err())
        self.h = C.c_void_p(h) if not isinstance(h, C.c_void_p) else h

    @staticmethod
    def from_bt(t: BTensor) -> "OTen":
        meta, data = t.to_flat()
        if data.size == 0:
            data = np.zeros(2)
        f = lib().oracle_ten_import_complex if t.is_complex else lib().oracle_ten_import
        return OTen(f(meta.ctypes.data_as(C.POINTER(C.c_int32)), data.ctypes.data_as(C.POINTER(C.c_double))))

    def to_bt(self) -> BTensor:
        nm, nd = C.c_int64(), C.c_int64()
        lib().oracle_ten_export_sizes(self.h, C.byref(nm), C.byref(nd))
        meta = np.zeros(nm.value, dtype=np.int32)
        data = np.zeros(max(nd.value, 1))
        lib().oracle_ten_export(self.h, meta.ctypes.data_as(C.POINTER(C.c_int32)),
                                data.ctypes.data_as(C.POINTER(C.c_double)))
        return BTensor.from_flat(meta, data[:nd.value], bool(lib().oracle_ten_is_complex(self.h)))

    def __del__(self):
        try:
            if self.h:
                lib().oracle_ten_free(self.h)
                self.h = None
        except Exception:
            pass


def contract(a: BTensor, b: BTensor, axes) -> BTensor:
    A, B = OTen.from_bt(a), OTen.from_bt(b)
    xa = np.asarray(axes[0], dtype=np.int32)
    xb = np.asarray(axes[1], dtype=np.int32)
    r = lib().oracle_contract(A.h, B.h, len(xa), xa.ctypes.data_as(C.POINTER(C.c_int32)),
                              xb.ctypes.data_as(C.POINTER(C.c_int32)))
    return OTen(r).to_bt()


def svd(t: BTensor, ldims, ldiv, trunc_err, dmin, dmax):
    T = OTen.from_bt(t)
    q = np.asarray(ldiv, dtype=np.int32)
    u, s, vt = C.c_void_p(), C.c_void_p(), C.c_void_p()
    err, D = C.c_double(), C.c_int64()
    rc = lib().oracle_svd(T.h, ldims, q.ctypes.data_as(C.POINTER(C.c_int32)), trunc_err, dmin, dmax,
                          C.byref(u), C.byref(s), C.byref(vt), C.byref(err), C.byref(D))
    if rc:
        raise RuntimeError("oracle svd: " + _err())
    return OTen(u).to_bt(), OTen(s).to_bt(), OTen(vt).to_bt(), err.value, D.value


def tridiag_lowest(a, b, want_vec=True):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b if len(b) else np.zeros(1), dtype=np.float64)
    ev = C.c_double()
    vec = np.zeros(len(a))
    rc = lib().oracle_tridiag_lowest(len(a), a.ctypes.data_as(C.POINTER(C.c_double)),
                                     b.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ev),
                                     vec.ctypes.data_as(C.POINTER(C.c_double)) if want_vec else None)
    if rc:
        raise RuntimeError("oracle tridiag failed")
    return ev.value, vec


def _h(t):
    return OTen.from_bt(t) if t is not None else OTen.from_bt(BTensor([], 1, {(): np.zeros(())}))


def heff_mul(L, W1, W2, R, psi, where):
    hs = [_h(x) for x in (L, W1, W2, R, psi)]
    r = lib().oracle_heff_mul(*[h.h for h in hs], where.encode())
    return OTen(r).to_bt()


def lanczos(L, W1, W2, R, init, where, error, max_iter):
    hs = [_h(x) for x in (L, W1, W2, R, init)]
    it, e, gs = C.c_int64(), C.c_double(), C.c_void_p()
    rc = lib().oracle_lanczos(*[h.h for h in hs], where.encode(), error, max_iter, C.byref(it), C.byref(e),
                              C.byref(gs))
    if rc:
        raise RuntimeError("oracle lanczos: " + _err())
    return it.value, e.value, OTen(gs).to_bt()


def centralize0(mps):
    hs = [OTen.from_bt(t) for t in mps]
    arr = (C.c_void_p * len(hs))(*[h.h for h in hs])
    for h in hs:
        h.h = None  # ownership passes to the array below
    rc = lib().oracle_centralize0(len(hs), arr)
    if rc:
        raise RuntimeError("oracle centralize: " + _err())
    return [OTen(arr[i]).to_bt() for i in range(len(hs))]


def dmrg(mps, mpo, sweeps, dmin, dmax, trunc_err, lanczos_error, lanczos_max_iter=200):
    """Returns dict(energy, energies, dims, iters, trunc_errs, mps)."""
    n = len(mps)
    hs = [OTen.from_bt(t) for t in mps]
    ws = [OTen.from_bt(t) for t in mpo]
    arr = (C.c_void_p * n)(*[h.h for h in hs])
    for h in hs:
        h.h = None
    warr = (C.c_void_p * n)(*[w.h for w in ws])
    nup = sweeps * 2 * (n - 1)
    en = np.zeros(nup)
    dims = np.zeros(nup, dtype=np.int64)
    its = np.zeros(nup, dtype=np.int64)
    te = np.zeros(nup)
    e0 = C.c_double()
    rc = lib().oracle_dmrg(n, arr, warr, sweeps, dmin, dmax, trunc_err, lanczos_error, lanczos_max_iter,
                           en.ctypes.data_as(C.POINTER(C.c_double)), dims.ctypes.data_as(C.POINTER(C.c_int64)),
                           its.ctypes.data_as(C.POINTER(C.c_int64)), te.ctypes.data_as(C.POINTER(C.c_double)),
                           C.byref(e0))
    if rc:
        raise RuntimeError("oracle dmrg: " + _err())
    out_mps = [OTen(arr[i]).to_bt() for i in range(n)]
    return dict(energy=e0.value, energies=en, dims=dims, iters=its, trunc_errs=te, mps=out_mps)


def expect(mps, ops):
    """<psi| prod_i ops[i] |psi> of the MPS as given (ops[i] None = identity)."""
    n = len(mps)
    hs = [OTen.from_bt(t) for t in mps]
    os_ = [OTen.from_bt(o) if o is not None else None for o in ops]
    arr = (C.c_void_p * n)(*[h.h for h in hs])
    oarr = (C.c_void_p * n)(*[o.h if o is not None else None for o in os_])
    out = C.c_double()
    rc = lib().oracle_expect(n, arr, oarr, C.byref(out))
    if rc:
        raise RuntimeError("oracle expect: " + _err())
    return out.value


def set_threads(n):
    return lib().oracle_set_threads(int(n))
