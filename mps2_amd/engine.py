"""ctypes binding of the device engine (libmps2_amd.so, include/mps2_amd.h).

Every operation here runs on the GPU through the C ABI.  There is no CPU
fallback: if the library or a device is missing, calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .flat import BTensor

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmps2_amd.so")
_lib = None

PREC_FP64, PREC_MIXED, PREC_MIXED_BF16X3 = 0, 1, 2
ORTH_LEFT, ORTH_RIGHT = 0, 1

_vp = C.c_void_p
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "mps2_version": (C.c_int, []),
    "mps2_last_error": (C.c_char_p, []),
    "mps2_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "mps2_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "mps2_ctx_destroy": (C.c_int, [_vp]),
    "mps2_ctx_sync": (C.c_int, [_vp]),
    "mps2_ctx_stats": (C.c_int, [_vp, _dp, _i64p, _i64p, _i64p]),
    "mps2_ctx_reset_stats": (C.c_int, [_vp]),
    "mps2_ctx_set_tuning": (C.c_int, [_vp, C.c_int]),
    "mps2_ten_import": (C.c_int, [_vp, _i32p, _dp, C.POINTER(_vp)]),
    "mps2_ten_import_complex": (C.c_int, [_vp, _i32p, _dp, C.POINTER(_vp)]),
    "mps2_ten_is_complex": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "mps2_ten_export_sizes": (C.c_int, [_vp, _i64p, _i64p]),
    "mps2_ten_export": (C.c_int, [_vp, _i32p, _dp]),
    "mps2_ten_free": (C.c_int, [_vp]),
    "mps2_ten_copy": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mps2_ten_dag": (C.c_int, [_vp, C.POINTER(_vp)]),
    "mps2_ten_div": (C.c_int, [_vp, _i32p]),
    "mps2_contract": (C.c_int, [_vp, _vp, C.c_int, _i32p, _i32p, C.POINTER(_vp)]),
    "mps2_linear_combine": (C.c_int, [C.c_int, _dp, C.POINTER(_vp), C.c_double, _vp]),
    "mps2_normalize": (C.c_int, [_vp, _dp]),
    "mps2_dot": (C.c_int, [_vp, _vp, _dp]),
    "mps2_svd": (C.c_int, [_vp, C.c_int, _i32p, C.c_double, C.c_int64, C.c_int64, C.c_int, C.POINTER(_vp),
                           C.POINTER(_vp), C.POINTER(_vp), _dp, _i64p]),
    "mps2_heff_mul": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_char_p, C.POINTER(_vp)]),
    "mps2_lanczos": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_char_p, C.c_double, C.c_int64, C.c_int, _i64p, _dp,
                               C.POINTER(_vp)]),
    "mps2_tridiag_lowest": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp]),
    "mps2_tridiag_lowest_noalloc": (C.c_int, [_vp, C.c_int, _vp, _vp, C.c_int, _vp]),
    "mps2_copy_bench": (C.c_int, [_vp, C.c_int64, C.c_int, _dp]),
    "mps2_heff_bench_shard": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, _dp]),
    "mps2_heff_bench": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_char_p, C.c_int, C.c_int, C.c_int, _dp]),
    "mps2_heff_create": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_char_p, C.c_int, C.POINTER(_vp)]),
    "mps2_heff_run": (C.c_int, [_vp, C.c_int]),
    "mps2_heff_info": (C.c_int, [_vp, _dp, _dp]),
    "mps2_heff_destroy": (C.c_int, [_vp]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class EngineError(RuntimeError):
    pass


def check(rc):
    if rc != 0:
        raise EngineError(f"mps2 error {rc}: {lib().mps2_last_error().decode()}")


def device_count() -> int:
    n = C.c_int()
    check(lib().mps2_device_count(C.byref(n)))
    return n.value


class Context:
    def __init__(self, device: int = 0):
        self.h = _vp()
        check(lib().mps2_ctx_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            check(lib().mps2_ctx_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().mps2_ctx_sync(self.h))

    def stats(self):
        f, n, live, peak = C.c_double(), C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().mps2_ctx_stats(self.h, C.byref(f), C.byref(n), C.byref(live), C.byref(peak)))
        return dict(gemm_flops=f.value, gemm_launches=n.value, live_bytes=live.value, peak_bytes=peak.value)

    def reset_stats(self):
        check(lib().mps2_ctx_reset_stats(self.h))

    # tensors --------------------------------------------------------------
    def put(self, t: BTensor) -> "DTen":
        return DTen(_import(self.h, t), self)

    def tridiag_lowest(self, a, b, want_vec=True):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b if len(b) else np.zeros(1), dtype=np.float64)
        ev = C.c_double()
        vec = np.zeros(len(a))
        check(lib().mps2_tridiag_lowest(self.h, len(a), a.ctypes.data_as(_dp), b.ctypes.data_as(_dp), C.byref(ev),
                                        vec.ctypes.data_as(_dp) if want_vec else None))
        return ev.value, vec


def _import(ctx_handle, t: BTensor):
    meta, data = t.to_flat()
    if data.size == 0:
        data = np.zeros(2)
    h = _vp()
    f = lib().mps2_ten_import_complex if t.is_complex else lib().mps2_ten_import
    check(f(ctx_handle, meta.ctypes.data_as(_i32p), data.ctypes.data_as(_dp), C.byref(h)))
    return h


def put_on(ctx_handle, t: BTensor) -> "DTen":
    """Import a tensor into the context behind a raw handle (e.g. the DMRG layer's)."""
    return DTen(_import(ctx_handle, t), None)


class DTen:
    """Owning handle of a device tensor."""

    def __init__(self, h, ctx):
        self.h = h
        self.ctx = ctx

    def free(self):
        if self.h:
            lib().mps2_ten_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def get(self) -> BTensor:
        nm, nd = C.c_int64(), C.c_int64()
        check(lib().mps2_ten_export_sizes(self.h, C.byref(nm), C.byref(nd)))
        meta = np.zeros(nm.value, dtype=np.int32)
        data = np.zeros(max(nd.value, 1))
        check(lib().mps2_ten_export(self.h, meta.ctypes.data_as(_i32p), data.ctypes.data_as(_dp)))
        return BTensor.from_flat(meta, data[:nd.value], self.is_complex())

    def is_complex(self) -> bool:
        v = C.c_int()
        check(lib().mps2_ten_is_complex(self.h, C.byref(v)))
        return bool(v.value)

    def copy(self):
        h = _vp()
        check(lib().mps2_ten_copy(self.h, C.byref(h)))
        return DTen(h, self.ctx)

    def dag(self):
        h = _vp()
        check(lib().mps2_ten_dag(self.h, C.byref(h)))
        return DTen(h, self.ctx)

    def normalize(self):
        n = C.c_double()
        check(lib().mps2_normalize(self.h, C.byref(n)))
        return n.value


def _h(t):
    return t.h if t is not None else None


def contract(a: DTen, b: DTen, axes) -> DTen:
    xa = np.asarray(axes[0], dtype=np.int32)
    xb = np.asarray(axes[1], dtype=np.int32)
    h = _vp()
    check(lib().mps2_contract(a.h, b.h, len(xa), xa.ctypes.data_as(_i32p), xb.ctypes.data_as(_i32p), C.byref(h)))
    return DTen(h, a.ctx)


def dot(a: DTen, b: DTen) -> float:
    r = C.c_double()
    check(lib().mps2_dot(a.h, b.h, C.byref(r)))
    return r.value


def linear_combine(coefs, ts, beta, out: DTen):
    c = np.ascontiguousarray(coefs, dtype=np.float64)
    arr = (_vp * len(ts))(*[t.h for t in ts])
    check(lib().mps2_linear_combine(len(ts), c.ctypes.data_as(_dp), arr, beta, out.h))


def svd(t: DTen, ldims, ldiv, trunc_err, dmin, dmax, orth=ORTH_LEFT):
    q = np.asarray(ldiv, dtype=np.int32)
    u, s, vt = _vp(), _vp(), _vp()
    err, D = C.c_double(), C.c_int64()
    check(lib().mps2_svd(t.h, ldims, q.ctypes.data_as(_i32p), trunc_err, dmin, dmax, orth, C.byref(u), C.byref(s),
                         C.byref(vt), C.byref(err), C.byref(D)))
    return DTen(u, t.ctx), DTen(s, t.ctx), DTen(vt, t.ctx), err.value, D.value


def heff_mul(L, W1, W2, R, psi, where):
    h = _vp()
    check(lib().mps2_heff_mul(_h(L), W1.h, W2.h, _h(R), psi.h, where.encode(), C.byref(h)))
    return DTen(h, psi.ctx)


def lanczos(L, W1, W2, R, init, where, error, max_iter=200, precision=PREC_FP64):
    it, e, gs = C.c_int64(), C.c_double(), _vp()
    check(lib().mps2_lanczos(_h(L), W1.h, W2.h, _h(R), init.h, where.encode(), error, max_iter, precision,
                             C.byref(it), C.byref(e), C.byref(gs)))
    return it.value, e.value, DTen(gs, init.ctx)


def heff_bench_shard(L, W1, W2, R, psi, where, precision, stage, nshard, shard, reps=5):
    """(ms per launch, share flops, stage flops) of rank `shard`'s output blocks of one matvec stage."""
    out = np.zeros(3)
    check(lib().mps2_heff_bench_shard(_h(L), W1.h, W2.h, _h(R), psi.h, where.encode(), precision, stage, nshard,
                                      shard, reps, out.ctypes.data_as(_dp)))
    return float(out[0]), float(out[1]), float(out[2])


def copy_bench(ctx_handle, nbytes, reps=5):
    """ms per plain device-to-device copy of nbytes (calib_copy_kernel)."""
    ms = C.c_double()
    check(lib().mps2_copy_bench(ctx_handle, int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def heff_bench(L, W1, W2, R, psi, where, precision=PREC_FP64, warmup=2, reps=10):
    out = np.zeros(12)
    check(lib().mps2_heff_bench(_h(L), W1.h, W2.h, _h(R), psi.h, where.encode(), precision, warmup, reps,
                                out.ctypes.data_as(_dp)))
    return dict(ms=out[0], flops=out[1], gemm_ms=out[2], dom_stage=int(out[3]), dom_flops=out[4],
                dom_ms=out[5], bytes_min=out[6], stage_ms=list(out[7:11]), dom_bytes=out[11])


SIGNATURES["mps2_contract_ex"] = (C.c_int, [_vp, _vp, C.c_int, _i32p, _i32p, C.c_int, C.POINTER(_vp)])


def contract_ex(a: DTen, b: DTen, axes, precision=PREC_FP64) -> DTen:
    xa = np.asarray(axes[0], dtype=np.int32)
    xb = np.asarray(axes[1], dtype=np.int32)
    h = _vp()
    check(lib().mps2_contract_ex(a.h, b.h, len(xa), xa.ctypes.data_as(_i32p), xb.ctypes.data_as(_i32p), precision,
                                 C.byref(h)))
    return DTen(h, a.ctx)


SIGNATURES["mps2_ctx_counter"] = (C.c_int, [_vp, C.c_char_p, _dp])
COUNTERS = ("svd_calls", "svd_ms", "svd_eigh_ms", "svd_sectors", "svd_max_sector_sum", "syncs", "lanczos_calls",
            "matvec_fp64", "matvec_fp32", "hip_malloc", "alloc_trim", "svd_refine_rounds", "svd_refine_sectors",
            "svd_refine_ms", "svd_refine_vec_ms", "svd_refine_eigh_ms", "tune_probes", "tune_decisions",
            "tune_changed", "tune_cache_hits", "eigh_trd_ms", "eigh_dc_ms", "eigh_back_ms", "eigh_trd_columns",
            "eigh_n3_sum_g", "gemm_narrow", "gemm_narrow_rowmajor_a", "plan_builds", "plan_build_ms",
            "svd_refine_settled")


def ctx_counters(ctx_handle, names=COUNTERS):
    """Read the named counters of a context handle (c_void_p)."""
    out = {}
    v = C.c_double()
    for n in names:
        check(lib().mps2_ctx_counter(ctx_handle, n.encode(), C.byref(v)))
        out[n] = v.value
    return out


class HeffOp:
    """The effective Hamiltonian bound once to (L, W1, W2, R) and a state
    layout at one precision (mps2_heff_create); run(n) applies it n times,
    stream-ordered, with no host sync."""

    def __init__(self, L, W1, W2, R, psi, where, precision=PREC_FP64):
        self.h = _vp()
        check(lib().mps2_heff_create(_h(L), _h(W1), _h(W2), _h(R), _h(psi), where.encode(), precision,
                                     C.byref(self.h)))
        f, b = C.c_double(), C.c_double()
        check(lib().mps2_heff_info(self.h, C.byref(f), C.byref(b)))
        self.flops, self.bytes = f.value, b.value

    def run(self, n):
        check(lib().mps2_heff_run(self.h, int(n)))

    def close(self):
        if self.h:
            check(lib().mps2_heff_destroy(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
