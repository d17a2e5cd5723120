"""ctypes binding of the gqmps2-compatible host layer (libgqmps2_amd.so,
include/gqmps2_capi.h): model -> MPOGenerator -> MPS init -> two-site sweeps.

The sweeps run on the GPU through the engine's C ABI; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import engine
from .flat import BTensor

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgqmps2_amd.so")
_lib = None

_vp = C.c_void_p
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)

SIGNATURES = {
    "gqmps2_model_create": (C.c_int, [C.c_int, C.c_int, _i32p, _i32p, _i32p, C.POINTER(_vp)]),
    "gqmps2_model_destroy": (C.c_int, [_vp]),
    "gqmps2_model_add_op": (C.c_int, [_vp, C.c_int, _dp, _i32p]),
    "gqmps2_model_add_term": (C.c_int, [_vp, C.c_double, C.c_int, _i32p, _i64p, C.c_int, _i32p]),
    "gqmps2_model_gen_mpo": (C.c_int, [_vp, _i64p]),
    "gqmps2_model_mpo_export_sizes": (C.c_int, [_vp, C.c_int, _i64p, _i64p]),
    "gqmps2_model_mpo_export": (C.c_int, [_vp, C.c_int, _i32p, _dp]),
    "gqmps2_mps_init": (C.c_int, [_vp, C.c_int, _i64p, C.c_int64, C.c_uint64, C.c_int]),
    "gqmps2_mps_export_sizes": (C.c_int, [_vp, C.c_int, _i64p, _i64p]),
    "gqmps2_mps_export": (C.c_int, [_vp, C.c_int, _i32p, _dp]),
    "gqmps2_mps_import": (C.c_int, [_vp, C.c_int, _i32p, _dp]),
    "gqmps2_mps_checkpoint": (C.c_int, [_vp]),
    "gqmps2_mps_restore": (C.c_int, [_vp]),
    "gqmps2_dmrg": (C.c_int, [_vp, C.c_int64, C.c_int64, C.c_int64, _i64p, C.c_double, C.c_double, C.c_int64,
                              C.c_int, C.c_int, _dp, _i64p, _i64p, _dp, _dp, _dp, _dp]),
    "gqmps2_set_bench": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int]),
    "gqmps2_get_bench": (C.c_int, [_vp, _dp]),
    "gqmps2_bench_tensors": (C.c_int, [_vp, C.POINTER(_vp)]),
    "gqmps2_bench_bond": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int]),
    "gqmps2_phase_times": (C.c_int, [_vp, _dp, C.c_int64]),
    "gqmps2_set_verbose": (C.c_int, [C.c_int]),
    "gqmps2_mps_truncate": (C.c_int, [_vp, C.c_double, C.c_int64, C.c_int64]),
    "gqmps2_measure_one_site": (C.c_int, [_vp, C.c_int, _i32p, C.POINTER(C.c_char_p), _dp]),
    "gqmps2_measure_multi_site": (C.c_int, [_vp, C.c_int64, _i32p, _i64p, _i32p, _i32p, C.c_char_p, _dp]),
    "gqmps2_engine_ctx": (_vp, []),
    "gqmps2_last_error": (C.c_char_p, []),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        engine.lib()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        # stdout stays machine-readable for FFI callers; set_verbose(True) restores the log lines
        L.gqmps2_set_verbose(0)
    return _lib


def set_verbose(on: bool):
    check(lib().gqmps2_set_verbose(1 if on else 0))


def check(rc):
    if rc != 0:
        msg = lib().gqmps2_last_error().decode()
        emsg = engine.lib().mps2_last_error().decode()
        raise RuntimeError(f"gqmps2 error {rc}: {msg or emsg}")


class Model:
    """Sites (list of [(qn tuple, dim), ...] per site), operators, terms."""

    def __init__(self, sites, qw=1):
        self.qw = qw
        self.N = len(sites)
        self.site_dims = [sum(d for _, d in s) for s in sites]
        nsect = np.asarray([len(s) for s in sites], dtype=np.int32)
        qns = np.asarray([q for s in sites for (qq, _) in s for q in qq], dtype=np.int32)
        dims = np.asarray([d for s in sites for (_, d) in s], dtype=np.int32)
        self.h = _vp()
        check(lib().gqmps2_model_create(qw, self.N, nsect.ctypes.data_as(_i32p), qns.ctypes.data_as(_i32p),
                                        dims.ctypes.data_as(_i32p), C.byref(self.h)))
        self._ops = {}

    def close(self):
        if self.h:
            lib().gqmps2_model_destroy(self.h)
            self.h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def op(self, site, dense):
        dense = np.ascontiguousarray(dense, dtype=np.float64)
        # the C side reads d x d elements for the site's local dimension d
        if 0 <= site < self.N and dense.shape != (self.site_dims[site],) * 2:
            d = self.site_dims[site]
            raise ValueError(f"operator for site {site} must be {d}x{d}, got shape {dense.shape}")
        key = (site, dense.tobytes())
        if key in self._ops:
            return self._ops[key]
        i = C.c_int32()
        check(lib().gqmps2_model_add_op(self.h, site, dense.ctypes.data_as(_dp), C.byref(i)))
        self._ops[key] = i.value
        return i.value

    def term(self, coef, ops, sites, inst=None):
        """ops: dense matrices (one per listed site); inst: insertion matrices."""
        ids = np.asarray([self.op(s, o) for o, s in zip(ops, sites)], dtype=np.int32)
        st = np.asarray(sites, dtype=np.int64)
        if inst:
            # insertion ops are defined on the sites they act on; register per site
            inst_ids = np.asarray([self.op(sites[0], o) for o in inst], dtype=np.int32)
            check(lib().gqmps2_model_add_term(self.h, coef, len(ids), ids.ctypes.data_as(_i32p),
                                              st.ctypes.data_as(_i64p), len(inst_ids),
                                              inst_ids.ctypes.data_as(_i32p)))
        else:
            check(lib().gqmps2_model_add_term(self.h, coef, len(ids), ids.ctypes.data_as(_i32p),
                                              st.ctypes.data_as(_i64p), 0, None))

    def gen_mpo(self):
        bd = np.zeros(self.N - 1, dtype=np.int64)
        check(lib().gqmps2_model_gen_mpo(self.h, bd.ctypes.data_as(_i64p)))
        return bd

    def _export(self, fn_sizes, fn, i):
        nm, nd = C.c_int64(), C.c_int64()
        check(fn_sizes(self.h, i, C.byref(nm), C.byref(nd)))
        meta = np.zeros(nm.value, dtype=np.int32)
        data = np.zeros(max(nd.value, 1))
        check(fn(self.h, i, meta.ctypes.data_as(_i32p), data.ctypes.data_as(_dp)))
        return BTensor.from_flat(meta, data[:nd.value])

    def mpo(self, i) -> BTensor:
        return self._export(lib().gqmps2_model_mpo_export_sizes, lib().gqmps2_model_mpo_export, i)

    def mps(self, i) -> BTensor:
        return self._export(lib().gqmps2_mps_export_sizes, lib().gqmps2_mps_export, i)

    def set_mps(self, i, t: BTensor):
        meta, data = t.to_flat()
        if data.size == 0:
            data = np.zeros(1)
        check(lib().gqmps2_mps_import(self.h, i, meta.ctypes.data_as(_i32p), data.ctypes.data_as(_dp)))

    def init_mps(self, stat_labs_sets, enlarged_dim=0, seed=0x5EED, centralize=True):
        labs = np.asarray(stat_labs_sets, dtype=np.int64).reshape(-1)
        check(lib().gqmps2_mps_init(self.h, len(stat_labs_sets), labs.ctypes.data_as(_i64p), int(enlarged_dim),
                                    int(seed), 1 if centralize else 0))

    def checkpoint(self):
        check(lib().gqmps2_mps_checkpoint(self.h))

    def restore(self):
        check(lib().gqmps2_mps_restore(self.h))

    def set_bench(self, site, warmup=2, reps=10):
        check(lib().gqmps2_set_bench(self.h, int(site), int(warmup), int(reps)))

    def bench_bond(self, site, warmup=2, reps=10):
        check(lib().gqmps2_bench_bond(self.h, int(site), int(warmup), int(reps)))
        return self.get_bench()

    def get_bench(self):
        out = np.zeros(24)
        check(lib().gqmps2_get_bench(self.h, out.ctypes.data_as(_dp)))

        def unpack(o):
            return dict(ms=o[0], flops=o[1], gemm_ms=o[2], dom_stage=int(o[3]), dom_flops=o[4], dom_ms=o[5],
                        bytes_min=o[6], stage_ms=list(o[7:11]), dom_bytes=o[11])
        return dict(fp64=unpack(out[:12]), mixed=unpack(out[12:]))

    def bench_tensors(self):
        """(L, W1, W2, R, psi) of the benchmarked bond as engine DTen handles."""
        arr = (_vp * 5)()
        check(lib().gqmps2_bench_tensors(self.h, arr))
        return [engine.DTen(_vp(arr[k]), None) if arr[k] else None for k in range(5)]

    def truncate_mps(self, trunc_err, dmin, dmax):
        """TruncateMPS on the resident MPS (device SVDs)."""
        check(lib().gqmps2_mps_truncate(self.h, float(trunc_err), int(dmin), int(dmax)))

    def measure_one_site(self, ops, basenames):
        """MeasureOneSiteOp with a list of dense operators (one result file
        each); returns an array [len(ops), N] of <op at site i>."""
        ids = np.asarray([self.op(0, o) for o in ops], dtype=np.int32)
        names = (C.c_char_p * len(ops))(*[str(b).encode() for b in basenames])
        out = np.zeros((len(ops), self.N))
        check(lib().gqmps2_measure_one_site(self.h, len(ops), ids.ctypes.data_as(_i32p), names,
                                            out.ctypes.data_as(_dp)))
        return out

    def measure_multi_site(self, events, basename):
        """MeasureMultiSiteOp.  events: list of (phys_ops, sites, inst_ops)
        with dense matrices; inst_ops has one entry per gap (None = the
        site's identity).  Returns the array of averages."""
        nphys, st, pid, iid = [], [], [], []
        for phys, sites, inst in events:
            assert len(phys) == len(sites) >= 2 and len(inst) == len(sites) - 1
            nphys.append(len(sites))
            st.extend(sites)
            pid.extend(self.op(s, o) for o, s in zip(phys, sites))
            iid.extend(-1 if o is None else self.op(min(sites[k] + 1, self.N - 1), o) for k, o in enumerate(inst))
        nphys = np.asarray(nphys, dtype=np.int32)
        st = np.asarray(st, dtype=np.int64)
        pid = np.asarray(pid, dtype=np.int32)
        iid = np.asarray(iid if iid else [0], dtype=np.int32)
        out = np.zeros(len(events))
        check(lib().gqmps2_measure_multi_site(self.h, len(events), nphys.ctypes.data_as(_i32p),
                                              st.ctypes.data_as(_i64p), pid.ctypes.data_as(_i32p),
                                              iid.ctypes.data_as(_i32p), str(basename).encode(),
                                              out.ctypes.data_as(_dp)))
        return out

    def dmrg(self, sweeps, dmin, dmax, trunc_err, lanczos_error, lanczos_max_iter=200,
             precision=engine.PREC_FP64, dmax_schedule=None, verbose=False):
        nup = sweeps * 2 * (self.N - 1)
        en = np.zeros(nup)
        dims = np.zeros(nup, dtype=np.int64)
        its = np.zeros(nup, dtype=np.int64)
        te = np.zeros(nup)
        ut = np.zeros(nup)
        st = np.zeros(sweeps)
        e0 = C.c_double()
        sched = None
        if dmax_schedule is not None:
            sched = np.asarray(dmax_schedule, dtype=np.int64)
            assert len(sched) == sweeps
        check(lib().gqmps2_dmrg(self.h, sweeps, dmin, dmax, sched.ctypes.data_as(_i64p) if sched is not None else None,
                                trunc_err, lanczos_error, lanczos_max_iter, precision, 1 if verbose else 0,
                                en.ctypes.data_as(_dp), dims.ctypes.data_as(_i64p), its.ctypes.data_as(_i64p),
                                te.ctypes.data_as(_dp), ut.ctypes.data_as(_dp), st.ctypes.data_as(_dp),
                                C.byref(e0)))
        ph = np.zeros(3 * nup)
        check(lib().gqmps2_phase_times(self.h, ph.ctypes.data_as(_dp), 3 * nup))
        return dict(energy=e0.value, energies=en, dims=dims, iters=its, trunc_errs=te, update_times=ut,
                    phase_times=ph.reshape(nup, 3),
                    sweep_times=st)


def engine_stats():
    ctx = lib().gqmps2_engine_ctx()
    f, n, live, peak = C.c_double(), C.c_int64(), C.c_int64(), C.c_int64()
    engine.check(engine.lib().mps2_ctx_stats(ctx, C.byref(f), C.byref(n), C.byref(live), C.byref(peak)))
    return dict(gemm_flops=f.value, gemm_launches=n.value, live_bytes=live.value, peak_bytes=peak.value)


def engine_reset_stats():
    ctx = lib().gqmps2_engine_ctx()
    engine.check(engine.lib().mps2_ctx_reset_stats(ctx))
