#!/usr/bin/env python3
"""Two-site DMRG sweep wall-time and H*psi TFLOP/s at D=4096, FP64 vs mixed.

Workload: J1-J2 Heisenberg model on a width-4 cylinder (4x24 = 96 sites,
periodic in y, snake ordering, J1 = 1, J2 = 0.5), U(1) Sz = 0, MPO from the
MPOGenerator, seeded random initial MPS (extended direct-product state).

Flow (one GPU):
  1. ramp sweeps (untimed, FP64) with Dmax 256 -> 512 -> 1024 -> 2048,
  2. `--warmup` untimed FP64 sweeps at Dmax = 4096, then a device checkpoint,
  3. `--steps` timed FP64 sweeps from the checkpoint, then the same number
     of timed mixed-precision sweeps from the same checkpoint,
  4. the centre-bond matvec benchmarked with HIP events (FP64 and mixed),
  5. the CPU FP64 oracle timed on that same centre-bond matvec (host cores).
A step is one full sweep (2(N-1) two-site updates) at Dmax = 4096.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sweep wall-time + H·ψ TFLOP/s at D=4096: FP64 vs mixed, |ΔE|/|E|≤1e-10"
PEAK = {"fp64": 78.6, "fp32": 157.3}   # MI355X dense matrix-core TFLOP/s


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(tensors, flops, budget_s):
    """Time the CPU FP64 oracle on the same centre-bond matvec."""
    from oracle import oracle_py as O
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cores = max(1, min(cores, 64))
    O.set_threads(cores)
    L, W1, W2, R, psi = [t.get() if t is not None else None for t in tensors]
    hs = [O.OTen.from_bt(x) for x in (L, W1, W2, R, psi)]
    lib = O.lib()
    reps, t_total = 0, 0.0
    while True:
        t0 = time.perf_counter()
        r = lib.oracle_heff_mul(hs[0].h, hs[1].h, hs[2].h, hs[3].h, hs[4].h, b"cent")
        t_total += time.perf_counter() - t0
        O.OTen(r)
        reps += 1
        if t_total >= budget_s or reps >= 5:
            break
    t = t_total / reps
    return dict(value=flops / t / 1e12, unit="TFLOP/s", cores=cores, kind="port",
                sample=f"{reps} centre-bond H*psi matvec(s) of the D=4096 state (same tensors as the GPU "
                       f"measurement), {t:.2f} s each, OpenMP over {cores} host threads")


def measured_traffic(args):
    """HBM bytes per launch of the roofline kernel, from the committed
    rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled
    by the same counter's reading for the calibration copy); None when no
    summary exists for this configuration."""
    path = os.path.join(ROOT, "profiles", "matvec_stage4_f32_j1j2_4x24_D4096_traffic.json")
    if args.dmax != 4096 or args.lx != 24 or args.width != 4 or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed sweeps per precision")
    ap.add_argument("--warmup", type=int, default=1, help="untimed sweeps at Dmax before timing")
    ap.add_argument("--lx", type=int, default=24)
    ap.add_argument("--width", type=int, default=4)
    ap.add_argument("--j2", type=float, default=0.5)
    ap.add_argument("--dmax", type=int, default=4096)
    ap.add_argument("--ramp", default="256,512,1024,2048")
    ap.add_argument("--lanczos-error", type=float, default=1e-10)
    ap.add_argument("--lanczos-max-iter", type=int, default=200)
    ap.add_argument("--matvec-reps", type=int, default=10)
    ap.add_argument("--cpu-budget-s", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--profile", action="store_true",
                    help="profiler runs: skip the timed sweeps and the CPU baseline (state, matvec bench and the "
                         "calibration copy as in the default run)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the DMRG sweep of this build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not an implemented path). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    from mps2_amd import engine
    from mps2_amd import gqmps2 as G
    from mps2_amd import hamiltonians as HM

    if engine.device_count() < 1:
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    Lx, W = args.lx, args.width
    N = Lx * W
    t0 = time.perf_counter()
    model = HM.j1j2_cylinder(Lx, W, 1.0, args.j2)
    w = int(max(model.bond_dims))
    log(f"model {W}x{Lx} J1-J2 (J2={args.j2}): N={N}, MPO bond dim max {w}, built in {time.perf_counter() - t0:.1f}s")
    labs = [[i % 2 for i in range(N)], [(i + 1) % 2 for i in range(N)]]
    model.init_mps(labs, enlarged_dim=8, seed=args.seed)
    ramp = [int(x) for x in args.ramp.split(",") if x]
    ramp = [d for d in ramp if d < args.dmax]
    le, lm = args.lanczos_error, args.lanczos_max_iter
    for d in ramp:
        t0 = time.perf_counter()
        r = model.dmrg(1, 1, d, 0.0, le, lm)
        log(f"ramp sweep Dmax={d}: E={r['energy']:.12f} maxD={r['dims'].max()} in {time.perf_counter() - t0:.1f}s")
    for k in range(args.warmup):
        t0 = time.perf_counter()
        r = model.dmrg(1, 1, args.dmax, 0.0, le, lm)
        log(f"warmup sweep {k + 1} at Dmax={args.dmax}: E={r['energy']:.12f} maxD={r['dims'].max()} "
            f"in {time.perf_counter() - t0:.1f}s")
    model.checkpoint()

    res = {}
    for name, prec in (() if args.profile else (("fp64", engine.PREC_FP64), ("mixed", engine.PREC_MIXED))):
        model.restore()
        engine.check(engine.lib().mps2_ctx_sync(G.lib().gqmps2_engine_ctx()))
        G.engine_reset_stats()
        t0 = time.perf_counter()
        r = model.dmrg(args.steps, 1, args.dmax, 0.0, le, lm, precision=prec)
        wall = time.perf_counter() - t0
        st = G.engine_stats()
        cnt = engine.ctx_counters(G.lib().gqmps2_engine_ctx())
        log(f"{name}: counters {cnt}")
        res[name] = dict(r=r, wall=wall, stats=st, counters=cnt)
        ph = r["phase_times"].sum(axis=0) / args.steps
        res[name]["phase"] = dict(lanczos=round(float(ph[0]), 3), svd=round(float(ph[1]), 3), env=round(float(ph[2]), 3))
        log(f"{name}: phase seconds per sweep {res[name]['phase']}")
        log(f"{name}: {args.steps} sweep(s) {wall:.2f}s, E={r['energy']:.14f}, maxD={r['dims'].max()}, "
            f"mean Lanczos iters {r['iters'].mean():.1f}, gemm flops {st['gemm_flops']:.3e}, "
            f"peak HBM {st['peak_bytes'] / 2**30:.1f} GiB")

    if res:
        e64, emx = res["fp64"]["r"]["energy"], res["mixed"]["r"]["energy"]
        de_rel = abs(emx - e64) / abs(e64)
        upd_rel = float(max(abs(a - b) / abs(b) for a, b in zip(res["mixed"]["r"]["energies"],
                                                               res["fp64"]["r"]["energies"])))
        dims_equal = bool((res["mixed"]["r"]["dims"] == res["fp64"]["r"]["dims"]).all())

    centre = N // 2 - 1
    t0 = time.perf_counter()
    mb = model.bench_bond(centre, warmup=2, reps=args.matvec_reps)
    log(f"centre-bond matvec bench ({time.perf_counter() - t0:.1f}s): fp64 {mb['fp64']['ms']:.3f} ms "
        f"({mb['fp64']['flops'] / mb['fp64']['ms'] / 1e9:.2f} TFLOP/s), mixed {mb['mixed']['ms']:.3f} ms "
        f"({mb['mixed']['flops'] / mb['mixed']['ms'] / 1e9:.2f} TFLOP/s), stage ms fp64 {mb['fp64']['stage_ms']} "
        f"mixed {mb['mixed']['stage_ms']}")

    # attainable copy bandwidth; the same kernel scales the profiler's byte counters
    copy_bytes = 1 << 30
    copy_ms = engine.copy_bench(G.lib().gqmps2_engine_ctx(), copy_bytes, reps=5)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"calibration copy: {copy_bytes >> 20} MiB in {copy_ms:.3f} ms = {copy_gbs:.0f} GB/s (read + write)")

    def tflops(b):
        return b["flops"] / (b["ms"] * 1e-3) / 1e12

    dom = mb["mixed"]
    dom_ach = dom["dom_flops"] / (dom["dom_ms"] * 1e-3) / 1e12 if dom["dom_ms"] > 0 else 0.0
    roofline = dict(bound="mfma", achieved=round(dom_ach, 3), peak=PEAK["fp32"], unit="TFLOP/s",
                    frac=round(dom_ach / PEAK["fp32"], 4), traffic=measured_traffic(args),
                    kernel=f"grouped_gemm_kernel<float, ..., 1> (stage {dom['dom_stage'] + 1} of the centre-bond "
                           f"matvec, run as the tagged instance)",
                    flops_per_launch=dom["dom_flops"], ms_per_launch=dom["dom_ms"])
    d64 = mb["fp64"]
    d64_ach = d64["dom_flops"] / (d64["dom_ms"] * 1e-3) / 1e12 if d64["dom_ms"] > 0 else 0.0

    if args.profile:
        print(json.dumps({"metric": METRIC, "profile": True, "matvec_ms": {k: mb[k]["ms"] for k in mb},
                          "roofline": roofline, "copy_gbs": round(copy_gbs, 1)}), flush=True)
        return

    L_, W1_, W2_, R_, psi_ = model.bench_tensors()
    ctxh = G.lib().gqmps2_engine_ctx()
    # one GPU's share: the dominant stage split over 8 owners of its output blocks
    st = dom["dom_stage"]
    rows = [engine.heff_bench_shard(L_, W1_, W2_, R_, psi_, "cent", engine.PREC_MIXED, st, 8, r, reps=3)
            for r in range(8)]
    sh_ms = [x[0] for x in rows]
    shard = dict(owners=8, stage=st + 1, precision="mixed", assignment="output blocks, longest-processing-time on flops",
                 ms_per_owner=[round(m, 4) for m in sh_ms], flops_per_owner=[x[1] for x in rows],
                 stage_flops=rows[0][2], full_stage_ms=round(dom["dom_ms"], 4), max_owner_ms=round(max(sh_ms), 4),
                 balance=round(sum(x[1] for x in rows) / 8 / max(x[1] for x in rows), 4),
                 owner_tflops=[round(x[1] / (x[0] * 1e-3) / 1e12, 2) for x in rows],
                 speedup_vs_one_gpu=round(dom["dom_ms"] / max(sh_ms), 3))
    log(f"8-owner share of stage {st + 1}: max {max(sh_ms):.3f} ms per owner vs {dom['dom_ms']:.3f} ms whole "
        f"(speedup {shard['speedup_vs_one_gpu']}, flop balance {shard['balance']})")
    # FP64 vs mixed Lanczos on the centre bond from one random start (seed 42)
    rng = np.random.default_rng(42)
    bt = psi_.get()
    for k in bt.blocks:
        bt.blocks[k] = rng.standard_normal(bt.blocks[k].shape)
    lz = {}
    for name, prec in (("fp64", engine.PREC_FP64), ("mixed", engine.PREC_MIXED)):
        init = engine.put_on(ctxh, bt)
        engine.check(engine.lib().mps2_ctx_sync(ctxh))
        t0 = time.perf_counter()
        it, e, gs = engine.lanczos(L_, W1_, W2_, R_, init, "cent", le, lm, prec)
        engine.check(engine.lib().mps2_ctx_sync(ctxh))
        lz[name] = dict(iters=it, energy=e, s=time.perf_counter() - t0, gs=gs)
    ov = abs(engine.dot(lz["fp64"]["gs"], lz["mixed"]["gs"])) / np.sqrt(
        engine.dot(lz["fp64"]["gs"], lz["fp64"]["gs"]) * engine.dot(lz["mixed"]["gs"], lz["mixed"]["gs"]))
    lanczos_cmp = dict(start="random normal (seed 42) in the centre-bond state's blocks",
                       iters={k: lz[k]["iters"] for k in lz}, energy={k: lz[k]["energy"] for k in lz},
                       seconds={k: round(lz[k]["s"], 3) for k in lz},
                       speedup_mixed_vs_fp64=round(lz["fp64"]["s"] / lz["mixed"]["s"], 3),
                       dE_rel=abs(lz["mixed"]["energy"] - lz["fp64"]["energy"]) / abs(lz["fp64"]["energy"]),
                       overlap=float(ov))
    log(f"centre-bond Lanczos from a random start: fp64 {lz['fp64']['iters']} iters {lz['fp64']['s']:.2f}s, "
        f"mixed {lz['mixed']['iters']} iters {lz['mixed']['s']:.2f}s, dE_rel {lanczos_cmp['dE_rel']:.2e}, "
        f"overlap {ov:.15f}")
    for v in lz.values():
        v["gs"].free()

    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.perf_counter()
        try:
            cpu = cpu_baseline(model.bench_tensors(), mb["fp64"]["flops"], args.cpu_budget_s)
            log(f"cpu baseline {cpu['value']:.4f} TFLOP/s on {cpu['cores']} cores ({time.perf_counter() - t0:.1f}s)")
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")

    step_ms = res["mixed"]["wall"] / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": round(tflops(mb["mixed"]), 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (FP32-MFMA matvec) + f64 (residual, Lanczos, SVD)",
        "data": "synthetic (seeded random initial MPS, D ramped to the target by sweeping)",
        "config": {"workload": f"J1-J2 Heisenberg {W}x{Lx} cylinder, snake order, J2={args.j2}, U(1) Sz=0, "
                               f"D={args.dmax}, two-site DMRG sweep",
                   "sites": N, "mpo_bond_dim": w, "dmax": args.dmax, "lanczos_error": le,
                   "lanczos_max_iter": lm, "step": "one full sweep (2(N-1) two-site updates)"},
        "sweep_s": {"fp64": round(res["fp64"]["wall"] / args.steps, 3),
                    "mixed": round(res["mixed"]["wall"] / args.steps, 3)},
        "sweep_phase_s": {"fp64": res["fp64"]["phase"], "mixed": res["mixed"]["phase"]},
        "sweep_speedup_mixed_vs_fp64": round(res["fp64"]["wall"] / res["mixed"]["wall"], 3),
        "matvec_tflops": {"fp64": round(tflops(mb["fp64"]), 4), "mixed": round(tflops(mb["mixed"]), 4)},
        "matvec_ms": {"fp64": round(mb["fp64"]["ms"], 4), "mixed": round(mb["mixed"]["ms"], 4)},
        "matvec_flops": mb["fp64"]["flops"],
        "matvec_frac_of_peak": {"fp64": round(tflops(mb["fp64"]) / PEAK["fp64"], 4),
                                "mixed": round(tflops(mb["mixed"]) / PEAK["fp32"], 4)},
        "roofline_fp64": dict(bound="mfma", achieved=round(d64_ach, 3), peak=PEAK["fp64"], unit="TFLOP/s",
                              frac=round(d64_ach / PEAK["fp64"], 4)),
        "energy": {"fp64": e64, "mixed": emx, "dE_rel_final": de_rel, "dE_rel_max_update": upd_rel,
                   "kept_dims_identical": dims_equal},
        "truncation_error": {k: {"max": float(res[k]["r"]["trunc_errs"].max()),
                                 "mean": float(res[k]["r"]["trunc_errs"].mean())} for k in res},
        "centre_bond_lanczos": lanczos_cmp,
        "one_gpu_share": shard,
        "lanczos_iters_mean": {"fp64": float(res["fp64"]["r"]["iters"].mean()),
                               "mixed": float(res["mixed"]["r"]["iters"].mean())},
        "peak_hbm_gib": round(max(res[k]["stats"]["peak_bytes"] for k in res) / 2**30, 2),
        "roofline": roofline,
        "hbm_copy_gbs": round(copy_gbs, 1),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
