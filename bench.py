"""Benchmark of the Dupire barrier pricing + AAD risk engine on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (one "step" = one AADriskOne-equivalent call over all paths):
Dupire local vol on the 30 x 36 grid (1,080 vols), up-and-out call K=120,
B=150, 5% smoothing, 3y weekly monitoring (156 log-Euler steps), 2^20 Sobol
paths, price + delta + 1,080 local-vol sensitivities, FP32 (FP64 sums).
Measured beside it: the same workload in FP64 and with the recompute
adjoint, the 2^19-path FP64/FP32 pricing run, the Black-Scholes European on
2^16 paths (one step), the product book on shared
paths (30 Europeans, 10 barriers, 5 Asians, aggregated risk, 2^20 paths) and a
replay of 512 uneven requests at 200/s through the request server (p50/p99
latency).  Inputs are synthetic (the local-vol formula of the
configuration); all tables are resident on the device before timing.

Prints one JSON line.  `value` is paths*steps per second of the FP32 AAD
call; `roofline` is the AAD path kernel's HBM rate: its algorithmic bytes
(the census: the 4-byte FP32 state of every path-step written in the forward
sweep and read back in the reverse, 8 B per path-step) over its HIP-event
duration, against 8 TB/s.  `traffic` is the kernel's measured HBM bytes per
launch (calibrated FETCH_SIZE + WRITE_SIZE, which also count the knot-bracket
bytes and the per-block adjoint grids).  `valu` gives each C3 and C2 kernel's
census FLOP rate against its vector peak.  The CPU baseline times the reference engine (oracle/_ref) on the
host cores, or the FP64 CPU restatement when the reference build is absent.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import compfinance_amd as cf  # noqa: E402
from compfinance_amd import workloads as W  # noqa: E402

METRIC = "Dupire price+1080 risks wall-ms & paths·steps/s FP32 vs FP64; serve p99 @QPS"
_BOOK_CLS = {"europeans": cf.Europeans, "barrier": cf.UOC, "asian": cf.Asian}

# Algorithmic work per path-step (FMA = 2 FLOP, FP32 transcendentals weighted
# 4, FP64 divide 8 and exp/log 20): forward 41/69, adjoint 26/46 FLOP.
FLOP_PER_PATH_STEP = {("aad", cf.FP32): 67.0, ("aad", cf.FP64): 115.0,
                      ("price", cf.FP32): 41.0, ("price", cf.FP64): 69.0}
PEAK_TFLOPS = {cf.FP32: 157.3, cf.FP64: 78.6}   # MI355X vector peaks
# Algorithmic HBM bytes of the checkpoint adjoint per path-step (the census):
# the state (4 B FP32 / 8 B FP64) written in the forward sweep and read back
# in the reverse sweep.  The knot-bracket byte stored beside each state and the
# per-block adjoint grids are the kernel's own additions: they show in the
# measured traffic, not here.
CKPT_BYTES_PER_PATH_STEP = {cf.FP32: 2 * 4, cf.FP64: 2 * 8}
HBM_PEAK_GBS = 8000.0                            # MI355X HBM3E
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic_aad_dupire_barrier_fp32.json")


def _time_calls(fn, steps, warmup):
    """Wall time of `steps` calls (nothing else in the timed region), then the
    path kernel's HIP-event time of each of `steps` further calls."""
    for _ in range(warmup):
        fn()
    cf.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    cf.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = []
    for _ in range(steps):
        fn()
        kernel_ms.append(cf.last_kernel_stats()["main_kernel_ms"])
    return wall, kernel_ms


def _cpu_baseline(n_steps):
    """Reference engine (oracle/_ref) on the host cores, bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    mdl = po.Model(**W.dupire_args(30))
    prd = po.Product(**W.barrier_args())
    cores = os.cpu_count() or 1
    try:
        cores = min(cores, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    if po.ref_available():
        n = 1 << 20
        try:
            out = po.ref_aad(mdl, prd, n, risk_index=0, threads=cores, repeats=5)
            ms = out["wall_ms_median"]
            return {"value": n * n_steps / (ms / 1e3), "unit": "paths*steps/s", "cores": cores,
                    "kind": "reference",
                    "sample": "reference AADriskOne (mcParallelSimulAAD, ThreadPool %d threads), 2^20 paths x %d "
                              "steps, 30x36 grid, median of 5: %.1f ms" % (cores, n_steps, ms)}, ms
        except (subprocess.CalledProcessError, OSError, ValueError):
            pass
    n = 1 << 14
    t0 = time.perf_counter()
    po.aad(mdl, prd, n, 0)
    dt = time.perf_counter() - t0
    return {"value": n * n_steps / dt, "unit": "paths*steps/s", "cores": 1, "kind": "port",
            "sample": "FP64 CPU restatement with tape AAD, 2^14 paths x %d steps, one thread: %.1f s" % (n_steps,
                                                                                                      dt)}, None


def _cpu_c4_parts(uoc_ms, weights, gpu_wall_ms):
    """The reference has no book and no Asian: beside the book, its
    AADriskAggregate on the book's 30 Europeans alone (the book's notionals),
    and its AADriskOne on one barrier (the C3 run: each of the book's 10
    barriers costs the same on the same grid and steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if uoc_ms is None or not po.ref_available():
        return None
    cores = os.cpu_count() or 1
    try:
        cores = min(cores, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    kind, kw = W.c4_parts()[0]
    assert kind == "europeans"
    nots = {i: float(weights[i]) for i in range(len(kw["strikes"]))}
    try:
        out = po.ref_aad(po.Model(**W.dupire_args(30)), po.Product(**kw), W.C4_PATHS, threads=cores, repeats=3,
                         notionals=nots)
    except (subprocess.CalledProcessError, OSError, ValueError):
        return None
    eu = out["wall_ms_median"]
    total = eu + 10 * uoc_ms
    return {"europeans_ms": eu, "uoc_ms": uoc_ms, "sum_of_parts_ms": total, "cores": cores, "kind": "reference",
            "book_fp32_wall_ms": gpu_wall_ms, "sum_of_parts_over_book_fp32": total / gpu_wall_ms,
            "sample": "reference AADriskAggregate on the 30 Europeans (2^20 paths, median of 3) plus 10 x its "
                      "AADriskOne on one barrier (median of 5); the 5 Asians have no reference class"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # (the device takes a few tens of calls to reach its steady clocks: with 3
    # warm-up calls the first 20 timed calls ran 12% slower than later ones)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the request-server replay")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine shards paths across GPUs only by running one "
              "process per GPU on disjoint path ranges (first_path), which this benchmark does not launch.",
              file=sys.stderr)
        sys.exit(2)

    mdl = cf.Dupire(**W.dupire_args(30))
    prd = cf.UOC(**W.barrier_args())
    n3, n2 = W.C3_PATHS, W.C2_PATHS
    n_steps = int(len(cf.simulation_timeline(mdl, prd)[0]) - 1)
    w = np.array([1.0, 0.0])

    def aad(prec):
        return lambda: cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(numPath=n3, precision=prec))

    def price(prec):
        return lambda: cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(numPath=n2, precision=prec), per_path=False)

    # headline: FP32 AAD
    wall32, k32 = _time_calls(aad(cf.FP32), args.steps, args.warmup)
    ms32 = 1e3 * wall32 / args.steps
    r32 = aad(cf.FP32)()
    st32 = cf.last_kernel_stats()
    # FP64 AAD form (fewer timed calls)
    # the other legs: as many timed calls as the headline, two warm-up calls
    s64, w2 = max(5, args.steps), 2
    wall64, k64 = _time_calls(aad(cf.FP64), s64, w2)
    ms64 = 1e3 * wall64 / s64
    r64 = aad(cf.FP64)()
    # pricing, C2
    wp32, kp32 = _time_calls(price(cf.FP32), s64, w2)
    wp64, kp64 = _time_calls(price(cf.FP64), s64, w2)
    p32 = price(cf.FP32)()
    p64 = price(cf.FP64)()

    # recompute adjoint (no per-step checkpoints)
    def aad_rc(prec):
        return lambda: cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(
            numPath=n3, precision=prec, aadMode=cf.AAD_RECOMPUTE))

    wr32, kr32 = _time_calls(aad_rc(cf.FP32), s64, w2)
    wr64, kr64 = _time_calls(aad_rc(cf.FP64), s64, w2)
    rr32 = aad_rc(cf.FP32)()

    # book on shared paths: 30 Europeans, 10 barriers, 5 Asians, aggregated risk
    book = cf.Book([_BOOK_CLS[k](**kw) for k, kw in W.c4_parts()])
    bw = W.c4_notionals(len(book.payoffLabels()))

    def c4(prec):
        return lambda: cf.mcParallelSimulAAD(book, mdl, weights=bw, num=cf.NumericalParam(
            numPath=W.C4_PATHS, precision=prec))

    # C1: Black-Scholes European call, 2^16 paths, one step (latency-bound)
    bs1, eu1 = cf.BlackScholes(100.0, 0.2), cf.European(100.0, 1.0)

    def c1(aad_call):
        n1 = cf.NumericalParam(numPath=1 << 16)
        if aad_call:
            return lambda: cf.mcParallelSimulAAD(eu1, bs1, num=n1)
        return lambda: cf.mcParallelSimul(eu1, bs1, num=n1, per_path=False)

    wc1, _ = _time_calls(c1(False), s64, w2)
    wc1a, _ = _time_calls(c1(True), s64, w2)
    c1_val, c1_risk = c1(False)(), c1(True)()

    wb32, kb32 = _time_calls(c4(cf.FP32), s64, w2)
    wb64, kb64 = _time_calls(c4(cf.FP64), s64, w2)
    b32, b64 = c4(cf.FP32)(), c4(cf.FP64)()
    nb_steps = int(len(cf.simulation_timeline(mdl, book)[0]) - 1)

    path_steps = n3 * n_steps
    kavg = float(np.mean(k32))
    alg_bytes = CKPT_BYTES_PER_PATH_STEP[cf.FP32] * path_steps
    gbs = alg_bytes / (kavg * 1e-3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    def valu(kind, prec, kernel_ms, paths):
        """Census FLOP rate of one kernel against its vector peak."""
        flop = FLOP_PER_PATH_STEP[(kind, prec)] * paths * n_steps
        tf = flop / (float(np.mean(kernel_ms)) * 1e-3) / 1e12
        return {"achieved": tf, "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s", "frac": tf / PEAK_TFLOPS[prec],
                "kernel_ms": float(np.mean(kernel_ms)), "flop_per_path_step": FLOP_PER_PATH_STEP[(kind, prec)],
                "path_steps": paths * n_steps}

    # FP32 vs FP64 in standard errors of the FP64 price (risk errors reported by the tests)
    se64 = float(p64.std_errs[0])
    out = {
        "metric": METRIC,
        "value": path_steps / (ms32 * 1e-3),
        "unit": "paths*steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (local-vol formula of the configuration)",
        "config": {"workload": "Dupire up-and-out call K=120 B=150 3y weekly, 30x36 local-vol grid, price + delta "
                               "+ 1080 local-vol risks (AADriskOne), FP32 with FP64 sums, checkpoint adjoint",
                   "paths": n3, "time_steps": n_steps, "risks": int(r32.risks.size)},
        "fp32_aad_wall_ms": ms32,
        "fp64_aad_wall_ms": ms64,
        "fp64_aad_paths_steps_per_s": path_steps / (ms64 * 1e-3),
        "fp32_over_fp64_aad": ms64 / ms32,
        "aad_kernel_ms": {"fp32": kavg, "fp64": float(np.mean(k64))},
        "fp32_aad_call_breakdown_ms": {"host_setup": st32["setup_ms"], "path_kernel": st32["main_kernel_ms"],
                                       "reductions": st32["reduce_ms"], "device_total": st32["total_ms"],
                                       "host_total": st32["host_ms"]},
        "price_c2": {"paths": n2, "fp32_wall_ms": 1e3 * wp32 / s64, "fp64_wall_ms": 1e3 * wp64 / s64,
                     "fp32_kernel_ms": float(np.mean(kp32)), "fp64_kernel_ms": float(np.mean(kp64)),
                     "fp64_value": float(p64.means[0]), "fp64_std_err": se64,
                     "fp32_minus_fp64_in_se": float((p32.means[0] - p64.means[0]) / se64)},
        "aad_recompute": {"fp32_wall_ms": 1e3 * wr32 / s64, "fp64_wall_ms": 1e3 * wr64 / s64,
                          "fp32_kernel_ms": float(np.mean(kr32)), "fp64_kernel_ms": float(np.mean(kr64)),
                          "fp32_max_risk_diff_vs_checkpoint": float(np.abs(rr32.risks - r32.risks).max()),
                          "fp32_risk_scale": float(np.abs(r32.risks).max())},
        "bs_european_c1": {"paths": 1 << 16, "time_steps": 1, "fp64_value_wall_ms": 1e3 * wc1 / s64,
                           "fp64_aad_wall_ms": 1e3 * wc1a / s64, "value": float(c1_val.means[0]),
                           "std_err": float(c1_val.std_errs[0]), "closed_form": 7.965567455405804,
                           "delta": float(c1_risk.risks[0]), "vega": float(c1_risk.risks[1])},
        "book_c4": {"paths": W.C4_PATHS, "time_steps": nb_steps, "payoffs": int(len(bw)),
                    "risks": int(b32.risks.size),
                    "fp32_wall_ms": 1e3 * wb32 / s64, "fp64_wall_ms": 1e3 * wb64 / s64,
                    "fp32_kernel_ms": float(np.mean(kb32)), "fp64_kernel_ms": float(np.mean(kb64)),
                    "fp32_paths_steps_per_s": W.C4_PATHS * nb_steps / (wb32 / s64),
                    "aggregate_fp32": float(b32.agg_mean), "aggregate_fp64": float(b64.agg_mean)},
        "aad_values": {"fp32": float(r32.agg_mean), "fp64": float(r64.agg_mean),
                       "delta_fp32": float(r32.risks[0]), "delta_fp64": float(r64.risks[0])},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_over_algorithmic": traffic / alg_bytes if traffic else None,
                     "kernel": "path_kernel<float, Dupire, barrier, checkpoint adjoint>", "kernel_ms": kavg,
                     "algorithmic_bytes": alg_bytes,
                     "bytes_per_path_step": CKPT_BYTES_PER_PATH_STEP[cf.FP32], "path_steps": path_steps,
                     "valu_frac": valu("aad", cf.FP32, k32, n3)["frac"]},
        "valu": {"c3_fp32_checkpoint": valu("aad", cf.FP32, k32, n3),
                 "c3_fp64_checkpoint": valu("aad", cf.FP64, k64, n3),
                 "c3_fp32_recompute": valu("aad", cf.FP32, kr32, n3),
                 "c3_fp64_recompute": valu("aad", cf.FP64, kr64, n3),
                 "c2_fp32_price": valu("price", cf.FP32, kp32, n2),
                 "c2_fp64_price": valu("price", cf.FP64, kp64, n2)},
    }
    if not args.no_serving:
        from compfinance_amd import serving
        out["serving_c5"] = serving.run_c5()
    if not args.no_cpu_baseline:
        out["cpu_baseline"], uoc_ms = _cpu_baseline(n_steps)
        parts = _cpu_c4_parts(uoc_ms, bw, out["book_c4"]["fp32_wall_ms"])
        if parts is not None:
            out["cpu_c4_parts"] = parts
    print(json.dumps(out))


if __name__ == "__main__":
    main()
