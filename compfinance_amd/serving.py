"""Serving workload: a seeded stream of uneven pricing and risk requests
replayed open-loop against the request server.

    python -m compfinance_amd.serving [--requests 512] [--rate 200] [--budget-gib 4]

Mix (one draw per request, numpy PCG64 seeded 20261204): Poisson arrivals
at `rate` per second; model Black-Scholes (sigma uniform [0.1, 0.4]) or Dupire
(grid of 20..60 spots uniform on [50, 200] x 12..60 monthly times, the local-vol
formula of the benchmark grid with each of its three coefficients scaled by a
uniform [0.5, 1.5] factor, maxDt = 1/52); product European call, up-and-out
call (barrier uniform 1.1..1.6 x max(K, 100), 5% smoothing, weekly monitoring)
or Asian call (monthly fixings moved to the weekly grid); strike uniform
[80, 130]; maturity uniform [0.5, 5] years; paths log-uniform on 2^14..2^20;
half the requests with full AAD risk (aggregate = the first payoff); FP32.
"""
import argparse
import json
import math
import threading
import time

import numpy as np

import compfinance_amd as cf

WEEK = 1.0 / 52.0


def c5_requests(n=512, rate=200.0, seed=20261204):
    """List of dicts: arrival (s), model, product, num, aad, path_steps."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        T = float(rng.uniform(0.5, 5.0))
        K = float(rng.uniform(80.0, 130.0))
        if rng.random() < 0.5:
            mdl = cf.BlackScholes(100.0, float(rng.uniform(0.1, 0.4)))
        else:
            ns, nt = int(rng.integers(20, 61)), int(rng.integers(12, 61))
            a, b, c = (0.15, 0.10, 0.05) * rng.uniform(0.5, 1.5, 3)
            spots = 50.0 + 150.0 * np.arange(ns) / (ns - 1)
            times = (np.arange(nt) + 1) / 12.0
            vols = a + b * np.log(spots / 100.0)[:, None] ** 2 + c * np.exp(-times)[None, :]
            mdl = cf.Dupire(100.0, spots, times, vols, WEEK)
        kind = int(rng.integers(0, 3))
        if kind == 0:
            prd = cf.European(K, T)
        elif kind == 1:
            prd = cf.UOC(K, float(rng.uniform(1.1, 1.6)) * max(K, 100.0), T, WEEK, 0.05)
        else:
            prd = cf.Asian(K, T, 1.0 / 12, WEEK)
        paths = int(round(2.0 ** rng.uniform(14.0, 20.0)))
        aad = bool(rng.random() < 0.5)
        steps = len(cf.simulation_timeline(mdl, prd)[0]) - 1
        out.append(dict(arrival=float(t[i]), model=mdl, product=prd, aad=aad, paths=paths,
                        num=cf.NumericalParam(numPath=paths, precision=cf.FP32), path_steps=paths * steps))
    return out


def replay(reqs, server, time_scale=1.0):
    """Submits each request at its arrival time (open loop) from one thread
    while the caller collects results; returns (results, wall seconds)."""
    ids = [None] * len(reqs)
    ready = [threading.Event() for _ in reqs]
    errors = []

    def submitter():
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            delay = t0 + r["arrival"] * time_scale - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            try:
                ids[i] = server.submit(r["product"], r["model"], r["num"], aad=r["aad"])
            except Exception as e:  # noqa: BLE001 - reported with the request
                errors.append((i, e))
            ready[i].set()

    t_start = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    results = []
    for i in range(len(reqs)):
        ready[i].wait()
        results.append(server.wait(ids[i]) if ids[i] is not None else None)
    th.join()
    wall = time.perf_counter() - t_start
    if errors:
        raise RuntimeError("submission failed: %r" % (errors[:3],))
    return results, wall


def summarize(reqs, results, wall, stats):
    lat = np.array([r["latency_ms"] for r in results])
    kern = np.array([r["kernel_ms"] for r in results])
    ps = sum(r["path_steps"] for r in reqs)
    return {
        "requests": len(reqs),
        "failed": int(stats["failed"]),
        "aad_requests": int(sum(r["aad"] for r in reqs)),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "throughput_req_per_s": len(reqs) / wall,
        "paths_steps_per_s": ps / wall,
        "offered_rate_per_s": len(reqs) / reqs[-1]["arrival"],
        "wall_s": wall,
        "device_kernel_ms_sum": float(kern.sum()),
        "peak_running": int(stats["peak_running"]),
        "admitted_while_busy": int(stats["admitted_busy"]),
        "memory_waits": int(stats["memory_waits"]),
        "recompute_switches": int(stats["recompute_switches"]),
        "arena_bytes": int(stats["arena_bytes"]),
        "peak_arena_bytes": int(stats["peak_arena_bytes"]),
        "recomputed_requests": int(sum(r["recomputed"] for r in results)),
        "graph_captures": int(stats["graph_captures"]),
        "graph_replays": int(stats["graph_replays"]),
        "packed_launches": int(stats["packed_launches"]),
        "packed_requests": int(stats["packed_requests"]),
    }


def run_c5(n=512, rate=200.0, slots=4, budget=4 << 30, seed=20261204, warmup=16, max_pack=8):
    """The C5 replay: n requests of the seeded mix at `rate` per second."""
    reqs = c5_requests(n, rate, seed)
    with cf.Server(slots=slots, hbmBudget=budget, maxPack=max_pack) as srv:
        # warm-up: kernel loading and first-touch of the arena, not timed
        for r in reqs[:warmup]:
            srv.wait(srv.submit(r["product"], r["model"], r["num"], aad=r["aad"]))
        base = srv.stats()
        results, wall = replay(reqs, srv)
        st = srv.stats()
        for k in ("admitted_busy", "memory_waits", "recompute_switches", "graph_captures", "graph_replays",
                  "packed_launches", "packed_requests"):
            st[k] -= base[k]
    return summarize(reqs, results, wall, st)


def run_burst(n=512, slots=4, budget=4 << 30, seed=20261204, max_pack=8):
    """The same mix submitted at once (saturation): throughput, latencies and
    how many requests shared packed launches."""
    return run_c5(n, 1e9, slots, budget, seed, max_pack=max_pack)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gib", type=float, default=4.0)
    ap.add_argument("--max-pack", type=int, default=8)
    a = ap.parse_args()
    print(json.dumps(run_c5(a.requests, a.rate, a.slots, int(a.budget_gib * (1 << 30)), max_pack=a.max_pack)))


if __name__ == "__main__":
    main()
