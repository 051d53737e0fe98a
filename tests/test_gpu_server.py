"""The request server: results in the mix equal to the same requests run
alone, overlap of requests on the device, admission by device footprint
(waits, switch to the recompute variant), errors at submission."""
import numpy as np
import pytest

import compfinance_amd as cf
from compfinance_amd import serving
from compfinance_amd import workloads as W

pytestmark = pytest.mark.gpu


def _alone(r):
    if r["aad"]:
        a = cf.mcParallelSimulAAD(r["product"], r["model"], num=r["num"])
        return a.payoff_means, a.risks, a.agg_mean
    a = cf.mcParallelSimul(r["product"], r["model"], num=r["num"], per_path=False)
    return a.means, None, None


def test_results_in_the_mix_equal_results_alone(engine):
    """24 uneven requests submitted at once to 4 slots: prices bit for bit,
    aggregates and risks to the rounding of the LDS-atomic sums (1e-12)."""
    reqs = serving.c5_requests(24, rate=1e9, seed=7)
    with cf.Server(slots=4, hbmBudget=8 << 30) as srv:
        ids = [srv.submit(r["product"], r["model"], r["num"], aad=r["aad"]) for r in reqs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == 24 and st["failed"] == 0
    assert st["peak_running"] >= 2 and st["admitted_busy"] >= 1     # requests overlapped
    assert st["recompute_switches"] == 0
    for r, got in zip(reqs, res):
        means, risks, agg = _alone(r)
        np.testing.assert_array_equal(got["means"], means)
        if r["aad"]:
            assert abs(got["agg_mean"] - agg) <= 1e-12 * max(abs(agg), 1e-300)
            np.testing.assert_allclose(got["risks"], risks, rtol=1e-12, atol=1e-12 * np.abs(risks).max())


def test_full_c5_stream_in_the_mix_equals_alone(engine):
    """The whole 512-request serving stream (seed 20261204), submitted at
    once: every request's prices equal its synchronous run bit for bit, its
    aggregate and risks to 1e-12."""
    reqs = serving.c5_requests()
    assert len(reqs) == 512
    with cf.Server(slots=4, hbmBudget=8 << 30) as srv:
        ids = [srv.submit(r["product"], r["model"], r["num"], aad=r["aad"]) for r in reqs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == 512 and st["failed"] == 0
    assert st["peak_running"] >= 2
    for r, got in zip(reqs, res):
        means, risks, agg = _alone(r)
        np.testing.assert_array_equal(got["means"], means)
        if r["aad"]:
            assert abs(got["agg_mean"] - agg) <= 1e-12 * max(abs(agg), 1e-300)
            np.testing.assert_allclose(got["risks"], risks, rtol=1e-12, atol=1e-12 * np.abs(risks).max())


def _small_mix():
    """Small requests of two kernels: FP32 Dupire barrier AAD on grids of
    20..40 spots (different LDS sizes) and Black-Scholes European prices."""
    rng = np.random.default_rng(5)
    out = []
    for k in range(16):
        if k % 2 == 0:
            ns = 20 + 2 * k
            spots = 50.0 + 150.0 * np.arange(ns) / (ns - 1)
            times = (np.arange(24) + 1) / 12.0
            vols = 0.15 + 0.10 * np.log(spots / 100.0)[:, None] ** 2 + 0.05 * np.exp(-times)[None, :]
            mdl = cf.Dupire(100.0, spots, times, vols, 1 / 52)
            prd = cf.UOC(float(rng.uniform(90, 120)), float(rng.uniform(140, 180)), 1.5, 1 / 52, 0.05)
            out.append((prd, mdl, cf.NumericalParam(numPath=int(2 ** rng.uniform(12, 15)), precision=cf.FP32), True))
        else:
            mdl = cf.BlackScholes(100.0, float(rng.uniform(0.1, 0.4)))
            prd = cf.European(float(rng.uniform(80, 130)), float(rng.uniform(0.5, 3.0)))
            out.append((prd, mdl, cf.NumericalParam(numPath=int(2 ** rng.uniform(12, 15)), precision=cf.FP32), False))
    return out


def test_small_requests_share_packed_launches(engine):
    """Sixteen small requests of two kernels queued behind a large one in a
    one-slot server run in packed launches (one path-kernel and one
    reduction launch for several requests of a kernel); each request's
    prices equal its synchronous run bit for bit, its risks to the rounding
    of the LDS-atomic sums.  With maxPack=1 nothing is packed."""
    mix = _small_mix()
    big = (cf.UOC(**W.barrier_args()), cf.Dupire(**W.dupire_args(30)),   # fills the device for ~15 ms
           cf.NumericalParam(numPath=1 << 23, precision=cf.FP64, aadMode=cf.AAD_RECOMPUTE), True)
    # the synchronous runs first: they also load the kernels, whose first use
    # can hold a submission until the device is idle
    refs = [cf.mcParallelSimulAAD(p, m, num=n) if a else cf.mcParallelSimul(p, m, num=n, per_path=False)
            for p, m, n, a in mix]
    for max_pack in (8, 1):
        with cf.Server(slots=1, hbmBudget=4 << 30, maxPack=max_pack) as srv:
            first = srv.submit(big[0], big[1], big[2], aad=big[3])
            ids = [srv.submit(p, m, n, aad=a) for p, m, n, a in mix]
            srv.wait(first)
            res = [srv.wait(i) for i in ids]
            st = srv.stats()
        assert st["completed"] == 17 and st["failed"] == 0
        if max_pack == 8:
            assert st["packed_launches"] >= 1 and st["packed_requests"] >= 2, (st, [g["queue_ms"] for g in res])
            assert max(g["packed"] for g in res) >= 2
        else:
            assert st["packed_launches"] == 0 and all(g["packed"] == 1 for g in res), st
        for (p, m, n, a), ref, got in zip(mix, refs, res):
            if a:
                np.testing.assert_array_equal(got["means"], ref.payoff_means)
                assert abs(got["agg_mean"] - ref.agg_mean) <= 1e-12 * max(abs(ref.agg_mean), 1e-300)
                np.testing.assert_allclose(got["risks"], ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())
            else:
                np.testing.assert_array_equal(got["means"], ref.means)


def test_admission_waits_for_memory_and_switches_to_recompute(engine):
    """A budget of one checkpoint block plus one recompute block: the second
    AAD request runs the recompute variant, the next ones wait for memory;
    each gives the results of its variant alone."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    num = cf.NumericalParam(numPath=1 << 20, precision=cf.FP32)
    ck = cf.aad_footprint(prd, mdl, num)
    rc = cf.aad_footprint(prd, mdl, cf.NumericalParam(numPath=1 << 20, precision=cf.FP32,
                                                      aadMode=cf.AAD_RECOMPUTE))
    budget = ck + rc + (1 << 20)
    with cf.Server(slots=4, hbmBudget=budget) as srv:
        ids = [srv.submit(prd, mdl, num, aad=True) for _ in range(6)]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == 6 and st["peak_arena_bytes"] <= budget
    assert st["memory_waits"] > 0 and st["recompute_switches"] > 0
    ref_ck = cf.mcParallelSimulAAD(prd, mdl, num=num)
    ref_rc = cf.mcParallelSimulAAD(prd, mdl, num=cf.NumericalParam(numPath=1 << 20, precision=cf.FP32,
                                                                   aadMode=cf.AAD_RECOMPUTE))
    for g in res:
        ref = ref_rc if g["recomputed"] else ref_ck
        np.testing.assert_array_equal(g["means"], ref.payoff_means)
        np.testing.assert_allclose(g["risks"], ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())


def test_book_request_and_errors(engine):
    mdl = cf.Dupire(**W.dupire_args(30))
    book = cf.Book([cf.Europeans([3.0] * 3, [90.0, 100.0, 110.0]), cf.UOC(**W.barrier_args()),
                    cf.Asian(100.0, 3.0, 1 / 12, 1 / 52)])
    num = cf.NumericalParam(numPath=1 << 16, precision=cf.FP32)
    w = np.array([1.0, 0.5, 2.0, 1.5, 0.0, 0.7])
    with cf.Server(slots=2, hbmBudget=64 << 20) as srv:
        rid = srv.submit(book, mdl, num, aad=True, weights=w)
        with pytest.raises(cf.CfError):
            srv.submit(cf.European(100.0, -1.0), mdl, num)            # invalid product
        with pytest.raises(cf.CfError):                               # beyond the arena
            srv.submit(cf.UOC(**W.barrier_args(maturity=5.0)), mdl, cf.NumericalParam(numPath=1 << 20),
                       aad=True)
        got = srv.wait(rid)
    ref = cf.mcParallelSimulAAD(book, mdl, weights=w, num=num)
    np.testing.assert_array_equal(got["means"], ref.payoff_means)
    np.testing.assert_allclose(got["risks"], ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())


def test_open_loop_replay_small(engine):
    """A short C5-style replay: every request completes, latencies positive."""
    reqs = serving.c5_requests(40, rate=400.0, seed=11)
    with cf.Server(slots=4, hbmBudget=4 << 30) as srv:
        res, wall = serving.replay(reqs, srv)
        st = srv.stats()
    s = serving.summarize(reqs, res, wall, st)
    assert s["requests"] == 40 and st["failed"] == 0
    assert 0 < s["p50_ms"] <= s["p99_ms"]


def test_repeated_requests_replay_a_captured_graph(engine):
    """One slot, the same two requests five times each, one at a time: each
    kind launches directly twice, is captured at its third sighting and
    replayed after; results equal the direct launches of a server without
    graphs and the synchronous entry points."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    num = cf.NumericalParam(numPath=1 << 16, precision=cf.FP32)
    bs, eu = cf.BlackScholes(100.0, 0.2), cf.European(100.0, 1.0)
    kinds = [(prd, mdl, num, True), (eu, bs, cf.NumericalParam(numPath=1 << 14), False)]
    out = {}
    for graphs in (True, False):
        with cf.Server(slots=1, hbmBudget=1 << 30, graphs=graphs) as srv:
            out[graphs] = [srv.wait(srv.submit(p, m, n, aad=a)) for _ in range(5) for p, m, n, a in kinds]
            st = srv.stats()
        if graphs:
            assert st["graph_captures"] == 2 and st["graph_replays"] == 4, st
        else:
            assert st["graph_captures"] == 0 and st["graph_replays"] == 0, st
    ref_aad = cf.mcParallelSimulAAD(prd, mdl, num=num)
    ref_val = cf.mcParallelSimul(eu, bs, num=kinds[1][2], per_path=False)
    for k, (g, d) in enumerate(zip(out[True], out[False])):
        np.testing.assert_array_equal(g["means"], d["means"])
        if k % 2 == 0:
            np.testing.assert_array_equal(g["means"], ref_aad.payoff_means)
            np.testing.assert_allclose(g["risks"], ref_aad.risks, rtol=1e-12,
                                       atol=1e-12 * np.abs(ref_aad.risks).max())
            np.testing.assert_allclose(g["risks"], d["risks"], rtol=1e-12, atol=1e-12 * np.abs(d["risks"]).max())
        else:
            np.testing.assert_array_equal(g["means"], ref_val.means)


def test_graphs_with_changing_seeds_and_weights(engine):
    """An mrg32k3a request with other seeds is another sequence; AAD
    requests that differ only in their weights share one graph and follow
    the weights."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    with cf.Server(slots=1, hbmBudget=1 << 30) as srv:
        seeds = [(12345, 1234), (12345, 1234), (777, 888), (12345, 1234), (12345, 1234)]
        mrg = [srv.wait(srv.submit(prd, mdl, cf.NumericalParam(numPath=1 << 14, useSobol=False, seed1=a, seed2=b)))
               for a, b in seeds]
        st0 = srv.stats()
        num = cf.NumericalParam(numPath=1 << 14, precision=cf.FP32)
        ws = [np.array([1.0, 0.0]), np.array([0.0, 1.0]), np.array([0.5, 2.0]), np.array([1.0, 0.0]),
              np.array([2.0, -1.0])]
        risk = [srv.wait(srv.submit(prd, mdl, num, aad=True, weights=w)) for w in ws]
        st1 = srv.stats()
    # captured at the third sighting, replayed after (with other weights too)
    assert st0["graph_captures"] == 1 and st0["graph_replays"] == 1, st0
    assert st1["graph_captures"] == 2 and st1["graph_replays"] == 3, st1
    for (a, b), g in zip(seeds, mrg):
        ref = cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(numPath=1 << 14, useSobol=False, seed1=a, seed2=b),
                                 per_path=False)
        np.testing.assert_array_equal(g["means"], ref.means)
    for w, g in zip(ws, risk):
        ref = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=num)
        assert abs(g["agg_mean"] - ref.agg_mean) <= 1e-12 * abs(ref.agg_mean)
        np.testing.assert_allclose(g["risks"], ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())


def test_second_wait_on_a_request_is_an_error(engine):
    """cf_server_wait releases a request: of two concurrent waits on one id
    exactly one gets the results, the other an error; a later wait is an
    error too."""
    import ctypes
    import threading
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    num = cf.NumericalParam(numPath=1 << 20, precision=cf.FP32)
    lib = cf._server_lib()
    with cf.Server(slots=1, hbmBudget=1 << 30) as srv:
        rid = srv.submit(prd, mdl, num, aad=True)
        means = [np.zeros(2), np.zeros(2)]
        outs = [cf.cf_result(), cf.cf_result()]
        rcs = [None, None]

        def waiter(k):
            outs[k].means, outs[k].cap_pay = cf._dptr(means[k]), 2
            rcs[k] = lib.cf_server_wait(srv._h, int(rid), ctypes.byref(outs[k]))

        ths = [threading.Thread(target=waiter, args=(k,)) for k in range(2)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        assert sorted(rc == 0 for rc in rcs) == [False, True], rcs
        got = means[rcs.index(0)]
        ref = cf.mcParallelSimulAAD(prd, mdl, num=num)
        np.testing.assert_array_equal(got, ref.payoff_means)
        assert lib.cf_server_wait(srv._h, int(rid), ctypes.byref(outs[0])) < 0   # released
