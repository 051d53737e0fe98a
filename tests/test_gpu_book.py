"""Books of products on shared paths and the Asian option on the GPU, against
the oracle (FP64 per path, risks against the oracle tape) and in FP32 within
0.1 Monte Carlo standard errors of FP64, through the C ABI."""
import numpy as np
import pytest

import compfinance_amd as cf
import pyoracle as po
from compfinance_amd import workloads as W
from test_gpu_parity import MODELS

pytestmark = pytest.mark.gpu

CLS = {"european": cf.European, "europeans": cf.Europeans, "barrier": cf.UOC, "asian": cf.Asian}

SMALL = [("europeans", dict(maturities=[3.0, 3.0, 1.5], strikes=[90.0, 120.0, 100.0])),
         ("barrier", W.barrier_args(strike=100.0, barrier=140.0)),
         ("barrier", W.barrier_args(strike=110.0, barrier=135.0, put=True)),
         ("european", dict(strike=105.0, exerciseDate=2.0)),
         ("asian", dict(strike=95.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52)),
         ("asian", dict(strike=105.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52, callPut=True)),
         ("asian", dict(strike=100.0, maturity=2.0, fixingFreq=0.25, snapFreq=1 / 52))]


def book(parts):
    return cf.Book([CLS[k](**kw) for k, kw in parts]), po.Product(parts=[kw for _, kw in parts])


def model(name):
    mk = MODELS[name]
    return (cf.Dupire(**mk) if "spots" in mk else cf.BlackScholes(**mk)), po.Model(**mk)


@pytest.mark.parametrize("mname", ["lv30", "bs", "bs_spot"])
def test_fp64_book_per_path_matches_oracle(engine, mname):
    mdl, om = model(mname)
    prd, op = book(SMALL)
    n = 1500
    res = cf.mcParallelSimul(prd, mdl, n)
    ref = po.value_paths(om, op, n)
    np.testing.assert_allclose(res.payoffs, ref, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("mode", [cf.AAD_CHECKPOINT, cf.AAD_RECOMPUTE])
@pytest.mark.parametrize("mname", ["lv30", "lv60", "bs", "bs_spot"])
def test_fp64_book_risks_match_oracle_tape(engine, mname, mode):
    mdl, om = model(mname)
    prd, op = book(SMALL)
    n = 1200
    npay = len(prd.payoffLabels())
    w = np.linspace(1.5, -0.5, npay)
    res = cf.mcParallelSimulAAD(prd, mdl, weights=w, per_path=True, num=cf.NumericalParam(numPath=n, aadMode=mode))
    pay, agg, risks = po.aad(om, op, n, weights=w)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.aggregated, agg, rtol=1e-10, atol=1e-10)
    scale = max(np.abs(risks).max(), 1e-300)
    assert np.abs(res.risks - risks).max() <= 1e-10 * scale, np.abs(res.risks - risks).max() / scale


@pytest.mark.parametrize("kw", [dict(strike=100.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52),
                                dict(strike=100.0, maturity=1.0, fixingFreq=0.1),
                                dict(strike=110.0, maturity=2.0, fixingFreq=1 / 12, callPut=True)])
@pytest.mark.parametrize("mname", ["lv30", "bs_spot"])
def test_asian_alone_matches_oracle(engine, mname, kw):
    mdl, om = model(mname)
    prd, op = cf.Asian(**kw), po.Product(**kw)
    n = 1000
    res = cf.mcParallelSimulAAD(prd, mdl, n, weights=[1.0], per_path=True)
    pay, _, risks = po.aad(om, op, n, weights=[1.0])
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    assert np.abs(res.risks - risks).max() <= 1e-10 * max(np.abs(risks).max(), 1e-300)


def test_c4_book_matches_oracle(engine):
    """The full C4 book: 30 Europeans, 10 barriers, 5 Asians (55 payoffs)."""
    mdl, om = model("lv30")
    prd, op = book(W.c4_parts())
    npay = len(prd.payoffLabels())
    assert npay == 55
    w = W.c4_notionals(npay)
    n = 700
    res = cf.mcParallelSimulAAD(prd, mdl, weights=w, per_path=True, num=cf.NumericalParam(numPath=n))
    pay, _, risks = po.aad(om, op, n, weights=w)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    assert np.abs(res.risks - risks).max() <= 1e-10 * np.abs(risks).max()


def test_c4_fp32_within_tenth_of_standard_error(engine):
    """FP32 book values and aggregated risks against FP64 on 2^20 paths, in
    standard errors from 32 disjoint batches."""
    mdl, _ = model("lv30")
    prd, _ = book(W.c4_parts())
    npay = len(prd.payoffLabels())
    w = W.c4_notionals(npay)
    n, batches = 1 << 20, 32
    full64 = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(numPath=n))
    full32 = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(numPath=n, precision=cf.FP32))
    m = n // batches
    per = [cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(numPath=m, firstPath=b * m))
           for b in range(batches)]
    risks = np.array([r.risks for r in per])
    vals = np.array([r.payoff_means for r in per])
    np.testing.assert_allclose(risks.mean(0), full64.risks, rtol=1e-9, atol=1e-14)
    se = risks.std(0, ddof=1) / np.sqrt(batches)
    live = se > 1e-6 * se.max()
    z = np.abs(full32.risks - full64.risks)[live] / se[live]
    assert z.max() <= 0.1, (z.max(), np.argmax(z))
    sev = vals.std(0, ddof=1) / np.sqrt(batches)
    ok = sev > 0
    assert np.all(np.abs(full32.payoff_means - full64.payoff_means)[ok] <= 0.1 * sev[ok])


def test_book_risk_is_the_weighted_sum_of_its_parts(engine):
    """Aggregated book risk = sum of notional x part risk, each part run alone
    on the same paths (parts sharing the book's weekly timeline)."""
    mdl, _ = model("lv30")
    parts = [p for p in W.c4_parts() if p[0] == "barrier"][:4] + [W.c4_parts()[0]]
    prd, _ = book(parts)
    npay = len(prd.payoffLabels())
    w = W.c4_notionals(npay)
    n = 1 << 16
    rb = cf.mcParallelSimulAAD(prd, mdl, n, weights=w)
    tot, col = 0.0, 0
    for k, kw in parts:
        p = CLS[k](**kw)
        npk = len(p.payoffLabels())
        tot = tot + cf.mcParallelSimulAAD(p, mdl, n, weights=w[col:col + npk]).risks
        col += npk
    np.testing.assert_allclose(rb.risks, tot, rtol=1e-9, atol=1e-12 * np.abs(tot).max())


@pytest.mark.parametrize("prec", [cf.FP64, cf.FP32])
@pytest.mark.parametrize("mname", ["lv30", "bs"])
def test_itemized_book_risks_are_the_one_hot_runs(engine, mname, prec):
    """mcParallelSimulAADMulti on a book (one forward kept, one reverse per
    payoff): column k equals the single AAD run with one-hot weights on
    payoff k (the same kernels on the same paths, to the rounding of the LDS
    atomic sums), and the means equal those runs' bit for bit."""
    mdl, _ = model(mname)
    prd, _ = book(SMALL)
    n = 3000                                   # ragged: not a multiple of 64
    num = cf.NumericalParam(numPath=n, precision=prec)
    means, risks = cf.mcParallelSimulAADMulti(prd, mdl, num=num)
    npay = len(prd.payoffLabels())
    assert risks.shape == (mdl.numParams(), npay)
    for k in range(npay):
        w = np.zeros(npay)
        w[k] = 1.0
        one = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=num)
        np.testing.assert_array_equal(means, one.payoff_means)
        ref = one.risks
        tol = (1e-12 if prec == cf.FP64 else 1e-6) * max(np.abs(ref).max(), 1e-300)
        assert np.abs(risks[:, k] - ref).max() <= tol, (k, np.abs(risks[:, k] - ref).max())


def test_itemized_c4_book_risks_sum_to_the_aggregate(engine):
    """The C4 book (55 payoffs, 2^20 paths, FP32): the notional-weighted sum
    of the itemized risks is the aggregated risk of the same paths."""
    mdl = cf.Dupire(**W.dupire_args(30))
    prd = cf.Book([CLS[k](**kw) for k, kw in W.c4_parts()])
    w = W.c4_notionals(len(prd.payoffLabels()))
    num = cf.NumericalParam(numPath=W.C4_PATHS, precision=cf.FP32)
    means, risks = cf.mcParallelSimulAADMulti(prd, mdl, num=num)
    agg = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=num)
    np.testing.assert_allclose(means, agg.payoff_means, rtol=1e-12, atol=0)
    combined = risks @ w
    assert np.abs(combined - agg.risks).max() <= 1e-5 * np.abs(agg.risks).max()
