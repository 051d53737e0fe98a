"""GPU parity of the path kernels against the CPU oracle, on every model x
product combination, in FP64 (per path, 1e-11) and FP32 (in Monte Carlo
standard errors of the FP64 result on the same paths), through the C ABI."""
import numpy as np
import pytest

import compfinance_amd as cf
import pyoracle as po
from compfinance_amd import workloads as W

pytestmark = pytest.mark.gpu

BS = dict(spot=100.0, vol=0.25, rate=0.03, div=0.01)
BS_SPOT = dict(spot=100.0, vol=0.25, rate=0.03, div=0.01, spotMeasure=True)
EUS = dict(maturities=[1.0, 1.0, 2.0, 3.0, 3.0, 3.0], strikes=[90.0, 110.0, 100.0, 80.0, 120.0, 140.0])

MODELS = {
    "bs": BS,
    "bs_spot": BS_SPOT,
    "lv30": W.dupire_args(30),
    "lv60": W.dupire_args(60),
}
PRODUCTS = {
    "european": dict(strike=100.0, exerciseDate=1.5),
    "european_settle": dict(strike=105.0, exerciseDate=1.0, settlementDate=1.25),
    "barrier": W.barrier_args(),
    "barrier_put": W.barrier_args(strike=110.0, barrier=140.0, maturity=2.0, put=True),
    "europeans": EUS,
}


def make(mk, pk):
    mdl = cf.Dupire(**mk) if "spots" in mk else cf.BlackScholes(**mk)
    if "barrier" in pk:
        prd = cf.UOC(**pk)
    elif "maturities" in pk:
        prd = cf.Europeans(**pk)
    else:
        prd = cf.European(**pk)
    return mdl, prd, po.Model(**mk), po.Product(**pk)


COMBOS = [(m, p) for m in MODELS for p in PRODUCTS]


@pytest.mark.parametrize("mname,pname", COMBOS)
def test_fp64_per_path_matches_oracle(engine, mname, pname):
    mdl, prd, om, op = make(MODELS[mname], PRODUCTS[pname])
    n = 3000   # not a multiple of the wave size
    res = cf.mcParallelSimul(prd, mdl, n)
    ref = po.value_paths(om, op, n)
    np.testing.assert_allclose(res.payoffs, ref, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.means, ref.mean(0), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("mname,pname", COMBOS)
def test_fp64_risks_match_oracle_tape(engine, mname, pname):
    mdl, prd, om, op = make(MODELS[mname], PRODUCTS[pname])
    n = 1500
    npay = len(prd.payoffLabels())
    w = np.linspace(1.0, 0.3, npay)   # an aggregate of every payoff
    res = cf.mcParallelSimulAAD(prd, mdl, n, weights=w, per_path=True)
    pay, agg, risks = po.aad(om, op, n, weights=w)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.aggregated, agg, rtol=1e-11, atol=1e-11)
    scale = max(np.abs(risks).max(), 1e-300)
    assert np.abs(res.risks - risks).max() <= 1e-10 * scale, np.abs(res.risks - risks).max() / scale


@pytest.mark.parametrize("mname,pname", [("lv30", "barrier"), ("lv60", "barrier"), ("lv30", "europeans"),
                                         ("bs", "barrier"), ("bs_spot", "european")])
def test_fp32_prices_within_tenth_of_standard_error(engine, mname, pname):
    mdl, prd, _, _ = make(MODELS[mname], PRODUCTS[pname])
    n = 1 << 18
    r64 = cf.mcParallelSimul(prd, mdl, n, per_path=False, precision=cf.FP64)
    r32 = cf.mcParallelSimul(prd, mdl, n, per_path=False, precision=cf.FP32)
    ok = r64.std_errs > 0
    assert np.all(np.abs(r32.means - r64.means)[ok] <= 0.1 * r64.std_errs[ok]), (r32.means, r64.means)


@pytest.mark.parametrize("mname", ["lv30", "lv60"])
def test_fp32_c2_price_at_full_size(engine, mname):
    """The C2 pricing run at its configured size, 2^19 Sobol paths x 156
    weekly steps: the FP32 barrier prices within 0.1 standard errors of FP64
    on the same paths."""
    mdl, prd, _, _ = make(MODELS[mname], PRODUCTS["barrier"])
    n = 1 << 19
    r64 = cf.mcParallelSimul(prd, mdl, n, per_path=False, precision=cf.FP64)
    r32 = cf.mcParallelSimul(prd, mdl, n, per_path=False, precision=cf.FP32)
    assert np.all(r64.std_errs > 0)
    assert np.all(np.abs(r32.means - r64.means) <= 0.1 * r64.std_errs), (r32.means, r64.means, r64.std_errs)


def test_one_hot_aggregate_selects_the_payoff(engine):
    mdl, prd, om, op = make(MODELS["lv30"], PRODUCTS["barrier"])
    n = 2048
    r1 = cf.mcParallelSimulAAD(prd, mdl, n, weights=[0.0, 1.0])
    _, _, risks = po.aad(om, op, n, 1)
    assert np.abs(r1.risks - risks).max() <= 1e-10 * np.abs(risks).max()
