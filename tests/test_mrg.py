"""mrg32k3a (mrg32k3a.h): the oracle against the reference build's golden
vectors (per-path payoffs, parallel means, AAD risks), and on the GPU the
engine's pre-generated normals against the oracle, per path and in risks,
with path ranges that start at even paths (the reference's batches)."""
import json
import os

import numpy as np
import pytest

import compfinance_amd as cf
import pyoracle as po

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "mrg.json")))


def _models(case):
    mk = GOLDEN["models"][case["model"]]
    mk = dict(mk, vols=np.array(mk["vols"])) if "vols" in mk else mk
    return mk, case["product"]


@pytest.mark.parametrize("case", GOLDEN["cases"], ids=[c["name"] for c in GOLDEN["cases"]])
def test_oracle_mrg_matches_reference(case):
    mk, pk = _models(case)
    m, p = po.Model(**mk), po.Product(**pk)
    seeds = tuple(case["seeds"])
    pp = po.value_paths(m, p, 128, mrg=seeds)
    np.testing.assert_array_equal(pp.ravel(), case["per_path"])
    means = po.value_paths(m, p, case["n_value"], mrg=seeds).mean(0)
    np.testing.assert_allclose(means, case["values"], rtol=1e-13, atol=1e-15)
    _, _, risks = po.aad(m, p, case["n_aad"], 0, mrg=seeds)
    np.testing.assert_allclose(risks, case["risks"], rtol=1e-12, atol=1e-14 * np.abs(risks).max())


def test_oracle_mrg_antithetic_pairs():
    mk, pk = _models(GOLDEN["cases"][1])
    m, p = po.Model(**mk), po.Product(strike=0.0, exerciseDate=1.0)
    pp = po.value_paths(m, p, 64, mrg=(5, 6))
    # one exact Black-Scholes step, zero strike: the pair's log-payoffs are
    # drift + std g and drift - std g, so their sums are all equal
    lp = np.log(pp[:, 0])
    sums = lp[0::2] + lp[1::2]
    assert np.ptp(sums) < 1e-12 and np.ptp(lp) > 0.1


def _engine(mk, pk):
    mdl = cf.Dupire(**mk) if "spots" in mk else cf.BlackScholes(**mk)
    if "barrier" in pk:
        prd = cf.UOC(**pk)
    elif "maturities" in pk:
        prd = cf.Europeans(**pk)
    else:
        prd = cf.European(**pk)
    return mdl, prd


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLDEN["cases"], ids=[c["name"] for c in GOLDEN["cases"]])
def test_gpu_mrg_matches_oracle(engine, case):
    mk, pk = _models(case)
    mdl, prd = _engine(mk, pk)
    seeds = tuple(case["seeds"])
    om, op = po.Model(**mk), po.Product(**pk)
    num = cf.NumericalParam(numPath=1000, useSobol=False, seed1=seeds[0], seed2=seeds[1])
    res = cf.mcParallelSimul(prd, mdl, num=num)
    np.testing.assert_allclose(res.payoffs, po.value_paths(om, op, 1000, mrg=seeds), rtol=1e-11, atol=1e-11)
    npay = len(prd.payoffLabels())
    w = np.linspace(1.0, 0.5, npay)
    for mode in (cf.AAD_CHECKPOINT, cf.AAD_RECOMPUTE):
        n2 = cf.NumericalParam(numPath=800, useSobol=False, seed1=seeds[0], seed2=seeds[1], aadMode=mode)
        r = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=n2)
        _, _, risks = po.aad(om, op, 800, weights=w, mrg=seeds)
        assert np.abs(r.risks - risks).max() <= 1e-10 * np.abs(risks).max()
    # ranges starting at even paths reproduce the full run
    full = cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(numPath=640, useSobol=False, seed1=seeds[0],
                                                               seed2=seeds[1])).payoffs
    parts = [cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(numPath=c, firstPath=f, useSobol=False,
                                                                seed1=seeds[0], seed2=seeds[1])).payoffs
             for f, c in [(0, 64), (64, 321), (385, 255)]]
    np.testing.assert_array_equal(np.concatenate(parts), full)
    # the reference's own means at 4096 paths
    v = cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(numPath=case["n_value"], useSobol=False,
                                                          seed1=seeds[0], seed2=seeds[1]), per_path=False)
    np.testing.assert_allclose(v.means, case["values"], rtol=1e-10, atol=1e-12)


@pytest.mark.gpu
def test_gpu_mrg_fp32_within_tenth_of_standard_error(engine):
    from compfinance_amd import workloads as W
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    kw = dict(numPath=1 << 18, useSobol=False, seed1=42, seed2=43)
    r64 = cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(**kw), per_path=False)
    r32 = cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(precision=cf.FP32, **kw), per_path=False)
    assert np.all(np.abs(r32.means - r64.means) <= 0.1 * r64.std_errs)
