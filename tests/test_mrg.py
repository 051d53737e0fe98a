"""mrg32k3a (mrg32k3a.h): the oracle against the reference build's golden
vectors (per-path payoffs, parallel means, AAD risks), and on the GPU the
kernels' own generators (stepped forward, and back in the reverse sweeps)
against the oracle, per path and in risks, over path ranges at any start,
in the book kernels and on models with zero local vols."""
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
    # path ranges reproduce the full run (the last one starts at an odd path)
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


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [cf.AAD_CHECKPOINT, cf.AAD_RECOMPUTE])
@pytest.mark.parametrize("mname", ["lv30", "bs_spot"])
def test_gpu_mrg_book_matches_oracle(engine, mname, mode):
    """The book kernels draw from the same generators: FP64 per-path payoffs
    and aggregated risks against the oracle tape."""
    from test_gpu_book import SMALL, book, model
    mdl, om = model(mname)
    prd, op = book(SMALL)
    n, seeds = 900, (12345, 67890)
    npay = len(prd.payoffLabels())
    w = np.linspace(1.5, -0.5, npay)
    res = cf.mcParallelSimulAAD(prd, mdl, weights=w, per_path=True,
                                num=cf.NumericalParam(numPath=n, useSobol=False, seed1=seeds[0], seed2=seeds[1],
                                                      aadMode=mode))
    pay, agg, risks = po.aad(om, op, n, weights=w, mrg=seeds)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.aggregated, agg, rtol=1e-10, atol=1e-10)
    assert np.abs(res.risks - risks).max() <= 1e-10 * np.abs(risks).max()


def _zero_vol_model():
    """The 30 x 36 local-vol grid with every node at or below spot 95 set to
    0: paths that fall there stop moving, and the reverse sweeps meet steps
    whose vol is exactly 0 (the Dupire checkpoint adjoint draws those normals
    again)."""
    from compfinance_amd import workloads as W
    mk = W.dupire_args(30)
    vols = np.array(mk["vols"], dtype=np.float64).copy()
    vols[np.asarray(mk["spots"]) <= 95.0, :] = 0.0
    return dict(mk, vols=vols)


@pytest.mark.gpu
@pytest.mark.parametrize("rng", ["sobol", "mrg"])
@pytest.mark.parametrize("mode", [cf.AAD_CHECKPOINT, cf.AAD_RECOMPUTE])
def test_gpu_zero_local_vols_match_oracle(engine, rng, mode):
    """FP64 risks and payoffs on a model with zero local vols against the
    oracle tape (Sobol and mrg32k3a), and the two FP32 adjoints against each
    other on the same paths."""
    from compfinance_amd import workloads as W
    mk = _zero_vol_model()
    pk = W.barrier_args(strike=100.0, barrier=140.0, maturity=2.0, put=True)
    mdl, prd = cf.Dupire(**mk), cf.UOC(**pk)
    om, op = po.Model(**mk), po.Product(**pk)
    seeds = (2024, 77)
    kw = dict(useSobol=False, seed1=seeds[0], seed2=seeds[1]) if rng == "mrg" else {}
    n = 1000
    w = [1.0, -0.5]   # the barrier option and its European
    res = cf.mcParallelSimulAAD(prd, mdl, weights=w, per_path=True,
                                num=cf.NumericalParam(numPath=n, aadMode=mode, **kw))
    pay, _, risks = po.aad(om, op, n, weights=w, mrg=seeds if rng == "mrg" else None)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    assert np.abs(res.risks - risks).max() <= 1e-10 * np.abs(risks).max()
    # the zero nodes carry risk: the redraw is exercised
    vega = np.asarray(risks[1:]).reshape(mk["vols"].shape)
    assert np.abs(vega[mk["vols"] == 0.0]).max() > 1e-3 * np.abs(vega).max()
    if mode != cf.AAD_CHECKPOINT:
        return
    # FP32: the checkpoint adjoint (derivatives recovered from the states,
    # normals drawn again below the threshold) against the recompute adjoint
    # (every normal drawn again) on the same FP32 paths, and the price against
    # FP64.  Two thirds of the paths here end within 0.5 of the 91.38 knot,
    # where the vol falls to 0 and the state stops; the vegas of that knot
    # amplify FP32's per-path differences (1e-4 of a payoff) to a few percent
    # against FP64, a property of the model and not of the sweeps.
    n32 = 1 << 16
    num = dict(numPath=n32, precision=cf.FP32, **kw)
    ck = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(aadMode=cf.AAD_CHECKPOINT, **num))
    rc = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(aadMode=cf.AAD_RECOMPUTE, **num))
    assert np.abs(ck.risks - rc.risks).max() <= 1e-5 * np.abs(rc.risks).max()
    r64 = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=cf.NumericalParam(numPath=n32, aadMode=mode, **kw))
    assert abs(ck.agg_mean - r64.agg_mean) <= 1e-5 * abs(r64.agg_mean)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [cf.AAD_CHECKPOINT, cf.AAD_RECOMPUTE])
def test_gpu_mrg_fp32_risks_follow_fp64(engine, mode):
    """FP32 mrg32k3a adjoints (the FP32 kernels' generators and segment
    draws) against FP64 on the same paths: C3's barrier at 2^18 paths."""
    from compfinance_amd import workloads as W
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    kw = dict(numPath=1 << 18, useSobol=False, seed1=42, seed2=43, aadMode=mode)
    r64 = cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.0], num=cf.NumericalParam(**kw))
    r32 = cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.0], num=cf.NumericalParam(precision=cf.FP32, **kw))
    assert abs(r32.agg_mean - r64.agg_mean) <= 1e-4 * abs(r64.agg_mean)
    assert np.abs(r32.risks - r64.risks).max() <= 2e-3 * np.abs(r64.risks).max()


def test_fp32_word_of_the_mrg_uniform():
    """The FP32 kernels map the uniform d / (m1 + 1), d in [1, m1], to the
    word w = d + floor(208 d / 2^32) of the FP32 inverse normal: within one
    unit of 2^32 u (exact rationals), never 0 (outside the table) and never
    past 2^32 - 1, at both ends and on random d."""
    from fractions import Fraction
    m1 = 4294967087
    rng = np.random.default_rng(5)
    ds = [1, 2, 208, 20650, 2 ** 31, m1 - 1, m1] + [int(v) for v in rng.integers(1, m1 + 1, 2000)]
    for d in ds:
        w = d + ((d * 208) >> 32)
        assert 1 <= w <= 2 ** 32 - 2
        assert abs(w - Fraction(d * 2 ** 32, m1 + 1)) <= 1
