"""The recompute AAD variant (no per-step checkpoints: Dupire states kept every
8 steps in FP32 / 4 in FP64 and segments re-simulated from the Sobol index;
Black-Scholes normals regenerated) against the oracle tape and against the
checkpoint variant on the same paths."""
import numpy as np
import pytest

import compfinance_amd as cf
import pyoracle as po
from compfinance_amd import workloads as W
from test_gpu_parity import COMBOS, MODELS, PRODUCTS, make

pytestmark = pytest.mark.gpu

REC = cf.AAD_RECOMPUTE


@pytest.mark.parametrize("mname,pname", COMBOS)
def test_fp64_recompute_risks_match_oracle_tape(engine, mname, pname):
    mdl, prd, om, op = make(MODELS[mname], PRODUCTS[pname])
    n = 1500
    npay = len(prd.payoffLabels())
    w = np.linspace(1.0, 0.3, npay)
    res = cf.mcParallelSimulAAD(prd, mdl, weights=w, per_path=True, num=cf.NumericalParam(numPath=n, aadMode=REC))
    pay, agg, risks = po.aad(om, op, n, weights=w)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.aggregated, agg, rtol=1e-11, atol=1e-11)
    scale = max(np.abs(risks).max(), 1e-300)
    assert np.abs(res.risks - risks).max() <= 1e-10 * scale, np.abs(res.risks - risks).max() / scale


@pytest.mark.parametrize("steps_case", ["five_year_weekly", "one_time", "european_today", "short"])
@pytest.mark.parametrize("prec", [cf.FP64, cf.FP32])
def test_partial_segments(engine, steps_case, prec):
    """Timelines whose step count is not a multiple of the segment length."""
    mk = W.dupire_args(30)
    pk = W.barrier_args()
    if steps_case == "five_year_weekly":
        pk = W.barrier_args(maturity=5.0)            # 260 steps
    elif steps_case == "one_time":
        mk = dict(mk, times=[1.0], vols=mk["vols"][:, :1])
    elif steps_case == "european_today":
        pk = dict(strike=95.0, exerciseDate=0.0)     # no step
    elif steps_case == "short":
        pk = W.barrier_args(maturity=5.0 / 52)       # 5 steps
    mdl = cf.Dupire(**mk)
    prd = cf.European(**pk) if "exerciseDate" in pk else cf.UOC(**pk)
    npay = len(prd.payoffLabels())
    n = 777
    ck = cf.mcParallelSimulAAD(prd, mdl, weights=np.ones(npay), num=cf.NumericalParam(numPath=n, precision=prec))
    rc = cf.mcParallelSimulAAD(prd, mdl, weights=np.ones(npay),
                               num=cf.NumericalParam(numPath=n, precision=prec, aadMode=REC))
    np.testing.assert_array_equal(rc.payoff_means, ck.payoff_means)
    scale = max(np.abs(ck.risks).max(), 1e-300)
    tol = 1e-11 if prec == cf.FP64 else 2e-4
    assert np.abs(rc.risks - ck.risks).max() <= tol * scale, np.abs(rc.risks - ck.risks).max() / scale
    if prec == cf.FP64:
        om, op = po.Model(**mk), po.Product(**pk)
        _, _, risks = po.aad(om, op, n, weights=np.ones(npay))
        assert np.abs(rc.risks - risks).max() <= 1e-10 * max(np.abs(risks).max(), 1e-300)


def test_fp32_recompute_risks_within_tenth_of_standard_error(engine):
    """FP32 recompute price, delta and vegas against FP64 on 2^20 paths, in
    standard errors from 32 disjoint batches (the checkpoint variant's bar)."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    n, batches = 1 << 20, 32
    full64 = cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.0], num=cf.NumericalParam(numPath=n, aadMode=REC))
    full32 = cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.0],
                                   num=cf.NumericalParam(numPath=n, precision=cf.FP32, aadMode=REC))
    m = n // batches
    per = np.array([cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.0],
                                          num=cf.NumericalParam(numPath=m, firstPath=b * m)).risks
                    for b in range(batches)])
    np.testing.assert_allclose(per.mean(0), full64.risks, rtol=1e-9, atol=1e-14)
    se = per.std(0, ddof=1) / np.sqrt(batches)
    live = se > 1e-6 * se.max()
    z = np.abs(full32.risks - full64.risks)[live] / se[live]
    assert z.max() <= 0.1, (z.max(), np.argmax(z))


def test_recompute_footprint_is_smaller(engine):
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    # checkpoints 8x (FP32) / 4x (FP64) smaller; the per-block local-vol
    # adjoint grids are the same in both
    for prec in (cf.FP32, cf.FP64):
        ck = cf.aad_footprint(prd, mdl, cf.NumericalParam(numPath=1 << 20, precision=prec))
        rc = cf.aad_footprint(prd, mdl, cf.NumericalParam(numPath=1 << 20, precision=prec, aadMode=REC))
        assert rc < ck / 3, (prec, ck, rc)
    bs = cf.BlackScholes(100.0, 0.2, False, 0.0, 0.0)
    eu = cf.European(100.0, 1.0)
    ck = cf.aad_footprint(eu, bs, cf.NumericalParam(numPath=1 << 20))
    rc = cf.aad_footprint(eu, bs, cf.NumericalParam(numPath=1 << 20, aadMode=REC))
    assert rc < ck


def test_store_entry_points_take_the_mode(engine):
    cf.clearStore()
    cf.putDupire(**W.dupire_args(30), store="lv")
    cf.putBarrier(**W.barrier_args(), store="uoc")
    a = cf.AADriskOne("lv", "uoc", cf.NumericalParam(numPath=1 << 14))
    b = cf.AADriskOne("lv", "uoc", cf.NumericalParam(numPath=1 << 14, aadMode=REC))
    np.testing.assert_allclose(a["risks"], b["risks"], rtol=1e-10, atol=1e-16)
    cf.clearStore()


@pytest.mark.parametrize("n_spots", [255, 256])
@pytest.mark.parametrize("prec", [cf.FP64, cf.FP32])
def test_checkpoint_bracket_bytes_at_the_knot_limit(engine, n_spots, prec):
    """The checkpoint variant reloads each state's knot bracket from a byte
    for grids of up to 255 knots and searches again above: both sides of the
    limit give the recompute variant's risks and, in FP64, the oracle tape's.
    Half a year of weekly steps keeps the (steps x spots) adjoint grid in LDS."""
    mk = W.dupire_args(n_spots)
    pk = W.barrier_args(maturity=0.5)
    mdl, prd = cf.Dupire(**mk), cf.UOC(**pk)
    n = 600
    ck = cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.5], num=cf.NumericalParam(numPath=n, precision=prec))
    rc = cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.5],
                               num=cf.NumericalParam(numPath=n, precision=prec, aadMode=REC))
    np.testing.assert_array_equal(rc.payoff_means, ck.payoff_means)
    scale = max(np.abs(ck.risks).max(), 1e-300)
    tol = 1e-11 if prec == cf.FP64 else 2e-4
    assert np.abs(rc.risks - ck.risks).max() <= tol * scale, np.abs(rc.risks - ck.risks).max() / scale
    if prec == cf.FP64:
        _, _, risks = po.aad(po.Model(**mk), po.Product(**pk), n, weights=[1.0, 0.5])
        assert np.abs(ck.risks - risks).max() <= 1e-10 * max(np.abs(risks).max(), 1e-300)
