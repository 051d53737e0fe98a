"""GPU parity of the path kernels against the CPU oracle (FP64) and against
the FP64 GPU form (FP32), through the C ABI."""
import numpy as np
import pytest

import pyoracle as po
from compfinance_amd import workloads as W

pytestmark = pytest.mark.gpu


def _c1(cf):
    m, p = W.c1()
    return cf.BlackScholes(**m), cf.European(**p), po.Model(**m), po.Product(**p)


def _dupire_uoc(cf, n_spots=30, **bar):
    m, p = W.dupire_args(n_spots), W.barrier_args(**bar)
    return cf.Dupire(**m), cf.UOC(**p), po.Model(**m), po.Product(**p)


def test_c1_black_scholes_european_fp64_per_path(engine):
    cf = engine
    mdl, prd, om, op = _c1(cf)
    n = 1 << 16
    res = cf.mcParallelSimul(prd, mdl, n)
    ref = po.value_paths(om, op, n)
    np.testing.assert_allclose(res.payoffs, ref, rtol=1e-12, atol=1e-12)
    assert abs(res.means[0] - ref.mean()) <= 1e-12 * abs(ref.mean())
    # closed form 7.965567455405804 within 3 standard errors
    assert abs(res.means[0] - W.C1_CLOSED_FORM) < 3 * res.std_errs[0] + 1e-3


@pytest.mark.parametrize("n_spots", [30, 60])
def test_c2_dupire_barrier_fp64_per_path(engine, n_spots):
    cf = engine
    mdl, prd, om, op = _dupire_uoc(cf, n_spots)
    n = 1 << 13
    res = cf.mcParallelSimul(prd, mdl, n)
    ref = po.value_paths(om, op, n)
    np.testing.assert_allclose(res.payoffs, ref, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.means, ref.mean(0), rtol=1e-12)


def test_c2_dupire_barrier_fp32_within_se(engine):
    cf = engine
    mdl, prd, om, op = _dupire_uoc(cf)
    n = 1 << 16
    r64 = cf.mcParallelSimul(prd, mdl, n, per_path=False, precision=cf.FP64)
    r32 = cf.mcParallelSimul(prd, mdl, n, per_path=False, precision=cf.FP32)
    # FP32 within 0.1 Monte Carlo standard errors of the FP64 result on the same paths
    assert np.all(np.abs(r32.means - r64.means) <= 0.1 * r64.std_errs), (r32.means, r64.means, r64.std_errs)


def test_c3_dupire_barrier_aad_fp64(engine):
    cf = engine
    mdl, prd, om, op = _dupire_uoc(cf)
    n = 1 << 12
    res = cf.mcParallelSimulAAD(prd, mdl, n, weights=[1.0, 0.0], per_path=True)
    pay, agg, risks = po.aad(om, op, n, 0)
    np.testing.assert_allclose(res.payoffs, pay, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(res.aggregated, agg, rtol=1e-11, atol=1e-11)
    scale = np.abs(risks).max()
    assert np.abs(res.risks - risks).max() <= 1e-10 * scale, np.abs(res.risks - risks).max() / scale


def test_c1_black_scholes_aad_fp64(engine):
    cf = engine
    mdl, prd, om, op = _c1(cf)
    n = 1 << 14
    res = cf.mcParallelSimulAAD(prd, mdl, n)
    pay, agg, risks = po.aad(om, op, n, 0)
    np.testing.assert_allclose(res.risks, risks, rtol=1e-10, atol=1e-12)
