"""cf_mc_simul_aad_noalloc (no counterpart in the reference; for callers that
cannot allocate device memory) through the C ABI: on two calls in a row, the
second with a smaller input, its payoff means, aggregate and risks (and per
path payoffs) equal those of cf_mc_simul_aad, which runs the same kernels on
the same launch geometry; a call beyond its scratch array is refused."""
import ctypes

import numpy as np
import pytest

import compfinance_amd as cf
from compfinance_amd import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(prd, mdl, num, weights, per_path=False):
    L = cf.lib()
    f = L.cf_mc_simul_aad_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.POINTER(cf.cf_model), ctypes.POINTER(cf.cf_product), ctypes.POINTER(cf.cf_num)] + \
        [ctypes.c_void_p] * 6
    dm, dp, n = mdl.descriptor(), prd.descriptor(), num.descriptor()
    npay, nparam = len(prd.payoffLabels()), mdl.numParams()
    w = np.ascontiguousarray(weights, dtype=np.float64)
    means, agg, risks = np.zeros(npay), np.zeros(1), np.zeros(nparam)
    pp = np.zeros((n.num_path, npay)) if per_path else None
    pa = np.zeros(n.num_path) if per_path else None
    rc = f(ctypes.byref(dm), ctypes.byref(dp), ctypes.byref(n), w.ctypes.data, means.ctypes.data, agg.ctypes.data,
           risks.ctypes.data, pp.ctypes.data if per_path else None, pa.ctypes.data if per_path else None)
    return rc, means, float(agg[0]), risks, pp, pa


def _check_equal(got, ref, per_path=False):
    rc, means, agg, risks, pp, pa = got
    assert rc == 0, cf.lib().cf_last_error().decode()
    np.testing.assert_allclose(means, ref.payoff_means, rtol=1e-13, atol=0)
    assert abs(agg - ref.agg_mean) <= 1e-13 * max(abs(ref.agg_mean), 1e-300)
    np.testing.assert_allclose(risks, ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())
    if per_path:
        np.testing.assert_array_equal(pp, ref.payoffs)
        np.testing.assert_array_equal(pa, ref.aggregated)


def test_noalloc_equals_the_allocating_entry_point_on_two_calls(engine):
    # 1: the benchmark risk run, 2^20 paths, FP32, checkpoint adjoint
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    num = cf.NumericalParam(numPath=1 << 20, precision=cf.FP32)
    w = np.array([1.0, 0.0])
    ref = cf.mcParallelSimulAAD(prd, mdl, weights=w, num=num)
    _check_equal(_noalloc(prd, mdl, num, w), ref)
    # 2: a smaller input right after: 60 x 36 grid, 5000 paths, FP64, per-path outputs
    mdl2 = cf.Dupire(**W.dupire_args(60))
    num2 = cf.NumericalParam(numPath=5000, precision=cf.FP64)
    w2 = np.array([0.7, 0.3])
    ref2 = cf.mcParallelSimulAAD(prd, mdl2, weights=w2, num=num2, per_path=True)
    _check_equal(_noalloc(prd, mdl2, num2, w2, per_path=True), ref2, per_path=True)


def test_noalloc_with_mrg32k3a(engine):
    """mrg32k3a needs no scratch beyond Sobol's (the generators run in the
    kernel): the benchmark risk run with mrg32k3a, then a smaller book call."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    num = cf.NumericalParam(numPath=1 << 20, precision=cf.FP32, useSobol=False, seed1=17, seed2=23)
    w = np.array([1.0, 0.0])
    _check_equal(_noalloc(prd, mdl, num, w), cf.mcParallelSimulAAD(prd, mdl, weights=w, num=num))
    cls = {"europeans": cf.Europeans, "barrier": cf.UOC, "asian": cf.Asian}
    book = cf.Book([cls[k](**kw) for k, kw in W.c4_parts()])
    wb = W.c4_notionals(len(book.payoffLabels()))
    num2 = cf.NumericalParam(numPath=3000, precision=cf.FP64, useSobol=False, seed1=17, seed2=23)
    ref2 = cf.mcParallelSimulAAD(book, mdl, weights=wb, num=num2, per_path=True)
    _check_equal(_noalloc(book, mdl, num2, wb, per_path=True), ref2, per_path=True)


def test_noalloc_refuses_a_call_beyond_its_scratch(engine):
    # FP64 checkpoints of 2^20 paths need more than the array holds
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    num = cf.NumericalParam(numPath=1 << 20, precision=cf.FP64)
    rc = _noalloc(prd, mdl, num, [1.0, 0.0])[0]
    assert rc == -5                                      # CF_ERR_CAPACITY
    assert "scratch" in cf.lib().cf_last_error().decode()
