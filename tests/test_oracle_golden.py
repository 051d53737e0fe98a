"""The CPU oracle against the golden vectors written by the reference build
(tests/golden/make_golden.py).  CPU only."""
import json
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _model(models, key):
    return po.Model(**models[key])


def test_sobol_points_bit_exact():
    pts = np.load(os.path.join(GOLDEN, "sobol_points.npz"))
    assert len(pts.files) == 12
    for key in pts.files:
        b = int(key.split("_")[0][1:])
        dims = int(key.split("_")[1][1:])
        ref = pts[key]
        mine = po.sobol_points(dims, b, ref.shape[0])
        assert np.array_equal(mine, ref), key


def test_sobol_first_points_are_the_reference_ones():
    # path 0 uses index 1: 1/2 in every dimension (sobol.h:77-100)
    p = po.sobol_points(1101, 0, 1)[0]
    assert np.all(p == 0x80000000)


def test_inverse_normal_matches_reference():
    d = np.load(os.path.join(GOLDEN, "inv_normal.npz"))
    z = po.inv_normal(d["u"])
    # same formula and libm: at most one ulp apart
    np.testing.assert_array_max_ulp(z, d["z"], maxulp=1)


def test_fill_data_matches_reference():
    import ctypes
    for case in _load("fill_data.json"):
        orig = np.asarray(case["original"], dtype=np.float64)
        add = np.asarray(case["add"], dtype=np.float64)
        out = np.zeros(4096)
        dp = ctypes.POINTER(ctypes.c_double)
        n = po.lib().or_fill_data(orig.ctypes.data_as(dp), len(orig), case["maxDx"], case["minDx"],
                                  add.ctypes.data_as(dp), len(add), out.ctypes.data_as(dp), len(out))
        assert n == len(case["filled"])
        assert np.array_equal(out[:n], np.asarray(case["filled"]))


@pytest.mark.parametrize("name", ["c1_bs_european", "bs_rates_european_settle", "bs_spot_measure_european",
                                  "bs_barrier", "bs_europeans", "c2_dupire30_barrier", "c2_dupire60_barrier",
                                  "dupire30_barrier_put", "dupire30_european", "dupire30_europeans"])
def test_per_path_payoffs_and_means(name):
    g = _load("values.json")
    case = g["cases"][name]
    m, p = _model(g["models"], case["model"]), po.Product(**case["product"])
    npay = po.num_payoffs(p)
    ref_pp = np.asarray(case["per_path"]).reshape(-1, npay)
    n = case["n_path"] if case["model"] in ("m0", "m1", "m2") else 1 << 13
    pp = po.value_paths(m, p, n)
    np.testing.assert_allclose(pp[: len(ref_pp)], ref_pp, rtol=1e-14, atol=1e-14)
    if n == case["n_path"]:
        np.testing.assert_allclose(pp.mean(0), case["values"], rtol=1e-13)


@pytest.mark.parametrize("name", ["c1_bs_european", "bs_rates_european_settle", "bs_spot_measure_european",
                                  "bs_barrier", "bs_europeans", "bs_europeans_aggregate"])
def test_black_scholes_risks_match_reference(name):
    g = _load("risks.json")
    case = g["cases"][name]
    m, p = _model(g["models"], case["model"]), po.Product(**case["product"])
    pay, agg, risks = po.aad(m, p, case["n_path"], case.get("risk_index", 0), case.get("weights"))
    np.testing.assert_allclose(risks, case["risks"], rtol=1e-12, atol=1e-14)
    assert abs(agg.mean() - case["risk_payoff_value"]) <= 1e-13 * max(1.0, abs(case["risk_payoff_value"]))


@pytest.mark.parametrize("name", ["c2_dupire30_barrier", "dupire30_barrier_put", "dupire30_european",
                                  "dupire30_europeans", "dupire30_europeans_aggregate",
                                  "dupire30_barrier_aggregate"])
def test_dupire_risks_match_reference(name):
    g = _load("risks.json")
    case = g["cases"][name]
    m, p = _model(g["models"], case["model"]), po.Product(**case["product"])
    pay, agg, risks = po.aad(m, p, case["n_path"], case.get("risk_index", 0), case.get("weights"))
    ref = np.asarray(case["risks"])
    assert np.abs(risks - ref).max() <= 1e-12 * np.abs(ref).max()
    np.testing.assert_allclose(pay.mean(0), case["payoff_values"], rtol=1e-13)
