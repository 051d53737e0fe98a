"""dupireSuperbucket end to end on the GPU (FP64 local-vol risks on the
calibrated model, then the chain to implied vols) against the reference's."""
import json
import os

import numpy as np
import pytest

import compfinance_amd as cf

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "calib.json")))
CLS = {"barrier": cf.putBarrier, "europeans": cf.putEuropeans}


@pytest.mark.parametrize("case", GOLDEN["superbuckets"], ids=[c["name"] for c in GOLDEN["superbuckets"]])
def test_superbucket_matches_reference(engine, case):
    g = GOLDEN["grid"]
    cf.clearStore()
    if case["name"] == "barrier":
        cf.putBarrier(**case["product"], store="p")
        prd = cf.UOC(**case["product"])
    else:
        cf.putEuropeans(case["product"]["maturities"], case["product"]["strikes"], "p")
        prd = cf.Europeans(**case["product"])
    labels = prd.payoffLabels()
    nots = {labels[int(i)]: w for i, w in case["notionals"].items()}
    spot, vol, ji, ja, js = g["ivs"]
    r = cf.dupireSuperbucket(spot, g["max_dt"], "p", nots, g["incl_spots"], g["max_ds"], g["incl_times"],
                             g["max_dt_vol"], g["strikes"], g["mats"], vol, ji, ja, js,
                             cf.NumericalParam(numPath=case["n_path"]))
    assert abs(r["value"] - case["value"]) <= 1e-10 * abs(case["value"])
    assert abs(r["delta"] - case["delta"]) <= 1e-10 * abs(case["delta"])
    ref = np.array(case["vega"])
    assert np.abs(r["vega"].ravel() - ref).max() <= 1e-5 * np.abs(ref).max()
    cf.clearStore()


@pytest.mark.parametrize("case", GOLDEN["superbuckets"], ids=[c["name"] for c in GOLDEN["superbuckets"]])
def test_superbucket_by_bumps_matches_reference(engine, case):
    """dupireSuperbucketBump (main.h:575-696): each risk-view spread bumped by
    1e-5 in turn, the local vols recalibrated and the book repriced on the GPU
    (FP64), against the reference's own bump superbucket on the same paths.
    The finite differences scale the prices' rounding by 1e5 (vega) and 1e8
    (delta)."""
    g = GOLDEN["grid"]
    cf.clearStore()
    if case["name"] == "barrier":
        cf.putBarrier(**case["product"], store="p")
        prd = cf.UOC(**case["product"])
    else:
        cf.putEuropeans(case["product"]["maturities"], case["product"]["strikes"], "p")
        prd = cf.Europeans(**case["product"])
    labels = prd.payoffLabels()
    nots = {labels[int(i)]: w for i, w in case["notionals"].items()}
    spot, vol, ji, ja, js = g["ivs"]
    r = cf.dupireSuperbucketBump(spot, g["max_dt"], "p", nots, g["incl_spots"], g["max_ds"], g["incl_times"],
                                 g["max_dt_vol"], g["strikes"], g["mats"], vol, ji, ja, js,
                                 cf.NumericalParam(numPath=case["n_path"]))
    ref = case["bump"]
    assert abs(r["value"] - ref["value"]) <= 1e-10 * abs(ref["value"])
    assert abs(r["delta"] - ref["delta"]) <= 1e-5 * max(abs(ref["delta"]), 1.0), (r["delta"], ref["delta"])
    rv = np.array(ref["vega"])
    err = np.abs(r["vega"].ravel() - rv).max() / np.abs(rv).max()
    assert err <= 1e-6, err
    cf.clearStore()
