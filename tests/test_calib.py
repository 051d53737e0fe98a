"""Local-vol calibration on Merton implied surfaces (host FP64) against the
reference's dupireCalib (golden vectors of its build), and the closed-form
checks: without jumps the surface is flat and the local vols equal its vol up
to the finite-difference error of Dupire's formula."""
import json
import os

import numpy as np
import pytest

import compfinance_amd as cf

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "calib.json")))


@pytest.mark.parametrize("case", GOLDEN["calibs"], ids=[c["name"] for c in GOLDEN["calibs"]])
def test_calibration_matches_reference(case):
    s, v, ji, ja, js = case["ivs"]
    r = cf.dupireCalib(case["incl_spots"], case["max_ds"], case["incl_times"], case["max_dt"], s, v, ji, ja, js)
    np.testing.assert_array_equal(r["spots"], case["spots"])
    np.testing.assert_array_equal(r["times"], case["times"])
    np.testing.assert_allclose(r["lVols"].ravel(), case["lvols"], rtol=1e-13, atol=0)


def test_flat_surface_gives_its_vol():
    r = cf.dupireCalib([60.0, 100.0, 160.0], 5.0, [0.5, 1.0, 2.0], 0.25, 100.0, 0.2)
    assert np.abs(r["lVols"] - 0.2).max() < 5e-3


@pytest.mark.parametrize("case", GOLDEN["superbuckets"], ids=[c["name"] for c in GOLDEN["superbuckets"]])
def test_superbucket_chain_matches_reference(case):
    """The chain from the local-vol risks (the reference's own microbucket) to
    the implied-vol risk view, against the reference's tape.  Both differentiate
    Dupire's finite-difference formula (steps 1e-4, second difference scaled by
    1e8), whose cancellation leaves rounding of about 1e-6 of the largest risk;
    the tolerance is 1e-5 of it."""
    import ctypes
    g = GOLDEN["grid"]
    L = cf._calib_lib()
    ivs = cf.cf_merton_ivs(*g["ivs"])
    s, t, k, m = (np.array(g[x], float) for x in ("incl_spots", "incl_times", "strikes", "mats"))
    mb = np.array(case["microbucket"])
    vega = np.zeros(len(k) * len(m))
    d = cf._dptr
    assert L.cf_superbucket_vega(ctypes.byref(ivs), d(s), len(s), g["max_ds"], d(t), len(t), g["max_dt_vol"], d(k),
                                 len(k), d(m), len(m), d(mb), len(mb), d(vega)) == 0
    ref = np.array(case["vega"])
    assert np.abs(vega - ref).max() <= 1e-5 * np.abs(ref).max(), np.abs(vega - ref).max()
