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


def test_calibration_with_a_zero_view_is_the_plain_calibration():
    """dupireCalib with a risk view of zero spreads gives the plain
    calibration bit for bit (mcMdlDupire.h:337-388, ivs.h:100-110)."""
    g = GOLDEN["grid"]
    s, v, ji, ja, js = g["ivs"]
    plain = cf.dupireCalib(g["incl_spots"], g["max_ds"], g["incl_times"], g["max_dt_vol"], s, v, ji, ja, js)
    view = cf.dupireCalibView(g["incl_spots"], g["max_ds"], g["incl_times"], g["max_dt_vol"], s, v, ji, ja, js,
                              g["strikes"], g["mats"], np.zeros((len(g["strikes"]), len(g["mats"]))))
    for k in ("spots", "times", "lVols"):
        np.testing.assert_array_equal(view[k], plain[k])


def test_constant_spread_is_a_vol_shift_of_a_flat_surface():
    """On a flat surface (no jumps) a constant spread c everywhere is the flat
    surface of vol + c (the smooth-step interpolation of equal spreads is the
    spread).  Dupire's finite-difference formula (steps 1e-4, the second
    difference scaled by 1e8) turns the implied-vol bisection's last digits
    into local-vol differences of about 1e-3, so the two agree to 1e-2 where
    both compute the local vol, against 15% between the 0.2 and 0.23
    surfaces.  The cut at 2.5 at-the-money standard deviations uses the
    surface without the view (mcMdlDupire.h:309), so outside the narrower cut
    the two extrapolate from different edges."""
    spots, times = [60.0, 100.0, 160.0], [0.5, 1.0, 2.0]
    k, m = [80.0, 100.0, 120.0], [0.5, 1.0, 2.0]
    shifted = cf.dupireCalibView(spots, 5.0, times, 0.25, 100.0, 0.2, 0.0, 0.0, 0.0, k, m, np.full((3, 3), 0.03))
    flat = cf.dupireCalib(spots, 5.0, times, 0.25, 100.0, 0.23)
    np.testing.assert_array_equal(shifted["spots"], flat["spots"])
    S, T = flat["spots"], flat["times"]
    inside = np.abs(S[:, None] - 100.0) < 0.95 * 2.5 * 100.0 * 0.2 * np.sqrt(T[None, :])
    assert inside.sum() > 50
    np.testing.assert_allclose(shifted["lVols"][inside], flat["lVols"][inside], rtol=1e-2)
    plain = cf.dupireCalib(spots, 5.0, times, 0.25, 100.0, 0.2)
    assert np.all(np.abs(plain["lVols"][inside] / flat["lVols"][inside] - 0.2 / 0.23) < 1e-2)


def test_one_spread_moves_the_local_vols_near_its_node_only():
    """A spread on one risk-view node moves the local vols whose Dupire
    points see it (smooth-step weights, flat outside the view) and leaves the
    others bit for bit: a bump at the (100, 1y) node of a 4 x 4 view."""
    g = GOLDEN["grid"]
    s, v, ji, ja, js = g["ivs"]
    args = (g["incl_spots"], g["max_ds"], g["incl_times"], g["max_dt_vol"], s, v, ji, ja, js, g["strikes"], g["mats"])
    sp = np.zeros((len(g["strikes"]), len(g["mats"])))
    base = cf.dupireCalibView(*args, sp)["lVols"]
    sp[1, 1] = 1e-3   # strike 100, maturity 1
    bumped = cf.dupireCalibView(*args, sp)["lVols"]
    moved = bumped != base
    assert moved.any()
    # only maturities strictly between the neighbouring view maturities (0.5, 2.0) see the node
    times = cf.dupireCalib(g["incl_spots"], g["max_ds"], g["incl_times"], g["max_dt_vol"], s, v, ji, ja, js)["times"]
    cols = np.nonzero(moved.any(axis=0))[0]
    assert np.all((times[cols] > 0.5 - 1e-3) & (times[cols] < 2.0 + 1e-3)), times[cols]
