"""The oracle's Asian and book products, pinned through the reference's own
products: a one-fixing Asian is a European, a zero-strike Asian is the average
of zero-strike Europeans at its fixing dates, a book is the concatenation of
its parts on shared paths and its aggregate risk is the notional-weighted sum
of the parts' risks.  Host-side planning of the engine (labels, timelines) is
checked against the oracle."""
import numpy as np
import pytest

import compfinance_amd as cf
import pyoracle as po
from compfinance_amd import workloads as W

LV = W.dupire_args(30)
BS = dict(spot=100.0, vol=0.25, rate=0.03, div=0.01)
ASIAN = dict(strike=100.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52)


def test_snapped_fixings_are_the_nearest_weekly_dates():
    t, _ = po.timeline(po.Model(**LV), po.Product(**ASIAN))
    grid, _ = po.timeline(po.Model(**LV), po.Product(**W.barrier_args()))
    fix = cf.Asian(100.0, 3.0, 1 / 12, 1 / 52).timeline()
    assert len(fix) == 36 and np.all(np.diff(fix) > 0)
    assert set(fix) <= set(grid)                       # on the barrier's weekly dates, bit for bit
    months = np.arange(1, 37) / 12
    for f, m in zip(fix, months):
        assert abs(f - m) <= 0.5 / 52 + 1e-12
    assert len(t) == 157                               # 156 steps, as the weekly barrier


@pytest.mark.parametrize("mk", [LV, BS])
def test_one_fixing_asian_is_a_european(mk):
    m = po.Model(**mk)
    a = po.value_paths(m, po.Product(strike=105.0, maturity=2.0, fixingFreq=2.0), 500)
    e = po.value_paths(m, po.Product(strike=105.0, exerciseDate=2.0), 500)
    np.testing.assert_array_equal(a, e)


def test_zero_strike_asian_averages_the_spot_at_its_fixings():
    m = po.Model(**LV)
    fix = cf.Asian(0.0, 3.0, 1 / 12, 1 / 52).timeline()
    a = po.value_paths(m, po.Product(parts=[dict(strike=0.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52),
                                            dict(maturities=list(fix), strikes=[0.0] * len(fix))]), 400)
    np.testing.assert_allclose(a[:, 0], a[:, 1:].mean(1), rtol=1e-14)


def _book_parts():
    return [dict(maturities=[3.0] * 4, strikes=[90.0, 100.0, 110.0, 120.0]),
            W.barrier_args(strike=100.0, barrier=140.0),
            W.barrier_args(strike=110.0, barrier=160.0),
            dict(strike=95.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52),
            dict(strike=105.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52, callPut=True)]


@pytest.mark.parametrize("mk", [LV, BS])
def test_book_is_the_concatenation_of_its_parts(mk):
    m = po.Model(**mk)
    parts = _book_parts()
    book = po.value_paths(m, po.Product(parts=parts), 300)
    # the barriers have the book's weekly timeline; so do, for Dupire
    # (maxDt = 1/52), the 3y Europeans, while Black-Scholes simulates those
    # alone in one step
    first = 0 if "spots" in mk else 1
    cols = [po.value_paths(m, po.Product(**p), 300) for p in parts[first:3]]
    np.testing.assert_array_equal(book[:, 4 * first:8], np.hstack(cols))


def test_book_aggregate_risk_is_the_weighted_sum_of_part_risks():
    m = po.Model(**LV)
    parts = _book_parts()[:3]
    w = np.array([0.5, 1.0, 1.5, 2.0, 0.7, -0.3, 1.2, 0.4])
    _, _, rb = po.aad(m, po.Product(parts=parts), 200, weights=w)
    r = po.aad(m, po.Product(**parts[0]), 200, weights=w[:4])[2]
    r = r + po.aad(m, po.Product(**parts[1]), 200, weights=w[4:6])[2]
    r = r + po.aad(m, po.Product(**parts[2]), 200, weights=w[6:8])[2]
    np.testing.assert_allclose(rb, r, rtol=1e-11, atol=1e-14 * np.abs(r).max())


def test_engine_book_plan_matches_oracle():
    parts = _book_parts()
    engine_parts = [cf.Europeans(**parts[0]), cf.UOC(**parts[1]), cf.UOC(**parts[2]),
                    cf.Asian(95.0, 3.0, 1 / 12, 1 / 52), cf.Asian(105.0, 3.0, 1 / 12, 1 / 52, True)]
    b = cf.Book(engine_parts)
    labels = b.payoffLabels()
    assert len(labels) == po.num_payoffs(po.Product(parts=parts)) == 10
    assert labels[0] == "0:call 3.00 90.00"
    assert labels[4] == "1:call 3.00 100.00 up and out 140.00 monitoring freq 0.02 smooth 0.05"
    assert labels[8] == "3:asian call 3.00 95.00" and labels[9] == "4:asian put 3.00 105.00"
    t, c = po.timeline(po.Model(**LV), po.Product(parts=parts))
    te = cf.simulation_timeline(cf.Dupire(**LV), b)
    np.testing.assert_array_equal(te[0], t)
    np.testing.assert_array_equal(te[1], c)


def test_book_limits():
    u = cf.UOC(**W.barrier_args())
    with pytest.raises(cf.CfError):
        cf.Book([u] * 17).payoffLabels()
    assert len(cf.Book([u] * 16).payoffLabels()) == 32
    with pytest.raises(cf.CfError):
        cf.Book([cf.European(100.0, 1.0, 1.5)]).payoffLabels()
    with pytest.raises(cf.CfError):
        cf.Book([cf.Book([u])]).payoffLabels()


@pytest.mark.parametrize("mk", [LV, BS])
def test_asian_put_call_parity_per_path(mk):
    """Per path, call - put = (average - K) / numeraire; the zero-strike call
    is average / numeraire, so call - put - (zero-strike call) = -K / numeraire,
    the same on every path (the numeraire is deterministic in both models
    here).  The risks of that combination (weights 1, -1, -1) vanish: the put's
    adjoint is -1{K > average} (AADExpr.h:571-583)."""
    m = po.Model(**mk)
    k = 100.0
    parts = [dict(strike=k, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52),
             dict(strike=k, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52, callPut=True),
             dict(strike=0.0, maturity=3.0, fixingFreq=1 / 12, snapFreq=1 / 52)]
    pay = po.value_paths(m, po.Product(parts=parts), 2000)
    c, p, a0 = pay[:, 0], pay[:, 1], pay[:, 2]
    assert (c > 0).sum() > 100 and (p > 0).sum() > 100           # both branches exercised
    const = c - p - a0
    disc = np.exp(-BS["rate"] * 3.0) if mk is BS else 1.0
    np.testing.assert_allclose(const, -k * disc, rtol=1e-12, atol=1e-12)
    _, _, risks = po.aad(m, po.Product(parts=parts), 2000, weights=[1.0, -1.0, -1.0])
    _, _, scale = po.aad(m, po.Product(parts=parts), 2000, weights=[0.0, 1.0, 0.0])
    if mk is LV:
        assert np.abs(risks).max() < 1e-12 * np.abs(scale).max()
    else:
        # Black-Scholes parameters (spot, vol, rate, div): only the rate moves -K exp(-rT)
        np.testing.assert_allclose(risks[[0, 1, 3]], 0.0, atol=1e-12 * np.abs(scale).max())
        np.testing.assert_allclose(risks[2], 3.0 * k * disc, rtol=1e-10)
