"""GPU results against the reference's own outputs (golden vectors of the
reference build) at the benchmark sizes: per-path payoffs, mean payoffs at
2^16 and 2^19 paths, AADriskOne / AADriskAggregate risk vectors at 2^16 and
2^20 paths, and bump-and-reprice of 20 seeded local-vol nodes."""
import json
import os

import numpy as np
import pytest

import compfinance_amd as cf

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _objects(models, case):
    mk, pk = dict(models[case["model"]]), dict(case["product"])
    mdl = cf.Dupire(**mk) if "spots" in mk else cf.BlackScholes(**mk)
    if "barrier" in pk:
        prd = cf.UOC(**pk)
    elif "maturities" in pk:
        prd = cf.Europeans(**pk)
    else:
        prd = cf.European(**pk)
    return mdl, prd


VALUE_CASES = list(_load("values.json")["cases"])
RISK_CASES = list(_load("risks.json")["cases"])


@pytest.mark.parametrize("name", VALUE_CASES)
def test_fp64_values_match_reference(engine, name):
    g = _load("values.json")
    case = g["cases"][name]
    mdl, prd = _objects(g["models"], case)
    res = cf.mcParallelSimul(prd, mdl, case["n_path"], per_path="per_path" in case)
    # the FP64 form's bar: 1e-10 relative on the means over the same paths
    np.testing.assert_allclose(res.means, case["values"], rtol=1e-10, atol=1e-15)
    if "per_path" in case:
        ref = np.asarray(case["per_path"]).reshape(-1, len(case["values"]))
        np.testing.assert_allclose(res.payoffs[: len(ref)], ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("name", RISK_CASES)
def test_fp64_risks_match_reference(engine, name):
    g = _load("risks.json")
    case = g["cases"][name]
    mdl, prd = _objects(g["models"], case)
    npay = len(case["payoff_values"])
    if "weights" in case:
        w = np.asarray(case["weights"])
    else:
        w = np.zeros(npay)
        w[case["risk_index"]] = 1.0
    res = cf.mcParallelSimulAAD(prd, mdl, case["n_path"], weights=w)
    ref = np.asarray(case["risks"])
    scale = np.abs(ref).max()
    err = np.abs(res.risks - ref)
    assert err.max() <= 1e-10 * scale, err.max() / scale
    big = np.abs(ref) > 1e-3 * scale
    assert np.all(err[big] <= 1e-9 * np.abs(ref[big]))
    np.testing.assert_allclose(res.payoff_means, case["payoff_values"], rtol=1e-10, atol=1e-15)
    assert abs(res.agg_mean - case["risk_payoff_value"]) <= 1e-10 * abs(case["risk_payoff_value"])


def test_bump_and_reprice_on_seeded_nodes(engine):
    b = _load("bumps.json")
    mk = dict(b["models"][b["model"]])
    prd = cf.UOC(**b["product"])
    n, h = b["n_path"], b["bump"]
    base = cf.mcParallelSimul(prd, cf.Dupire(**mk), n, per_path=False).means[0]
    assert abs(base - b["base"]) <= 1e-10 * abs(b["base"])
    aad = cf.mcParallelSimulAAD(prd, cf.Dupire(**mk), n, weights=[1.0, 0.0])
    scale = np.abs(aad.risks[1:]).max()
    vols = np.asarray(mk["vols"])
    for k, fd_ref in zip(b["nodes"], b["fd"]):
        v = vols.copy().ravel()
        v[k] += h
        up = cf.mcParallelSimul(prd, cf.Dupire(**dict(mk, vols=v.reshape(vols.shape))), n, per_path=False).means[0]
        fd = (up - base) / h
        # same paths as the reference's bump: equal up to FP64 rounding / h
        assert abs(fd - fd_ref) <= 1e-8, (k, fd, fd_ref)
        # 1bp forward differences carry the second-order term of the bump
        assert abs(aad.risks[1 + k] - fd) <= 1e-4 * scale + 5e-3 * abs(fd), (k, aad.risks[1 + k], fd)
