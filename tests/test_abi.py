"""The C ABI library without a GPU: it loads, exports every function the
headers declare, and its host logic (labels, timelines, validation, store,
errors) matches the reference.  CPU only; no simulation is run."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

import compfinance_amd as cf
import pyoracle as po
from compfinance_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(ROOT, "include", fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(cf_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(cf.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in cf.lib().cf_version()


def _golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def _engine_objects(models, case):
    mk = dict(models[case["model"]])
    pk = dict(case["product"])
    if "spots" in mk:
        mdl = cf.Dupire(**mk)
    else:
        mdl = cf.BlackScholes(**mk)
    if "barrier" in pk:
        prd = cf.UOC(**pk)
    elif "maturities" in pk:
        prd = cf.Europeans(**pk)
    else:
        prd = cf.European(**pk)
    return mdl, prd, po.Model(**mk), po.Product(**pk)


def test_payoff_labels_match_reference():
    g = _golden("values.json")
    for name, case in g["cases"].items():
        mdl, prd, _, _ = _engine_objects(g["models"], case)
        assert prd.payoffLabels() == case["labels"], name


def test_parameter_labels_match_reference():
    g = _golden("values.json")
    labels = _golden("param_labels.json")
    for key, ref in labels.items():
        mk = g["models"][key]
        mdl = cf.Dupire(**mk) if "spots" in mk else cf.BlackScholes(**mk)
        assert mdl.parameterLabels() == ref, key


def test_simulation_timelines_match_oracle():
    g = _golden("values.json")
    for name, case in g["cases"].items():
        mdl, prd, om, op = _engine_objects(g["models"], case)
        t, c = cf.simulation_timeline(mdl, prd)
        to, co = po.timeline(om, op)
        assert np.array_equal(t, to), name
        assert np.array_equal(c, co), name


def test_product_timelines():
    t = cf.UOC(**W.barrier_args()).timeline()
    assert len(t) == 157 and t[0] == 0.0 and t[-1] == 3.0
    e = cf.Europeans([3.0, 1.0, 3.0], [100.0, 90.0, 110.0])
    assert list(e.timeline()) == [1.0, 3.0]
    assert e.payoffLabels() == ["call 1.00 90.00", "call 3.00 100.00", "call 3.00 110.00"]


@pytest.mark.parametrize("bad, msg", [
    (dict(spot=-1.0), "spot must be positive"),
    (dict(spots=[100.0, 50.0, 200.0]), "spots must increase"),
    (dict(times=[1.0, 0.5, 2.0]), "times must increase"),
    (dict(maxDt=0.0), "maxDt must be positive"),
])
def test_model_validation(bad, msg):
    args = dict(spot=100.0, spots=[50.0, 100.0, 200.0], times=[0.5, 1.0, 2.0], vols=np.full((3, 3), 0.2),
                maxDt=0.25)
    args.update(bad)
    with pytest.raises(cf.CfError, match=msg) as e:
        cf.Dupire(**args).parameterLabels()
    assert e.value.code == -1


def test_nan_inputs_are_rejected():
    v = np.full((3, 3), 0.2)
    v[1, 1] = np.nan
    with pytest.raises(cf.CfError, match="finite"):
        cf.Dupire(100.0, [50.0, 100.0, 200.0], [0.5, 1.0, 2.0], v).parameterLabels()
    with pytest.raises(cf.CfError, match="finite"):
        cf.UOC(120.0, float("inf"), 3.0, 1 / 52, 0.05).payoffLabels()


def test_product_validation():
    with pytest.raises(cf.CfError, match="monitoring frequency"):
        cf.UOC(120.0, 150.0, 3.0, 0.0, 0.05).payoffLabels()
    with pytest.raises(cf.CfError, match="maturity must be positive"):
        cf.UOC(120.0, 150.0, -1.0, 0.1, 0.05).payoffLabels()
    with pytest.raises(cf.CfError, match="exercise date"):
        cf.European(100.0, -1.0).payoffLabels()
    with pytest.raises(ValueError):
        cf.Europeans([1.0, 2.0], [100.0])


def test_too_many_steps_for_sobol_table():
    # 1,101 Sobol dimensions: 30y weekly needs 1,560 steps
    mdl = cf.Dupire(**W.dupire_args(30))
    with pytest.raises(cf.CfError, match="Sobol dimensions") as e:
        cf.simulation_timeline(mdl, cf.UOC(120.0, 150.0, 30.0, 1 / 52, 0.05))
    assert e.value.code == -5


def test_store_errors_follow_reference_messages():
    cf.clearStore()
    cf.putDupire(**W.dupire_args(30), store="lv")
    cf.putBarrier(**W.barrier_args(), store="uoc")
    cf.putBlackScholes(100.0, 0.2, False, 0.0, 0.0, "bs")
    num = cf.NumericalParam(numPath=64)
    with pytest.raises(cf.CfError, match=r"value\(\) : Could not retrieve model and product") as e:
        cf.value("nope", "uoc", num)
    assert e.value.code == -2
    with pytest.raises(cf.CfError, match=r"AADrisk\(\) : Could not retrieve model and product"):
        cf.AADriskOne("lv", "nope", num)
    with pytest.raises(cf.CfError, match=r"AADriskOne\(\) : payoff not found"):
        cf.AADriskOne("lv", "uoc", num, "no such payoff")
    with pytest.raises(cf.CfError, match=r"AADriskAggregate\(\) : payoff not found"):
        cf.AADriskAggregate("lv", "uoc", {"no such payoff": 1.0}, num)
    with pytest.raises(cf.CfError, match=r"dupireAADRisk\(\) : Model not a Dupire"):
        cf.dupireAADRisk("bs", "uoc", {}, num)
    with pytest.raises(cf.CfError, match=r"dupireAADRisk\(\) : Model not found"):
        cf.dupireAADRisk("nope", "uoc", {}, num)
    cf.clearStore()


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_simulation_without_gpu_fails_loudly():
    mdl, prd = cf.BlackScholes(100.0, 0.2), cf.European(100.0, 1.0)
    with pytest.raises(cf.CfError) as e:
        cf.mcParallelSimul(prd, mdl, 64)
    assert e.value.code == -4


@pytest.mark.parametrize("kw, code, msg", [
    (dict(numPath=0), -1, "numPath must be positive"),
    (dict(numPath=-5), -1, "numPath must be positive"),
    (dict(numPath=64, firstPath=-1), -1, "first path must be non-negative"),
    (dict(numPath=64, precision=7), -1, "unknown precision"),
    (dict(numPath=64, aadMode=9), -1, "unknown AAD mode"),
])
def test_numerical_parameters_are_validated_before_the_device(kw, code, msg):
    """Empty and negative path ranges and unknown options fail with an
    argument error before any device work (the reference's int path counts
    keep the last Sobol index below 2^32 - 1)."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    with pytest.raises(cf.CfError, match=msg) as e:
        cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(**kw), per_path=False)
    assert e.value.code == code
    with pytest.raises(cf.CfError, match=msg):
        cf.mcParallelSimulAAD(prd, mdl, num=cf.NumericalParam(**kw))


def test_aggregation_weights_need_one_entry_per_payoff():
    """The barrier has two payoffs (the option and its European): a weight
    vector of another length is an error in the binding and in the oracle,
    checked before anything reaches the device."""
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    assert len(prd.payoffLabels()) == 2
    for w in ([1.0], [1.0, 0.0, 0.0]):
        with pytest.raises(ValueError, match="one entry per payoff"):
            cf.mcParallelSimulAAD(prd, mdl, 64, weights=w)
        with pytest.raises(ValueError, match="one entry per payoff"):
            po.aad(po.Model(**W.dupire_args(30)), po.Product(**W.barrier_args()), 8, weights=w)
