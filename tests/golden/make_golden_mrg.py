"""Writes tests/golden/mrg.json from the reference build: per-path payoffs,
mean payoffs and AADriskOne risks with the mrg32k3a generator (mrg32k3a.h)
for seeded runs.

    python tests/golden/make_golden_mrg.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, ROOT)

import pyoracle as po  # noqa: E402
from compfinance_amd import workloads as W  # noqa: E402

CASES = [
    ("dupire30_barrier", W.dupire_args(30), W.barrier_args(), (12345, 1234)),
    ("bs_rates_european", dict(spot=100.0, vol=0.25, rate=0.03, div=0.01), dict(strike=100.0, exerciseDate=1.0),
     (777, 4242)),
    ("dupire30_europeans", W.dupire_args(30), dict(maturities=[1.0, 2.0, 3.0], strikes=[90.0, 100.0, 120.0]),
     (1, 2)),
]


def main():
    if not po.ref_available():
        raise SystemExit("reference build missing: run oracle/build_ref.sh")
    out = []
    for name, mk, pk, seeds in CASES:
        m, p = po.Model(**mk), po.Product(**pk)
        v = po.ref_value(m, p, 4096, threads=4, n_out=128, mrg=seeds)
        r = po.ref_aad(m, p, 1024, 0, mrg=seeds)
        out.append(dict(name=name, model=name.split("_")[0], seeds=list(seeds), product=pk, n_value=4096,
                        values=v["values"], per_path=v["per_path"], n_aad=1024, risks=r["risks"]))
    models = {"dupire30": {k: (np.asarray(x).tolist() if hasattr(x, "shape") else x)
                           for k, x in W.dupire_args(30).items()},
              "bs": dict(spot=100.0, vol=0.25, rate=0.03, div=0.01)}
    with open(os.path.join(HERE, "mrg.json"), "w") as f:
        json.dump(dict(models=models, cases=out), f)
    print("written", os.path.join(HERE, "mrg.json"))


if __name__ == "__main__":
    main()
