"""A reference-style C++ caller (tests/cpp/drop_in.cpp) compiled against the
mirror headers and libcfamd.so: host-side results without a GPU, simulation
results on the GPU equal to the Python binding's, and the mirror's
bump-and-reprice risks against the AAD risks."""
import json
import os
import subprocess

import numpy as np
import pytest

import compfinance_amd as cf
from compfinance_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def drop_in(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("cpp") / "drop_in")
    lib_dir = os.path.join(ROOT, "compfinance_amd")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "drop_in.cpp"), "-L", lib_dir, "-lcfamd",
                    "-Wl,-rpath," + lib_dir, "-o", exe], check=True)
    return exe


def run(exe, mode):
    return json.loads(subprocess.run([exe, mode], capture_output=True, text=True, check=True).stdout)


def test_host_side(drop_in):
    out = run(drop_in, "host")
    book = cf.Book([cf.Europeans([3.0, 3.0, 2.0], [100.0, 120.0, 110.0]), cf.UOC(**W.barrier_args()),
                    cf.Asian(100.0, 3.0, 1 / 12, 1 / 52)])
    assert out["book_labels"] == book.payoffLabels()
    np.testing.assert_array_equal(out["book_timeline"], book.timeline())
    cal = cf.dupireCalib([50.0, 100.0, 200.0], 5.0, [0.25, 1.0, 3.0], 0.1, 100.0, 0.15, 0.05, -0.15, 0.1)
    np.testing.assert_array_equal(out["calib_lvols"], cal["lVols"].ravel())
    assert "Could not retrieve" in out["missing_model_error"] or "not" in out["missing_model_error"]


@pytest.mark.gpu
def test_device_side(engine, drop_in):
    out = run(drop_in, "device")
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    n = 1 << 14
    ref = cf.mcParallelSimulAAD(prd, mdl, n, weights=[1.0, 0.0])
    np.testing.assert_array_equal(out["uoc_values"], ref.payoff_means)
    np.testing.assert_allclose(out["uoc_risks"], ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())
    np.testing.assert_allclose(out["template_risks"], ref.risks, rtol=1e-12, atol=1e-12 * np.abs(ref.risks).max())
    book = cf.Book([cf.Europeans([3.0, 3.0, 2.0], [100.0, 120.0, 110.0]), cf.UOC(**W.barrier_args()),
                    cf.Asian(100.0, 3.0, 1 / 12, 1 / 52)])
    w = np.zeros(6)
    w[0], w[3], w[5] = 1.0, -0.5, 2.0
    rb = cf.mcParallelSimulAAD(book, mdl, n, weights=w)
    np.testing.assert_allclose(out["book_risks"], rb.risks, rtol=1e-12, atol=1e-12 * np.abs(rb.risks).max())
    # bumpRisk (main.h:316-359): one-sided 1e-8 bumps of every parameter on the
    # same paths; its vegas equal the AAD vegas to the finite difference's
    # rounding (a 1e-8 bump leaves about 1e-8 of the price's last digits)
    bump = np.array(out["bump_risks"])
    assert bump.shape == ref.risks.shape
    assert np.abs(bump[1:] - ref.risks[1:]).max() <= 1e-5 * np.abs(ref.risks[1:]).max()
