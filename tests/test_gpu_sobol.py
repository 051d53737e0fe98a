"""Sobol points drawn on the device by the path kernels' own code
(cf_sobol_points: sobol_stage_lo, sobol_block, sobol_low_word and
inv_normal, shared with the path kernels) against the reference's points
(tests/golden/sobol_points.npz, written by the reference build: skipTo and
next, sobol.h:77-151), bit for bit, in the FULL (rows and quads) and LEAN
forms; ranges with
and without the merge of the last block of 64 indices against the oracle; the
normals against the oracle's Moro function (FP64, gaussians.h:47-87) and the
FP32 table's accuracy bound."""
import os

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
pytestmark = pytest.mark.gpu

ONEOVER2POW32 = 2.3283064365387E-10   # sobol.h:26


def _layouts(dim):
    return [0, 1, 2] if dim <= 261 else [1]


def test_device_words_equal_the_reference_points(engine):
    cf = engine
    pts = np.load(os.path.join(GOLDEN, "sobol_points.npz"))
    checked = 0
    for key in pts.files:
        b, dims = int(key.split("_")[0][1:]), int(key.split("_")[1][1:])
        ref = pts[key]
        for layout in _layouts(dims):
            words, _ = cf.sobol_points(b, ref.shape[0], dims, cf.FP64, layout)
            assert np.array_equal(words, ref), (key, layout)
            checked += 1
    assert checked == 32


@pytest.mark.parametrize("first,count", [(0, 1), (0, 1000), (0, 65536), (5, 1000), (1065, 3835), (65535, 4097)])
@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("dim", [156, 157])
def test_device_word_ranges_equal_the_oracle(engine, first, count, layout, dim):
    """Path ranges from path 0 (Sobol index 0 idle), with and without the
    merged last block, in 156 dimensions and in 157 (a last group of four
    dimensions with one dimension)."""
    cf = engine
    words, _ = cf.sobol_points(first, count, dim, cf.FP64, layout)
    assert np.array_equal(words, po.sobol_points(dim, first, count))


def test_fp64_normals_are_the_moro_values(engine):
    cf = engine
    words, z = cf.sobol_points(0, 1 << 14, 156, cf.FP64, 0)
    ref = po.inv_normal(ONEOVER2POW32 * words.astype(np.float64).ravel()).reshape(words.shape)
    # the same formula and branches; the device contracts the rational function's
    # multiply-adds into FMAs (the oracle's build does not) and its tail logarithm
    # is within two ulps of libm's: a few tens of ulps at most
    np.testing.assert_allclose(z, ref, rtol=1e-14, atol=1e-15)
    # the branch between the central and the tail form is the reference's: |u - 1/2| < 0.42
    u = ONEOVER2POW32 * words.astype(np.float64)
    tail = np.abs(np.minimum(u, 1.0 - u) - 0.5) >= 0.42
    assert 0.1 < tail.mean() < 0.2


@pytest.mark.parametrize("layout", [0, 1, 2])
def test_fp32_normals_within_the_table_bound(engine, layout):
    """The FP32 kernels' table form: within 5e-7 of the FP64 Moro values (the
    bound the CPU restatement of the table pins at every row boundary), odd
    about u = 1/2, and the words themselves equal the FP64 path's."""
    cf = engine
    words, z = cf.sobol_points(0, 1 << 14, 156, cf.FP32, layout)
    w64, z64 = cf.sobol_points(0, 1 << 14, 156, cf.FP64, layout)
    assert np.array_equal(words, w64)
    err = np.abs(z.astype(np.float64) - z64)
    assert err.max() < 5e-7
    assert err.mean() < 1e-7
    assert np.all(z[words == 0x80000000] == 0.0)     # u = 1/2


def test_rejects_bad_arguments(engine):
    cf = engine
    for args in [(0, 0, 10), (-1, 10, 10), (0, 10, 0), (0, 10, 1102)]:
        with pytest.raises(cf.CfError):
            cf.sobol_points(*args)
