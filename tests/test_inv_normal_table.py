"""The FP32 kernels' inverse-normal table (compfinance_amd/csrc/inv_normal_f32.inc):
its fit target is the oracle's Moro function, the committed table is what
the generator writes, and its FP32 evaluation (the kernel's arithmetic,
restated in numpy float32) stays within one FP32 ulp of the far tail of the
FP64 Moro values.  CPU only."""
import os
import sys

import numpy as np

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "compfinance_amd", "csrc")
sys.path.insert(0, CSRC)
import gen_inv_normal_f32 as gen  # noqa: E402


def _committed():
    rows = []
    for line in open(os.path.join(CSRC, "inv_normal_f32.inc")):
        if line.strip().startswith("//"):
            continue
        rows.append([float(x.strip().rstrip("f")) for x in line.split("//")[0].split(",") if x.strip()])
    return np.array(rows, dtype=np.float32)


def _kernel_eval(tab, xi):
    """The kernel's FP32 steps: exact tail mass, row from the float bits,
    t from the low 20 bits, Horner in float32, sign."""
    xi = np.asarray(xi, dtype=np.uint64)
    sup = xi >= 2 ** 31
    xs = np.where(sup, 2 ** 32 - xi, xi)
    bits = xs.astype(np.float32).view(np.uint32)
    row = (bits >> 20).astype(np.int64) - (127 << 3)
    t = ((bits & 0xFFFFF).astype(np.float32) * np.float32(1.0 / 524288.0) + np.float32(-1.0)).astype(np.float32)
    c = tab[row]
    p = (c[:, 3] * t + c[:, 2]).astype(np.float32)
    p = (p * t + c[:, 1]).astype(np.float32)
    p = (p * t + c[:, 0]).astype(np.float32)
    return np.where(sup, -p, p), row


def test_fit_target_is_the_oracle_moro_function():
    u = np.concatenate([np.linspace(1e-9, 0.5, 20001), [2.0 ** -32, 0.0799, 0.08, 0.0801, 0.25, 0.5]])
    np.testing.assert_allclose(gen.moro_lower(u), po.inv_normal(u), rtol=1e-14, atol=1e-15)


def test_committed_table_is_the_generated_one():
    rows, worst = gen.table()
    assert worst < 2e-7
    tab = _committed()
    assert tab.shape == (249, 4)
    np.testing.assert_array_equal(tab, rows.astype(np.float32))


def test_fp32_evaluation_within_an_ulp_of_the_tail():
    tab = _committed()
    rng = np.random.default_rng(20261202)
    edges = [1, 2, 3, 255, 256, 2 ** 24 - 1, 2 ** 24 + 1, 2 ** 31 - 1, 2 ** 31, 2 ** 31 + 1, 2 ** 32 - 2, 2 ** 32 - 1]
    # every row boundary from below and above
    bounds = [int((1.0 + q / 8.0) * 2.0 ** e) for e in range(31) for q in range(8)]
    near = sorted({v for b in bounds for v in (b - 1, b, b + 1) if v >= 1})    # 0: see the next test
    xi = np.concatenate([rng.integers(1, 2 ** 32, size=400_000, dtype=np.uint64),
                         np.array(edges + near, dtype=np.uint64)])
    z, row = _kernel_eval(tab, xi)
    assert row.min() >= 0 and row.max() <= 248
    u = xi.astype(np.float64) * 2.0 ** -32
    ref = po.inv_normal(u)
    err = np.abs(z - ref)
    assert err.max() < 5e-7, (err.max(), u[np.argmax(err)])
    assert err.mean() < 1e-7
    assert z[xi == 2 ** 31][0] == 0.0                    # u = 1/2
    np.testing.assert_array_equal(_kernel_eval(tab, 2 ** 32 - xi[:1000])[0], -z[:1000])   # odd symmetry


def test_zero_word_never_reaches_the_table():
    """Only Sobol index 0 (no path: path p uses point p + 1) has a zero word,
    and word 0 maps to row -1016, outside the table.  The kernel gives that
    idle lane words of 2^31 instead (its low words are 0 and its high row is
    filled with 2^31), which land on the table's middle row: z = 0."""
    tab = _committed()
    row0 = int(np.float32(0).view(np.uint32) >> 20) - (127 << 3)
    assert row0 == -1016
    z, row = _kernel_eval(tab, np.array([2 ** 31], dtype=np.uint64))
    assert row[0] == 248 and z[0] == 0.0
    # the reference's points of paths (Sobol index >= 1) have no zero word
    pts = np.load(os.path.join(ROOT, "tests", "golden", "sobol_points.npz"))
    for key in pts.files:
        assert (pts[key] != 0).all(), key
