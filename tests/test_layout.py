"""Launch arithmetic shared by the host engine and the kernels, checked on the
CPU: the merge of the last block of 64 Sobol indices (against brute force
over path ranges), the shared-memory layout (alignment, the 160 KiB LDS of a
CU, the bank offset of the second adjoint grid, every region the kernel
addresses inside its region and inside the launch's LDS size, the FP32
inverse-normal rows, the interleaved Sobol and checkpoint tables) and the knot search (against upper_bound, interp.h:40)."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LDS_PER_CU = 160 * 1024


def test_merge_and_layout(tmp_path):
    exe = str(tmp_path / "layout_check")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "cpp", "layout_check.cpp"), "-o", exe],
                   check=True)
    out = json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert out["wrong"] == 0, out
    assert 0 < out["merged"] < out["cases"]
    assert out["misaligned"] == 0
    assert out["fp32_c3"] <= LDS_PER_CU and out["fp64_c3"] <= LDS_PER_CU and out["lean_60x60_5y_fp64"] <= LDS_PER_CU
    assert out["copy_offset_mod128"] == 64
    assert out["regions"] > 55000 and out["out_of_region"] == 0
    assert out["row_probes"] > 700 and out["bad_rows"] == 0
    assert out["il_probes"] > 2_000_000 and out["il_bad"] == 0


def test_knot_search_equals_upper_bound(tmp_path):
    """The kernels' knot search and its table builder (one definition in
    cf_internal.h) give upper_bound's count on the benchmark grids and on 600
    random grids (even, random, clustered and repeated knots; up to 300 knots
    in one bucket), at every knot, its neighbouring floats, random states and
    +-inf, in FP32 and FP64."""
    exe = str(tmp_path / "knot_search_check")
    subprocess.run(["g++", "-std=c++17", "-O1", os.path.join(ROOT, "tests", "cpp", "knot_search_check.cpp"), "-o",
                    exe], check=True)
    out = json.loads(subprocess.run([exe], capture_output=True, text=True, check=True).stdout)
    assert out["probes"] > 1_000_000
    assert out["wrong_fp32"] == 0 and out["wrong_fp64"] == 0, out
    assert out["max_fix_fp32"] > 2 and out["max_fix_fp64"] > 2      # the comparison loop is exercised
    assert out["c3_fp32_fix"] <= 2                                   # the benchmark grid needs no loop
