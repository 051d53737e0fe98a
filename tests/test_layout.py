"""Launch arithmetic shared by the host engine and the kernels, checked on the
CPU: the merge of the last block of 64 Sobol indices (against brute force
over path ranges) and the shared-memory layout (alignment, the 160 KiB LDS
of a CU, the bank offset of the second adjoint grid)."""
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
