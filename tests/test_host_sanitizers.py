"""The library's host code (planning, validation, store, calibration,
server set-up: cf_engine.cpp, cf_store.cpp, cf_calib.cpp, cf_server.cpp)
built with AddressSanitizer and UndefinedBehaviorSanitizer on the host side
only (the kernels' objects are the regular build's), driven through its C
ABI by tests/cpp/host_calls.cpp.  CPU only; without a device the entry points
that need one must fail cleanly."""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "compfinance_amd", "csrc")
BUILD = os.path.join(ROOT, "compfinance_amd", "_build")
HOST_SOURCES = ["cf_engine.cpp", "cf_store.cpp", "cf_calib.cpp", "cf_server.cpp"]
# the device-code objects of the regular build (every .hip unit it compiles)
from compfinance_amd import build as _build  # noqa: E402
KERNEL_OBJECTS = [o + ".o" for src, o, _ in _build.SOURCES if src.endswith(".hip")]
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    pytest.skip("hipcc not found")


def test_host_code_under_asan_and_ubsan(tmp_path):
    objs = [os.path.join(BUILD, o) for o in KERNEL_OBJECTS]
    if not all(os.path.exists(o) for o in objs):
        pytest.skip("kernel objects not built (python -m compfinance_amd.build)")
    hipcc = _hipcc()
    flags = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include")] + SAN

    def compile_one(src):
        out = str(tmp_path / (os.path.splitext(os.path.basename(src))[0] + ".o"))
        subprocess.run([hipcc] + flags + ["-c", src, "-o", out], check=True, capture_output=True, text=True)
        return out

    srcs = [os.path.join(CSRC, s) for s in HOST_SOURCES] + [os.path.join(ROOT, "tests", "cpp", "host_calls.cpp")]
    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        host_objs = list(ex.map(compile_one, srcs))
    exe = str(tmp_path / "host_calls")
    subprocess.run([hipcc, "--offload-arch=gfx950"] + SAN + ["-o", exe] + host_objs + objs, check=True,
                   capture_output=True, text=True)
    syms = subprocess.run(["nm", exe], capture_output=True, text=True).stdout
    assert "__asan_init" in syms and "__ubsan_handle" in syms      # the sanitizers are in the build
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    res = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    out = res.stdout + res.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
    assert res.returncode == 0 and res.stdout.strip().endswith("ok"), out[-4000:]
