"""Build libcfamd.so (HIP kernels for gfx950 + C++17 host engine) in place.

    python -m compfinance_amd.build

The library is written next to this file; nothing is installed.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcfamd.so")

# (source, object, defines): the path kernel families compile in parallel
SOURCES = ([("cf_kern_inst.hip", "cf_kern_r%d_a%d_b%d" % (r, a, b),
             ["-DCF_INST_REAL=%d" % r, "-DCF_INST_ADJ=%d" % a, "-DCF_INST_SRC=%d" % b])
            for r in (0, 1) for a in (0, 1, 2) for b in (0, 1)] +
           [("cf_kern_inst.hip", "cf_kern_r%d_a1_b%d" % (r, b),
             ["-DCF_INST_REAL=%d" % r, "-DCF_INST_ADJ=1", "-DCF_INST_SRC=%d" % b])
            for r in (0, 1) for b in (2, 4)] +
           [("cf_kern_inst.hip", "cf_kern_r%d_a%d_b3" % (r, a),
             ["-DCF_INST_REAL=%d" % r, "-DCF_INST_ADJ=%d" % a, "-DCF_INST_SRC=3"])
            for r in (0, 1) for a in (0, 1, 2)] +

           [(s, os.path.splitext(s)[0], []) for s in ("cf_kernels.hip", "cf_engine.cpp", "cf_store.cpp",
                                                      "cf_server.cpp", "cf_calib.cpp", "cf_noalloc.hip")])

# Packed FP32 instructions are disabled for all device code.  Multiply-adds
# are fused only within one source expression (-ffp-contract=on), so every
# kernel instantiation rounds a shared function (the Euler step, the local-vol
# interpolation, the payoffs) the same way: a call gives the same bits in the
# checkpoint and recompute kernels, alone and packed.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compiles every source (concurrently) and links the library."""
    tmpdir = os.path.join(HERE, "_build")
    os.makedirs(tmpdir, exist_ok=True)
    jobs = []
    for src, name, defs in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(tmpdir, name + ".o")
        cmd = [hipcc()] + HIP_FLAGS + defs + ["-I", os.path.join(ROOT, "include"), "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        jobs.append((cmd, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, obj, proc in jobs:
        out = _filter(proc.communicate()[0])
        if proc.returncode != 0:
            failed.append("build failed: " + " ".join(cmd) + "\n" + out)
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("\n".join(failed))
    tmp_lib = LIB + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib] + [j[1] for j in jobs], verbose)
    os.replace(tmp_lib, LIB)
    return LIB


def _filter(text):
    return "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = _filter(res.stdout + res.stderr)
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    if verbose and out.strip():
        print(out)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
