"""Runs benchmark workloads a fixed number of times, for profilers.

    python -m compfinance_amd.profile_run --workload c3_fp32_aad --calls 3
    rocprofv3 --kernel-trace --stats -f csv -d OUT -- python3 compfinance_amd/profile_run.py --workload all

Workloads: c3_fp32_aad, c3_fp64_aad (Dupire barrier, 30x36 grid, 2^20 paths,
156 steps, price + 1,081 risks, checkpoint adjoint), c3_fp32_rec, c3_fp64_rec
(the same with the recompute adjoint), c4_fp32, c4_fp64 (the book of 30
Europeans, 10 barriers and 5 Asians, aggregated risk, 2^20 paths), c2_fp32,
c2_fp64 (pricing, 2^19 paths),
copy (a 1 GiB device-to-device copy by the HIP runtime's copy kernel, the
byte-counter calibration),
all (every one of them).
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import compfinance_amd as cf  # noqa: E402
from compfinance_amd import workloads as W  # noqa: E402

COPY_BYTES = 1 << 30


def _device_copy(nbytes):
    """Device-to-device copy of nbytes through the HIP runtime (its copy
    kernel): nbytes read and nbytes written, the known traffic that calibrates
    the byte counters."""
    import ctypes
    hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError("%s failed: hip error %d" % (what, rc))

    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    ok(hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)), "hipMalloc")
    ok(hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)), "hipMalloc")
    ok(hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)), "hipMemset")
    ok(hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3), "hipMemcpy")   # 3: device to device
    ok(hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
    ok(hip.hipFree(a), "hipFree")
    ok(hip.hipFree(b), "hipFree")


def run(name, calls):
    mdl, prd = cf.Dupire(**W.dupire_args(30)), cf.UOC(**W.barrier_args())
    for _ in range(calls):
        if name == "copy":
            _device_copy(COPY_BYTES)
        elif name.startswith("c3"):
            prec = cf.FP32 if "fp32" in name else cf.FP64
            mode = cf.AAD_RECOMPUTE if name.endswith("_rec") else cf.AAD_CHECKPOINT
            cf.mcParallelSimulAAD(prd, mdl, weights=[1.0, 0.0],
                                  num=cf.NumericalParam(numPath=W.C3_PATHS, precision=prec, aadMode=mode))
        elif name.startswith("c4"):
            prec = cf.FP32 if "fp32" in name else cf.FP64
            cls = {"europeans": cf.Europeans, "barrier": cf.UOC, "asian": cf.Asian}
            book = cf.Book([cls[k](**kw) for k, kw in W.c4_parts()])
            w = W.c4_notionals(len(book.payoffLabels()))
            cf.mcParallelSimulAAD(book, mdl, weights=w, num=cf.NumericalParam(numPath=W.C4_PATHS, precision=prec))
        elif name.startswith("c2"):
            prec = cf.FP32 if "fp32" in name else cf.FP64
            cf.mcParallelSimul(prd, mdl, num=cf.NumericalParam(numPath=W.C2_PATHS, precision=prec), per_path=False)
        else:
            raise SystemExit("unknown workload " + name)
        stats = cf.last_kernel_stats() if name != "copy" else {}
        print(name, stats.get("main_kernel_ms", ""), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="all")
    ap.add_argument("--calls", type=int, default=3)
    a = ap.parse_args()
    names = (["c3_fp32_aad", "c3_fp64_aad", "c3_fp32_rec", "c3_fp64_rec", "c4_fp32", "c4_fp64", "c2_fp32", "c2_fp64",
              "copy"]
             if a.workload == "all" else [a.workload])
    for n in names:
        run(n, a.calls)


if __name__ == "__main__":
    main()
