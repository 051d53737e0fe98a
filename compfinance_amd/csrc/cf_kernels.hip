// Cross-block reductions, the diagnostic copy kernel, and the dispatch of
// the path kernels (instantiated per family in cf_kern_inst.hip; the kernel
// itself is in cf_path_kernel.h).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "cf_internal.h"

namespace cfk {

// Column sums of a row-major [rows][cols] matrix.  A block takes 64 columns;
// wave w sums rows w, w + 16, ... in order and the 16 partial sums are added
// in wave order, so the result does not depend on scheduling.
__global__ void __launch_bounds__(1024) reduce_rows_kernel(const double* in, int rows, int cols, double* out)
{
    __shared__ double part[16][kWave];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int c = blockIdx.x * kWave + lane;
    double s = 0.0;
    if (c < cols)
        for (int r = w; r < rows; r += 16) s += in[size_t(r) * cols + c];
    part[w][lane] = s;
    __syncthreads();
    if (w == 0 && c < cols) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 16; ++k) t += part[k][lane];
        out[c] = t;
    }
}

// Plain streaming copy, 16 bytes per lane per iteration (diagnostics).
__global__ void copy_kernel(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n)
{
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

int launch_copy(const void* in, void* out, size_t bytes, void* stream)
{
    const size_t n = bytes / 16;
    hipLaunchKernelGGL(copy_kernel, dim3(2048), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint4*>(in), static_cast<uint4*>(out), n);
    return int(hipGetLastError());
}

// ------------------------------------------------------------------ dispatch

static const void* pick(const KernelConfig& c)
{
    const int adj = !c.adjoint ? ADJ_NONE : c.recompute ? ADJ_RECOMPUTE : ADJ_CHECKPOINT;
    if (c.precision == 0) {
        switch (adj) {
        case ADJ_NONE: return CF_PICK_NAME(0, 0)(c.model, c.product, c.lean);
        case ADJ_CHECKPOINT: return CF_PICK_NAME(0, 1)(c.model, c.product, c.lean);
        default: return CF_PICK_NAME(0, 2)(c.model, c.product, c.lean);
        }
    }
    switch (adj) {
    case ADJ_NONE: return CF_PICK_NAME(1, 0)(c.model, c.product, c.lean);
    case ADJ_CHECKPOINT: return CF_PICK_NAME(1, 1)(c.model, c.product, c.lean);
    default: return CF_PICK_NAME(1, 2)(c.model, c.product, c.lean);
    }
}

int launch_path_kernel(const KernelConfig& cfg, const LaunchParams& p, int grid, int block, size_t smem, void* stream)
{
    const void* f = pick(cfg);
    if (!f) return int(hipErrorInvalidDeviceFunction);
    void* args[] = {const_cast<LaunchParams*>(&p)};
    hipError_t e = hipLaunchKernel(f, dim3(grid), dim3(block), args, smem, static_cast<hipStream_t>(stream));
    return int(e);
}

int launch_reduce_rows(const double* in, int rows, int cols, double* out, void* stream)
{
    const int threads = 1024;
    const int blocks = (cols + kWave - 1) / kWave;
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(blocks), dim3(threads), 0, static_cast<hipStream_t>(stream), in, rows,
                       cols, out);
    return int(hipGetLastError());
}

int path_kernel_occupancy(const KernelConfig& cfg, int block, size_t smem, int* blocks_per_cu)
{
    const void* f = pick(cfg);
    if (!f) return int(hipErrorInvalidDeviceFunction);
    // the dynamic LDS limit and the occupancy of a kernel are queried once per size
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, size_t>, int> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(f, block, smem);
    auto it = cache.find(key);
    if (it != cache.end()) {
        *blocks_per_cu = it->second;
        return 0;
    }
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return int(e);
    }
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, block, smem);
    if (e == hipSuccess) cache[key] = *blocks_per_cu;
    return int(e);
}

}  // namespace cfk
