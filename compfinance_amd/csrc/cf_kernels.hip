// Cross-block reductions, the Sobol points kernel, and the dispatch of
// the path kernels (instantiated per family in cf_kern_inst.hip; the kernel
// itself is in cf_path_kernel.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "cf_path_kernel.h"

namespace cfk {

// Column sums of two row-major matrices with the same row count,
// [rows][cols0] and [rows][cols1] (cols1 may be 0), in one launch: blocks
// [0, reduce_blocks(cols0)) take the first, the rest the second.  A block of
// 1024 threads takes kReduceCols (16) columns, so a grid of local-vol
// adjoints spreads over every compute unit; lane l of wave w works column
// l % 16 and sums rows p, p + 64, ... in order (p = 4w + l / 16, the row
// phase), and the 64 phase sums of a column are added in phase order, so the
// result does not depend on scheduling.
__device__ __forceinline__ void reduce_rows_block(const double* in0, int cols0, double* out0, const double* in1,
                                                  int cols1, double* out1, int rows, int bid)
{
    static_assert(kReduceCols == 16, "four row phases of 16 columns per wave");
    __shared__ double part[64][kReduceCols];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int b0 = reduce_blocks(cols0);
    const bool first = bid < b0;
    const double* in = first ? in0 : in1;
    double* out = first ? out0 : out1;
    const int cols = first ? cols0 : cols1;
    const int col = lane & (kReduceCols - 1), phase = 4 * w + (lane >> 4);
    const int c = (first ? bid : bid - b0) * kReduceCols + col;
    double s = 0.0;
    if (c < cols)
        for (int r = phase; r < rows; r += 64) s += in[size_t(r) * cols + c];
    part[phase][col] = s;
    __syncthreads();
    if (w == 0 && lane < kReduceCols && c < cols) {
        double t = 0.0;
#pragma unroll 16
        for (int k = 0; k < 64; ++k) t += part[k][col];
        out[c] = t;
    }
}

__global__ void __launch_bounds__(1024) reduce_rows_kernel(const double* in0, int cols0, double* out0,
                                                           const double* in1, int cols1, double* out1, int rows)
{
    reduce_rows_block(in0, cols0, out0, in1, cols1, out1, rows, int(blockIdx.x));
}

// The reductions of a packed launch: blocks [first[r], first[r + 1]) sum call
// r's rows exactly as reduce_rows_kernel sums them alone.
__global__ void __launch_bounds__(1024) reduce_packed_kernel(const ReducePack R)
{
    const int b = int(blockIdx.x);
    int r = 0;
    while (r + 1 < R.n && b >= R.first[r + 1]) ++r;
    r = __builtin_amdgcn_readfirstlane(r);
    reduce_rows_block(R.in0[r], R.cols0[r], R.out0[r], R.in1[r], R.cols1[r], R.out1[r], R.rows[r], b - R.first[r]);
}

// ------------------------------------------------------------------ Sobol directions

// Direction numbers [32][kSobolDims] of every device, in a module variable
// (no allocation): filled once per device by the engine.
__device__ uint32_t g_sobol_dirs[32 * kSobolDims];

int upload_sobol_dirs(const uint32_t* host, uint32_t** device)
{
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(g_sobol_dirs));
    if (e != hipSuccess) return int(e);
    e = hipMemcpy(p, host, sizeof(g_sobol_dirs), hipMemcpyHostToDevice);
    if (e != hipSuccess) return int(e);
    *device = static_cast<uint32_t*>(p);
    return 0;
}

// ------------------------------------------------------------------ Sobol points

// Words and normals of paths [first, first + npath) in dim dimensions, formed
// exactly as the path kernels form them (sobol_stage_lo, sobol_block,
// sobol_low_word / sobol_words4, inv_normal); written [path][dim].
// 256-thread blocks; LDS: LO (rows, quads or LEAN), one high row per wave,
// the tail row, and (FP32 FULL) the inverse-normal table.  The quads layout
// reads dimensions four at a time up to the last whole group, then one at a
// time, as the FP32 adjoint kernels' forward sweep does.
template <typename R, int LO>
__global__ void __launch_bounds__(256) sobol_points_kernel(const uint32_t* dirs, int dim, int64_t first, int64_t npath,
                                                           int merge_tail, uint32_t* words, R* normals)
{
    constexpr bool LEAN = LO == kLoLean;
    extern __shared__ __align__(16) unsigned char smem[];
    const int waves = blockDim.x >> 6, tid = threadIdx.x, nthr = blockDim.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const size_t o_hi = cf_align16(sobol_lo_words(dim, LEAN) * 4);
    const size_t o_tail = cf_align16(o_hi + size_t(waves) * sobol_hi_stride(dim) * 4);
    const size_t o_invn = cf_align16(o_tail + size_t(sobol_hi_stride(dim)) * 4);
    uint32_t* lo_tab = reinterpret_cast<uint32_t*>(smem);
    uint32_t* hi = reinterpret_cast<uint32_t*>(smem + o_hi) + wave * sobol_hi_stride(dim);
    uint32_t* hi_tail = reinterpret_cast<uint32_t*>(smem + o_tail);
    float* invn = reinterpret_cast<float*>(smem + o_invn);
    sobol_stage_lo<LO>(lo_tab, dirs, dim, tid, nthr);
    if constexpr (sizeof(R) == 4 && !LEAN)
        for (int e = tid; e < kInvNormalRows * 4; e += nthr) invn[e] = kInvNormalF32[e];
    __syncthreads();
    const R* tab;
    if constexpr (LEAN && sizeof(R) == 4) tab = reinterpret_cast<const R*>(kInvNormalF32);
    else tab = reinterpret_cast<const R*>(invn);
    tab -= 4 * kInvNormalRowBias;
    const SobolLaneMasks lm = sobol_lane_masks(lane);
    const int64_t nlo = first + 1, nhi = first + npath;
    const int64_t q0 = nlo >> 6, q1 = nhi >> 6;
    const bool merge = merge_tail != 0;
    const int64_t qe = merge ? q1 - 1 : q1;
    for (int64_t q = q0 + int64_t(blockIdx.x) * waves + wave; q <= qe; q += int64_t(gridDim.x) * waves) {
        const SobolLane sl = sobol_block(hi, hi_tail, dirs, dim, q, q0, q1, nlo, nhi, merge, lane, true);
        if (sl.active) {
            auto put = [&](int i, uint32_t w) {
                if (words) words[size_t(sl.pidx) * dim + i] = w;
                if (normals) normals[size_t(sl.pidx) * dim + i] = inv_normal(w, tab);
            };
            int i = 0;
            if constexpr (LO == kLoQuads) {
                for (; i + 4 <= dim; i += 4) {
                    uint32_t w4[4];
                    sobol_words4(lo_tab, sl.hl, lane, i, w4);
                    for (int k = 0; k < 4; ++k) put(i + k, w4[k]);
                }
            }
            for (; i < dim; ++i) put(i, sl.hl[i] ^ sobol_low_word<LO>(lo_tab, lm, lane, i));
        }
        // the rows are rewritten for the wave's next block
        wave_sync_lds();
    }
}

size_t sobol_points_smem(int real_bytes, bool lean, int dim, int block)
{
    const size_t o_hi = cf_align16(sobol_lo_words(dim, lean) * 4);
    const size_t o_tail = cf_align16(o_hi + size_t(block / kWave) * sobol_hi_stride(dim) * 4);
    const size_t o_invn = cf_align16(o_tail + size_t(sobol_hi_stride(dim)) * 4);
    return o_invn + (real_bytes == 4 && !lean ? size_t(kInvNormalRows) * 16 : 0);
}

int launch_sobol_points(int real_bytes, int layout, const uint32_t* dirs, int dim, int64_t first, int64_t npath,
                        uint32_t* words, void* normals, void* stream)
{
    const int block = 256;
    const bool lean = layout == kLoLean;
    const size_t smem = sobol_points_smem(real_bytes, lean, dim, block);
    const bool merge = merge_tail_blocks(first, npath);
    const int64_t blocks64 = ((first + npath) >> 6) - ((first + 1) >> 6) + 1 - (merge ? 1 : 0);
    const int grid = int(std::min<int64_t>((blocks64 + block / kWave - 1) / (block / kWave), 4096));
    const void* f;
    auto pick = [&](auto real) -> const void* {
        using T = decltype(real);
        if (layout == kLoLean) return reinterpret_cast<const void*>(&sobol_points_kernel<T, kLoLean>);
        if (layout == kLoQuads) return reinterpret_cast<const void*>(&sobol_points_kernel<T, kLoQuads>);
        return reinterpret_cast<const void*>(&sobol_points_kernel<T, kLoRows>);
    };
    f = real_bytes == 4 ? pick(0.0f) : pick(0.0);
    if (smem > 65536) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
        if (e != hipSuccess) return int(e);
    }
    int mt = merge ? 1 : 0;
    void* args[] = {const_cast<uint32_t**>(&dirs), &dim, &first, &npath, &mt, &words, &normals};
    return int(hipLaunchKernel(f, dim3(grid), dim3(block), args, smem, static_cast<hipStream_t>(stream)));
}

// ------------------------------------------------------------------ dispatch

static const void* pick(const KernelConfig& c)
{
    using Pick = const void* (*)(int, int, bool);
    static const Pick table[2][3][2] = {
        {{pick_r0_a0_b0, pick_r0_a0_b1}, {pick_r0_a1_b0, pick_r0_a1_b1}, {pick_r0_a2_b0, pick_r0_a2_b1}},
        {{pick_r1_a0_b0, pick_r1_a0_b1}, {pick_r1_a1_b0, pick_r1_a1_b1}, {pick_r1_a2_b0, pick_r1_a2_b1}}};
    const int adj = !c.adjoint ? ADJ_NONE : c.recompute ? ADJ_RECOMPUTE : ADJ_CHECKPOINT;
    if (c.item) {
        if (adj != ADJ_CHECKPOINT || c.packed) return nullptr;
        if (c.mrg) return (c.precision == 0 ? pick_r0_a1_b4 : pick_r1_a1_b4)(c.model, c.product, c.lean);
        return (c.precision == 0 ? pick_r0_a1_b2 : pick_r1_a1_b2)(c.model, c.product, c.lean);
    }
    if (c.packed) {
        if (c.mrg) return nullptr;
        static const Pick packed[2][3] = {{pick_r0_a0_b3, pick_r0_a1_b3, pick_r0_a2_b3},
                                          {pick_r1_a0_b3, pick_r1_a1_b3, pick_r1_a2_b3}};
        return packed[c.precision == 0 ? 0 : 1][adj](c.model, c.product, c.lean);
    }
    return table[c.precision == 0 ? 0 : 1][adj][c.mrg ? 1 : 0](c.model, c.product, c.lean);
}

int launch_path_kernel(const KernelConfig& cfg, const LaunchParams& p, int grid, int block, size_t smem, void* stream)
{
    if (cfg.packed) return int(hipErrorInvalidValue);
    const void* f = pick(cfg);
    if (!f) return int(hipErrorInvalidDeviceFunction);
    void* args[] = {const_cast<LaunchParams*>(&p)};
    hipError_t e = hipLaunchKernel(f, dim3(grid), dim3(block), args, smem, static_cast<hipStream_t>(stream));
    return int(e);
}

int launch_path_kernel_packed(const KernelConfig& cfg, const PackArgs& A, int block, size_t smem, void* stream)
{
    if (!cfg.packed || cfg.item || cfg.mrg || A.n < 1 || A.n > kMaxPack) return int(hipErrorInvalidValue);
    const void* f = pick(cfg);
    if (!f) return int(hipErrorInvalidDeviceFunction);
    void* args[] = {const_cast<PackArgs*>(&A)};
    return int(hipLaunchKernel(f, dim3(A.first[A.n]), dim3(block), args, smem, static_cast<hipStream_t>(stream)));
}

int launch_reduce_packed(const ReducePack& R, void* stream)
{
    if (R.n < 1 || R.n > kMaxPack) return int(hipErrorInvalidValue);
    if (R.first[R.n] == 0) return 0;
    hipLaunchKernelGGL(reduce_packed_kernel, dim3(R.first[R.n]), dim3(1024), 0, static_cast<hipStream_t>(stream), R);
    return int(hipGetLastError());
}

int launch_reduce_rows(const double* in0, int cols0, double* out0, const double* in1, int cols1, double* out1,
                       int rows, void* stream)
{
    const int threads = 1024;
    const int blocks = reduce_blocks(cols0) + reduce_blocks(cols1);
    if (blocks == 0) return 0;
    hipLaunchKernelGGL(reduce_rows_kernel, dim3(blocks), dim3(threads), 0, static_cast<hipStream_t>(stream), in0,
                       cols0, out0, in1, cols1, out1, rows);
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
