// Host-side checks of the launch arithmetic shared with the kernels
// (compfinance_amd/csrc/cf_internal.h): the merge of the last block of 64
// Sobol indices into the first block's idle lanes, and the shared-memory
// layout: every region the path kernel addresses, at the largest index its
// code forms, lies inside its own region and inside the dynamic LDS size the
// host passes to the launch (smem_layout(...).total), for each layout.
// Prints one JSON object; tests/test_layout.py reads it.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../compfinance_amd/csrc/cf_internal.h"

using namespace cfk;

// One region: [off, off + used) must lie in [off, next region's offset) and in
// [0, total).  `used` is one past the largest byte the kernel's code addresses.
static long g_out_of_region = 0, g_regions = 0;
static void region(size_t off, size_t used, size_t next, size_t total)
{
    ++g_regions;
    if (used == 0) return;
    if (off + used > next || off + used > total) ++g_out_of_region;
}

// The kernel's largest index of each region (cf_path_kernel.h), for a block of
// `waves` waves on n steps, m knots, ns samples, npay payoffs.
static void kernel_extents(int n, int m, int ns, int npay, int waves, int rb, bool dup, bool eus, bool ad, bool lean,
                           int nb, int copies)
{
    const SmemLayout L = smem_layout(n, m, ns, npay, waves, rb, dup, eus, ad, lean, nb, copies);
    const size_t T = L.total;
    const int mp = lv_cols(m);
    // stage loops: lo_tab[e], e < n * 8 (LEAN) or e < 4 ceil(n / 4) * 64 (FULL,
    // interleaved by four dimensions, read 16 bytes at a time);
    // hi_all[wave * stride + d], hi_tail[d], d < n (16-byte reads of d < 4 floor(n / 4))
    const size_t hs = size_t(sobol_hi_stride(n));
    region(L.lo, (lean ? size_t(n) * 8 : size_t((n + 3) / 4) * 4 * kWave) * 4, L.hi, T);
    region(L.hi, (size_t(waves - 1) * hs + size_t(n)) * 4, L.hi_tail, T);
    region(L.hi_tail, size_t(n) * 4, L.step_sample, T);
    // step_sample[e], e <= n; the FP32 checkpoint reverse sweep reads the
    // aligned group of four holding step n - 1, up to e = 4 floor((n - 1) / 4) + 3
    const size_t ss_used = size_t(rb == 4 && ad && dup && n > 0 ? std::max(n, 4 * ((n - 1) / 4) + 3) + 1 : n + 1);
    region(L.step_sample, ss_used * 4, L.invn, T);
    // FP32 FULL: the table itself, rows 0 .. kInvNormalRows - 1 of 16 bytes
    region(L.invn, rb == 4 && !lean ? size_t(kInvNormalRows) * 16 : 0, L.knots, T);
    // interval pairs kxi[2c], kxi[2c + 1], c <= m; knot search reads kxi[2 + 2 (m - 1)]
    region(L.knots, dup ? size_t(2 * m + 2) * rb : 0, L.bucket, T);
    region(L.bucket, dup && !lean ? size_t(nb) * (rb == 4 ? sizeof(BucketEntry<float>) : sizeof(BucketEntry<double>)) : 0,
           L.iv, T);
    // iv[i * mp + c + 1], i < n, c <= m
    region(L.iv, dup && !lean ? (size_t(n - 1) * mp + m + 2) * rb : 0, L.drift, T);
    region(L.drift, dup ? 0 : size_t(n) * rb, L.stdv, T);
    region(L.stdv, dup ? 0 : size_t(n) * rb, L.samples, T);
    region(L.samples, dup ? 0 : size_t(3 * ns) * 8, L.pay_first, T);
    region(L.pay_first, eus ? size_t(ns + 1) * 4 : 0, L.pay_strike, T);
    region(L.pay_strike, eus ? size_t(npay) * 8 : 0, L.acc, T);
    region(L.acc, size_t(waves) * out_stride(npay) * 8, L.tails, T);
    // FP64: tails[wave * row + l], l < row (the inverse normal's tail draws)
    region(L.tails, rb == 8 ? size_t(waves) * tail_row(lean) * 8 : 0, L.ivbar, T);
    // ivbar: copy c at c * ivbar_copy_stride, node (i, j + 1) with j <= m
    // (padded column m + 1), and the end columns 0 and m + 1 on write-out
    const size_t last = size_t(copies - 1) * ivbar_copy_stride(n, m) + size_t(n - 1) * mp + m + 1;
    region(L.ivbar, ad && dup ? (last + 1) * 8 : 0, T, T);
}

int main()
{
    // merge_tail_blocks against brute force: paths [f, f + n) use Sobol indices
    // f + 1 .. f + n; the last aligned block of 64 can ride in the first one
    // when it differs from it and each of its lanes is idle in the first
    long cases = 0, wrong = 0, merged = 0;
    for (int f = 0; f < 260; ++f) {
        for (int n = 1; n < 400; ++n) {
            const long lo = f + 1, hi = f + n, q0 = lo / 64, q1 = hi / 64;
            bool fits = q1 > q0;
            for (long i = q1 * 64; fits && i <= hi; ++i) {
                const long lane = i % 64;
                const long in_first = q0 * 64 + lane;   // the same lane of the first block
                if (in_first >= lo && in_first <= hi) fits = false;
            }
            ++cases;
            merged += fits ? 1 : 0;
            if (fits != merge_tail_blocks(f, n)) ++wrong;
        }
    }
    // layouts: every region 16-byte aligned; the benchmark's adjoint blocks fit
    // 160 KiB; the second adjoint grid starts 64 bytes into a 128-byte bank row
    long misaligned = 0;
    size_t fp32_c3 = 0, fp64_c3 = 0, lean_60x60 = 0;
    int copy_offset_mod128 = -1;
    for (int real = 4; real <= 8; real += 4) {
        for (int lean = 0; lean < 2; ++lean) {
            for (int copies = 1; copies <= 2; ++copies) {
                const SmemLayout L = smem_layout(156, 30, 157, 55, 16, real, true, true, true, lean != 0, 64,
                                                 copies);
                const size_t off[] = {L.lo, L.hi, L.hi_tail, L.step_sample, L.invn, L.knots, L.bucket, L.iv,
                                      L.drift, L.stdv, L.samples, L.pay_first, L.pay_strike, L.acc,
                                      L.tails, L.ivbar, L.total};
                for (size_t o : off) misaligned += (o % 16) ? 1 : 0;
            }
        }
    }
    // region extents over layouts, precisions, models, products and adjoints
    for (int rb = 4; rb <= 8; rb += 4)
        for (int lean = 0; lean < 2; ++lean)
            for (int ad = 0; ad < 2; ++ad)
                for (int dup = 0; dup < 2; ++dup)
                    for (int eus = 0; eus < 2; ++eus)
                        for (int copies = 1; copies <= 2; ++copies)
                            for (int n : {1, 2, 156, 261})
                                for (int m : {1, 2, 30, 60, 255})
                                    for (int waves : {1, 8, 16})
                                        kernel_extents(n, m, n + 1, eus ? 31 : 2, waves, rb, dup != 0, eus != 0,
                                                       ad != 0, lean != 0, 2 * m + 3, copies);
    // FP32 inverse normal: the row the kernel reads, (bits(float(xs)) >> 18 &
    // 0x3ffc) / 4 - kInvNormalRowBias, is inside the table for every xs in
    // [1, 2^31] (each binade's ends and every row boundary, from both sides)
    long bad_rows = 0, row_probes = 0;
    {
        auto row_of = [](uint32_t xs) {
            const float f = float(xs);
            uint32_t bits;
            std::memcpy(&bits, &f, 4);
            return int((bits >> 18) & 0x3ffcu) / 4 - (127 << 3);
        };
        for (int e = 0; e <= 31; ++e)
            for (int q = 0; q < 8; ++q) {
                const double b = (1.0 + q / 8.0) * double(1ull << e);
                for (double d : {-1.0, 0.0, 1.0}) {
                    const double v = b + d;
                    if (v < 1.0 || v > 2147483648.0) continue;
                    const int r = row_of(uint32_t(v));
                    ++row_probes;
                    if (r < 0 || r >= kInvNormalRows) ++bad_rows;
                }
            }
    }
    // interleaved tables: the Sobol LO word of (dimension d, lane r) and the
    // FP32 checkpoint element of (step i, lane r) lie inside their tables, and
    // the kernel's staging order (word e -> dimension 4 (e >> 8) + (e & 3),
    // lane (e >> 2) & 63) inverts the read index
    long il_probes = 0, il_bad = 0;
    for (int n = 1; n <= 300; ++n)
        for (int d = 0; d < n; ++d)
            for (int r = 0; r < kWave; ++r) {
                const int e = sobol_lo_el(d, r);
                ++il_probes;
                if (e < 0 || size_t(e) >= sobol_lo_words(n, false)) ++il_bad;
                if (((e >> 8) << 2) + (e & 3) != d || ((e >> 2) & 63) != r) ++il_bad;
                if (e >= ckpt_steps(n, 4) * kWave) ++il_bad;
            }
    fp32_c3 = smem_layout(156, 30, 157, 2, 16, 4, true, false, true, false, 64, 2).total;
    fp64_c3 = smem_layout(156, 30, 157, 2, 16, 8, true, false, true, false, 64, 1).total;
    lean_60x60 = smem_layout(261, 60, 262, 2, 16, 8, true, false, true, true, 64, 1).total;
    copy_offset_mod128 = int((size_t(ivbar_copy_stride(156, 30)) * 8) % 128);
    std::printf("{\"cases\": %ld, \"wrong\": %ld, \"merged\": %ld, \"misaligned\": %ld, \"fp32_c3\": %zu, "
                "\"fp64_c3\": %zu, \"lean_60x60_5y_fp64\": %zu, \"copy_offset_mod128\": %d, \"regions\": %ld, "
                "\"out_of_region\": %ld, \"row_probes\": %ld, \"bad_rows\": %ld, \"il_probes\": %ld, "
                "\"il_bad\": %ld}\n",
                cases, wrong, merged, misaligned, fp32_c3, fp64_c3, lean_60x60, copy_offset_mod128, g_regions,
                g_out_of_region, row_probes, bad_rows, il_probes, il_bad);
    return 0;
}
