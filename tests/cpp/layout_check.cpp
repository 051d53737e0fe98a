// Host-side checks of the launch arithmetic shared with the kernels
// (compfinance_amd/csrc/cf_internal.h): the merge of the last block of 64
// Sobol indices into the first block's idle lanes, and the shared-memory
// layout.  Prints one JSON object; tests/test_layout.py reads it.
#include <cstdio>

#include "../../compfinance_amd/csrc/cf_internal.h"

int main()
{
    using namespace cfk;
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
                const SmemLayout L = smem_layout(156, 30, 157, 55, 16, real, true, true, true, lean != 0, 64, 1000,
                                                 copies);
                const size_t off[] = {L.lo, L.hi, L.hi_tail, L.step_sample, L.invn, L.knots, L.bucket, L.iv,
                                      L.drift, L.stdv, L.samples, L.pay_first, L.pay_strike, L.book, L.acc,
                                      L.ivbar, L.total};
                for (size_t o : off) misaligned += (o % 16) ? 1 : 0;
            }
        }
    }
    fp32_c3 = smem_layout(156, 30, 157, 2, 16, 4, true, false, true, false, 64, 0, 2).total;
    fp64_c3 = smem_layout(156, 30, 157, 2, 16, 8, true, false, true, false, 64, 0, 1).total;
    lean_60x60 = smem_layout(261, 60, 262, 2, 16, 8, true, false, true, true, 64, 0, 1).total;
    copy_offset_mod128 = int((size_t(ivbar_copy_stride(156, 30)) * 8) % 128);
    std::printf("{\"cases\": %ld, \"wrong\": %ld, \"merged\": %ld, \"misaligned\": %ld, \"fp32_c3\": %zu, "
                "\"fp64_c3\": %zu, \"lean_60x60_5y_fp64\": %zu, \"copy_offset_mod128\": %d}\n",
                cases, wrong, merged, misaligned, fp32_c3, fp64_c3, lean_60x60, copy_offset_mod128);
    return 0;
}
