// Launch parameters and shared-memory layout of the path kernels, shared by
// the host engine and the HIP kernels.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <vector>

#ifdef __HIPCC__
#define CF_HD __host__ __device__
#else
#define CF_HD
#endif

namespace cfk {

enum ModelKind { MODEL_BS = 0, MODEL_DUPIRE = 1 };
enum ProductKind { PRD_EUROPEAN = 0, PRD_BARRIER = 1, PRD_EUROPEANS = 2, PRD_BOOK = 3 };
enum AdjointKind { ADJ_NONE = 0, ADJ_CHECKPOINT = 1, ADJ_RECOMPUTE = 2 };

constexpr int kWave = 64;

// Doubles per wave in the FP64 kernels' LDS row of inverse-normal tail draws.
CF_HD constexpr int tail_row(bool lean) { return lean ? 32 : 64; }
constexpr int kSobolDims = 1101;     // direction-number table stride

// Recompute variant: the Dupire state is kept at every kSeg-th timeline
// point (8 steps in FP32, 4 in FP64: the re-simulated segment lives in
// registers); Black-Scholes keeps nothing (normals are regenerated).
constexpr int kSegF32 = 8, kSegF64 = 4;

// The checkpoint adjoint recovers d x_{i+1} / d vol = g - vol from the two
// stored states as (x_{i+1} - x_i) / vol - vol / 2.  The difference is exact,
// but it carries the rounding of x_{i+1} (half an ulp of the state), divided
// by vol; below these vols the reverse sweep draws the normal again instead
// (FP64: the error stays below 1e-12 for states under 8 in magnitude).
constexpr double kRecoverVolF64 = 0x1p-10, kRecoverVolF32 = 0x1p-8;

// Book of products on shared paths (also an Asian on its own): vanilla calls
// attached to samples, up-and-out barriers and arithmetic-average groups.  Each
// barrier and each distinct fixing schedule keeps per-lane state in registers,
// hence the bounds; the book kernels run 512-thread blocks for the registers.
constexpr int kMaxBookBarriers = 16, kMaxBookGroups = 4;
constexpr int kBookBlock = 512;

// Threads per block of a path kernel: 1024, except the book kernels whose
// registers need 512 (FP64, and the FP32 recompute variant).
constexpr int book_block_limit(int prd, int real_bytes, int adj)
{
    return prd == PRD_BOOK && (real_bytes == 8 || adj == ADJ_RECOMPUTE) ? kBookBlock : 1024;
}

// Byte offsets of the book tables inside one blob (staged whole in LDS).
struct BookLayout {
    int n_van, n_bar, n_grp, n_asian;
    int bytes;
    int van_strike;   // [n_van] double
    int van_col;      // [n_van] int: payoff column
    int masks;        // [n_samples][4] uint32, 16-byte rows: barriers observing at s (bits 0-15)
                      // and maturing at s (16-31), schedules fixing at s (bits 0-3) and ending
                      // at s (4-7), first vanilla of s, first vanilla of s + 1
    int bar_par;      // [n_bar][4] double: strike, B + w, B - w, 1 / 2w
    int bar_par_f;    // [n_bar][4] the same in float
    int bar_col;      // [n_bar] int: column of the barrier payoff (vanilla next)
    int bar_put;      // [n_bar] int
    int grp_inv;      // [n_grp] double: 1 / number of fixings
    int grp_inv_f;    // [n_grp] float
    int grp_first;    // [n_grp + 1] int: Asians of each schedule
    int as_strike;    // [n_asian] double
    int as_col;       // [n_asian] int
    int as_put;       // [n_asian] int
};

// Rows of 64 reals in a wave's checkpoint slot.  The Dupire checkpoint
// variant also keeps, per step and lane, one byte: the knot count of the
// state's bracket (grids of up to kMaxStoredKnots knots).
// The Dupire checkpoint slot starts with kCkptGuardRows rows that the reverse
// sweep's prefetch may read below step 0 (values never used).
// FP32 Dupire checkpoints are interleaved by groups of four steps: the state
// of step i of lane l is element ((i >> 2) * 64 + l) * 4 + (i & 3) (a lane's
// four states in one 16-byte word), its bracket byte the same element of the
// byte rows that follow the states; the steps are padded to a multiple of four.
// FP64 keeps one row of 64 per step.
constexpr int kMaxStoredKnots = 255;
constexpr int kCkptGuardRows = 4;
CF_HD inline int ckpt_steps(int n_steps, int real_bytes) { return real_bytes == 4 ? (n_steps + 3) & ~3 : n_steps; }
CF_HD inline int ckpt_rows(int n_steps, bool dupire, bool recompute, int real_bytes)
{
    const int ns = ckpt_steps(n_steps, real_bytes);
    if (!recompute) return dupire ? kCkptGuardRows + ns + (ns + real_bytes - 1) / real_bytes : n_steps;
    if (!dupire) return 0;
    const int seg = real_bytes == 4 ? kSegF32 : kSegF64;
    return (n_steps + seg - 1) / seg;
}

// Rows of 64 reals of a path's end state kept between the forward-only and
// the reverse-only launches (ck_mode 1, 2): the state, the product state
// (alive, pnz, nzero, dead, euro, final forward, final numeraire) and, per
// book barrier and fixing schedule, two values (1 each for other products).
CF_HD inline int state_rows(bool book) { return 8 + 2 * (book ? kMaxBookBarriers : 1) + 2 * (book ? kMaxBookGroups : 1); }

// Per-block accumulator row: payoff sums, payoff sums of squares, aggregate
// sum, aggregate sum of squares, then up to kMaxParamAdj parameter adjoints
// accumulated in registers (Dupire: spot; Black-Scholes: spot, vol, rate, div).
constexpr int kMaxParamAdj = 4;
CF_HD inline int out_stride(int n_pay) { return 2 * n_pay + 2 + kMaxParamAdj; }

struct LaunchParams {
    // paths [first_path, first_path + n_path); path p uses Sobol point p + 1
    int64_t first_path;
    int64_t n_path;
    int merge_tail;              // 1: the last block of 64 Sobol indices runs in the first block's
                                 // idle low lanes (see merge_tail_blocks)
    int n_steps;                 // simulation steps = Sobol dimension
    const uint32_t* dirs;        // [32][kSobolDims] direction numbers

    // model tables (real type of the kernel)
    int n_spots;                 // Dupire knots
    // interval tables indexed by the knot count c = #{knots <= x} in [0, m]:
    const void* knots;           // [n_spots + 1] first knot, then the knots: log spots
                                 // (FP64) / log(spot/S0) (FP32); interval c starts at [c]
    const void* inv_dx;          // [n_spots + 1] 0, 1 / (knot[j+1] - knot[j]) for j < m - 1, 0
    const void* iv;              // [n_steps][n_spots + 2] sqrt(dt)-scaled local vols, each
                                 // row padded with its first and last value
    const void* bucket;          // [n_buckets] BucketEntry<real>: the knot search table
    int n_buckets, bucket_fix;   // table size, knots per bucket at most
    double bucket_invh, bucket_c0;   // bucket_pos coefficients, values of the kernel's real type
    const void* drift;           // Black-Scholes [n_steps]
    const void* stdv;            // Black-Scholes [n_steps]
    double spot;                 // S0
    double x0;                   // initial state: log(S0) (FP64), 0 (FP32)
    int spot_measure;
    double bs_vol;               // Black-Scholes volatility (adjoint only)
    const double* bs_aux;        // Black-Scholes adjoint data: [n_samples] sample
                                 // time, [n_samples] forward maturity - sample time,
                                 // [n_samples] discount maturity - sample time,
                                 // [n_steps] step length

    // product samples on the simulation timeline
    int n_samples;
    const int* step_sample;      // [n_steps + 1] sample index at each timeline point, -1 if none
    const double* ff;            // [n_samples] forward factor (Black-Scholes)
    const double* disc;          // [n_samples] discount factor
    const double* num;           // [n_samples] numeraire (spot measure: divided by spot)

    // product
    int n_pay;
    int put;
    double strike;
    double bar_smooth, minus_smooth, two_smooth, inv_two_smooth;
    double bar_xlo;              // Dupire barrier: at or below this state the spot is under the
                                 // smoothing zone (with a margin over exp's rounding)
    const int* pay_first;        // Europeans [n_samples + 1]: payoffs of sample s
    const double* pay_strike;    // Europeans [n_pay]
    const void* book;            // book tables, layout bk
    BookLayout bk;
    double book_lo;              // the lowest B - w of the book's barriers (+inf without one)

    // mrg32k3a (null: Sobol): jump matrices [32][2][9], the two component
    // matrices raised to n_steps * 2^j (mod m1, m2), and the seeds as residues
    const uint64_t* mrg_jump;
    uint32_t mrg_seed1, mrg_seed2;

    // outputs
    double* block_out;           // [gridDim][out_stride(n_pay)]
    double* per_path;            // [n_path][n_pay] or null

    // adjoint
    const double* weights;       // [n_pay] aggregation weights
    void* ckpt;                  // [slots][ckpt_rows][64] path states or normals,
                                 // then (Dupire) [n_steps][64] bracket bytes
    int ivbar_copies;            // LDS local-vol adjoint grids per block: even and odd lanes (2) or one
    int iv_recover;              // 1: every local-vol table entry is at least kRecoverVol (the
                                 // reverse sweep needs no normal)
    double* ivbar_blocks;        // [gridDim][n_steps][n_spots] local-vol adjoints
    double* per_path_agg;        // [n_path] or null
    // checkpoint adjoint shared across payoffs (itemized risk): 0 one call
    // (forward and reverse, checkpoint slot per wave); 1 forward only, the
    // checkpoint slot per block of 64 Sobol indices kept and each path's end
    // state written to `state`; 2 reverse only, from those slots and states
    int ck_mode;
    int rev_point;               // the reverse starts at this timeline point's step (n_steps: all;
                                 // itemized risk: the one payoff's last observation)
    void* state;                 // [blocks of 64 paths][state_rows(product)][64] reals
};

// Calls packed into one launch of a path kernel: call r runs on blocks
// [first[r], first[r + 1]).  Passed by value (kernel arguments stay under
// 4 KiB).
constexpr int kMaxPack = 8;

// Cross-block column sums: a block of the reduction takes kReduceCols
// columns of one matrix (launch_reduce_rows, the packed reduction).
constexpr int kReduceCols = 16;
CF_HD inline int reduce_blocks(int cols) { return (cols + kReduceCols - 1) / kReduceCols; }
struct PackArgs {
    LaunchParams p[kMaxPack];
    int first[kMaxPack + 1];
    int n;
};
static_assert(sizeof(PackArgs) <= 3584, "packed kernel arguments");

// The cross-block reductions of the calls of a packed launch, one launch:
// call r's column sums run on blocks [first[r], first[r + 1]).
struct ReducePack {
    const double* in0[kMaxPack];
    double* out0[kMaxPack];
    const double* in1[kMaxPack];
    double* out1[kMaxPack];
    int cols0[kMaxPack], cols1[kMaxPack], rows[kMaxPack];
    int first[kMaxPack + 1];
    int n;
};

// Shared-memory carve of one block.  Offsets in bytes, 16-byte aligned.
struct SmemLayout {
    size_t lo, hi, hi_tail, step_sample, invn, knots, bucket, iv, drift, stdv, samples, pay_first, pay_strike, acc,
        tails, ivbar, total;
};

CF_HD inline size_t cf_align16(size_t x) { return (x + 15) & ~size_t(15); }

// Sobol words in LDS.  LO, the low words of the 64 lanes of a block of 64
// Sobol indices, is interleaved by groups of four dimensions: dimension d of
// lane l is word ((d >> 2) * 64 + l) * 4 + (d & 3), so a lane reads four
// consecutive dimensions in one 16-byte word (dimensions padded to a multiple
// of four).  The LEAN layout keeps only direction numbers 0..5 of each
// dimension (8 words).  A wave's row of high words holds the dimensions at a
// stride of a multiple of four words, 16-byte aligned.
CF_HD inline int sobol_dims4(int n) { return (n + 3) & ~3; }
CF_HD inline size_t sobol_lo_words(int n, bool lean) { return size_t(lean ? n * 8 : sobol_dims4(n) * kWave); }
CF_HD inline int sobol_hi_stride(int n) { return sobol_dims4(n); }
CF_HD inline int sobol_lo_el(int d, int lane) { return ((d >> 2) << 8) + (lane << 2) + (d & 3); }

// Paths [first, first + n) use Sobol indices [first + 1, first + n], taken in
// blocks of 64 aligned indices.  When the last block's indices (lanes 0..(hi
// mod 64)) all fall below the first block's first lane (lo mod 64), the last
// block runs in the first block's idle lanes.  Blocks of 64 then number
// (hi >> 6) - (lo >> 6) instead of one more.
CF_HD inline bool merge_tail_blocks(int64_t first, int64_t n)
{
    const int64_t lo = first + 1, hi = first + n;
    return (hi >> 6) > (lo >> 6) && (hi & 63) < (lo & 63);
}

// Two layouts.  FULL keeps the per-lane low Sobol words (one 32-bit word per
// dimension and lane) and the local-vol table in LDS.  LEAN, for long
// timelines or large grids, keeps only direction numbers 0..5 of each
// dimension (the low word is formed in registers) and reads the local-vol
// table from global memory through the caches.
// Entry of the knot search table: bucket b holds lo, the number of knots in
// earlier buckets, and the next two knots (NaN past the last knot, so that
// their comparisons fail), so one LDS access brings all a search compares
// against.
template <typename R>
struct alignas(16) BucketEntry {
    R k0, k1;
    int lo, pad;
};

// ---- knot search (upper_bound on the knots, interp.h:40), one definition for
// the kernels and for the host that builds the table.  The bucket of state x
// is floor(clamp(fma(x, invh, c0), 0, nb - 1)): monotone in x, so the knots
// of earlier buckets are all <= x and those of later buckets all > x.
// Bucket b holds lo, the number of knots in earlier buckets, and its first two
// knots (NaN past the last knot); at most `fix` knots share a bucket.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ inline float clamp_bucket(float b, float top) { return __builtin_amdgcn_fmed3f(b, 0.0f, top); }
#else
inline float clamp_bucket(float b, float top)
{
    b = b > 0.0f ? b : 0.0f;
    return b < top ? b : top;
}
#endif
CF_HD inline double clamp_bucket(double b, double top)
{
    b = b > 0.0 ? b : 0.0;
    return b < top ? b : top;
}
CF_HD inline float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
CF_HD inline double fma_r(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Bucket of x, as a real in [0, top] (NaN: converts to bucket 0 on the device).
template <typename R>
CF_HD inline R bucket_pos(R x, R invh, R c0, R top)
{
    return clamp_bucket(fma_r(x, invh, c0), top);
}

template <typename R>
struct KnotSearch {
    const R* knots;     // knot c at knots[stride c]
    const BucketEntry<R>* table;
    int m, fix, stride;
    R invh, c0, top;    // top = number of buckets - 1
};

// #{knots <= x}, exactly as upper_bound (interp.h:40): states below the first
// knot map to bucket 0, whose count is 0 and whose comparisons all fail.
// WIDE: tables with more than two knots in a bucket (fix > 2); the FULL layout
// is only chosen for tables with fix <= 2 and compiles without the loop.
template <bool WIDE = true, typename R>
CF_HD inline int count_le(const KnotSearch<R>& s, R x)
{
    // one read: the bucket's knot count and its first two candidate knots
    // (NaN past the last knot); the knots increase, so each one <= x adds one
    const BucketEntry<R> e = s.table[int(bucket_pos(x, s.invh, s.c0, s.top))];
    const int top = s.m - 1;
    int c = e.lo + int(e.k0 <= x) + int(e.k1 <= x);
    if (WIDE) {
        for (int k = 2; k < s.fix; ++k) {   // grids with more than two knots in a bucket
            const R kc = s.knots[s.stride * (c < top ? c : top)];
            c += int((c <= top) & (kc <= x));
        }
    }
    return c;
}

// Host: the knot search table of increasing knots (at least one), in the
// kernel's real type.  Buckets of width at most half the smallest knot gap
// (at most 16384 of them), so that normally no more than two knots share one.
template <typename R>
struct KnotTable {
    std::vector<BucketEntry<R>> entries;
    int fix = 1;
    R invh = R(0), c0 = R(0);
};

template <typename R>
inline KnotTable<R> build_knot_table(const R* knots, int m)
{
    KnotTable<R> t;
    int nb = 1;
    const R x0 = knots[0];
    const double range = double(knots[m - 1]) - double(x0);
    double dmin = range;
    for (int j = 0; j + 1 < m; ++j) dmin = dmin < double(knots[j + 1]) - double(knots[j])
                                               ? dmin : double(knots[j + 1]) - double(knots[j]);
    if (m >= 2 && range > 0.0 && dmin > 0.0) {
        const double want = std::ceil(2.0 * range / dmin) + 1.0;
        nb = int(want < 16384.0 ? want : 16384.0);
    }
    t.invh = nb > 1 ? R(double(nb) / range) : R(0);
    t.c0 = R(-double(x0) * double(t.invh));
    const R top = R(nb - 1);
    std::vector<int> cnt(size_t(nb), 0);
    for (int k = 0; k < m; ++k) cnt[size_t(int(bucket_pos(knots[k], t.invh, t.c0, top)))]++;
    t.entries.resize(size_t(nb));
    const R none = R(NAN);
    int acc = 0;
    for (int b = 0; b < nb; ++b) {
        BucketEntry<R>& e = t.entries[size_t(b)];
        e.lo = acc;
        e.pad = 0;
        e.k0 = acc < m ? knots[acc] : none;
        e.k1 = acc + 1 < m ? knots[acc + 1] : none;
        acc += cnt[size_t(b)];
        t.fix = t.fix > cnt[size_t(b)] ? t.fix : cnt[size_t(b)];
    }
    return t;
}

// Rows of the FP32 inverse-normal table (four coefficients each): eighths of
// the 31 binades of 2^32 min(u, 1 - u), and its value at 1/2.
constexpr int kInvNormalRows = 31 * 8 + 1;

// Columns of a padded local-vol row: interval c of a row spans columns c and
// c + 1, flat outside the knots (the first and last value repeated).
CF_HD inline int lv_cols(int n_spots) { return n_spots + 2; }

// Doubles between the starts of the two LDS local-vol adjoint grids: the grid
// plus 8 doubles (64 bytes, half the banks), so that a node of the second grid
// shares no bank with the same node of the first (an even and an odd lane on
// one node in the same atomic instruction).
constexpr int kIvbarCopyPad = 8;
CF_HD inline int ivbar_copy_stride(int n_steps, int n_spots) { return n_steps * lv_cols(n_spots) + kIvbarCopyPad; }

CF_HD inline SmemLayout smem_layout(int n_steps, int n_spots, int n_samples, int n_pay, int waves,
                                    int real_bytes, bool dupire, bool europeans, bool adjoint, bool lean,
                                    int n_buckets, int ivbar_copies)
{
    SmemLayout L{};
    size_t o = 0;
    L.lo = o;          o = cf_align16(o + sobol_lo_words(n_steps, lean) * 4);
    L.hi = o;          o = cf_align16(o + size_t(waves) * sobol_hi_stride(n_steps) * 4);
    L.hi_tail = o;     o = cf_align16(o + size_t(sobol_hi_stride(n_steps)) * 4);
    L.step_sample = o; o = cf_align16(o + size_t(n_steps + 1) * 4);
    // LEAN reads the inverse-normal table and the bucket entries, like the
    // local-vol table, from global memory through L1
    L.invn = o;        o = cf_align16(o + (real_bytes == 4 && !lean ? size_t(kInvNormalRows) * 4 * 4 : 0));
    // interval (start, inverse width) pairs, interleaved
    L.knots = o;       o = cf_align16(o + (dupire ? size_t(2 * (n_spots + 1)) * real_bytes : 0));
    L.bucket = o;      o = cf_align16(o + (dupire && !lean ? size_t(n_buckets) * (real_bytes == 4 ? sizeof(BucketEntry<float>)
                                                                                            : sizeof(BucketEntry<double>))
                                                           : 0));
    L.iv = o;          o = cf_align16(o + (dupire && !lean ? size_t(n_steps) * lv_cols(n_spots) * real_bytes : 0));
    L.drift = o;       o = cf_align16(o + (dupire ? 0 : size_t(n_steps) * real_bytes));
    L.stdv = o;        o = cf_align16(o + (dupire ? 0 : size_t(n_steps) * real_bytes));
    L.samples = o;     o = cf_align16(o + (dupire ? 0 : size_t(n_samples) * 3 * 8));
    L.pay_first = o;   o = cf_align16(o + (europeans ? size_t(n_samples + 1) * 4 : 0));
    L.pay_strike = o;  o = cf_align16(o + (europeans ? size_t(n_pay) * 8 : 0));
    L.acc = o;         o = cf_align16(o + size_t(waves) * out_stride(n_pay) * 8);
    // FP64: one row per wave for the inverse normal's tail draws, 64 doubles
    // (FULL) or 32 (LEAN, beside the largest grids' tables)
    L.tails = o;       o = cf_align16(o + (real_bytes == 8 ? size_t(waves) * tail_row(lean) * 8 : 0));
    L.ivbar = o;       o = cf_align16(o + (adjoint && dupire ? size_t(ivbar_copies) * ivbar_copy_stride(n_steps, n_spots) * 8
                                                             : 0));
    L.total = o;
    return L;
}

// Host-side launch entry points, defined in cf_kernels.hip.
struct KernelConfig {
    int precision;   // 0 = FP64, 1 = FP32
    int model;
    int product;
    bool adjoint;
    bool lean;       // LEAN shared-memory layout
    bool recompute;  // adjoint by re-simulation instead of per-step checkpoints
    bool mrg;        // normals from mrg32k3a generators in the kernel instead of Sobol
    bool item;       // checkpoint kernel with the itemized-risk modes (LaunchParams::ck_mode)
    bool packed;     // path_kernel_packed (several calls, PackArgs)
};

// Path kernel families (precision r: 0 FP64, 1 FP32; adjoint a: AdjointKind;
// b: 0 Sobol, 1 mrg32k3a, 2 itemized-risk checkpoint kernels (4 with mrg32k3a), 3 Sobol
// kernels of packed launches), one per build of cf_kern_inst.hip: kernel for
// a model, product and layout.
#define CF_PICK_CAT(r, a, b) pick_r##r##_a##a##_b##b
#define CF_PICK_NAME(r, a, b) CF_PICK_CAT(r, a, b)
#define CF_PICK_DECL(r, a)                                                                                            \
    const void* pick_r##r##_a##a##_b0(int model, int prd, bool lean);                                                  \
    const void* pick_r##r##_a##a##_b1(int model, int prd, bool lean);
CF_PICK_DECL(0, 0)
CF_PICK_DECL(0, 1)
CF_PICK_DECL(0, 2)
CF_PICK_DECL(1, 0)
CF_PICK_DECL(1, 1)
CF_PICK_DECL(1, 2)
#undef CF_PICK_DECL
const void* pick_r0_a1_b2(int model, int prd, bool lean);
const void* pick_r1_a1_b2(int model, int prd, bool lean);
const void* pick_r0_a1_b4(int model, int prd, bool lean);
const void* pick_r1_a1_b4(int model, int prd, bool lean);
#define CF_PICK_DECL_PACKED(r)                                                                                       \
    const void* pick_r##r##_a0_b3(int model, int prd, bool lean);                                                  \
    const void* pick_r##r##_a1_b3(int model, int prd, bool lean);                                                  \
    const void* pick_r##r##_a2_b3(int model, int prd, bool lean);
CF_PICK_DECL_PACKED(0)
CF_PICK_DECL_PACKED(1)
#undef CF_PICK_DECL_PACKED

// Returns 0 on success, else a hipError_t value.
int launch_path_kernel(const KernelConfig& cfg, const LaunchParams& p, int grid, int block,
                       size_t smem, void* stream);
// Column sums of [rows][cols0] into out0 and of [rows][cols1] into out1
// (cols1 may be 0), one launch.
int launch_reduce_rows(const double* in0, int cols0, double* out0, const double* in1, int cols1, double* out1,
                       int rows, void* stream);
// A packed launch (cfg.packed: Sobol, not itemized) of A.n calls over grid =
// A.first[A.n] blocks, and the calls' reductions in one launch.
int launch_path_kernel_packed(const KernelConfig& cfg, const PackArgs& A, int block, size_t smem, void* stream);
int launch_reduce_packed(const ReducePack& R, void* stream);
int path_kernel_occupancy(const KernelConfig& cfg, int block, size_t smem, int* blocks_per_cu);

// Copies the direction numbers [32][kSobolDims] into the device's module
// variable (the current device) and returns its address there.
int upload_sobol_dirs(const uint32_t* host, uint32_t** device);

// Sobol words ([npath][dim], nullable) and normals ([npath][dim] of the real
// type, nullable) of paths [first, first + npath), as the path kernels form
// them, in the FULL or LEAN form.
size_t sobol_points_smem(int real_bytes, bool lean, int dim, int block);
// layout: 0 FULL rows, 1 LEAN, 2 FULL quads (SobolLoLayout)
int launch_sobol_points(int real_bytes, int layout, const uint32_t* dirs, int dim, int64_t first, int64_t npath,
                        uint32_t* words, void* normals, void* stream);

// Sets the thread-local message returned by cf_last_error() (cf_engine.cpp).
void set_last_error(const char* msg);

}  // namespace cfk
