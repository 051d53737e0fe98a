// HIP path kernels for gfx950 (MI355X): Sobol -> inverse normal -> model
// step -> product, one path per lane, forward (pricing) and forward +
// hand-written adjoint (risk).
//
// Reference semantics followed: sobol.h:55-151, gaussians.h:47-87,
// interp.h:26-63, mcMdlDupire.h:238-280, mcMdlBS.h:321-350,
// mcPrd.h:113-125, :235-288, :374-399; the adjoint reproduces the derivative
// conventions of the reference tape (AADExpr.h:115-253, :364-598).
//
// Block organisation
//   * a persistent grid; wave w of the grid takes Sobol blocks q = w, w + W, ...
//     of 64 consecutive Sobol indices n = 64q + lane (path p = n - 1);
//   * Gray code splits as gray(64q + r) = gray6(r) ^ (q & 1) << 5 ^ gray(q) << 6,
//     so the point of lane r in dimension d is  LO[d][r] ^ HI_q[d]:  LO is a
//     per-block LDS table built once, HI_q a per-wave LDS row built per block
//     of paths;
//   * local-vol table, interval table, knot search table, inverse-normal
//     table and timeline sample map in LDS (the LEAN layout of long timelines
//     and large grids reads the local-vol, knot search and inverse-normal
//     tables from global memory through L1 instead);
//   * FP64 draws four steps at a time, the inverse normal's tail form
//     evaluated once per wave for the lanes that need it (a per-wave LDS row
//     of 64 doubles, 32 in the LEAN layout);
//   * mrg32k3a instead of Sobol: one generator per lane in registers, jumped
//     to its path's block of numbers and stepped back in the reverse sweeps;
//   * payoffs reduced per wave with DPP cross-lane moves into per-wave LDS
//     accumulators, then per block, then across blocks by a separate kernel
//     in fixed order;
//   * adjoint: path states checkpointed in a per-wave HBM slot (reused by the
//     wave for each block of paths) with the knot bracket of each state in a
//     byte, reverse sweep over the Euler steps,
//     local-vol adjoints accumulated with LDS atomics on a per-block
//     (step x spot) grid.
// This header holds the device helpers and the path kernel template; each
// cf_kern_inst.hip build instantiates one (precision, adjoint, source of
// normals: Sobol or mrg32k3a) family.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "cf_internal.h"

namespace cfk {

// ------------------------------------------------------------------ helpers

// One DPP move of an FP64 value (both halves, the same lane pattern); lanes
// the pattern does not write, or whose source lies outside the row, read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, int(u & 0xffffffffu), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, int(u >> 32), CTRL, ROW_MASK, 0xf, true);
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo));
}

// The wave's sum lands in this lane.
constexpr int kSumLane = kWave - 1;

// FP64 sum over the wave by data-parallel primitives in registers (no LDS):
// prefix sums of 2, 4, 8 and 16 lanes within each row of 16 (row_shr), then
// row 0's sum into row 1 and row 2's into row 3 (row_bcast:15), then rows
// 0-1 into rows 2-3 (row_bcast:31).  The total is valid in lane kSumLane
// only; the order of the additions is fixed, so the sum is reproducible.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_f64<0x111, 0xf>(v);   // row_shr:1
    v += dpp_f64<0x112, 0xf>(v);   // row_shr:2
    v += dpp_f64<0x114, 0xf>(v);   // row_shr:4
    v += dpp_f64<0x118, 0xf>(v);   // row_shr:8
    v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Read-only global data read with wave-uniform indices: a pointer into the
// constant address space, which the compiler reads with scalar loads through
// the scalar cache into scalar registers.
template <typename T>
using CPtr = const __attribute__((address_space(4))) T*;
template <typename T>
__device__ __forceinline__ CPtr<T> cptr(const void* p)
{
    return (CPtr<T>)(p);
}

// a / b in FP64 from the hardware reciprocal refined by two Newton steps and a
// residual correction (within an ulp of the IEEE quotient)
__device__ __forceinline__ double fast_div(double a, double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(fma(-b, r, 1.0), r, r);
    r = fma(fma(-b, r, 1.0), r, r);
    const double q = a * r;
    return fma(fma(-b, q, a), r, q);
}

// Natural logarithm of a positive normal double: x = 2^e m with m in
// [1/sqrt(2), sqrt(2)), ln(m) = 2 atanh(s) with s = (m - 1) / (m + 1),
// |s| <= 0.1716, summed to s^21 (truncation below 1e-17 relative), and
// e ln(2) added in two parts.  About two ulps; used where the argument is
// known to be positive and finite (the inverse-normal tail).
__device__ __forceinline__ double log_pos(double x)
{
    double m = __builtin_amdgcn_frexp_mant(x);   // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);       // x = m 2^e
    const bool lo = m < 0.70710678118654752440;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const double s = fast_div(m - 1.0, m + 1.0);
    const double s2 = s * s;
    double p = 1.0 / 21.0;
    p = fma(p, s2, 1.0 / 19.0);
    p = fma(p, s2, 1.0 / 17.0);
    p = fma(p, s2, 1.0 / 15.0);
    p = fma(p, s2, 1.0 / 13.0);
    p = fma(p, s2, 1.0 / 11.0);
    p = fma(p, s2, 1.0 / 9.0);
    p = fma(p, s2, 1.0 / 7.0);
    p = fma(p, s2, 1.0 / 5.0);
    p = fma(p, s2, 1.0 / 3.0);
    const double ed = double(e);
    const double ln2_hi = 6.93147180369123816490e-01;   // 0x3FE62E42FEE00000, exact in ed * ln2_hi
    const double ln2_lo = 1.90821492927058770002e-10;
    // ed * ln2_hi is exact (32-bit constant times an 11-bit integer)
    return fma(ed, ln2_hi, fma(2.0 * s * s2, p, fma(ed, ln2_lo, 2.0 * s)));
}

// Beasley-Springer-Moro inverse normal (gaussians.h:47-87), FP64, same
// branches and constants as the reference: the central rational form of
// x = min(p, 1 - p) - 1/2 for |x| < 0.42, else the tail polynomial in
// log(-log(min(p, 1 - p))); both halves are also evaluated apart
// (inv_normal4_f64) and give the same values.
__device__ __forceinline__ double moro_central(double x)
{
    const double r = x * x;
    return x * (((-25.44106049637 * r + 41.39119773534) * r + -18.61500062529) * r + 2.50662823884) /
           ((((3.13082909833 * r + -21.06224101826) * r + 23.08336743743) * r + -8.47351093090) * r + 1.0);
}
__device__ __forceinline__ double moro_tail(double up)
{
    double r = log_pos(-log_pos(up));
    r = 0.3374754822726147 +
        r * (0.9761690190917186 +
        r * (0.1607979714918209 +
        r * (0.0276438810333863 +
        r * (0.0038405729373609 +
        r * (0.0003951896511919 +
        r * (0.0000321767881768 +
        r * (0.0000002888167364 + r * 0.0000003960315187)))))));
    return r;
}
__device__ __forceinline__ double inv_normal_p(double p)
{
    const bool sup = p > 0.5;
    const double up = sup ? 1.0 - p : p;
    const double x = up - 0.5;
    if (fabs(x) < 0.42) {
        const double z = moro_central(x);
        return sup ? -z : z;
    }
    const double r = moro_tail(up);
    return sup ? r : -r;
}

// of the Sobol uniform u = ONEOVER2POW32 * xi (sobol.h:26, :110-116)
__device__ __forceinline__ double inv_normal(uint32_t xi, double)
{
    return inv_normal_p(2.3283064365387E-10 * double(xi));
}

// FP32 form: v = min(u, 1 - u) is formed exactly in the integer domain as
// xs = 2^32 v in [1, 2^31]; the IEEE bits of float(xs) give the table row
// (exponent and top three mantissa bits: one eighth of a binade of v) and
// t in [-1, 1) (the low 20 mantissa bits).  Each row is a cubic in t fitted
// to the same Moro function in FP64 (inv_normal_f32.inc, written by
// gen_inv_normal_f32.py).  No logarithm, no branch: one 16-byte table read
// and three FMAs.  Differences to the FP64 Moro values are below 5e-7 (the FP32
// rounding of the result in the far tail).
// `tabm` is the table's address less 1016 rows (kInvNormalRowBias): row r of
// the table is at tabm + 4 (r + 1016), so the row address is formed from the
// float's bits by one mask and one add.
constexpr int kInvNormalRowBias = 127 << 3;
__device__ __forceinline__ float inv_normal(uint32_t xi, const float* tabm)
{
    const uint32_t sgn = xi & 0x80000000u;            // u >= 1/2 for the reference scaling
    // 2^32 - xi, exact.  xi = 0 would give row -1016: the kernel never passes
    // it (only Sobol index 0 has a zero word, and its lane reads 2^31 instead)
    // (as signed integers xi < 0 exactly when u >= 1/2, and xi = 2^31 maps to itself)
    const uint32_t xs = uint32_t(max(int(xi), int(0u - xi)));
    const uint32_t bits = __float_as_uint(float(xs));
    // row (bits >> 20) - 1016 in [0, kInvNormalRows - 1], 16 bytes a row
    const float4 c = *reinterpret_cast<const float4*>(tabm + ((bits >> 18) & 0x3ffcu));
    const float t = fmaf(float(bits & 0xfffffu), 1.0f / 524288.0f, -1.0f);
    const float p = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
    return __uint_as_float(__float_as_uint(p) ^ sgn);   // -p for u >= 1/2
}

// FP64: the reference's function; the table pointer is unused
__device__ __forceinline__ double inv_normal(uint32_t xi, const double*) { return inv_normal(xi, 0.0); }

// Four FP64 normals of this lane's Sobol words, with Moro's tail evaluated
// once per wave.  About 16% of draws need the tail (two logarithms and a
// polynomial), so a wave of 64 lanes almost always has one that does and,
// draw by draw, every lane would run it.  Here every lane runs the central
// form of its four draws; the tail draws of the wave, counted by ballots, go
// to `row` (this wave's ROW doubles of LDS) by their rank, each lane evaluates
// at most one, and the owners read their values back (in rounds of ROW when
// more draws need the tail).  The values equal inv_normal_p's, draw by draw.
template <int ROW>
__device__ __forceinline__ void inv_normal4_p(const double p[4], double z[4], double* row, int lane)
{
    double up[4];
    bool sup[4], tail[4];
    uint64_t m[4];
    int base[4];
    int total = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        sup[k] = p[k] > 0.5;
        up[k] = sup[k] ? 1.0 - p[k] : p[k];
        const double x = up[k] - 0.5;
        tail[k] = !(fabs(x) < 0.42);
        const double c = moro_central(x);
        z[k] = sup[k] ? -c : c;
        m[k] = __builtin_amdgcn_ballot_w64(tail[k]);
        base[k] = total;
        total += __builtin_popcountll(m[k]);
    }
    static_assert(ROW >= 1 && ROW <= kWave, "a row of at most one double per lane");
    for (int r0 = 0; r0 < total; r0 += ROW) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int pos = base[k] +
                            int(__builtin_amdgcn_mbcnt_hi(uint32_t(m[k] >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m[k]), 0u))) - r0;
            if (tail[k] && pos >= 0 && pos < ROW) row[pos] = up[k];
        }
        wave_sync_lds();
        if (lane < ROW && lane < total - r0) row[lane] = moro_tail(row[lane]);
        wave_sync_lds();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int pos = base[k] +
                            int(__builtin_amdgcn_mbcnt_hi(uint32_t(m[k] >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m[k]), 0u))) - r0;
            if (tail[k] && pos >= 0 && pos < ROW) {
                const double t = row[pos];
                z[k] = sup[k] ? t : -t;
            }
        }
        wave_sync_lds();
    }
}

// Of Sobol words (sobol.h:26, :110-116)
template <int ROW>
__device__ __forceinline__ void inv_normal4_f64(const uint32_t w[4], double z[4], double* row, int lane)
{
    double p[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) p[k] = 2.3283064365387E-10 * double(w[k]);
    inv_normal4_p<ROW>(p, z, row, lane);
}

// FP32 inverse-normal table: [rows][4 monomial coefficients]
static __constant__ float kInvNormalF32[kInvNormalRows * 4] = {
#include "inv_normal_f32.inc"
};

// ------------------------------------------------------------------ mrg32k3a

// L'Ecuyer's combined multiple recursive generator as the reference runs it
// (mrg32k3a.h:41-86), one generator per lane in registers:
//   x_n = a12 x_{n-2} - a13 x_{n-3}  mod m1,   y_n = a21 y_{n-1} - a23 y_{n-3}  mod m2,
// uniform (x_n - y_n) / (m1 + 1) folded into (0, 1).  Both moduli are
// 2^32 - c with small c, so a product r < 2^64 reduces by folds
// r -> (r >> 32) c + (r mod 2^32) (one 32 x 32 + 64 multiply-add each) and a
// last conditional subtraction.  Both recursions are invertible (a13, a23
// are units modulo the primes m1, m2), so the generator also steps back.
namespace mrg {
constexpr uint32_t kM1 = 4294967087u, kM2 = 4294944443u;
constexpr uint32_t kC1 = 209u, kC2 = 22853u;            // 2^32 - m
constexpr uint32_t kA12 = 1403580u, kA13 = 810728u;
constexpr uint32_t kA21 = 527612u, kA23 = 1370589u;
constexpr uint64_t pow_mod(uint64_t a, uint64_t e, uint64_t m)
{
    uint64_t r = 1;
    for (a %= m; e; e >>= 1, a = a * a % m)
        if (e & 1) r = r * a % m;
    return r;
}
constexpr uint32_t kInv13 = uint32_t(pow_mod(kA13, kM1 - 2, kM1));   // a13^-1 mod m1 (Fermat)
constexpr uint32_t kInv23 = uint32_t(pow_mod(kA23, kM2 - 2, kM2));   // a23^-1 mod m2
static_assert(uint64_t(kInv13) * kA13 % kM1 == 1 && uint64_t(kInv23) * kA23 % kM2 == 1, "inverses");

// r mod (2^32 - C) for any r < 2^64 (C < 2^15): the three folds leave
// r < 2^32 + C < 2 (2^32 - C)
template <uint32_t C>
__device__ __forceinline__ uint32_t reduce(uint64_t r)
{
    constexpr uint32_t M = uint32_t(0u - C);
#pragma unroll
    for (int k = 0; k < 3; ++k) r = uint64_t(uint32_t(r >> 32)) * C + uint32_t(r);
    return uint32_t(r >= M ? r - M : r);
}

// v = A v mod M, A a 3 x 3 matrix of residues (row-major), v residues
template <uint32_t C>
__device__ __forceinline__ void mat_vec(const uint64_t* __restrict__ A, uint32_t v[3])
{
    uint32_t r[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        uint32_t acc = 0;
#pragma unroll
        for (int l = 0; l < 3; ++l) acc = reduce<C>(uint64_t(uint32_t(A[3 * j + l])) * v[l] + acc);
        r[j] = acc;
    }
    v[0] = r[0];
    v[1] = r[1];
    v[2] = r[2];
}

// One lane's generator.  x[0], y[0] are the pos-th numbers of the lane's
// block (x[1], x[2] the two before); number pos feeds draw pos - 1.
struct Lane {
    uint32_t x[3], y[3];
    uint32_t x0[3], y0[3];   // the block's start state (pos 0)
    int pos;
    bool neg;                // odd path: the antithetic (negated) normals of its pair

    // Path p of a call (absolute index) draws from block p >> 1 of `dim`
    // numbers, negated for odd p (mrg32k3a.h:147-180); its start state is the
    // seed state advanced by (p >> 1) dim numbers through the jump matrices
    // A^(dim 2^j) of the set bits j (mrg32k3a.h:187-381; jump: 2 x 9 per j,
    // x then y).  Seeds are residues.
    __device__ __forceinline__ void start(const uint64_t* __restrict__ jump, uint32_t s1, uint32_t s2, int64_t p)
    {
        uint32_t xs[3] = {s1, s1, s1}, ys[3] = {s2, s2, s2};
        const uint64_t k = uint64_t(p) >> 1;
        for (int j = 0; j < 32 && (k >> j); ++j) {   // p < 2^32 (checked on the host)
            if ((k >> j) & 1) {
                mat_vec<kC1>(jump + size_t(18 * j), xs);
                mat_vec<kC2>(jump + size_t(18 * j + 9), ys);
            }
        }
#pragma unroll
        for (int l = 0; l < 3; ++l) {
            x[l] = x0[l] = xs[l];
            y[l] = y0[l] = ys[l];
        }
        pos = 0;
        neg = (p & 1) != 0;
    }
    __device__ __forceinline__ void forward()
    {
        // a12 x1 + a13 (m1 - x2) = x_n mod m1: both terms below 2^53
        const uint32_t xn = reduce<kC1>(uint64_t(kA12) * x[1] + uint64_t(kA13) * (kM1 - x[2]));
        const uint32_t yn = reduce<kC2>(uint64_t(kA21) * y[0] + uint64_t(kA23) * (kM2 - y[2]));
        x[2] = x[1]; x[1] = x[0]; x[0] = xn;
        y[2] = y[1]; y[1] = y[0]; y[0] = yn;
        ++pos;
    }
    __device__ __forceinline__ void back()
    {
        // x_{n-3} = (a12 x_{n-2} - x_n) / a13,  y_{n-3} = (a21 y_{n-1} - y_n) / a23
        const uint32_t dx = reduce<kC1>(uint64_t(kA12) * x[2] + (kM1 - x[0]));
        const uint32_t dy = reduce<kC2>(uint64_t(kA21) * y[1] + (kM2 - y[0]));
        const uint32_t xp = reduce<kC1>(uint64_t(dx) * kInv13);
        const uint32_t yp = reduce<kC2>(uint64_t(dy) * kInv23);
        x[0] = x[1]; x[1] = x[2]; x[2] = xp;
        y[0] = y[1]; y[1] = y[2]; y[2] = yp;
        --pos;
    }
    // Draw i is number i + 1 of the block, reached from the current position
    // by stepping forward or back, or forward from the start when that is
    // shorter.
    __device__ __forceinline__ void seek(int i)
    {
        const int t = i + 1;
        if (t < pos - t) {
#pragma unroll
            for (int l = 0; l < 3; ++l) { x[l] = x0[l]; y[l] = y0[l]; }
            pos = 0;
        }
        while (pos < t) forward();
        while (pos > t) back();
    }
    // FP64: the reference's uniform (mrg32k3a.h:84-86) and inverse normal
    // (gaussians.h:47-87)
    __device__ __forceinline__ double uniform(int i)
    {
        seek(i);
        const double diff = double(x[0]) - double(y[0]);
        return x[0] > y[0] ? diff / 4294967088.0 : (diff + 4294967087.0) / 4294967088.0;
    }
    __device__ __forceinline__ double normal(int i)
    {
        const double z = inv_normal_p(uniform(i));
        return neg ? -z : z;
    }
    // FP32: the uniform d / (m1 + 1), d in [1, m1], as a word of 32 bits,
    // w = d + floor(208 d / 2^32) in [1, 2^32 - 2] (within one unit of
    // 2^32 u), for the FP32 inverse normal of Sobol words
    __device__ __forceinline__ float normal_f32(int i, const float* tabm)
    {
        seek(i);
        const uint32_t d = x[0] > y[0] ? x[0] - y[0] : x[0] - y[0] + kM1;
        const uint32_t w = d + uint32_t((uint64_t(d) * 208u) >> 32);
        const float z = inv_normal(w, tabm);
        return neg ? -z : z;
    }
};
}   // namespace mrg

// Spot from the path state: FP64 state is log(S) (reference), FP32 state is
// log(S / S0), which keeps more significant bits in single precision.
__device__ __forceinline__ double state_spot(double x, double) { return exp(x); }
__device__ __forceinline__ float state_spot(float x, double s0) { return float(s0) * __expf(x); }

__device__ __forceinline__ double rexp(double x) { return exp(x); }
__device__ __forceinline__ float rexp(float x) { return __expf(x); }

// Local vol at state x on one timeline row (interp.h:26-63): upper_bound on
// the knots, linear in between, flat outside.  Returns the two grid nodes
// touched and their weights for the adjoint.
template <typename R>
struct Bracket {
    R vol;
    R slope;      // d vol / d x
    int j;        // padded column of the left node (the row's columns j, j + 1)
    R wa, wb;     // adjoint weights of nodes j and j + 1
};

// Interval c = count_le(ks, x), the number of knots <= x, found by the
// caller (the reverse sweep reloads it from the forward's).  The interval
// tables make the flat extrapolation an ordinary interval: for c = 0 and
// c = m the padded row repeats the end value and the inverse width is 0, so
// t = 0, vol = the end value, slope = 0; inside, t = (x - k[c-1]) / (k[c] -
// k[c-1]) and vol = y[c-1] + (y[c] - y[c-1]) t.  No clamps, no selects.
// kxi: [m + 1] (interval start, inverse width) pairs, one aligned read;
// row: [m + 2].
template <typename R>
__device__ __forceinline__ Bracket<R> local_vol_at(const R* kxi, const R* row, R x, int c)
{
    const R y1 = row[c], y2 = row[c + 1];
    const R ih = kxi[2 * c + 1];
    const R t = (x - kxi[2 * c]) * ih;
    const R dy = y2 - y1;
    Bracket<R> b;
    b.j = c;
    b.vol = y1 + dy * t;
    b.slope = dy * ih;
    b.wa = R(1) - t;
    b.wb = t;
    return b;
}

template <bool LEAN, typename R>
__device__ __forceinline__ Bracket<R> local_vol(const KnotSearch<R>& ks, const R* kxi, const R* row, R x)
{
    return local_vol_at(kxi, row, x, count_le<LEAN>(ks, x));
}

__device__ __forceinline__ double recip_mul(double a, double b) { return fast_div(a, b); }
__device__ __forceinline__ float recip_mul(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Log-Euler step of the Dupire state (mcMdlDupire.h:271); one definition for
// the forward sweep and the recompute variant's re-simulation, so both give
// the same states.
template <typename R>
__device__ __forceinline__ R euler_step(R x, R vol, R g)
{
    return x + vol * (R(-0.5) * vol + g);
}

// ------------------------------------------------------------------ Sobol

// Sobol points of a block of 64 indices n = 64 q + r (sobol.h:55-151):
// gray(64q + r) = gray6(r) ^ (q & 1) << 5 ^ gray(q) << 6, so the word of lane r
// in dimension d is LO[d][r] ^ HI_q[d].  The path kernels and the points
// kernel share these functions.

// LO staged in LDS by a block, in one of three layouts: rows (the 64 lanes'
// low words of each dimension, [d][lane]), quads (the same words interleaved
// by groups of four dimensions, [d / 4][lane][d % 4], for kernels that read a
// lane's four consecutive dimensions in one 16-byte word; a lane reading one
// word at a time would stride 16 bytes and conflict on the LDS banks), or
// LEAN (only direction numbers 0..5: the low word is formed in registers from
// the lane's masks).
enum SobolLoLayout { kLoRows = 0, kLoLean = 1, kLoQuads = 2 };

template <int LO>
__device__ __forceinline__ void sobol_stage_lo(uint32_t* lo_tab, const uint32_t* dirs, int n, int tid, int nthr)
{
    if constexpr (LO == kLoLean) {
        for (int e = tid; e < n * 8; e += nthr) {
            const int d = e >> 3, k = e & 7;
            lo_tab[e] = k < 6 ? dirs[k * kSobolDims + d] : 0u;
        }
    } else {
        // word e: rows dimension e >> 6, lane e & 63; quads dimension
        // 4 (e >> 8) + (e & 3), lane (e >> 2) & 63
        for (int e = tid; e < sobol_dims4(n) * kWave; e += nthr) {
            const int d = LO == kLoQuads ? ((e >> 8) << 2) + (e & 3) : e >> 6;
            const int r = LO == kLoQuads ? (e >> 2) & 63 : e & 63;
            if (d >= n) continue;
            const uint32_t g = uint32_t(r ^ (r >> 1));
            uint32_t x = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if ((g >> k) & 1u) x ^= dirs[k * kSobolDims + d];
            lo_tab[e] = x;
        }
    }
}

// LEAN: masks of the bits of gray6(lane), selecting direction numbers 0..5
struct SobolLaneMasks {
    uint32_t m[6];
};
__device__ __forceinline__ SobolLaneMasks sobol_lane_masks(int lane)
{
    SobolLaneMasks lm;
    const uint32_t g6 = uint32_t(lane ^ (lane >> 1));
#pragma unroll
    for (int k = 0; k < 6; ++k) lm.m[k] = ((g6 >> k) & 1u) ? 0xffffffffu : 0u;
    return lm;
}

// low word of dimension i for this lane
template <int LO>
__device__ __forceinline__ uint32_t sobol_low_word(const uint32_t* lo_tab, const SobolLaneMasks& lm, int lane, int i)
{
    if constexpr (LO == kLoLean) {
        const uint32_t* v = lo_tab + i * 8;
        return (lm.m[0] & v[0]) ^ (lm.m[1] & v[1]) ^ (lm.m[2] & v[2]) ^ (lm.m[3] & v[3]) ^ (lm.m[4] & v[4]) ^
               (lm.m[5] & v[5]);
    } else if constexpr (LO == kLoQuads) {
        return lo_tab[sobol_lo_el(i, lane)];
    } else {
        return lo_tab[i * kWave + lane];
    }
}

// Quads layout: this lane's words of dimensions i .. i + 3 (i a multiple of
// four, below the padded dimension count), one 16-byte read of LO and one of
// the lane's high row (16-byte aligned rows)
__device__ __forceinline__ void sobol_words4(const uint32_t* lo_tab, const uint32_t* hl, int lane, int i, uint32_t* w)
{
    const uint4 lw = *reinterpret_cast<const uint4*>(lo_tab + sobol_lo_el(i, lane));
    const uint4 hw = *reinterpret_cast<const uint4*>(hl + i);
    w[0] = lw.x ^ hw.x;
    w[1] = lw.y ^ hw.y;
    w[2] = lw.z ^ hw.z;
    w[3] = lw.w ^ hw.w;
}

// HI_qb of the n dimensions into a wave's LDS row, formed by the wave's lanes
__device__ __forceinline__ void sobol_fill_hi(uint32_t* row, const uint32_t* dirs, int n, int64_t qb, int lane)
{
    const uint32_t gq = uint32_t(qb ^ (qb >> 1));
    for (int d = lane; d < n; d += kWave) {
        uint32_t x = (qb & 1) ? dirs[5 * kSobolDims + d] : 0u;
        uint32_t b = gq;
        while (b) {
            const int j = __builtin_ctz(b);
            x ^= dirs[(j + 6) * kSobolDims + d];
            b &= b - 1u;
        }
        row[d] = x;
    }
}

// One lane's share of a wave's block of 64 Sobol indices: its high-word row,
// Sobol index, whether it holds a path of [nlo, nhi] and which.
struct SobolLane {
    const uint32_t* hl;
    int64_t sidx;
    bool active;
    int64_t pidx;
};

// Fills the wave's high rows for block q (and, for the first block when the
// last block is merged into its idle low lanes, the tail row of block q1) and
// selects this lane's.  Sobol index 0 (lane 0 of block 0, never a path) has
// every word 0, outside the FP32 inverse normal's domain; without a merge the
// tail row is free, so that lane reads words of 2^31 from it instead (its low
// words are 0) and its draws stay inside the table.  fill: the rows are
// formed (false when the words are not used).
__device__ __forceinline__ SobolLane sobol_block(uint32_t* hi, uint32_t* hi_tail, const uint32_t* dirs, int n,
                                                 int64_t q, int64_t q0, int64_t q1, int64_t nlo, int64_t nhi,
                                                 bool merge, int lane, bool fill)
{
    const bool tail_wave = merge && q == q0;
    if (fill) {
        sobol_fill_hi(hi, dirs, n, q, lane);
        if (tail_wave) sobol_fill_hi(hi_tail, dirs, n, q1, lane);
        if (!merge && q == 0)
            for (int d = lane; d < n; d += kWave) hi_tail[d] = 0x80000000u;
        wave_sync_lds();
    }
    // lanes 0..(nhi mod 64) of the first block take the last block's paths
    const bool tail = tail_wave && lane <= int(nhi & (kWave - 1));
    SobolLane sl;
    sl.sidx = (tail ? q1 : q) * kWave + lane;
    sl.hl = (tail || sl.sidx == 0) ? hi_tail : hi;
    sl.active = sl.sidx >= nlo && sl.sidx <= nhi;
    sl.pidx = sl.sidx - nlo;
    return sl;
}

// ------------------------------------------------------------------ products

// Per-lane product state carried along the path.
template <typename R>
struct ProductState {
    R alive;        // barrier: product of smoothing factors (before a breach)
    R pnz;          // barrier: after an exactly-zero smoothing factor, the product of the non-zero ones
    int nzero;      // barrier: number of exactly-zero factors
    bool dead;      // barrier: breached (alive forced to 0, no derivative)
    R euro;         // barrier: vanilla payoff at maturity
    R fin_f;        // barrier: final forward
    R fin_num;      // barrier: final numeraire
    R aggp;         // running aggregate of payoffs emitted at samples
};

// ------------------------------------------------------------------ kernel

// The kernel of one call, run by block bid of the call's nblk blocks (the
// whole grid, or the call's share of a packed launch).
template <typename R, int MODEL, int PRD, int ADJ, bool LEAN, bool MRG, bool ITEM>
__device__ __forceinline__ void path_body(const LaunchParams& P, const int bid, const int nblk)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool DUP = MODEL == MODEL_DUPIRE;
    constexpr bool AD = ADJ != ADJ_NONE;
    constexpr bool RECOMP = ADJ == ADJ_RECOMPUTE;
    constexpr int kSeg = sizeof(R) == 4 ? kSegF32 : kSegF64;   // recompute segment
    constexpr bool EUS = PRD == PRD_EUROPEANS;
    const int waves = blockDim.x >> 6;
    const int n = P.n_steps, m = P.n_spots, ns = P.n_samples, npay = P.n_pay;
    const int mp = lv_cols(m);   // padded local-vol row
    const int stride = out_stride(npay);
    constexpr bool BOOK = PRD == PRD_BOOK;
    const SmemLayout L = smem_layout(n, m, ns, npay, waves, int(sizeof(R)), DUP, EUS, AD, LEAN, P.n_buckets,
                                     P.ivbar_copies);

    uint32_t* lo_tab = reinterpret_cast<uint32_t*>(smem + L.lo);
    uint32_t* hi_all = reinterpret_cast<uint32_t*>(smem + L.hi);
    int* step_sample = reinterpret_cast<int*>(smem + L.step_sample);
    R* kxi = reinterpret_cast<R*>(smem + L.knots);   // interval (start, inverse width) pairs
    BucketEntry<R>* bucket = reinterpret_cast<BucketEntry<R>*>(smem + L.bucket);
    R* iv_lds = reinterpret_cast<R*>(smem + L.iv);
    R* drift = reinterpret_cast<R*>(smem + L.drift);
    R* stdv = reinterpret_cast<R*>(smem + L.stdv);
    double* s_ff = reinterpret_cast<double*>(smem + L.samples);
    double* s_disc = s_ff + ns;
    double* s_num = s_disc + ns;
    int* pay_first = reinterpret_cast<int*>(smem + L.pay_first);
    double* pay_strike = reinterpret_cast<double*>(smem + L.pay_strike);
    double* acc = reinterpret_cast<double*>(smem + L.acc);
    constexpr int TR = tail_row(LEAN);
    double* tails = reinterpret_cast<double*>(smem + L.tails);   // FP64: [waves][TR]
    double* ivbar = reinterpret_cast<double*>(smem + L.ivbar);
    R* invn = reinterpret_cast<R*>(smem + L.invn);   // FP32 inverse-normal table

    const int tid = threadIdx.x, nthr = blockDim.x;
    // the inverse-normal table (FP32) and the knot search table are read from
    // global memory through the L1 cache, not LDS, by the LEAN kernels and by
    // the FP32 recompute kernels, whose re-simulation otherwise loads the LDS
    // most (measured: C3 FP32 recompute 0.719 -> 0.643 ms; the pricing kernels
    // are faster with both in LDS).  The FP32 Dupire checkpoint kernels other
    // than the book, bound by the LDS pipe, read the inverse-normal table
    // through L1 (C3 0.587 -> 0.576 ms; with the knot search table as well
    // they lose: the vector memory path saturates instead)
    constexpr bool GTAB = LEAN || (RECOMP && sizeof(R) == 4);
    constexpr bool GINV = GTAB || (sizeof(R) == 4 && ADJ == ADJ_CHECKPOINT && DUP && !BOOK);

    // Sobol LO layout: the FP32 Dupire adjoint kernels other than the book
    // read four dimensions at once (quads); the others one at a time (rows)
    constexpr bool QUADS = sizeof(R) == 4 && AD && DUP && !BOOK && !LEAN && !MRG;
    constexpr int LOL = LEAN ? kLoLean : (QUADS ? kLoQuads : kLoRows);
    // ---- stage the block's tables in LDS
    sobol_stage_lo<LOL>(lo_tab, P.dirs, n, tid, nthr);
    for (int e = tid; e <= n; e += nthr) step_sample[e] = P.step_sample[e];
    if constexpr (sizeof(R) == 4 && !GINV)
        for (int e = tid; e < kInvNormalRows * 4; e += nthr) invn[e] = kInvNormalF32[e];
    if constexpr (DUP) {
        const R* gk = static_cast<const R*>(P.knots);
        const R* gi = static_cast<const R*>(P.inv_dx);
        const R* gv = static_cast<const R*>(P.iv);
        for (int e = tid; e <= m; e += nthr) { kxi[2 * e] = gk[e]; kxi[2 * e + 1] = gi[e]; }
        if constexpr (!GTAB)
            for (int e = tid; e < P.n_buckets; e += nthr) bucket[e] = static_cast<const BucketEntry<R>*>(P.bucket)[e];
        if constexpr (!LEAN)
            for (int e = tid; e < n * mp; e += nthr) iv_lds[e] = gv[e];
    } else {
        const R* gd = static_cast<const R*>(P.drift);
        const R* gs = static_cast<const R*>(P.stdv);
        for (int e = tid; e < n; e += nthr) { drift[e] = gd[e]; stdv[e] = gs[e]; }
        for (int e = tid; e < ns; e += nthr) { s_ff[e] = P.ff[e]; s_disc[e] = P.disc[e]; s_num[e] = P.num[e]; }
    }
    if constexpr (EUS) {
        for (int e = tid; e <= ns; e += nthr) pay_first[e] = P.pay_first[e];
        for (int e = tid; e < npay; e += nthr) pay_strike[e] = P.pay_strike[e];
    }
    for (int e = tid; e < waves * stride; e += nthr) acc[e] = 0.0;
    if constexpr (AD && DUP)
        for (int e = tid; e < P.ivbar_copies * ivbar_copy_stride(n, m); e += nthr) ivbar[e] = 0.0;
    __syncthreads();

    // the wave index is uniform: block-level offsets derived from it stay in
    // scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    // local-vol table: LDS (FULL) or global memory (LEAN)
    const R* iv;
    if constexpr (LEAN) iv = static_cast<const R*>(P.iv); else iv = iv_lds;
    const SobolLaneMasks lmask = sobol_lane_masks(lane);
    auto low_word = [&](int i) -> uint32_t { return sobol_low_word<LOL>(lo_tab, lmask, lane, i); };
    uint32_t* hi = hi_all + wave * sobol_hi_stride(n);
    double* accw = acc + wave * stride;
    const int64_t nlo = P.first_path + 1, nhi = P.first_path + P.n_path;   // Sobol indices
    const int64_t q0 = nlo >> 6, q1 = nhi >> 6;
    // the last block's paths ride in the first block's idle low lanes when
    // they fit there (e.g. 2^k paths from path 0: Sobol indices 1..2^k), so no
    // wave makes an extra pass for them; those lanes read their high words
    // from the block's tail row
    const bool merge = P.merge_tail != 0;
    const int64_t qe = merge ? q1 - 1 : q1;
    uint32_t* hi_tail = reinterpret_cast<uint32_t*>(smem + L.hi_tail);
    const int slot = bid * waves + wave;
    // checkpoint slot of the wave: [n][64] states (Dupire) or normals
    // (Black-Scholes); recompute variant: [ceil(n / kSeg)][64] Dupire states
    const int ck_rows = ckpt_rows(n, DUP, RECOMP, int(sizeof(R)));
    // bracket bytes after the states: the knot count of x_i at each step
    constexpr bool CKB = DUP && ADJ == ADJ_CHECKPOINT;
    // (the Dupire checkpoint slot starts with kCkptGuardRows guard rows);
    // a wave's slot, or with ck_mode 1 and 2 the slot of each block of 64
    // Sobol indices (set per block below)
    // (only the ITEM checkpoint kernels read ck_mode)
    const int ck_mode = ITEM && ADJ == ADJ_CHECKPOINT ? P.ck_mode : 0;
    R* ck = AD ? static_cast<R*>(P.ckpt) + (size_t(slot) * ck_rows + (CKB ? kCkptGuardRows : 0)) * kWave : nullptr;
    // (the FULL layout is only chosen for grids of at most kMaxStoredKnots knots)
    const bool ck_lo = CKB && (!LEAN || m <= kMaxStoredKnots);
    // FP32: states and bracket bytes interleaved by groups of four steps (a
    // lane's four in one 16-byte / 4-byte word); FP64: one row per step
    constexpr bool IL = CKB && sizeof(R) == 4;
    const int nck = ckpt_steps(n, int(sizeof(R)));
    uint8_t* ckl = CKB ? reinterpret_cast<uint8_t*>(ck + size_t(nck) * kWave) : nullptr;
    // element of step i of this lane in the states (and bracket bytes)
    auto ck_el = [&](int i) -> int { return IL ? ((i >> 2) << 8) + (lane << 2) + (i & 3) : i * kWave + lane; };
    constexpr int kStateRows = BOOK ? 8 + 2 * kMaxBookBarriers + 2 * kMaxBookGroups : 12;
    static_assert(kStateRows == (BOOK ? 8 + 2 * kMaxBookBarriers + 2 * kMaxBookGroups : 8 + 2 + 2),
                  "state rows as state_rows() counts them");
    KnotSearch<R> ks;
    ks.knots = kxi + 2;     // the knots themselves: interval starts from interval 1
    ks.stride = 2;
    if constexpr (GTAB) ks.table = static_cast<const BucketEntry<R>*>(P.bucket); else ks.table = bucket;
    ks.m = m;
    ks.fix = P.bucket_fix;
    ks.invh = R(P.bucket_invh);
    ks.c0 = R(P.bucket_c0);
    ks.top = R(P.n_buckets - 1);
    const double s0 = P.spot;
    const int last = ns - 1;

    // this lane's local-vol adjoint grid: odd lanes use the second copy, which
    // halves the lanes contending for one address in an atomic
    double* ivbl = ivbar + ((P.ivbar_copies > 1 && (lane & 1)) ? ivbar_copy_stride(n, m) : 0);
    // per-lane parameter adjoints, summed over the lane's paths
    double padj[kMaxParamAdj] = {0.0, 0.0, 0.0, 0.0};

    // Book tables and aggregation weights: wave-uniform values, read from
    // global memory through the scalar cache (constant address space) into
    // scalar registers, not staged in LDS
    const unsigned char* bkb = static_cast<const unsigned char*>(P.book);
    CPtr<double> bk_van_strike = cptr<double>(bkb + P.bk.van_strike);
    CPtr<int> bk_van_col = cptr<int>(bkb + P.bk.van_col);
    // per sample: barriers observing (bits 0-15) and maturing (16-31),
    // schedules fixing (bits 0-3) and ending (4-7), and the sample's vanillas
    // [van_lo, van_hi)
    struct alignas(16) SampleRow {
        uint32_t bars, grps;
        int van_lo, van_hi;
    };
    CPtr<SampleRow> bk_rows_c = cptr<SampleRow>(bkb + P.bk.masks);
    auto bk_rows = [&](int s) -> SampleRow {
        SampleRow r;
        r.bars = bk_rows_c[s].bars;
        r.grps = bk_rows_c[s].grps;
        r.van_lo = bk_rows_c[s].van_lo;
        r.van_hi = bk_rows_c[s].van_hi;
        return r;
    };
    CPtr<int> bk_grp_first = cptr<int>(bkb + P.bk.grp_first);
    CPtr<double> bk_as_strike = cptr<double>(bkb + P.bk.as_strike);
    CPtr<int> bk_as_col = cptr<int>(bkb + P.bk.as_col);
    CPtr<int> bk_as_put = cptr<int>(bkb + P.bk.as_put);
    CPtr<int> bk_bar_col = cptr<int>(bkb + P.bk.bar_col);
    CPtr<int> bk_bar_put = cptr<int>(bkb + P.bk.bar_put);
    // fixing weights in the kernel's precision
    CPtr<R> bk_ginv = cptr<R>(bkb + (sizeof(R) == 4 ? P.bk.grp_inv_f : P.bk.grp_inv));
    // barrier b's (strike, B + w, B - w, 1 / 2w) in the kernel's precision
    struct alignas(16) BarPar {
        R k, up, lo, i2s;
    };
    CPtr<BarPar> bk_par_c = cptr<BarPar>(bkb + (sizeof(R) == 4 ? P.bk.bar_par_f : P.bk.bar_par));
    auto bar_par = [&](int b) -> BarPar {
        BarPar r;
        r.k = bk_par_c[b].k;
        r.up = bk_par_c[b].up;
        r.lo = bk_par_c[b].lo;
        r.i2s = bk_par_c[b].i2s;
        return r;
    };
    CPtr<double> wts = cptr<double>(P.weights);
    const R bk_lo = R(P.book_lo);   // the lowest B - w (rounding keeps it the lowest)
    // Barrier smoothing factors (B + w - F) / 2w: FP64 tracks exactly-zero
    // factors (F = B + w) as the tape does.  FP32 keeps B + w - F at least
    // tiny_d instead, so a factor is never 0 and the product of the others,
    // the tape's derivative at a zero factor, stays the product over the
    // factor; the price moves by at most 1e-30 of the payoff there.
    constexpr bool ZT = sizeof(R) == 8;
    const R tiny_d = ZT ? R(0) : max(R(P.two_smooth) * R(1e-30), R(1e-37));
    constexpr int NBAR = BOOK ? kMaxBookBarriers : 1, NGRP = BOOK ? kMaxBookGroups : 1;
    // Book barrier state in one real per barrier:
    //   v in [0, 1]   no exactly-zero smoothing factor yet, v = product of the
    //                 factors (the survival factor)
    //   v in [-1, 0)  one zero factor, -v = product of the others (FP64)
    //   v = -2        two or more zero factors (FP64)
    //   v = -4        breached (survival 0, no derivative)
    // FP32 keeps every factor at least 1e-30 (as the single barrier keeps
    // B + w - F above a floor), so its states are only [0, 1] or -4.
    constexpr float kTwoZeros = -2.0f, kBreached = -4.0f;

    for (int64_t q = q0 + int64_t(bid) * waves + wave; q <= qe; q += int64_t(nblk) * waves) {
        // ---- high Sobol words of this block of 64 indices (and of the
        // merged last block); this lane's row, index and path
        const SobolLane sl = sobol_block(hi, hi_tail, P.dirs, n, q, q0, q1, nlo, nhi, merge, lane, !MRG);
        const uint32_t* hl = sl.hl;
        const bool active = sl.active;
        const int64_t pidx = sl.pidx;
        // itemized risk: the checkpoint slot and the end state of this block of 64 paths
        R* stq = nullptr;
        if (ck_mode != 0) {
            const size_t sb = size_t(q - q0);
            ck = static_cast<R*>(P.ckpt) + (sb * ck_rows + (CKB ? kCkptGuardRows : 0)) * kWave;
            if constexpr (CKB) ckl = reinterpret_cast<uint8_t*>(ck + size_t(nck) * kWave);
            stq = static_cast<R*>(P.state) + sb * kStateRows * kWave + lane;
        }
        // normal of step i: from the Sobol point, or from this lane's
        // mrg32k3a generator (idle lanes run the call's first path)
        mrg::Lane gen;
        if constexpr (MRG) gen.start(P.mrg_jump, P.mrg_seed1, P.mrg_seed2, P.first_path + (active ? pidx : 0));
        // the inverse-normal table (LDS, or for LEAN the constant array
        // itself) less kInvNormalRowBias rows
        const R* invn_tab;
        if constexpr (GINV && sizeof(R) == 4) invn_tab = reinterpret_cast<const R*>(kInvNormalF32);
        else invn_tab = invn;
        invn_tab -= 4 * kInvNormalRowBias;
        auto draw = [&](int i) -> R {
            if constexpr (MRG) {
                if constexpr (sizeof(R) == 4) return gen.normal_f32(i, invn_tab);
                else return gen.normal(i);
            }
            else return inv_normal(hl[i] ^ low_word(i), invn_tab);
        };

        ProductState<R> ps;
        ps.alive = R(1); ps.pnz = R(1); ps.nzero = 0; ps.dead = false;
        bool nzw = false;   // barrier: some lane of the wave has an exactly-zero smoothing factor
        ps.euro = R(0); ps.fin_f = R(0); ps.fin_num = R(1); ps.aggp = R(0);
        // book: per barrier the state above and (after its maturity) weight x
        // vanilla, the adjoint of its survival factor; per fixing schedule
        // the running sum, replaced at its last fixing by the adjoint of each
        // fixing, and the weighted intrinsic value (numeraire adjoint)
        R bst[NBAR], babar[NBAR], gsum[NGRP], gq[NGRP];
#pragma unroll
        for (int b = 0; b < NBAR; ++b) { bst[b] = R(1); babar[b] = R(0); }
#pragma unroll
        for (int g = 0; g < NGRP; ++g) { gsum[g] = R(0); gq[g] = R(0); }

        // payoff `col` of this lane's path: wave sums, per-path output
        auto emit = [&](int col, R pay) {
            const double v = active ? double(pay) : 0.0;
            const double vs = wave_sum(v), vq = wave_sum(v * v);
            if (lane == kSumLane) { accw[col] += vs; accw[npay + col] += vq; }
            if (P.per_path && active) P.per_path[pidx * npay + col] = v;
            if constexpr (AD) ps.aggp = ps.aggp + R(wts[col]) * pay;
        };

        // observation of product sample s with spot S (and for Black-Scholes
        // the sample's forward factor / numeraire / discount)
        auto observe = [&](int s, R S) {
            R F = S, numv = R(1), discv = R(1);
            if constexpr (!DUP) {
                F = S * R(s_ff[s]);
                numv = P.spot_measure ? R(s_num[s]) * S : R(s_num[s]);
                discv = R(s_disc[s]);
            }
            if constexpr (PRD == PRD_BARRIER) {
                // mcPrd.h:256-287, branch free: a breach kills the path, the
                // smoothing zone multiplies in (B + w - F) / 2w
                const bool live = !ps.dead;
                const bool brk = live && F > R(P.bar_smooth);
                const bool zone = live && !brk && F > R(P.minus_smooth);
                const R a_prev = ps.alive;
                ps.dead = ps.dead || brk;
                if constexpr (ZT) {
                    const R f = (R(P.bar_smooth) - F) * R(P.inv_two_smooth);
                    if (zone) ps.alive = a_prev * f;
                    // exactly-zero factors (F = B + w), rare: from the first one on,
                    // pnz holds the product of the non-zero factors
                    const bool fz = zone && f == R(0);
                    nzw = nzw | (__builtin_amdgcn_ballot_w64(fz) != 0);
                    if (nzw) {
                        if (fz) {
                            if (ps.nzero == 0) ps.pnz = a_prev;
                            ps.nzero++;
                        } else if (zone && ps.nzero > 0) {
                            ps.pnz = ps.pnz * f;
                        }
                    }
                } else {
                    const R f = max(R(P.bar_smooth) - F, tiny_d) * R(P.inv_two_smooth);
                    ps.alive = a_prev * (zone ? f : R(1));
                }
                // the vanilla of the last sample: after the path (the last
                // sample is the last timeline point)
            } else if constexpr (PRD == PRD_EUROPEAN) {
                // mcPrd.h:113-125
                const R k = R(P.strike);
                const R pay = (F - k > R(0) ? F - k : R(0)) * discv / numv;
                const double v = active ? double(pay) : 0.0;
                const double vs = wave_sum(v), vq = wave_sum(v * v);
                if (lane == kSumLane) { accw[0] += vs; accw[npay] += vq; }
                if (P.per_path && active) P.per_path[pidx * npay] = v;
                if constexpr (AD) ps.aggp = ps.aggp + R(wts[0]) * pay;
            } else if constexpr (PRD == PRD_BOOK) {
                const SampleRow sm = bk_rows(s);
                // vanilla calls at this sample (European / Europeans parts)
                for (int k = sm.van_lo; k < sm.van_hi; ++k) {
                    const R kk = R(bk_van_strike[k]);
                    emit(bk_van_col[k], (F - kk > R(0) ? F - kk : R(0)) / numv);
                }
                // up-and-out barriers (mcPrd.h:256-287, one state per barrier),
                // branch free: breach, smoothing zone, exactly-zero factor;
                // no state changes while no lane is at or above the lowest zone
                const uint32_t bm0 = sm.bars & 0xffffu, bl = sm.bars >> 16;
                const uint32_t bm = bm0 && __any(F > bk_lo) ? bm0 : 0u;
                if (bm | bl) {
#pragma unroll
                    for (int b = 0; b < NBAR; ++b) {
                        if (!(((bm | bl) >> b) & 1u)) continue;
                        const BarPar bp = bar_par(b);
                        // the state changes only at or above the zone and only
                        // before a breach: skipped when no live lane of the wave
                        // is there
                        if (((bm >> b) & 1u) && __any(F > bp.lo && bst[b] > R(kBreached))) {
                            const R bs = bp.up, ms = bp.lo;
                            const R v = bst[b];
                            const bool live = v > R(kBreached);
                            const bool zone = live && F > ms && !(F > bs);
                            R nv = v;
                            if constexpr (ZT) {
                                const R f = (bs - F) * bp.i2s;
                                nv = zone && f != R(0) && v >= R(-1) ? v * f : nv;          // one more factor
                                nv = zone && f == R(0) ? (v >= R(0) ? -v : R(kTwoZeros)) : nv;   // a zero factor
                            } else {
                                // FP32: a factor at least 1e-30, never exactly 0
                                // (the state stays in [0, 1] until a breach)
                                const R f = max((bs - F) * bp.i2s, R(1e-30));
                                nv = v * (zone ? f : R(1));
                            }
                            nv = live && F > bs ? R(kBreached) : nv;                    // breach
                            bst[b] = nv;
                        }
                        if ((bl >> b) & 1u) {
                            const R k = bp.k;
                            const R intrinsic = bk_bar_put[b] ? (k - F > R(0) ? k - F : R(0))
                                                                         : (F - k > R(0) ? F - k : R(0));
                            const R euro = intrinsic / numv;
                            const R alive = bst[b] > R(0) ? bst[b] : R(0);
                            const int col = bk_bar_col[b];
                            emit(col, alive * euro);
                            emit(col + 1, euro);
                            // weight x vanilla: the adjoint of the survival
                            // factor (the ITEM kernels keep the vanilla and
                            // weigh it before the reverse sweep)
                            if constexpr (AD) babar[b] = ITEM ? euro : R(wts[col]) * euro;
                        }
                    }
                }
                // arithmetic averages: running sums per fixing schedule
                const uint32_t gm = sm.grps & 0xfu, gl = (sm.grps >> 4) & 0xfu;
                if (gm) {
#pragma unroll
                    for (int g = 0; g < NGRP; ++g) {
                        if (!((gm >> g) & 1u)) continue;
                        gsum[g] = gsum[g] + F;
                        if ((gl >> g) & 1u) {
                            const R inv_n = bk_ginv[g];
                            const R avg = gsum[g] * inv_n;
                            R dsum = R(0), qsum = R(0);
                            for (int a = bk_grp_first[g]; a < bk_grp_first[g + 1]; ++a) {
                                const R k = R(bk_as_strike[a]);
                                const bool put = bk_as_put[a] != 0;
                                const R intrinsic = put ? (k - avg > R(0) ? k - avg : R(0))
                                                        : (avg - k > R(0) ? avg - k : R(0));
                                const int col = bk_as_col[a];
                                emit(col, intrinsic / numv);
                                if constexpr (AD && !ITEM) {
                                    const R w = R(wts[col]);
                                    const R d = put ? (k > avg ? R(-1) : R(0)) : (avg > k ? R(1) : R(0));
                                    dsum = dsum + w * d;
                                    qsum = qsum + w * intrinsic;
                                }
                            }
                            // the adjoint of each fixing and the weighted
                            // intrinsic values (the numeraire adjoint); the ITEM
                            // kernels keep the average and the numeraire and
                            // form both before the reverse sweep
                            if constexpr (AD) {
                                gsum[g] = ITEM ? avg : dsum * inv_n / numv;
                                gq[g] = ITEM ? numv : qsum;
                            }
                        }
                    }
                }
            } else {
                // mcPrd.h:374-399
                for (int k = pay_first[s]; k < pay_first[s + 1]; ++k) {
                    const R kk = R(pay_strike[k]);
                    const R pay = (F - kk > R(0) ? F - kk : R(0)) / numv;
                    const double v = active ? double(pay) : 0.0;
                    const double vs = wave_sum(v), vq = wave_sum(v * v);
                    if (lane == kSumLane) { accw[k] += vs; accw[npay + k] += vq; }
                    if (P.per_path && active) P.per_path[pidx * npay + k] = v;
                    if constexpr (AD) ps.aggp = ps.aggp + R(wts[k]) * pay;
                }
            }
        };

        // Dupire barrier: a sample other than the last one changes nothing
        // while the spot is under the smoothing zone, so the spot is not
        // formed there (forward and reverse)
        constexpr bool UNDER = DUP && PRD == PRD_BARRIER;
        const R x_lo = UNDER ? R(P.bar_xlo) : R(0);
        // Dupire barrier: whether timeline points base .. base + 63 are
        // samples, one wave-uniform 64-bit mask from one ballot (bit b: point
        // base + b); the sweeps test a bit per step instead of reading the
        // sample map
        auto sample_bits = [&](int base) -> uint64_t {
            const int p = base + lane;
            return __builtin_amdgcn_ballot_w64(p <= n && step_sample[min(p, n)] >= 0);
        };
        R x;
        if (ck_mode == 2) {
            // reverse only: the end state the forward-only launch kept
            x = stq[0];
            ps.alive = stq[1 * kWave];
            ps.pnz = stq[2 * kWave];
            ps.nzero = int(stq[3 * kWave]);
            ps.dead = stq[4 * kWave] != R(0);
            ps.euro = stq[5 * kWave];
            ps.fin_f = stq[6 * kWave];
            ps.fin_num = stq[7 * kWave];
#pragma unroll
            for (int b = 0; b < NBAR; ++b) {
                bst[b] = stq[(8 + b) * kWave];
                babar[b] = stq[(8 + NBAR + b) * kWave];
            }
#pragma unroll
            for (int g = 0; g < NGRP; ++g) {
                gsum[g] = stq[(8 + 2 * NBAR + g) * kWave];
                gq[g] = stq[(8 + 2 * NBAR + NGRP + g) * kWave];
            }
            nzw = __builtin_amdgcn_ballot_w64(ps.nzero != 0) != 0;
        } else {
            // ---- forward sweep
            // adjoint kernels: waves in the forward (vector-ALU bound) take issue
            // priority over waves in the reverse (bound by LDS atomics)
            if constexpr (AD) __builtin_amdgcn_s_setprio(1);
            if constexpr (DUP) x = R(P.x0); else x = R(s0);   // Dupire: log-spot state; BS: spot
            {
                const int s = step_sample[0];
                if (s >= 0) observe(s, DUP ? state_spot(x, s0) : x);
            }
            // step i0 + k, with every per-step address a fixed offset (k) from a
            // base formed once per group of steps at i0.  GRP: a step of an
            // aligned group of four whose state and bracket the caller stores
            // at once (interleaved FP32 checkpoints), else stored here.
            R gx[4];
            uint32_t glo = 0;
            R gz[4];          // QW: the group's four normals, drawn before its steps
            uint64_t fbits = 0;   // UNDER: bit b, point 64 (i >> 6) + 1 + b is a sample
            auto fwd_step = [&](int i0, int k, auto grp, auto qw) {
                constexpr bool GRP = decltype(grp)::value, QW = decltype(qw)::value;
                const int i = i0 + k;
                R g;
                if constexpr (QW) g = gz[k];
                else if constexpr (MRG || LEAN) g = draw(i);
                else g = inv_normal(hl[i] ^ low_word(i), invn_tab);
                if constexpr (DUP) {
                    if constexpr (GRP) gx[k] = x;
                    else if constexpr (IL) ck[ck_el(i)] = x;
                    else if constexpr (ADJ == ADJ_CHECKPOINT) (ck + i0 * kWave + lane)[k * kWave] = x;
                    if constexpr (RECOMP)
                        if ((i & (kSeg - 1)) == 0) ck[(i / kSeg) * kWave + lane] = x;
                    const int lo = count_le<LEAN>(ks, x);
                    if constexpr (GRP) glo |= uint32_t(lo) << (8 * k);
                    else if constexpr (IL) { if (ck_lo) ckl[ck_el(i)] = uint8_t(lo); }
                    else if constexpr (CKB) { if (ck_lo) (ckl + i0 * kWave + lane)[k * kWave] = uint8_t(lo); }
                    const Bracket<R> b = local_vol_at(kxi, (iv + i0 * mp) + k * mp, x, lo);
                    x = euler_step(x, b.vol, g);
                } else {
                    if constexpr (ADJ == ADJ_CHECKPOINT) ck[i * kWave + lane] = g;
                    x = x * rexp(drift[i] + stdv[i] * g);      // mcMdlBS.h:344-345
                }
                // a wave-uniform test: the sample's code runs with all lanes (a lane
                // under the zone changes nothing there)
                if constexpr (UNDER) {
                    // (a group of four steps lies in one window of 64)
                    if ((QW ? k == 0 : true) && (i & 63) == 0) fbits = sample_bits(i + 1);
                    if (((fbits >> (i & 63)) & 1u) && __builtin_amdgcn_ballot_w64(x > x_lo) != 0)
                        observe(0, state_spot(x, s0));
                } else {
                    const int s = __builtin_amdgcn_readfirstlane((step_sample + i0 + 1)[k]);
                    if (s >= 0) observe(s, DUP ? state_spot(x, s0) : x);
                }
            };
            // FP32 Dupire adjoint kernels other than the book: groups of four steps
            // (the book's sample code, four times over, and FP64 measured slower)
            constexpr int UF = AD && DUP && !BOOK && sizeof(R) == 4 ? 4 : 1;
            int i = 0;
            if constexpr (UF > 1) {
                // (UF = 4 kernels with interleaved checkpoints: one 16-byte
                // store of the four states and one word of brackets per group)
                constexpr bool GRP = IL && UF == 4;
                // (the quads layout: the group's four Sobol words in two 16-byte reads)
                constexpr bool QW = QUADS && UF == 4;
                for (; i + UF <= n; i += UF) {
                    glo = 0;
                    if constexpr (QW) {
                        uint32_t gw[4];
                        sobol_words4(lo_tab, hl, lane, i, gw);
#pragma unroll
                        for (int k = 0; k < 4; ++k) gz[k] = inv_normal(gw[k], invn_tab);
                    }
    #pragma unroll
                    for (int k = 0; k < UF; ++k)
                        fwd_step(i, k, std::integral_constant<bool, GRP>{}, std::integral_constant<bool, QW>{});
                    if constexpr (GRP) {
                        *reinterpret_cast<float4*>(ck + ck_el(i)) = make_float4(gx[0], gx[1], gx[2], gx[3]);
                        if (ck_lo) *reinterpret_cast<uint32_t*>(ckl + ck_el(i)) = glo;
                    }
                }
            }
            // FP64: four draws at a time with the inverse normal's tail once
            // per wave (inv_normal4_p), the steps one by one (not the book:
            // its four steps' code measured slower there)
            if constexpr (sizeof(R) == 8 && !BOOK) {
                double* trow = tails + wave * TR;
                for (; i + 4 <= n; i += 4) {
                    if constexpr (MRG) {
                        double pu[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) pu[k] = gen.uniform(i + k);
                        inv_normal4_p<TR>(pu, gz, trow, lane);
                        if (gen.neg) {
#pragma unroll
                            for (int k = 0; k < 4; ++k) gz[k] = -gz[k];
                        }
                    } else {
                        uint32_t gw[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) gw[k] = hl[i + k] ^ low_word(i + k);
                        inv_normal4_f64<TR>(gw, gz, trow, lane);
                    }
#pragma unroll
                    for (int k = 0; k < 4; ++k) fwd_step(i, k, std::false_type{}, std::true_type{});
                }
            }
            if constexpr (sizeof(R) == 8 && BOOK) {
                // the book, FP64: the same four-draw batches, the steps
                // in a loop that is not unrolled (each takes its draw by a
                // select, so the book's step is compiled once)
                double* trow = tails + wave * TR;
                double gb[4] = {0.0, 0.0, 0.0, 0.0};
                const int n4 = n & ~3;
                for (; i < n4; ++i) {
                    const int k = i & 3;
                    if (k == 0) {
                        if constexpr (MRG) {
                            double pu[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) pu[j] = gen.uniform(i + j);
                            inv_normal4_p<TR>(pu, gb, trow, lane);
                            if (gen.neg) {
#pragma unroll
                                for (int j = 0; j < 4; ++j) gb[j] = -gb[j];
                            }
                        } else {
                            uint32_t gw[4];
#pragma unroll
                            for (int j = 0; j < 4; ++j) gw[j] = hl[i + j] ^ low_word(i + j);
                            inv_normal4_f64<TR>(gw, gb, trow, lane);
                        }
                    }
                    gz[0] = k == 0 ? gb[0] : k == 1 ? gb[1] : k == 2 ? gb[2] : gb[3];
                    fwd_step(i, 0, std::false_type{}, std::true_type{});
                }
            }
            for (; i < n; ++i) fwd_step(i, 0, std::false_type{}, std::false_type{});

            // ---- payoffs of the barrier (known at the end of the path)
            if constexpr (PRD == PRD_BARRIER) {
                if (last >= 0) {   // mcPrd.h:281-287, at the last sample (timeline point n)
                    const R S = DUP ? state_spot(x, s0) : x;
                    const R F = DUP ? S : S * R(s_ff[last]);
                    const R numv = DUP ? R(1) : (P.spot_measure ? R(s_num[last]) * S : R(s_num[last]));
                    const R k = R(P.strike);
                    const R intrinsic = P.put ? (k - F > R(0) ? k - F : R(0)) : (F - k > R(0) ? F - k : R(0));
                    ps.euro = intrinsic / numv;
                    ps.fin_f = F;
                    ps.fin_num = numv;
                }
                const R p0 = ps.dead ? R(0) : ps.alive * ps.euro;
                const R p1 = ps.euro;
                const double v0 = active ? double(p0) : 0.0, v1 = active ? double(p1) : 0.0;
                const double a0 = wave_sum(v0), a1 = wave_sum(v1), b0 = wave_sum(v0 * v0), b1 = wave_sum(v1 * v1);
                if (lane == kSumLane) { accw[0] += a0; accw[1] += a1; accw[npay] += b0; accw[npay + 1] += b1; }
                if (P.per_path && active) { P.per_path[pidx * 2] = v0; P.per_path[pidx * 2 + 1] = v1; }
                if constexpr (AD) ps.aggp = R(wts[0]) * p0 + R(wts[1]) * p1;
            }
        }
        if constexpr (AD) {
            __builtin_amdgcn_s_setprio(0);
            if (ck_mode != 2) {
                const double agg = active ? double(ps.aggp) : 0.0;
                const double as = wave_sum(agg), aq = wave_sum(agg * agg);
                if (lane == kSumLane) { accw[2 * npay] += as; accw[2 * npay + 1] += aq; }
                if (P.per_path_agg && active) P.per_path_agg[pidx] = agg;
            }
            if (ck_mode == 1) {
                // forward only: keep the end state for the reverse-only launches
                stq[0] = x;
                stq[1 * kWave] = ps.alive;
                stq[2 * kWave] = ps.pnz;
                stq[3 * kWave] = R(ps.nzero);
                stq[4 * kWave] = ps.dead ? R(1) : R(0);
                stq[5 * kWave] = ps.euro;
                stq[6 * kWave] = ps.fin_f;
                stq[7 * kWave] = ps.fin_num;
#pragma unroll
                for (int b = 0; b < NBAR; ++b) {
                    stq[(8 + b) * kWave] = bst[b];
                    stq[(8 + NBAR + b) * kWave] = babar[b];
                }
#pragma unroll
                for (int g = 0; g < NGRP; ++g) {
                    stq[(8 + 2 * NBAR + g) * kWave] = gsum[g];
                    stq[(8 + 2 * NBAR + NGRP + g) * kWave] = gq[g];
                }
                continue;
            }
            if constexpr (BOOK && ITEM) {
                // the weighted adjoints: each barrier's weight x vanilla (the
                // adjoint of its survival factor); per fixing schedule, from
                // the average and the numeraire, the adjoint of each fixing
                // sum_a w_a d max(.)/d avg / n / numeraire and the weighted
                // intrinsic values (the numeraire adjoint)
#pragma unroll
                for (int b = 0; b < NBAR; ++b)
                    if (b < P.bk.n_bar) babar[b] = R(wts[bk_bar_col[b]]) * babar[b];
#pragma unroll
                for (int g = 0; g < NGRP; ++g) {
                    if (g >= P.bk.n_grp) continue;
                    const R avg = gsum[g], numv = gq[g], inv_n = bk_ginv[g];
                    R dsum = R(0), qsum = R(0);
                    for (int a = bk_grp_first[g]; a < bk_grp_first[g + 1]; ++a) {
                        const R k = R(bk_as_strike[a]);
                        const bool put = bk_as_put[a] != 0;
                        const R intrinsic = put ? (k - avg > R(0) ? k - avg : R(0)) : (avg - k > R(0) ? avg - k : R(0));
                        const R w = R(wts[bk_as_col[a]]);
                        const R d = put ? (k > avg ? R(-1) : R(0)) : (avg > k ? R(1) : R(0));
                        dsum = dsum + w * d;
                        qsum = qsum + w * intrinsic;
                    }
                    gsum[g] = dsum * inv_n / numv;
                    gq[g] = qsum;
                }
            }

            // adjoints of the forward, numeraire and discount at sample s
            const R w0 = R(wts[0]);
            const R w1 = npay > 1 ? R(wts[1]) : R(0);
            const R alive_eff = ps.dead ? R(0) : ps.alive;
            const R euro_bar = w0 * alive_eff + w1;       // barrier: adjoint of the vanilla
            const R alive_bar = w0 * ps.euro;             // barrier: adjoint of alive
            // at_last: s is the last sample (the last timeline point); the
            // sweeps pass a constant false inside their step loops
            auto sample_adj = [&](int s, R F, R numv, R discv, R& Fb, R& nb, R& db, bool at_last) {
                Fb = R(0); nb = R(0); db = R(0);
                if constexpr (PRD == PRD_BARRIER) {
                    // d alive / dF = -(product of the other factors) / 2w: alive / f
                    // without a zero factor; with one, pnz for the zero factor
                    // and 0 for the others; 0 with two or more
                    const bool zone = !ps.dead && F > R(P.minus_smooth);
                    const R f = (ZT ? R(P.bar_smooth) - F : max(R(P.bar_smooth) - F, tiny_d)) * R(P.inv_two_smooth);
                    R others = recip_mul(ps.alive, f);
                    if (ZT && __builtin_expect(nzw, 0))
                        others = ps.nzero == 0 ? others : ((ps.nzero == 1 && f == R(0)) ? ps.pnz : R(0));
                    if (zone) Fb -= alive_bar * others * R(P.inv_two_smooth);
                    if (at_last) {
                        const R k = R(P.strike);
                        const R dmax = P.put ? (k > F ? R(-1) : R(0)) : (F > k ? R(1) : R(0));
                        Fb += euro_bar * dmax / numv;
                        nb += -euro_bar * ps.euro / numv;
                    }
                } else if constexpr (PRD == PRD_EUROPEAN) {
                    const R k = R(P.strike);
                    const R ind = F > k ? R(1) : R(0);
                    const R intrinsic = F > k ? F - k : R(0);
                    Fb = w0 * ind * discv / numv;
                    db = w0 * intrinsic / numv;
                    nb = -w0 * intrinsic * discv / (numv * numv);
                } else if constexpr (PRD == PRD_BOOK) {
                    const SampleRow sm = bk_rows(s);
                    for (int k = sm.van_lo; k < sm.van_hi; ++k) {
                        const R kk = R(bk_van_strike[k]);
                        const R wk = R(wts[bk_van_col[k]]);
                        if (F > kk) {
                            Fb += wk / numv;
                            nb += -wk * (F - kk) / (numv * numv);
                        }
                    }
                    // no lane at or above the lowest zone: no survival adjoint
                    const uint32_t bm0 = sm.bars & 0xffffu, bl = sm.bars >> 16;
                    const uint32_t bm = bm0 && __any(F > bk_lo) ? bm0 : 0u;
                    if (bm | bl) {
#pragma unroll
                        for (int b = 0; b < NBAR; ++b) {
                            if (!(((bm | bl) >> b) & 1u)) continue;
                            const BarPar bp = bar_par(b);
                            if ((bl >> b) & 1u) {
                                const int col = bk_bar_col[b];
                                const R k = bp.k;
                                const bool put = bk_bar_put[b] != 0;
                                const R intrinsic = put ? (k - F > R(0) ? k - F : R(0))
                                                        : (F - k > R(0) ? F - k : R(0));
                                const R dmax = put ? (k > F ? R(-1) : R(0)) : (F > k ? R(1) : R(0));
                                const R alive = bst[b] > R(0) ? bst[b] : R(0);
                                const R euro_bar = R(wts[col]) * alive + R(wts[col + 1]);
                                Fb += euro_bar * dmax / numv;
                                nb += -euro_bar * (intrinsic / numv) / numv;
                            }
                            if ((bm >> b) & 1u) {
                                // d alive / dF = -(product of the other factors) / 2w
                                const R v = bst[b];
                                if constexpr (ZT) {
                                    if (v > R(kBreached) && F > bp.lo) {   // smoothing zone
                                        const R i2s = bp.i2s;
                                        const R f = (bp.up - F) * i2s;
                                        // product of the other factors: v / f without a zero
                                        // factor, the non-zero ones for the only zero factor
                                        // (v >= 0 with f == 0 only after the product underflowed
                                        // to zero: the other factors' product is 0)
                                        const R others = v >= R(0) ? (f != R(0) ? recip_mul(v, f) : R(0))
                                                                   : ((v >= R(-1) && f == R(0)) ? -v : R(0));
                                        Fb -= babar[b] * others * i2s;
                                    }
                                } else if (v > R(kBreached) && F > bp.lo) {
                                    // FP32 (factors at least 1e-30): the others' product
                                    // is v / f (a wave with no lane in the zone skips it)
                                    const R f = max((bp.up - F) * bp.i2s, R(1e-30));
                                    Fb -= babar[b] * bp.i2s * recip_mul(v, f);
                                }
                            }
                        }
                    }
                    const uint32_t gm = sm.grps & 0xfu, gl = (sm.grps >> 4) & 0xfu;
                    if (gm) {
#pragma unroll
                        for (int g = 0; g < NGRP; ++g) {
                            if (!((gm >> g) & 1u)) continue;
                            Fb += gsum[g];
                            if ((gl >> g) & 1u) nb += -gq[g] / (numv * numv);
                        }
                    }
                } else {
                    for (int k = pay_first[s]; k < pay_first[s + 1]; ++k) {
                        const R kk = R(pay_strike[k]);
                        const R wk = R(wts[k]);
                        if (F > kk) {
                            Fb += wk / numv;
                            nb += -wk * (F - kk) / (numv * numv);
                        }
                    }
                }
            };

            // Dupire barrier, a sample before the last: S x d payoff / dF for spot S,
            // = -alive_bar (product of the other factors) S / 2w in the zone.
            // Without a zero factor the other factors' product is alive / f with
            // f = (B + w - S) / 2w, so the adjoint is kz S / (B + w - S),
            // kz = -alive_bar alive; with one, pnz at the zero factor
            const R kz = -alive_bar * (ps.dead ? R(0) : ps.alive);
            auto barrier_sf = [&](R S) -> R {
                const bool zone = !ps.dead && S > R(P.minus_smooth);
                const R d = ZT ? R(P.bar_smooth) - S : max(R(P.bar_smooth) - S, tiny_d);
                R sf = kz * recip_mul(S, d);
                if (ZT && __builtin_expect(nzw, 0))
                    sf = ps.nzero == 0 ? sf
                                       : ((ps.nzero == 1 && d * R(P.inv_two_smooth) == R(0))
                                              ? -alive_bar * ps.pnz * R(P.inv_two_smooth) * S : R(0));
                return zone ? sf : R(0);
            };

            if constexpr (DUP && RECOMP) {
                // ---- reverse sweep, recompute variant: segments of kSeg steps
                // from the last one back; each segment is first re-simulated
                // from its stored state (Sobol point, inverse normal, local
                // vol), keeping in registers the step's local derivatives
                //   abar_i = abar_{i+1} (1 + (g - vol) dvol/dx) + sample adjoint
                //   vol adjoint = abar_{i+1} (g - vol), split on the two nodes
                R abar = R(0);
                {
                    const int s = step_sample[n];
                    if (s >= 0) {
                        const R S = state_spot(x, s0);
                        R Fb, nb, db;
                        sample_adj(s, S, R(1), R(1), Fb, nb, db, true);   // step_sample[n] is the last sample
                        abar = S * Fb;
                    }
                }
                const int nseg = (n + kSeg - 1) / kSeg;
                for (int sg = nseg - 1; sg >= 0; --sg) {
                    const int i0 = sg * kSeg;
                    R xi = ck[sg * kWave + lane];
                    R ca[kSeg], cb[kSeg], cw[kSeg], cv[kSeg];
                    int cj[kSeg];
                    // quads layout: the segment's Sobol words, four at a time
                    // (a group from step n on is not read)
                    uint32_t sw[kSeg];
                    if constexpr (QUADS) {
                        static_assert(kSeg % 4 == 0, "segments of whole groups of four");
#pragma unroll
                        for (int q4 = 0; q4 < kSeg; q4 += 4)
                            if (i0 + q4 < n) sobol_words4(lo_tab, hl, lane, i0 + q4, sw + q4);
                    }
                    // mrg32k3a: the segment's normals drawn top down, so the
                    // generator only steps back over the reverse sweep; FP64
                    // Sobol: the segment's four draws with the inverse
                    // normal's tail once per wave (steps from n on draw 1/2)
                    R mz[kSeg];
                    if constexpr (MRG && sizeof(R) == 8) {
                        static_assert(kSeg == 4, "FP64 segments of four draws");
                        double pu[4];
#pragma unroll
                        for (int k = 3; k >= 0; --k) pu[k] = i0 + k < n ? gen.uniform(i0 + k) : 0.5;
                        inv_normal4_p<TR>(pu, mz, tails + wave * TR, lane);
                        if (gen.neg) {
#pragma unroll
                            for (int k = 0; k < 4; ++k) mz[k] = -mz[k];
                        }
                    } else if constexpr (MRG) {
#pragma unroll
                        for (int k = kSeg - 1; k >= 0; --k) mz[k] = i0 + k < n ? draw(i0 + k) : R(0);
                    } else if constexpr (sizeof(R) == 8) {
                        static_assert(kSeg == 4, "FP64 segments of four draws");
                        uint32_t gw[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) gw[k] = i0 + k < n ? (hl[i0 + k] ^ low_word(i0 + k)) : 0x80000000u;
                        inv_normal4_f64<TR>(gw, mz, tails + wave * TR, lane);
                    }
#pragma unroll
                    for (int k = 0; k < kSeg; ++k) {
                        const int i = i0 + k;
                        ca[k] = R(1);
                        cb[k] = R(0);
                        cw[k] = R(0);
                        cv[k] = R(0);
                        cj[k] = 0;
                        if (i >= n) continue;
                        R g;
                        if constexpr (QUADS) g = inv_normal(sw[k], invn_tab);
                        else if constexpr (MRG || sizeof(R) == 8) g = mz[k];
                        else g = draw(i);
                        const Bracket<R> b = local_vol<LEAN>(ks, kxi, iv + i * mp, xi);
                        const R ws = g - b.vol;    // d x_{i+1} / d vol
                        ca[k] = R(1) + ws * b.slope;
                        cw[k] = ws * b.wa;
                        cv[k] = ws * b.wb;
                        cj[k] = b.j;
                        const int smp = __builtin_amdgcn_readfirstlane(step_sample[i]);
                        if constexpr (UNDER) {
                            if (smp >= 0 && __builtin_amdgcn_ballot_w64(xi > x_lo) != 0)
                                cb[k] = barrier_sf(state_spot(xi, s0));
                        } else if (smp >= 0) {
                            const R S = state_spot(xi, s0);
                            R Fb, nb, db;
                            sample_adj(smp, S, R(1), R(1), Fb, nb, db, false);
                            cb[k] = S * Fb;
                        }
                        xi = euler_step(xi, b.vol, g);
                    }
#pragma unroll
                    for (int k = kSeg - 1; k >= 0; --k) {
                        const int i = i0 + k;
                        if (i >= n) continue;
                        if (active) {
                            double* row = ivbl + i * mp;
                            atomicAdd(row + cj[k], double(abar * cw[k]));
                            atomicAdd(row + cj[k] + 1, double(abar * cv[k]));
                        }
                        abar = abar * ca[k] + cb[k];
                    }
                }
                if (active) padj[0] += double(abar) / s0;      // d log(S0) / d S0
            } else if constexpr (DUP) {
                // ---- reverse sweep over the Euler steps (mcMdlDupire.h:259-279)
                R abar = R(0);   // adjoint of the log-spot state at the current point
                {
                    const int s = step_sample[n];
                    if (s >= 0) {
                        const R S = state_spot(x, s0);
                        R Fb, nb, db;
                        sample_adj(s, S, R(1), R(1), Fb, nb, db, true);   // step_sample[n] is the last sample
                        abar = S * Fb;
                    }
                }
                R xnext = x;     // state at point i + 1
                // every local-vol node at least the recovery threshold: no
                // normal is drawn again
                const bool iv_recover = P.iv_recover != 0;
                const R vmin = R(sizeof(R) == 8 ? kRecoverVolF64 : kRecoverVolF32);
                // groups of U steps, top down to top - U + 1; the checkpoint
                // loads of the next group are issued before the current group
                // is worked (FP64 and the book hold more registers: U = 1).
                // FP32 groups of four are the aligned groups of the interleaved
                // checkpoints (steps 4g .. 4g + 3), one 16-byte load of states
                // and one word of brackets each; the top group may be partial
                // (steps from n on are skipped).  The prefetch below step 0
                // reads the slot's guard rows (never used).
                constexpr int U = sizeof(R) == 8 ? 1 : (BOOK ? 1 : 4);
                constexpr bool VG = IL && U == 4;
                static_assert(U == 1 || VG, "groups of four steps read interleaved checkpoints");
                static_assert(U <= kCkptGuardRows, "the prefetch reads at most kCkptGuardRows rows below 0");
                R buf[U];
                int lbuf[U];
                int rwin = -1;          // UNDER: window of 64 points held in rbits
                uint64_t rbits = 0;
                // states and brackets of steps top, top - 1, ..., top - U + 1
                auto load_group = [&](int top, R* xb, int* lb) {
                    if constexpr (VG) {
                        const int e = ck_el(top - 3);   // a 16-byte aligned group
                        const float4 v = *reinterpret_cast<const float4*>(ck + e);
                        xb[0] = v.w;
                        xb[1] = v.z;
                        xb[2] = v.y;
                        xb[3] = v.x;
                        const uint32_t w = ck_lo ? *reinterpret_cast<const uint32_t*>(ckl + e) : 0u;
#pragma unroll
                        for (int k = 0; k < 4; ++k) lb[k] = int((w >> (8 * (3 - k))) & 0xffu);
                    } else {
                        const int e = ck_el(top);
                        xb[0] = ck[e];
                        lb[0] = ck_lo ? int(ckl[e]) : 0;
                    }
                };
                auto group = [&](int top, auto partial) {
                    constexpr bool PART = decltype(partial)::value;
                    R nbuf[U];
                    int nlbuf[U];
                    load_group(top - U, nbuf, nlbuf);
                    // phase 1, independent of the adjoint chain: brackets, the
                    // derivative of each step to its vol, the sample adjoints
                    R ws[U], sl[U], wa[U], wb[U], sf[U];
                    int js[U];
                    // the sample index of each step (aligned groups: one 16-byte
                    // read); the Dupire barrier tests the window's sample bits
                    // (a group lies in one window of 64 points)
                    int smps[U];
                    if constexpr (UNDER) {
                        if ((top >> 6) != rwin) {
                            rwin = top >> 6;
                            rbits = sample_bits(rwin << 6);
                        }
#pragma unroll
                        for (int k = 0; k < U; ++k) smps[k] = int(((rbits >> ((top - k) & 63)) & 1u)) - 1;
                    } else if constexpr (VG) {
                        const int4 v = *reinterpret_cast<const int4*>(step_sample + (top - 3));
                        smps[0] = v.w;
                        smps[1] = v.z;
                        smps[2] = v.y;
                        smps[3] = v.x;
                    } else {
                        smps[0] = step_sample[top];
                    }
#pragma unroll
                    for (int k = 0; k < U; ++k) {
                        const int i = top - k;
                        const int smp = UNDER ? smps[k] : __builtin_amdgcn_readfirstlane(smps[k]);
                        ws[k] = R(0);
                        sl[k] = R(0);
                        wa[k] = R(0);
                        wb[k] = R(0);
                        sf[k] = R(0);
                        js[k] = 0;
                        if (PART && i >= n) continue;
                        const R xi = buf[k];
                        const int lo = ck_lo ? lbuf[k] : count_le<LEAN>(ks, xi);
                        const Bracket<R> b = local_vol_at(kxi, iv + i * mp, xi, lo);
                        // d x_{i+1} / d vol = g - vol, recovered from the two states:
                        // x_{i+1} - x_i = vol (g - vol / 2) up to the rounding of
                        // x_{i+1}; where a local vol is below the recovery
                        // threshold (rare; a wave-uniform test) the normal is
                        // drawn again
                        ws[k] = recip_mul(xnext - xi, b.vol) - R(0.5) * b.vol;
                        if (!iv_recover && __builtin_amdgcn_ballot_w64(b.vol < vmin) != 0)
                            if (b.vol < vmin) ws[k] = draw(i) - b.vol;
                        sl[k] = b.slope;
                        wa[k] = b.wa;
                        wb[k] = b.wb;
                        js[k] = b.j;
                        if constexpr (UNDER) {
                            if (smp >= 0 && __builtin_amdgcn_ballot_w64(xi > x_lo) != 0)
                                sf[k] = barrier_sf(state_spot(xi, s0));
                        } else {
                            if (smp >= 0) {
                                const R S = state_spot(xi, s0);
                                R Fb, nb, db;
                                sample_adj(smp, S, R(1), R(1), Fb, nb, db, false);
                                sf[k] = S * Fb;
                            }
                        }
                        xnext = xi;
                    }
                    // phase 2: the adjoint of the state, step by step backwards
                    R vb[U];
#pragma unroll
                    for (int k = 0; k < U; ++k) {
                        vb[k] = abar * ws[k];
                        abar = abar + vb[k] * sl[k];
                        abar = abar + sf[k];
                    }
                    // phase 3: scatter onto the two bracketing local-vol nodes
                    // (rows top - U + 1 .. top at fixed offsets from one base)
                    if (active) {
                        double* rows = ivbl + (top - (U - 1)) * mp;
#pragma unroll
                        for (int k = 0; k < U; ++k) {
                            if (PART && top - k >= n) continue;
                            double* row = rows + (U - 1 - k) * mp;
                            atomicAdd(row + js[k], double(vb[k] * wa[k]));
                            atomicAdd(row + js[k] + 1, double(vb[k] * wb[k]));
                        }
                    }
#pragma unroll
                    for (int k = 0; k < U; ++k) { buf[k] = nbuf[k]; lbuf[k] = nlbuf[k]; }
                };
                if (n > 0) {
                    // the sweep starts at the last point whose sample carries
                    // weight (one payoff's reverse of itemized risk: later
                    // steps have no adjoint), the next state from its checkpoint
                    const int last = min(n - 1, P.rev_point);
                    int top = VG ? ((last >> 2) << 2) + 3 : last;
                    if (top + 1 < n) xnext = ck[ck_el(top + 1)];
                    load_group(top, buf, lbuf);
                    if (VG && top >= n) {
                        group(top, std::true_type{});
                        top -= U;
                    }
                    for (; top >= U - 1; top -= U) group(top, std::false_type{});
                }
                if (active) padj[0] += double(abar) / s0;      // d log(S0) / d S0
            } else {
                // ---- Black-Scholes: forward tangents of the spot with respect
                // to (spot, vol, rate, div), replaying the checkpointed normals
                // (mcMdlBS.h:203-225, :240-275, :288-349)
                const double* tl = P.bs_aux;     // sample times
                const double* tf = tl + ns;      // forward maturity - sample time
                const double* td = tf + ns;      // discount maturity - sample time
                const double* dt = td + ns;      // step lengths
                const double vol = P.bs_vol;
                R spot = R(s0);
                R tau[4] = {R(1.0 / s0), R(0), R(0), R(0)};   // d spot / spot
                auto accumulate = [&](int s, R S) {
                    const R ff = R(s_ff[s]);
                    const R F = S * ff;
                    const R numv = P.spot_measure ? R(s_num[s]) * S : R(s_num[s]);
                    const R discv = R(s_disc[s]);
                    R Fb, nb, db;
                    sample_adj(s, F, numv, discv, Fb, nb, db, s == last);
                    // dF = F (tau + dff/ff), ff = exp(mu (Tf - t))
                    const R tfs = R(tf[s]), tds = R(td[s]), ts = R(tl[s]);
                    R dF[4] = {F * tau[0], F * tau[1], F * (tau[2] + tfs), F * (tau[3] - tfs)};
                    R dD[4] = {R(0), R(0), -tds * discv, R(0)};
                    R dN[4];
                    if (P.spot_measure) {
                        // num = exp(div t) / S0 * spot
                        dN[0] = numv * (tau[0] - R(1.0 / s0));
                        dN[1] = numv * tau[1];
                        dN[2] = numv * tau[2];
                        dN[3] = numv * (tau[3] + ts);
                    } else {
                        dN[0] = R(0); dN[1] = R(0); dN[2] = numv * ts; dN[3] = R(0);
                    }
                    if (active)
                        for (int c = 0; c < 4; ++c) padj[c] += double(Fb * dF[c] + db * dD[c] + nb * dN[c]);
                };
                {
                    const int s = step_sample[0];
                    if (s >= 0) accumulate(s, spot);
                }
                const int nrev = min(n, P.rev_point);   // samples after it carry no weight
                for (int i = 0; i < nrev; ++i) {
                    R g;
                    if constexpr (RECOMP) g = draw(i);   // regenerated
                    else g = ck[i * kWave + lane];
                    spot = spot * rexp(drift[i] + stdv[i] * g);
                    const R dti = R(dt[i]);
                    const R sq = R(sqrt(dt[i]));
                    tau[1] = tau[1] + (P.spot_measure ? R(vol) * dti : -R(vol) * dti) + sq * g;
                    tau[2] = tau[2] + dti;
                    tau[3] = tau[3] - dti;
                    const int s = step_sample[i + 1];
                    if (s >= 0) accumulate(s, spot);
                }
            }
        }
    }

    // ---- parameter adjoints: wave, then per-wave accumulators
    if constexpr (AD) {
        for (int c = 0; c < kMaxParamAdj; ++c) {
            const double v = wave_sum(padj[c]);
            if (lane == kSumLane) accw[2 * npay + 2 + c] += v;
        }
    }
    __syncthreads();
    for (int c = tid; c < stride; c += nthr) {
        double s = 0.0;
        for (int w = 0; w < waves; ++w) s += acc[w * stride + c];
        P.block_out[size_t(bid) * stride + c] = s;
    }
    if constexpr (AD && DUP) {
        double* dst = P.ivbar_blocks + size_t(bid) * n * m;
        // padded columns back onto the grid nodes: column j + 1 is node j,
        // the end columns 0 and m + 1 belong to the first and last node
        for (int e = tid; e < n * m; e += nthr) {
            const int i = e / m, j = e - i * m;
            double v = 0.0;
            for (int c = 0; c < P.ivbar_copies; ++c) {
                const double* r = ivbar + size_t(c) * ivbar_copy_stride(n, m) + size_t(i) * mp;
                v += r[j + 1];
                if (j == 0) v += r[0];
                if (j == m - 1) v += r[m + 1];
            }
            dst[e] = v;
        }
    }
}

#ifndef CF_PATH_KERNEL_ATTR
#define CF_PATH_KERNEL_ATTR
#endif

template <typename R, int MODEL, int PRD, int ADJ, bool LEAN, bool MRG, bool ITEM>
__global__ void __launch_bounds__(book_block_limit(PRD, int(sizeof(R)), ADJ)) CF_PATH_KERNEL_ATTR
path_kernel(const LaunchParams P)
{
    path_body<R, MODEL, PRD, ADJ, LEAN, MRG, ITEM>(P, int(blockIdx.x), int(gridDim.x));
}

// Several calls of one kernel in one launch: blocks [first[r], first[r + 1])
// run call r (Sobol normals, no itemized modes).  The calls' parameters sit in
// the kernel arguments, read with scalar loads like the single call's.
template <typename R, int MODEL, int PRD, int ADJ, bool LEAN>
__global__ void __launch_bounds__(book_block_limit(PRD, int(sizeof(R)), ADJ)) CF_PATH_KERNEL_ATTR
path_kernel_packed(const PackArgs A)
{
    const int b = int(blockIdx.x);
    int r = 0;
    while (r + 1 < A.n && b >= A.first[r + 1]) ++r;
    r = __builtin_amdgcn_readfirstlane(r);
    path_body<R, MODEL, PRD, ADJ, LEAN, false, false>(A.p[r], b - A.first[r], A.first[r + 1] - A.first[r]);
}

}  // namespace cfk
