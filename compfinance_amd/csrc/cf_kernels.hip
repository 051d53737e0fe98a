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
//   * local-vol table, log-spot knots, timeline sample map in LDS;
//   * payoffs reduced per wave with cross-lane shuffles into per-wave LDS
//     accumulators, then per block, then across blocks by a separate kernel
//     in fixed order;
//   * adjoint: path states checkpointed in a per-wave HBM slot (reused by the
//     wave for each block of paths), reverse sweep over the Euler steps,
//     local-vol adjoints accumulated with LDS atomics on a per-block
//     (step x spot) grid.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <tuple>

#include "cf_internal.h"

namespace cfk {

// ------------------------------------------------------------------ helpers

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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

// Beasley-Springer-Moro inverse normal of u = ONEOVER2POW32 * xi
// (gaussians.h:47-87, sobol.h:26, :110-116), FP64, same branches and
// constants as the reference.
__device__ __forceinline__ double inv_normal(uint32_t xi, double)
{
    const double p = 2.3283064365387E-10 * double(xi);
    const bool sup = p > 0.5;
    const double up = sup ? 1.0 - p : p;
    const double x = up - 0.5;
    if (fabs(x) < 0.42) {
        const double r = x * x;
        const double z = x * (((-25.44106049637 * r + 41.39119773534) * r + -18.61500062529) * r + 2.50662823884) /
                         ((((3.13082909833 * r + -21.06224101826) * r + 23.08336743743) * r + -8.47351093090) * r + 1.0);
        return sup ? -z : z;
    }
    double r = log_pos(-log_pos(up));
    r = 0.3374754822726147 +
        r * (0.9761690190917186 +
        r * (0.1607979714918209 +
        r * (0.0276438810333863 +
        r * (0.0038405729373609 +
        r * (0.0003951896511919 +
        r * (0.0000321767881768 +
        r * (0.0000002888167364 + r * 0.0000003960315187)))))));
    return sup ? r : -r;
}

// FP32 form: Giles' single-precision erfinv approximation (M. Giles,
// "Approximating the erfinv function", GPU Computing Gems Jade Edition,
// 2011), z = sqrt(2) erfinv(2u - 1).  Its central branch covers 99.7% of the
// draws (Moro's 84%), so a wave rarely runs the tail branch.  The tail mass
// min(u, 1 - u) is formed exactly in the integer domain (2^32 - xi) before
// scaling, so no precision is lost near u = 1.  Differences to the FP64
// Moro values are of the order of FP32 rounding.
__device__ __forceinline__ float inv_normal(uint32_t xi, float)
{
    const bool sup = xi >= 0x80000000u;              // u > 1/2 for the reference scaling
    const uint32_t xs = sup ? (0u - xi) : xi;         // 2^32 - xi, exact
    const float up = float(xs) * 2.3283064e-10f;      // min(u, 1 - u)
    const float ax = 1.0f - 2.0f * up;                // |2u - 1|
    const float ln2 = 0.693147180559945f;
    // w = -ln((1 - x)(1 + x)) = -ln(4 up (1 - up)); the argument is a normal number
    float w = -__builtin_amdgcn_logf(4.0f * up * (1.0f - up)) * ln2;
    float p;
    if (w < 5.0f) {
        w = w - 2.5f;
        p = 2.81022636e-08f;
        p = 3.43273939e-07f + p * w;
        p = -3.5233877e-06f + p * w;
        p = -4.39150654e-06f + p * w;
        p = 0.00021858087f + p * w;
        p = -0.00125372503f + p * w;
        p = -0.00417768164f + p * w;
        p = 0.246640727f + p * w;
        p = 1.50140941f + p * w;
    } else {
        w = __builtin_amdgcn_sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = 0.000100950558f + p * w;
        p = 0.00134934322f + p * w;
        p = -0.00367342844f + p * w;
        p = 0.00573950773f + p * w;
        p = -0.0076224613f + p * w;
        p = 0.00943887047f + p * w;
        p = 1.00167406f + p * w;
        p = 2.83297682f + p * w;
    }
    const float z = 1.41421356237f * p * ax;
    return sup ? z : -z;
}

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
    int j;        // left node
    R wa, wb;     // adjoint weights of nodes j and j + 1
};

// upper_bound on the knots (interp.h:40) through a uniform bucket table:
// bucket b = min(floor((x - x0) * invh), nb - 1) holds the number of knots
// in earlier buckets, a lower bound of the answer; at most `fix` knots fall in
// one bucket and are resolved by comparisons.  The host builds the table with
// the kernel's own arithmetic, so the answer is exactly upper_bound's.
template <typename R>
struct KnotSearch {
    const R* knots;
    const uint16_t* table;
    int m, nb, fix;
    R x0, invh;
};

template <typename R>
__device__ __forceinline__ int count_le(const KnotSearch<R>& s, R x)
{
    // states below the first knot map to bucket 0, whose count is 0 and whose
    // comparisons all fail; NaN maps there too
    R bf = (x - s.x0) * s.invh;
    bf = bf > R(0) ? bf : R(0);
    bf = bf < R(s.nb - 1) ? bf : R(s.nb - 1);
    int lo = s.table[int(bf)];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int c = lo < s.m ? lo : s.m - 1;
        lo += (lo < s.m && s.knots[c] <= x) ? 1 : 0;
    }
    for (int k = 2; k < s.fix; ++k) {   // grids with more than two knots in a bucket
        const int c = lo < s.m ? lo : s.m - 1;
        lo += (lo < s.m && s.knots[c] <= x) ? 1 : 0;
    }
    return lo;
}

// Branch-free: every load is in range (nodes clamped), the extrapolation
// cases are selected afterwards.
template <typename R>
__device__ __forceinline__ Bracket<R> local_vol(const KnotSearch<R>& ks, const R* inv_dx, const R* row, R x)
{
    const int m = ks.m;
    const int lo = count_le(ks, x);   // number of knots <= x
    int j = lo - 1;
    j = j > m - 2 ? m - 2 : j;
    j = j < 0 ? 0 : j;
    const int j2 = j + 1 < m ? j + 1 : j;
    const R y1 = row[j], y2 = row[j2];
    const R x1 = ks.knots[j], ih = inv_dx[j];
    const bool right = lo >= m && m > 1;
    const bool left = !right && (lo == 0 || m == 1);
    const R t = (x - x1) * ih;
    const R dy = y2 - y1;
    Bracket<R> b;
    b.j = j;
    b.vol = left ? y1 : (right ? y2 : y1 + dy * t);
    b.slope = (left || right) ? R(0) : dy * ih;
    b.wa = left ? R(1) : (right ? R(0) : R(1) - t);
    b.wb = left ? R(0) : (right ? R(1) : t);
    return b;
}

__device__ __forceinline__ double recip_mul(double a, double b) { return fast_div(a, b); }
__device__ __forceinline__ float recip_mul(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// ------------------------------------------------------------------ products

// Per-lane product state carried along the path.
template <typename R>
struct ProductState {
    R alive;        // barrier: product of smoothing factors (before a breach)
    R pnz;          // barrier: product of the non-zero smoothing factors
    int nzero;      // barrier: number of exactly-zero factors
    bool dead;      // barrier: breached (alive forced to 0, no derivative)
    R euro;         // barrier: vanilla payoff at maturity
    R fin_f;        // barrier: final forward
    R fin_num;      // barrier: final numeraire
    R aggp;         // running aggregate of payoffs emitted at samples
};

// ------------------------------------------------------------------ kernel

template <typename R, int MODEL, int PRD, bool ADJ, bool LEAN>
__global__ void __launch_bounds__(1024) path_kernel(const LaunchParams P)
{
    extern __shared__ __align__(16) unsigned char smem[];
    constexpr bool DUP = MODEL == MODEL_DUPIRE;
    constexpr bool EUS = PRD == PRD_EUROPEANS;
    const int waves = blockDim.x >> 6;
    const int n = P.n_steps, m = P.n_spots, ns = P.n_samples, npay = P.n_pay;
    const int stride = out_stride(npay);
    const SmemLayout L = smem_layout(n, m, ns, npay, waves, int(sizeof(R)), DUP, EUS, ADJ, LEAN, P.n_buckets);

    uint32_t* lo_tab = reinterpret_cast<uint32_t*>(smem + L.lo);
    uint32_t* hi_all = reinterpret_cast<uint32_t*>(smem + L.hi);
    int* step_sample = reinterpret_cast<int*>(smem + L.step_sample);
    R* knots = reinterpret_cast<R*>(smem + L.knots);
    R* inv_dx = reinterpret_cast<R*>(smem + L.inv_dx);
    uint16_t* bucket = reinterpret_cast<uint16_t*>(smem + L.bucket);
    R* iv_lds = reinterpret_cast<R*>(smem + L.iv);
    R* drift = reinterpret_cast<R*>(smem + L.drift);
    R* stdv = reinterpret_cast<R*>(smem + L.stdv);
    double* s_ff = reinterpret_cast<double*>(smem + L.samples);
    double* s_disc = s_ff + ns;
    double* s_num = s_disc + ns;
    int* pay_first = reinterpret_cast<int*>(smem + L.pay_first);
    double* pay_strike = reinterpret_cast<double*>(smem + L.pay_strike);
    double* acc = reinterpret_cast<double*>(smem + L.acc);
    double* ivbar = reinterpret_cast<double*>(smem + L.ivbar);

    const int tid = threadIdx.x, nthr = blockDim.x;

    // ---- stage the block's tables in LDS
    if constexpr (LEAN) {
        for (int e = tid; e < n * 8; e += nthr) {
            const int d = e >> 3, k = e & 7;
            lo_tab[e] = k < 6 ? P.dirs[k * kSobolDims + d] : 0u;
        }
    } else {
        for (int e = tid; e < n * kWave; e += nthr) {
            const int d = e >> 6, r = e & 63;
            const uint32_t g = uint32_t(r ^ (r >> 1));
            uint32_t x = 0;
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if ((g >> k) & 1u) x ^= P.dirs[k * kSobolDims + d];
            lo_tab[e] = x;
        }
    }
    for (int e = tid; e <= n; e += nthr) step_sample[e] = P.step_sample[e];
    if constexpr (DUP) {
        const R* gk = static_cast<const R*>(P.knots);
        const R* gi = static_cast<const R*>(P.inv_dx);
        const R* gv = static_cast<const R*>(P.iv);
        for (int e = tid; e < m; e += nthr) { knots[e] = gk[e]; inv_dx[e] = gi[e]; }
        for (int e = tid; e < P.n_buckets; e += nthr) bucket[e] = P.bucket[e];
        if constexpr (!LEAN)
            for (int e = tid; e < n * m; e += nthr) iv_lds[e] = gv[e];
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
    if constexpr (ADJ && DUP)
        for (int e = tid; e < n * m; e += nthr) ivbar[e] = 0.0;
    __syncthreads();

    const int wave = tid >> 6, lane = tid & 63;
    // local-vol table: LDS (FULL) or global memory (LEAN)
    const R* iv;
    if constexpr (LEAN) iv = static_cast<const R*>(P.iv); else iv = iv_lds;
    // LEAN: masks of the bits of gray6(lane), selecting direction numbers 0..5
    uint32_t lmask[6];
    {
        const uint32_t g6 = uint32_t(lane ^ (lane >> 1));
#pragma unroll
        for (int k = 0; k < 6; ++k) lmask[k] = ((g6 >> k) & 1u) ? 0xffffffffu : 0u;
    }
    // low Sobol word of dimension i for this lane
    auto low_word = [&](int i) -> uint32_t {
        if constexpr (LEAN) {
            const uint32_t* v = lo_tab + i * 8;
            return (lmask[0] & v[0]) ^ (lmask[1] & v[1]) ^ (lmask[2] & v[2]) ^ (lmask[3] & v[3]) ^
                   (lmask[4] & v[4]) ^ (lmask[5] & v[5]);
        } else {
            return lo_tab[i * kWave + lane];
        }
    };
    uint32_t* hi = hi_all + wave * n;
    double* accw = acc + wave * stride;
    const int64_t nlo = P.first_path + 1, nhi = P.first_path + P.n_path;   // Sobol indices
    const int64_t q0 = nlo >> 6, q1 = nhi >> 6;
    const int slot = blockIdx.x * waves + wave;
    R* ck = ADJ ? static_cast<R*>(P.ckpt) + size_t(slot) * n * kWave : nullptr;
    KnotSearch<R> ks;
    ks.knots = knots;
    ks.table = bucket;
    ks.m = m;
    ks.nb = P.n_buckets;
    ks.fix = P.bucket_fix;
    ks.x0 = R(P.bucket_x0);
    ks.invh = R(P.bucket_invh);
    const double s0 = P.spot;
    const int last = ns - 1;

    // per-lane parameter adjoints, summed over the lane's paths
    double padj[kMaxParamAdj] = {0.0, 0.0, 0.0, 0.0};

    for (int64_t q = q0 + int64_t(blockIdx.x) * waves + wave; q <= q1; q += int64_t(gridDim.x) * waves) {
        // ---- high Sobol words of this block of 64 indices
        {
            const uint32_t gq = uint32_t(q ^ (q >> 1));
            for (int d = lane; d < n; d += kWave) {
                uint32_t x = (q & 1) ? P.dirs[5 * kSobolDims + d] : 0u;
                uint32_t b = gq;
                while (b) {
                    const int j = __builtin_ctz(b);
                    x ^= P.dirs[(j + 6) * kSobolDims + d];
                    b &= b - 1u;
                }
                hi[d] = x;
            }
            wave_sync_lds();
        }
        const int64_t sidx = q * kWave + lane;
        const bool active = sidx >= nlo && sidx <= nhi;
        const int64_t pidx = sidx - nlo;

        ProductState<R> ps;
        ps.alive = R(1); ps.pnz = R(1); ps.nzero = 0; ps.dead = false;
        ps.euro = R(0); ps.fin_f = R(0); ps.fin_num = R(1); ps.aggp = R(0);

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
                // mcPrd.h:256-287
                if (!ps.dead) {
                    if (F > R(P.bar_smooth)) {
                        ps.dead = true;
                    } else if (F > R(P.minus_smooth)) {
                        const R f = (R(P.bar_smooth) - F) * R(P.inv_two_smooth);
                        ps.alive = ps.alive * f;
                        if (f == R(0)) ps.nzero++; else ps.pnz = ps.pnz * f;
                    }
                }
                if (s == last) {
                    const R k = R(P.strike);
                    const R intrinsic = P.put ? (k - F > R(0) ? k - F : R(0)) : (F - k > R(0) ? F - k : R(0));
                    ps.euro = intrinsic / numv;
                    ps.fin_f = F;
                    ps.fin_num = numv;
                }
            } else if constexpr (PRD == PRD_EUROPEAN) {
                // mcPrd.h:113-125
                const R k = R(P.strike);
                const R pay = (F - k > R(0) ? F - k : R(0)) * discv / numv;
                const double v = active ? double(pay) : 0.0;
                const double vs = wave_sum(v), vq = wave_sum(v * v);
                if (lane == 0) { accw[0] += vs; accw[npay] += vq; }
                if (P.per_path && active) P.per_path[pidx * npay] = v;
                if constexpr (ADJ) ps.aggp = ps.aggp + R(P.weights[0]) * pay;
            } else {
                // mcPrd.h:374-399
                for (int k = pay_first[s]; k < pay_first[s + 1]; ++k) {
                    const R kk = R(pay_strike[k]);
                    const R pay = (F - kk > R(0) ? F - kk : R(0)) / numv;
                    const double v = active ? double(pay) : 0.0;
                    const double vs = wave_sum(v), vq = wave_sum(v * v);
                    if (lane == 0) { accw[k] += vs; accw[npay + k] += vq; }
                    if (P.per_path && active) P.per_path[pidx * npay + k] = v;
                    if constexpr (ADJ) ps.aggp = ps.aggp + R(P.weights[k]) * pay;
                }
            }
        };

        // ---- forward sweep
        R x;
        if constexpr (DUP) x = R(P.x0); else x = R(s0);   // Dupire: log-spot state; BS: spot
        {
            const int s = step_sample[0];
            if (s >= 0) observe(s, DUP ? state_spot(x, s0) : x);
        }
        for (int i = 0; i < n; ++i) {
            const uint32_t u = hi[i] ^ low_word(i);
            const R g = inv_normal(u, R());
            if constexpr (DUP) {
                if constexpr (ADJ) ck[i * kWave + lane] = x;
                const Bracket<R> b = local_vol(ks, inv_dx, iv + i * m, x);
                x = x + b.vol * (R(-0.5) * b.vol + g);     // mcMdlDupire.h:271
            } else {
                if constexpr (ADJ) ck[i * kWave + lane] = g;
                x = x * rexp(drift[i] + stdv[i] * g);      // mcMdlBS.h:344-345
            }
            const int s = step_sample[i + 1];
            if (s >= 0) observe(s, DUP ? state_spot(x, s0) : x);
        }

        // ---- payoffs of the barrier (known at the end of the path)
        if constexpr (PRD == PRD_BARRIER) {
            const R p0 = ps.dead ? R(0) : ps.alive * ps.euro;
            const R p1 = ps.euro;
            const double v0 = active ? double(p0) : 0.0, v1 = active ? double(p1) : 0.0;
            const double a0 = wave_sum(v0), a1 = wave_sum(v1), b0 = wave_sum(v0 * v0), b1 = wave_sum(v1 * v1);
            if (lane == 0) { accw[0] += a0; accw[1] += a1; accw[npay] += b0; accw[npay + 1] += b1; }
            if (P.per_path && active) { P.per_path[pidx * 2] = v0; P.per_path[pidx * 2 + 1] = v1; }
            if constexpr (ADJ) ps.aggp = R(P.weights[0]) * p0 + R(P.weights[1]) * p1;
        }

        if constexpr (ADJ) {
            const double agg = active ? double(ps.aggp) : 0.0;
            const double as = wave_sum(agg), aq = wave_sum(agg * agg);
            if (lane == 0) { accw[2 * npay] += as; accw[2 * npay + 1] += aq; }
            if (P.per_path_agg && active) P.per_path_agg[pidx] = agg;

            // adjoints of the forward, numeraire and discount at sample s
            const R w0 = R(P.weights[0]);
            const R w1 = npay > 1 ? R(P.weights[1]) : R(0);
            const R alive_eff = ps.dead ? R(0) : ps.alive;
            const R euro_bar = w0 * alive_eff + w1;       // barrier: adjoint of the vanilla
            const R alive_bar = w0 * ps.euro;             // barrier: adjoint of alive
            auto sample_adj = [&](int s, R F, R numv, R discv, R& Fb, R& nb, R& db) {
                Fb = R(0); nb = R(0); db = R(0);
                if constexpr (PRD == PRD_BARRIER) {
                    if (!ps.dead && F > R(P.minus_smooth)) {
                        const R f = (R(P.bar_smooth) - F) * R(P.inv_two_smooth);
                        R others;
                        if (ps.nzero == 0) others = recip_mul(ps.alive, f);
                        else others = (ps.nzero == 1 && f == R(0)) ? ps.pnz : R(0);
                        Fb -= alive_bar * others * R(P.inv_two_smooth);
                    }
                    if (s == last) {
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
                } else {
                    for (int k = pay_first[s]; k < pay_first[s + 1]; ++k) {
                        const R kk = R(pay_strike[k]);
                        const R wk = R(P.weights[k]);
                        if (F > kk) {
                            Fb += wk / numv;
                            nb += -wk * (F - kk) / (numv * numv);
                        }
                    }
                }
            };

            if constexpr (DUP) {
                // ---- reverse sweep over the Euler steps (mcMdlDupire.h:259-279)
                R abar = R(0);   // adjoint of the log-spot state at the current point
                {
                    const int s = step_sample[n];
                    if (s >= 0) {
                        const R S = state_spot(x, s0);
                        R Fb, nb, db;
                        sample_adj(s, S, R(1), R(1), Fb, nb, db);
                        abar = S * Fb;
                    }
                }
                R xnext = x;     // state at point i + 1
                // checkpoint loads run kPrefetch steps ahead of their use
                constexpr int kPrefetch = sizeof(R) == 8 ? 1 : 4;   // FP64 holds twice the registers
                R buf[kPrefetch];
#pragma unroll
                for (int k = 0; k < kPrefetch; ++k) buf[k] = n - 1 - k >= 0 ? ck[(n - 1 - k) * kWave + lane] : R(0);
                for (int i0 = n - 1; i0 >= 0; i0 -= kPrefetch) {
                    // phase 1, independent of the adjoint chain: brackets, the
                    // derivative of each step to its vol, the sample adjoints
                    R ws[kPrefetch], sl[kPrefetch], wa[kPrefetch], wb[kPrefetch], sf[kPrefetch];
                    int js[kPrefetch];
#pragma unroll
                    for (int k = 0; k < kPrefetch; ++k) {
                        const int i = i0 - k;
                        ws[k] = R(0);
                        sl[k] = R(0);
                        wa[k] = R(0);
                        wb[k] = R(0);
                        sf[k] = R(0);
                        js[k] = 0;
                        if (i < 0) continue;
                        const R xi = buf[k];
                        buf[k] = i - kPrefetch >= 0 ? ck[(i - kPrefetch) * kWave + lane] : R(0);
                        const Bracket<R> b = local_vol(ks, inv_dx, iv + i * m, xi);
                        // d x_{i+1} / d vol = g - vol, recovered from the two states:
                        // x_{i+1} - x_i = vol (g - vol / 2), exact difference in FP64
                        if (b.vol != R(0)) {
                            ws[k] = recip_mul(xnext - xi, b.vol) - R(0.5) * b.vol;
                        } else {
                            const uint32_t u = hi[i] ^ low_word(i);
                            ws[k] = inv_normal(u, R());
                        }
                        sl[k] = b.slope;
                        wa[k] = b.wa;
                        wb[k] = b.wb;
                        js[k] = b.j;
                        const int smp = step_sample[i];
                        if (smp >= 0) {
                            const R S = state_spot(xi, s0);
                            R Fb, nb, db;
                            sample_adj(smp, S, R(1), R(1), Fb, nb, db);
                            sf[k] = S * Fb;
                        }
                        xnext = xi;
                    }
                    // phase 2: the adjoint of the state, step by step backwards
                    R vb[kPrefetch];
#pragma unroll
                    for (int k = 0; k < kPrefetch; ++k) {
                        vb[k] = abar * ws[k];
                        abar = abar + vb[k] * sl[k];
                        abar = abar + sf[k];
                    }
                    // phase 3: scatter onto the two bracketing local-vol nodes
                    if (active) {
#pragma unroll
                        for (int k = 0; k < kPrefetch; ++k) {
                            const int i = i0 - k;
                            if (i < 0) break;
                            double* row = ivbar + i * m;
                            atomicAdd(row + js[k], double(vb[k] * wa[k]));
                            if (m > 1) atomicAdd(row + js[k] + 1, double(vb[k] * wb[k]));
                        }
                    }
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
                    sample_adj(s, F, numv, discv, Fb, nb, db);
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
                for (int i = 0; i < n; ++i) {
                    const R g = ck[i * kWave + lane];
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
    if constexpr (ADJ) {
        for (int c = 0; c < kMaxParamAdj; ++c) {
            const double v = wave_sum(padj[c]);
            if (lane == 0) accw[2 * npay + 2 + c] += v;
        }
    }
    __syncthreads();
    for (int c = tid; c < stride; c += nthr) {
        double s = 0.0;
        for (int w = 0; w < waves; ++w) s += acc[w * stride + c];
        P.block_out[size_t(blockIdx.x) * stride + c] = s;
    }
    if constexpr (ADJ && DUP) {
        double* dst = P.ivbar_blocks + size_t(blockIdx.x) * n * m;
        for (int e = tid; e < n * m; e += nthr) dst[e] = ivbar[e];
    }
}

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

template <typename R, int MODEL, int PRD, bool ADJ>
static const void* kernel_ptr(bool lean)
{
    return lean ? reinterpret_cast<const void*>(&path_kernel<R, MODEL, PRD, ADJ, true>)
                : reinterpret_cast<const void*>(&path_kernel<R, MODEL, PRD, ADJ, false>);
}

template <typename R, int MODEL, bool ADJ>
static const void* pick_product(int prd, bool lean)
{
    switch (prd) {
    case PRD_EUROPEAN: return kernel_ptr<R, MODEL, PRD_EUROPEAN, ADJ>(lean);
    case PRD_BARRIER: return kernel_ptr<R, MODEL, PRD_BARRIER, ADJ>(lean);
    case PRD_EUROPEANS: return kernel_ptr<R, MODEL, PRD_EUROPEANS, ADJ>(lean);
    }
    return nullptr;
}

template <typename R, bool ADJ>
static const void* pick_model(int model, int prd, bool lean)
{
    return model == MODEL_DUPIRE ? pick_product<R, MODEL_DUPIRE, ADJ>(prd, lean)
                                 : pick_product<R, MODEL_BS, ADJ>(prd, lean);
}

static const void* pick(const KernelConfig& c)
{
    if (c.precision == 0)
        return c.adjoint ? pick_model<double, true>(c.model, c.product, c.lean)
                         : pick_model<double, false>(c.model, c.product, c.lean);
    return c.adjoint ? pick_model<float, true>(c.model, c.product, c.lean)
                     : pick_model<float, false>(c.model, c.product, c.lean);
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
