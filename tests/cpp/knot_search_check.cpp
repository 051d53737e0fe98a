// The knot search of the path kernels (cfk::count_le over the table of
// cfk::build_knot_table, compfinance_amd/csrc/cf_internal.h: the same code the
// kernels and the host engine compile) against std::upper_bound, the
// reference's search (interp.h:40), on the benchmark grids and on random
// grids, in FP32 and FP64.  Probes: every knot, its neighbouring floating
// point values, random states inside and outside the grid, and +-infinity.
// Prints one JSON object; tests/test_layout.py reads it.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

#include "../../compfinance_amd/csrc/cf_internal.h"

using namespace cfk;

template <typename R>
long check(const std::vector<R>& knots, std::mt19937_64& rng, long* probes, int* max_fix)
{
    const int m = int(knots.size());
    const KnotTable<R> t = build_knot_table(knots.data(), m);
    KnotSearch<R> s;
    s.knots = knots.data();
    s.table = t.entries.data();
    s.m = m;
    s.fix = t.fix;
    s.stride = 1;
    s.invh = t.invh;
    s.c0 = t.c0;
    s.top = R(int(t.entries.size()) - 1);
    *max_fix = std::max(*max_fix, t.fix);
    std::vector<R> xs;
    const R inf = std::numeric_limits<R>::infinity();
    for (R k : knots) {
        xs.push_back(k);
        xs.push_back(std::nextafter(k, inf));
        xs.push_back(std::nextafter(k, -inf));
    }
    const double lo = double(knots.front()) - 1.0, hi = double(knots.back()) + 1.0;
    std::uniform_real_distribution<double> u(lo, hi);
    for (int i = 0; i < 2000; ++i) xs.push_back(R(u(rng)));
    xs.push_back(inf);
    xs.push_back(-inf);
    long wrong = 0;
    for (R x : xs) {
        const int want = int(std::upper_bound(knots.begin(), knots.end(), x) - knots.begin());
        if (count_le<true>(s, x) != want) ++wrong;
        // the FULL layout's form (no loop) on the tables it is chosen for
        if (t.fix <= 2 && count_le<false>(s, x) != want) ++wrong;
        ++*probes;
    }
    return wrong;
}

template <typename R>
std::vector<R> bench_grid(int n, bool shifted)
{
    // spots 50 + 150 i / (n - 1); FP64 state log S, FP32 state log(S / S0)
    std::vector<R> k(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        const double s = n > 1 ? 50.0 + 150.0 * i / (n - 1) : 100.0;
        k[size_t(i)] = R(shifted ? std::log(s) - std::log(100.0) : std::log(s));
    }
    return k;
}

template <typename R>
long random_grids(std::mt19937_64& rng, long* probes, int* max_fix)
{
    long wrong = 0;
    std::uniform_int_distribution<int> nm(1, 300);
    std::uniform_real_distribution<double> gap(0.0, 1.0);
    for (int g = 0; g < 300; ++g) {
        const int m = nm(rng);
        std::vector<R> k(static_cast<size_t>(m));
        double x = std::log(1.0 + 200.0 * gap(rng));
        const int kind = g % 4;   // even, random, clustered, with repeated knots
        for (int j = 0; j < m; ++j) {
            k[size_t(j)] = R(x);
            double d = kind == 0 ? 0.05 : gap(rng) * 0.2;
            if (kind == 2 && j % 7 != 0) d *= 1e-5;
            if (kind == 3 && j % 5 == 0) d = 0.0;
            x += d;
        }
        std::sort(k.begin(), k.end());
        wrong += check(k, rng, probes, max_fix);
    }
    return wrong;
}

int main()
{
    std::mt19937_64 rng(20261205);
    long probes = 0, wrong_f = 0, wrong_d = 0;
    int fix_f = 0, fix_d = 0;
    for (int n : {1, 2, 20, 30, 60}) {
        wrong_f += check(bench_grid<float>(n, true), rng, &probes, &fix_f);
        wrong_d += check(bench_grid<double>(n, false), rng, &probes, &fix_d);
    }
    wrong_f += random_grids<float>(rng, &probes, &fix_f);
    wrong_d += random_grids<double>(rng, &probes, &fix_d);
    const KnotTable<float> c3 = build_knot_table(bench_grid<float>(30, true).data(), 30);
    std::printf("{\"probes\": %ld, \"wrong_fp32\": %ld, \"wrong_fp64\": %ld, \"max_fix_fp32\": %d, "
                "\"max_fix_fp64\": %d, \"c3_fp32_fix\": %d, \"c3_fp32_buckets\": %zu}\n",
                probes, wrong_f, wrong_d, fix_f, fix_d, c3.fix, c3.entries.size());
    return 0;
}
