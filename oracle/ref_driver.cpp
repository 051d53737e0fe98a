/*
 * Driver for the reference engine, compiled together with the reference
 * sources by build_ref.sh into oracle/_ref/ref_driver.  TEST INFRASTRUCTURE
 * ONLY: it writes golden vectors (tests/golden/make_golden.py) and times the
 * reference on the host cores (bench.py, CPU baseline).  It calls the
 * reference's own store and entry points (store.h:38-205, main.h:43-254).
 *
 * Usage:  ref_driver <command> [args] < spec
 *   sobol <dims> <b> <n_next>        states after skipTo(b) and n_next next()
 *   invnorm <n>                      n uniforms on stdin -> invNormalCdf
 *   filldata <maxDx> <minDx> <n> <x...> <nAdd> <a...>   (args on stdin)
 *   value <nPath> <nOut> <threads> <repeats>            spec on stdin
 *   aad <nPath> <riskPayoff|-1> <threads> <repeats> <nOut>   spec on stdin
 *   calib                            dupireCalib (main.h:414-438), on stdin:
 *        <spot> <vol> <jmpIntens> <jmpAverage> <jmpStd> <nS> <spots> <maxDs> <nT> <times> <maxDt>
 *   superbucket <nPath>              dupireSuperbucket (main.h:453-569), on stdin:
 *        <spot> <maxDt> <nS> <spots> <maxDs> <nT> <times> <maxDtVol>
 *        <nK> <strikes> <nM> <mats> <vol> <jmpIntens> <jmpAverage> <jmpStd>
 *        then a product spec and notionals as for aad
 *   superbucket_bump <nPath>         dupireSuperbucketBump (main.h:575-696), the same input
 * The spec is whitespace separated:
 *   model bs <spot> <vol> <qspot> <rate> <div>
 *   model dupire <spot> <maxDt> <nS> <spots> <nT> <times> <vols, spot major>
 *   product european <K> <Tex> <Tset>
 *   product barrier <K> <B> <T> <freq> <smooth> <put 0/1>
 *   product europeans <n> <maturities> <strikes>
 *   notionals <n> <payoff index, notional>...   (aad: AADriskAggregate)
 * Output is one JSON object on stdout.
 */
#include "main.h"

#include <chrono>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

static void fail(const std::string& m)
{
    std::fprintf(stderr, "ref_driver: %s\n", m.c_str());
    std::exit(2);
}

static std::vector<double> readVec(std::istream& in, size_t n)
{
    std::vector<double> v(n);
    for (auto& x : v) in >> x;
    return v;
}

static void printVec(const char* name, const std::vector<double>& v, bool comma = true)
{
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s\n", comma ? "," : "");
}

struct Spec {
    std::vector<std::pair<std::string, double>> notionals;
    bool aggregate = false;
    bool mrg = false;            // rng mrg <seed1> <seed2>: mrg32k3a instead of Sobol
    int seed1 = 0, seed2 = 0;
};

static Spec readSpec()
{
    Spec spec;
    std::string tok;
    while (std::cin >> tok) {
        if (tok == "model") {
            std::string kind;
            std::cin >> kind;
            if (kind == "bs") {
                double spot, vol, q, r, d;
                std::cin >> spot >> vol >> q >> r >> d;
                putBlackScholes(spot, vol, q > 0, r, d, "m");
            } else if (kind == "dupire") {
                double spot, maxDt;
                size_t nS, nT;
                std::cin >> spot >> maxDt >> nS;
                auto spots = readVec(std::cin, nS);
                std::cin >> nT;
                auto times = readVec(std::cin, nT);
                matrix<double> vols(nS, nT);
                for (size_t i = 0; i < nS; ++i)
                    for (size_t j = 0; j < nT; ++j) std::cin >> vols[i][j];
                putDupire(spot, spots, times, vols, maxDt, "m");
            } else fail("unknown model " + kind);
        } else if (tok == "product") {
            std::string kind;
            std::cin >> kind;
            if (kind == "european") {
                double K, Tex, Tset;
                std::cin >> K >> Tex >> Tset;
                putEuropean(K, Tex, Tset, "p");
            } else if (kind == "barrier") {
                double K, B, T, f, s, put;
                std::cin >> K >> B >> T >> f >> s >> put;
                putBarrier(K, B, T, f, s, put > 0, "p");
            } else if (kind == "europeans") {
                size_t n;
                std::cin >> n;
                auto mats = readVec(std::cin, n);
                auto strikes = readVec(std::cin, n);
                putEuropeans(mats, strikes, "p");
            } else fail("unknown product " + kind);
        } else if (tok == "rng") {
            std::string kind;
            std::cin >> kind >> spec.seed1 >> spec.seed2;
            if (kind != "mrg") fail("unknown rng " + kind);
            spec.mrg = true;
        } else if (tok == "notionals") {
            size_t n;
            std::cin >> n;
            const auto* labels = getPayoffLabels("p");
            if (!labels) fail("notionals before product");
            for (size_t i = 0; i < n; ++i) {
                size_t idx;
                double w;
                std::cin >> idx >> w;
                spec.notionals.emplace_back((*labels)[idx], w);
            }
            spec.aggregate = true;
        } else fail("bad token " + tok);
    }
    return spec;
}

static double msSince(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static double median(std::vector<double> v)
{
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

int main(int argc, char** argv)
{
    if (argc < 2) fail("no command");
    const std::string cmd = argv[1];

    if (cmd == "sobol") {
        // points of paths b, b+1, ..., b+n-1 (mcBase.h:375-381: skipTo(first)
        // then one nextU per path); the uint32 state behind each uniform is
        // recovered exactly from u = ONEOVER2POW32 * state (sobol.h:102-108)
        const int dims = std::atoi(argv[2]);
        const unsigned b = unsigned(std::strtoul(argv[3], nullptr, 10));
        const int n = std::atoi(argv[4]);
        Sobol s;
        s.init(dims);
        s.skipTo(b);
        std::vector<double> u(dims);
        std::printf("{\"states\": [\n");
        for (int k = 0; k < n; ++k) {
            s.nextU(u);
            std::printf("%s[", k ? ",\n" : "");
            for (int d = 0; d < dims; ++d) {
                long long i = std::llround(u[d] / ONEOVER2POW32);
                for (long long c : {i, i - 1, i + 1})
                    if (ONEOVER2POW32 * (unsigned long)c == u[d]) { i = c; break; }
                if (ONEOVER2POW32 * (unsigned long)i != u[d]) fail("state not recoverable");
                std::printf("%s%lld", d ? ", " : "", i);
            }
            std::printf("]");
        }
        std::printf("]}\n");
        return 0;
    }

    if (cmd == "invnorm") {
        const long n = std::atol(argv[2]);
        std::vector<double> out(n);
        for (long i = 0; i < n; ++i) {
            double u;
            std::cin >> u;
            out[i] = invNormalCdf(u);
        }
        std::printf("{");
        printVec("z", out, false);
        std::printf("}\n");
        return 0;
    }

    if (cmd == "filldata") {
        double maxDx, minDx;
        size_t n, nAdd;
        std::cin >> maxDx >> minDx >> n;
        auto x = readVec(std::cin, n);
        std::cin >> nAdd;
        auto a = readVec(std::cin, nAdd);
        auto f = nAdd ? fillData(x, maxDx, minDx, a.data(), a.data() + nAdd) : fillData(x, maxDx, minDx);
        std::printf("{");
        printVec("filled", f, false);
        std::printf("}\n");
        return 0;
    }

    if (cmd == "value" || cmd == "aad") {
        const int nPath = std::atoi(argv[2]);
        const bool aad = cmd == "aad";
        const int riskIdx = aad ? std::atoi(argv[3]) : -1;
        const int threads = std::atoi(argv[aad ? 4 : 3]);
        const int repeats = std::max(1, std::atoi(argv[aad ? 5 : 4]));
        const int nOut = std::atoi(argv[aad ? 6 : 5]);
        Spec spec = readSpec();
        NumericalParam num;
        num.parallel = threads > 1;
        num.useSobol = !spec.mrg;
        num.seed1 = spec.seed1;
        num.seed2 = spec.seed2;
        num.numPath = nPath;
        if (threads > 1) ThreadPool::getInstance()->start(threads - 1);

        std::vector<double> times;
        std::printf("{\"threads\": %d, \"nPath\": %d,\n", threads, nPath);
        if (!aad) {
            decltype(value("m", "p", num)) res;
            for (int r = 0; r < repeats; ++r) {
                auto t0 = std::chrono::steady_clock::now();
                res = value("m", "p", num);
                times.push_back(msSince(t0));
            }
            printVec("values", res.values);
            // first nOut per-path payoffs, via the serial simulator on the same
            // Sobol sequence (mcBase.h:267-308; identical paths to the parallel one)
            if (nOut > 0) {
                std::unique_ptr<RNG> rng;
                if (spec.mrg) rng = std::make_unique<mrg32k3a>(spec.seed1, spec.seed2);
                else rng = std::make_unique<Sobol>();
                auto mat = mcSimul(*getProduct<double>("p"), *getModel<double>("m"), *rng, nOut);
                std::vector<double> flat;
                for (auto& row : mat) flat.insert(flat.end(), row.begin(), row.end());
                printVec("per_path", flat);
            }
            std::printf("\"labels\": [");
            for (size_t i = 0; i < res.identifiers.size(); ++i)
                std::printf("%s\"%s\"", i ? ", " : "", res.identifiers[i].c_str());
            std::printf("],\n");
        } else {
            const auto& labels = *getPayoffLabels("p");
            std::vector<double> payoffValues, risks;
            double riskValue = 0.0;
            std::vector<std::string> paramIds;
            for (int r = 0; r < repeats; ++r) {
                auto t0 = std::chrono::steady_clock::now();
                if (spec.aggregate) {
                    std::map<std::string, double> nots(spec.notionals.begin(), spec.notionals.end());
                    auto res = AADriskAggregate("m", "p", nots, num);
                    times.push_back(msSince(t0));
                    payoffValues = res.payoffValues; risks = res.risks; riskValue = res.riskPayoffValue;
                    paramIds = res.paramIds;
                } else {
                    auto res = AADriskOne("m", "p", num, riskIdx >= 0 ? labels[riskIdx] : std::string());
                    times.push_back(msSince(t0));
                    payoffValues = res.payoffValues; risks = res.risks; riskValue = res.riskPayoffValue;
                    paramIds = res.paramIds;
                }
            }
            printVec("payoff_values", payoffValues);
            std::printf("\"risk_payoff_value\": %.17g,\n", riskValue);
            printVec("risks", risks);
            std::printf("\"param_ids\": [");
            for (size_t i = 0; i < paramIds.size(); ++i)
                std::printf("%s\"%s\"", i ? ", " : "", paramIds[i].c_str());
            std::printf("],\n");
            (void)nOut;
        }
        printVec("wall_ms", times);
        std::printf("\"wall_ms_median\": %.6f}\n", median(times));
        if (threads > 1) ThreadPool::getInstance()->stop();
        return 0;
    }
    if (cmd == "calib") {
        double spot, vol, ji, ja, js, maxDs, maxDt;
        size_t nS, nT;
        std::cin >> spot >> vol >> ji >> ja >> js >> nS;
        auto spots = readVec(std::cin, nS);
        std::cin >> maxDs >> nT;
        auto times = readVec(std::cin, nT);
        std::cin >> maxDt;
        auto res = dupireCalib(spots, maxDs, times, maxDt, spot, vol, ji, ja, js);
        std::vector<double> lv;
        for (size_t i = 0; i < res.lVols.rows(); ++i)
            for (size_t j = 0; j < res.lVols.cols(); ++j) lv.push_back(res.lVols[i][j]);
        std::printf("{");
        printVec("spots", res.spots);
        printVec("times", res.times);
        printVec("lvols", lv, false);
        std::printf("}\n");
        return 0;
    }

    if (cmd == "superbucket" || cmd == "superbucket_bump") {
        const int nPath = std::atoi(argv[2]);
        double spot, maxDt, maxDs, maxDtVol, vol, ji, ja, js;
        size_t nS, nT, nK, nM;
        std::cin >> spot >> maxDt >> nS;
        auto spots = readVec(std::cin, nS);
        std::cin >> maxDs >> nT;
        auto times = readVec(std::cin, nT);
        std::cin >> maxDtVol >> nK;
        auto strikes = readVec(std::cin, nK);
        std::cin >> nM;
        auto mats = readVec(std::cin, nM);
        std::cin >> vol >> ji >> ja >> js;
        Spec spec = readSpec();
        NumericalParam num;
        num.parallel = false;
        num.useSobol = true;
        num.numPath = nPath;
        std::map<std::string, double> nots(spec.notionals.begin(), spec.notionals.end());
        if (cmd == "superbucket_bump") {
            auto res = dupireSuperbucketBump(spot, maxDt, "p", nots, spots, maxDs, times, maxDtVol, strikes, mats,
                                             vol, ji, ja, js, num);
            std::vector<double> vega;
            for (size_t i = 0; i < res.vega.rows(); ++i)
                for (size_t j = 0; j < res.vega.cols(); ++j) vega.push_back(res.vega[i][j]);
            std::printf("{\"value\": %.17g, \"delta\": %.17g,\n", res.value, res.delta);
            printVec("vega", vega, false);
            std::printf("}\n");
            return 0;
        }
        auto res = dupireSuperbucket(spot, maxDt, "p", nots, spots, maxDs, times, maxDtVol, strikes, mats, vol, ji,
                                     ja, js, num);
        std::vector<double> vega;
        for (size_t i = 0; i < res.vega.rows(); ++i)
            for (size_t j = 0; j < res.vega.cols(); ++j) vega.push_back(res.vega[i][j]);
        // the microbucket of the calibrated model, stored by the superbucket
        // as "superbucket" (main.h:519-534), recomputed on the same paths
        auto micro = dupireAADRisk("superbucket", "p", nots, num);
        std::vector<double> mb(micro.vega.begin(), micro.vega.end());
        std::printf("{\"value\": %.17g, \"delta\": %.17g,\n", res.value, res.delta);
        printVec("microbucket", mb);
        printVec("vega", vega, false);
        std::printf("}\n");
        return 0;
    }

    fail("unknown command " + cmd);
}
