// C++17 mirror of the reference simulation framework (mcBase.h:37-753) on the
// MI355X engine.  Same interface names and signatures for what callers use:
// Product<T>, Model<T>, RNG, Sobol, mcSimul / mcParallelSimul,
// AADSimulResults, mcSimulAAD / mcParallelSimulAAD.  The per-path virtuals of
// the reference (generatePath, payoffs, nextG) are replaced by descriptor():
// models and products flatten themselves into the C ABI's POD descriptors and
// the whole simulation runs in the HIP kernels of libcfamd.so.
#pragma once

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../cfamd.h"

namespace cfamd {

using std::string;
using std::unique_ptr;
using std::vector;

using Time = double;

// Engine options of the calling thread: precision and AAD variant.
struct EngineOptions {
    int precision = CF_FP64;
    int aadMode = CF_AAD_CHECKPOINT;
};
inline EngineOptions& engineOptions()
{
    thread_local EngineOptions opts;
    return opts;
}

// Throws std::runtime_error with the engine's message on a negative status.
inline int cfCheck(int rc)
{
    if (rc < 0) throw std::runtime_error(cf_last_error());
    return rc;
}

// AAD number (AADExpr.h:856-1134) reduced to what callers read: a value and
// its adjoint.  The engine computes adjoints itself; no tape is recorded.
class Number {
    double myValue = 0.0;
    double myAdjoint = 0.0;

public:
    Number() = default;
    Number(double v) : myValue(v) {}
    double& value() { return myValue; }
    double value() const { return myValue; }
    double& adjoint() { return myAdjoint; }
    double adjoint() const { return myAdjoint; }
    explicit operator double() const { return myValue; }
    Number& operator=(double v)
    {
        myValue = v;
        return *this;
    }
};

inline double toDouble(double x) { return x; }
inline double toDouble(const Number& x) { return x.value(); }

// SampleDef (mcBase.h:45-65): what a product needs on each event date.
struct SampleDef {
    bool numeraire = true;
    struct RateDef {
        Time start;
        Time end;
        string curve;
        RateDef(const Time s, const Time e, const string& c) : start(s), end(e), curve(c) {}
    };
    vector<Time> discountMats;
    vector<RateDef> liborDefs;
    vector<vector<Time>> forwardMats;
};

// Product<T> (mcBase.h:122-152)
template <class T>
class Product {
    inline static const vector<string> defaultAssetNames = {"spot"};

public:
    virtual const vector<Time>& timeline() const = 0;
    virtual const vector<SampleDef>& defline() const = 0;
    virtual const size_t numAssets() const { return 1; }
    virtual const vector<string>& assetNames() const { return defaultAssetNames; }
    virtual const vector<string>& payoffLabels() const = 0;
    virtual unique_ptr<Product<T>> clone() const = 0;
    // Flattened form for the engine; pointers stay valid while *this lives.
    virtual cf_product descriptor() const = 0;
    virtual ~Product() {}
};

// Model<T> (mcBase.h:157-223)
template <class T>
class Model {
    inline static const vector<string> defaultAssetNames = {"spot"};

public:
    virtual const size_t numAssets() const { return 1; }
    virtual const vector<string>& assetNames() const { return defaultAssetNames; }
    virtual unique_ptr<Model<T>> clone() const = 0;
    virtual ~Model() {}
    virtual const vector<T*>& parameters() = 0;
    virtual const vector<string>& parameterLabels() const = 0;
    size_t numParams() const { return const_cast<Model*>(this)->parameters().size(); }
    // Flattened form for the engine; pointers stay valid while *this lives.
    virtual cf_model descriptor() const = 0;
    // Simulation dimension for a product (Model::simDim after allocate)
    size_t simDim(const Product<T>& prd) const
    {
        const cf_model m = descriptor();
        const cf_product p = prd.descriptor();
        return size_t(cfCheck(cf_sim_timeline(&m, &p, nullptr, nullptr, 0)) - 1);
    }
};

// RNG (mcBase.h:228-246).  The engine draws the numbers on the device; an
// RNG here only selects the sequence, its seeds and the first path (skipTo).
class RNG {
public:
    virtual unique_ptr<RNG> clone() const = 0;
    virtual ~RNG() {}
    virtual void init(const size_t) {}
    virtual void skipTo(const unsigned b) = 0;
    virtual bool isSobol() const = 0;
    virtual unsigned firstPath() const = 0;
    virtual unsigned seed1() const { return 0; }
    virtual unsigned seed2() const { return 0; }
};

// Sobol (sobol.h:30-152): path p of a simulation uses point p + 1 of the
// sequence; skipTo(b) starts the simulation at path b.
class Sobol : public RNG {
    unsigned myFirst = 0;

public:
    unique_ptr<RNG> clone() const override { return std::make_unique<Sobol>(*this); }
    void skipTo(const unsigned b) override
    {
        if (b) myFirst = b;   // skipTo(0) is a no-op in the reference (sobol.h:122)
    }
    bool isSobol() const override { return true; }
    unsigned firstPath() const override { return myFirst; }
};

// mrg32k3a (mrg32k3a.h:25-395) with seeds: path 2k draws block k of the
// sequence and path 2k + 1 its negation (antithetic pairs); skipTo(b) with b
// even starts at path b, as the reference's parallel batches do.
class mrg32k3a : public RNG {
    unsigned myA, myB, myFirst = 0;

public:
    mrg32k3a(const unsigned a = 12345, const unsigned b = 12346) : myA(a), myB(b) {}
    unique_ptr<RNG> clone() const override { return std::make_unique<mrg32k3a>(*this); }
    void skipTo(const unsigned b) override { myFirst = b; }
    bool isSobol() const override { return false; }
    unsigned firstPath() const override { return myFirst; }
    unsigned seed1() const override { return myA; }
    unsigned seed2() const override { return myB; }
};

// Compatibility of model and product (mcBase.h:254-260)
template <class T>
inline bool checkCompatiblity(const Product<T>& prd, const Model<T>& mdl)
{
    return prd.assetNames() == mdl.assetNames();
}

inline cf_num makeNum(const RNG& rng, size_t nPath)
{
    if (nPath == 0 || nPath > size_t(0x7fffffff)) throw std::runtime_error("number of paths out of range");
    cf_num n{};
    n.parallel = 1;
    n.use_sobol = rng.isSobol() ? 1 : 0;
    n.num_path = int(nPath);
    n.seed1 = rng.isSobol() ? 12345 : int(rng.seed1());
    n.seed2 = rng.isSobol() ? 1234 : int(rng.seed2());
    n.precision = engineOptions().precision;
    n.aad_mode = engineOptions().aadMode;
    n.first_path = int(rng.firstPath());
    return n;
}

// mcParallelSimul (mcBase.h:314-400): matrix (0..nPath-1, 0..nPay-1) of payoffs.
inline vector<vector<double>> mcParallelSimul(const Product<double>& prd, const Model<double>& mdl, const RNG& rng,
                                              const size_t nPath)
{
    if (!checkCompatiblity(prd, mdl)) throw std::runtime_error("Model and product are not compatible");
    const cf_model m = mdl.descriptor();
    const cf_product p = prd.descriptor();
    const cf_num num = makeNum(rng, nPath);
    const size_t nPay = prd.payoffLabels().size();
    vector<double> flat(nPath * nPay), means(nPay);
    cfCheck(cf_mc_simul(&m, &p, &num, means.data(), nullptr, flat.data()));
    vector<vector<double>> results(nPath, vector<double>(nPay));
    for (size_t i = 0; i < nPath; ++i) std::copy_n(flat.begin() + i * nPay, nPay, results[i].begin());
    return results;
}

// mcSimul (mcBase.h:267-308): the serial form yields the same paths.
inline vector<vector<double>> mcSimul(const Product<double>& prd, const Model<double>& mdl, const RNG& rng,
                                      const size_t nPath)
{
    return mcParallelSimul(prd, mdl, rng, nPath);
}

// AADSimulResults (mcBase.h:405-422)
struct AADSimulResults {
    AADSimulResults(const size_t nPath, const size_t nPay, const size_t nParam)
        : payoffs(nPath, vector<double>(nPay)), aggregated(nPath), risks(nParam)
    {
    }
    vector<vector<double>> payoffs;
    vector<double> aggregated;
    vector<double> risks;
};

// Default aggregator = first payoff (mcBase.h:425)
const auto defaultAggregator = [](const vector<Number>& v) { return v[0]; };

// The engine differentiates aggregates linear in the payoffs, which covers the
// reference's aggregators (main.h:135, :210-213).  The weights are read by
// evaluating the aggregator on unit vectors; linearity is checked on a dense
// combination and a non-linear aggregator is rejected.
template <class F>
inline vector<double> linearWeights(const F& aggFun, size_t nPay)
{
    auto eval = [&](const vector<double>& x) {
        vector<Number> v(x.begin(), x.end());
        return toDouble(aggFun(v));
    };
    vector<double> e(nPay, 0.0), w(nPay);
    const double c = eval(e);
    for (size_t k = 0; k < nPay; ++k) {
        e[k] = 1.0;
        w[k] = eval(e) - c;
        e[k] = 0.0;
    }
    vector<double> probe(nPay);
    double lin = c, scale = std::fabs(c);
    for (size_t k = 0; k < nPay; ++k) {
        probe[k] = 0.5 + 0.37 * double(k % 7) - 0.11 * double(k % 3);
        lin += w[k] * probe[k];
        scale += std::fabs(w[k] * probe[k]);
    }
    if (c != 0.0 || std::fabs(eval(probe) - lin) > 1e-12 * (1.0 + scale))
        throw std::runtime_error("mcParallelSimulAAD: the aggregator must be linear in the payoffs");
    return w;
}

// mcParallelSimulAAD (mcBase.h:564-753)
template <class F = decltype(defaultAggregator)>
inline AADSimulResults mcParallelSimulAAD(const Product<Number>& prd, const Model<Number>& mdl, const RNG& rng,
                                          const size_t nPath, const F& aggFun = defaultAggregator)
{
    if (!checkCompatiblity(prd, mdl)) throw std::runtime_error("Model and product are not compatible");
    const size_t nPay = prd.payoffLabels().size();
    const size_t nParam = mdl.numParams();
    const vector<double> w = linearWeights(aggFun, nPay);
    const cf_model m = mdl.descriptor();
    const cf_product p = prd.descriptor();
    const cf_num num = makeNum(rng, nPath);
    AADSimulResults results(nPath, nPay, nParam);
    vector<double> flat(nPath * nPay), means(nPay);
    double agg = 0.0;
    cfCheck(cf_mc_simul_aad(&m, &p, &num, w.data(), means.data(), &agg, results.risks.data(), flat.data(),
                            results.aggregated.data()));
    for (size_t i = 0; i < nPath; ++i) std::copy_n(flat.begin() + i * nPay, nPay, results.payoffs[i].begin());
    // the reference leaves the summed adjoints on the model parameters (mcBase.h:743)
    auto& params = const_cast<Model<Number>&>(mdl).parameters();
    for (size_t j = 0; j < nParam; ++j) params[j]->adjoint() = results.risks[j] * double(nPath);
    return results;
}

template <class F = decltype(defaultAggregator)>
inline AADSimulResults mcSimulAAD(const Product<Number>& prd, const Model<Number>& mdl, const RNG& rng,
                                  const size_t nPath, const F& aggFun = defaultAggregator)
{
    return mcParallelSimulAAD(prd, mdl, rng, nPath, aggFun);
}

}  // namespace cfamd
