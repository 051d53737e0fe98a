// C++17 mirror of the reference entry points (main.h:33-411) on the MI355X
// engine: NumericalParam, value, AADriskOne, AADriskAggregate, bumpRisk and
// dupireAADRisk, with the reference's signatures, result layouts and error
// messages.  NumericalParam carries two more fields, the precision and the
// AAD variant, defaulting to the reference's FP64 results.
//
// A reference caller switches by including this header instead of main.h;
// `using namespace cfamd;` brings the names into scope as the reference's
// global definitions were.
#pragma once

#include <map>
#include <numeric>

#include "store.h"

namespace cfamd {

// main.h:33-40 plus engine options
struct NumericalParam {
    bool parallel = true;
    bool useSobol = true;
    int numPath = 0;
    int seed1 = 12345;
    int seed2 = 1234;
    int precision = CF_FP64;
    int aadMode = CF_AAD_CHECKPOINT;
};

// main.h:51-52: Sobol, or mrg32k3a with the seeds
inline unique_ptr<RNG> makeRng(const NumericalParam& num)
{
    if (!num.useSobol) return std::make_unique<mrg32k3a>(unsigned(num.seed1), unsigned(num.seed2));
    return std::make_unique<Sobol>();
}

struct OptionsScope {
    EngineOptions saved;
    explicit OptionsScope(const NumericalParam& num) : saved(engineOptions())
    {
        engineOptions().precision = num.precision;
        engineOptions().aadMode = num.aadMode;
    }
    ~OptionsScope() { engineOptions() = saved; }
};

struct ValueResults {
    vector<string> identifiers;
    vector<double> values;
};

// main.h:43-77: mean payoffs, reduced on the device (the per-path matrix is
// not transferred)
inline ValueResults value(const Model<double>& model, const Product<double>& product, const NumericalParam& num)
{
    if (!checkCompatiblity(product, model)) throw std::runtime_error("Model and product are not compatible");
    OptionsScope scope(num);
    auto rng = makeRng(num);
    const cf_model m = model.descriptor();
    const cf_product p = product.descriptor();
    const cf_num n = makeNum(*rng, size_t(num.numPath));
    ValueResults results;
    results.identifiers = product.payoffLabels();
    results.values.resize(results.identifiers.size());
    cfCheck(cf_mc_simul(&m, &p, &n, results.values.data(), nullptr, nullptr));
    return results;
}

// main.h:80-96
inline ValueResults value(const string& modelId, const string& productId, const NumericalParam& num)
{
    const Model<double>* model = getModel<double>(modelId);
    const Product<double>* product = getProduct<double>(productId);
    if (!model || !product) throw std::runtime_error("value() : Could not retrieve model and product");
    return value(*model, *product, num);
}

struct RiskResults {
    vector<string> payoffIds;
    vector<double> payoffValues;
    double riskPayoffValue = 0.0;
    vector<string> paramIds;
    vector<double> risks;
};

inline RiskResults aadRisk(const Model<Number>& model, const Product<Number>& product, const NumericalParam& num,
                           const vector<double>& weights)
{
    if (!checkCompatiblity(product, model)) throw std::runtime_error("Model and product are not compatible");
    OptionsScope scope(num);
    auto rng = makeRng(num);
    const cf_model m = model.descriptor();
    const cf_product p = product.descriptor();
    const cf_num n = makeNum(*rng, size_t(num.numPath));
    RiskResults results;
    results.payoffIds = product.payoffLabels();
    results.payoffValues.resize(results.payoffIds.size());
    results.paramIds = model.parameterLabels();
    results.risks.resize(results.paramIds.size());
    cfCheck(cf_mc_simul_aad(&m, &p, &n, weights.data(), results.payoffValues.data(), &results.riskPayoffValue,
                            results.risks.data(), nullptr, nullptr));
    return results;
}

// main.h:99-173
inline RiskResults AADriskOne(const string& modelId, const string& productId, const NumericalParam& num,
                              const string& riskPayoff = "")
{
    const Model<Number>* model = getModel<Number>(modelId);
    const Product<Number>* product = getProduct<Number>(productId);
    if (!model || !product) throw std::runtime_error("AADrisk() : Could not retrieve model and product");
    const vector<string>& allPayoffs = product->payoffLabels();
    size_t idx = 0;
    if (!riskPayoff.empty()) {
        auto it = std::find(allPayoffs.begin(), allPayoffs.end(), riskPayoff);
        if (it == allPayoffs.end()) throw std::runtime_error("AADriskOne() : payoff not found");
        idx = size_t(std::distance(allPayoffs.begin(), it));
    }
    vector<double> w(allPayoffs.size(), 0.0);
    w[idx] = 1.0;
    return aadRisk(*model, *product, num, w);
}

// main.h:176-254
inline RiskResults AADriskAggregate(const string& modelId, const string& productId,
                                    const std::map<string, double>& notionals, const NumericalParam& num)
{
    const Model<Number>* model = getModel<Number>(modelId);
    const Product<Number>* product = getProduct<Number>(productId);
    if (!model || !product) throw std::runtime_error("AADriskAggregate() : Could not retrieve model and product");
    const vector<string>& allPayoffs = product->payoffLabels();
    vector<double> vnots(allPayoffs.size(), 0.0);
    for (const auto& notional : notionals) {
        auto it = std::find(allPayoffs.begin(), allPayoffs.end(), notional.first);
        if (it == allPayoffs.end()) throw std::runtime_error("AADriskAggregate() : payoff not found");
        vnots[size_t(std::distance(allPayoffs.begin(), it))] = notional.second;
    }
    return aadRisk(*model, *product, num, vnots);
}

// main.h:260-266
struct RiskReports {
    vector<string> payoffs;
    vector<string> params;
    vector<double> values;
    matrix<double> risks;
};

// main.h:316-359: one-sided finite differences of every payoff to every
// parameter, bumping by 1e-8, on the same Sobol paths
inline RiskReports bumpRisk(const string& modelId, const string& productId, const NumericalParam& num,
                            const double bump = 1.0e-08)
{
    auto* orig = getModel<double>(modelId);
    const Product<double>* product = getProduct<double>(productId);
    if (!orig || !product) throw std::runtime_error("bumpRisk() : Could not retrieve model and product");
    RiskReports results;
    auto baseRes = value(*orig, *product, num);
    results.payoffs = baseRes.identifiers;
    results.values = baseRes.values;
    auto model = orig->clone();
    results.params = model->parameterLabels();
    const vector<double*> parameters = model->parameters();
    const size_t n = parameters.size(), m = results.payoffs.size();
    results.risks.resize(n, m);
    for (size_t i = 0; i < n; ++i) {
        *parameters[i] += bump;
        auto bumpRes = value(*model, *product, num);
        *parameters[i] -= bump;
        // (1.0 / bump is exactly 1e8 for the default bump: the reference's scaling)
        for (size_t j = 0; j < m; ++j) results.risks[i][j] = (bumpRes.values[j] - baseRes.values[j]) * (1.0 / bump);
    }
    return results;
}

// main.h:269-312: itemized AAD risk, payoffs in columns, parameters in rows
inline RiskReports AADriskMulti(const string& modelId, const string& productId, const NumericalParam& num)
{
    const Model<Number>* model = getModel<Number>(modelId);
    const Product<Number>* product = getProduct<Number>(productId);
    if (!model || !product) throw std::runtime_error("AADrisk() : Could not retrieve model and product");
    if (!checkCompatiblity(*product, *model)) throw std::runtime_error("Model and product are not compatible");
    OptionsScope scope(num);
    auto rng = makeRng(num);
    const cf_model m = model->descriptor();
    const cf_product p = product->descriptor();
    const cf_num n = makeNum(*rng, size_t(num.numPath));
    RiskReports results;
    results.payoffs = product->payoffLabels();
    results.params = model->parameterLabels();
    results.values.resize(results.payoffs.size());
    results.risks.resize(results.params.size(), results.payoffs.size());
    cfCheck(cf_mc_simul_aad_multi(&m, &p, &n, results.values.data(), &*results.risks.begin()));
    return results;
}

struct DupireRisk {
    double value = 0.0;
    double delta = 0.0;
    matrix<double> vega;
};

// main.h:364-411
inline DupireRisk dupireAADRisk(const string& modelId, const string& productId,
                                const std::map<string, double>& notionals, const NumericalParam& num)
{
    const Model<Number>* model = getModel<Number>(modelId);
    if (!model) throw std::runtime_error("dupireAADRisk() : Model not found");
    const Dupire<Number>* dupire = dynamic_cast<const Dupire<Number>*>(model);
    if (!dupire) throw std::runtime_error("dupireAADRisk() : Model not a Dupire");
    DupireRisk results;
    auto simulResults = AADriskAggregate(modelId, productId, notionals, num);
    results.value = simulResults.riskPayoffValue;
    results.delta = simulResults.risks[0];
    results.vega.resize(dupire->spots().size(), dupire->times().size());
    std::copy(std::next(simulResults.risks.begin()), simulResults.risks.end(), results.vega.begin());
    return results;
}

// main.h:414-438: local vols on a Merton implied-vol surface (FP64, host)
struct CalibResults {
    vector<double> spots;
    vector<Time> times;
    matrix<double> lVols;
};

inline CalibResults dupireCalib(const vector<double>& inclSpots, const double maxDs, const vector<Time>& inclTimes,
                                const double maxDt, const double spot, const double vol, const double jmpIntens = 0.0,
                                const double jmpAverage = 0.0, const double jmpStd = 0.0)
{
    const cf_merton_ivs ivs{spot, vol, jmpIntens, jmpAverage, jmpStd};
    int nS = 0, nT = 0;
    cfCheck(cf_dupire_calib(&ivs, inclSpots.data(), int(inclSpots.size()), maxDs, inclTimes.data(),
                            int(inclTimes.size()), maxDt, nullptr, nullptr, nullptr, &nS, &nT, 0));
    CalibResults r;
    r.spots.resize(nS);
    r.times.resize(nT);
    r.lVols.resize(nS, nT);
    cfCheck(cf_dupire_calib(&ivs, inclSpots.data(), int(inclSpots.size()), maxDs, inclTimes.data(),
                            int(inclTimes.size()), maxDt, r.spots.data(), r.times.data(), &*r.lVols.begin(), &nS,
                            &nT, nS * nT));
    return r;
}

struct SuperbucketResults {
    double value = 0.0;
    double delta = 0.0;
    vector<double> strikes;
    vector<Time> mats;
    matrix<double> vega;
};

// main.h:453-569: calibration, local-vol risks on the GPU (dupireAADRisk),
// then the chain to the implied-vol risk view
inline SuperbucketResults dupireSuperbucket(const double spot, const double maxDt, const string& productId,
                                            const std::map<string, double>& notionals,
                                            const vector<double>& inclSpots, const double maxDs,
                                            const vector<Time>& inclTimes, const double maxDtVol,
                                            const vector<double>& strikes, const vector<Time>& mats,
                                            const double vol, const double jmpIntens, const double jmpAverage,
                                            const double jmpStd, const NumericalParam& num)
{
    auto cal = dupireCalib(inclSpots, maxDs, inclTimes, maxDtVol, spot, vol, jmpIntens, jmpAverage, jmpStd);
    putDupire(spot, cal.spots, cal.times, cal.lVols, maxDt, "superbucket");
    auto micro = dupireAADRisk("superbucket", productId, notionals, num);
    SuperbucketResults r;
    r.value = micro.value;
    r.delta = micro.delta;
    r.strikes = strikes;
    r.mats = mats;
    r.vega.resize(strikes.size(), mats.size());
    const cf_merton_ivs ivs{spot, vol, jmpIntens, jmpAverage, jmpStd};
    cfCheck(cf_superbucket_vega(&ivs, inclSpots.data(), int(inclSpots.size()), maxDs, inclTimes.data(),
                                int(inclTimes.size()), maxDtVol, strikes.data(), int(strikes.size()), mats.data(),
                                int(mats.size()), &*micro.vega.begin(), int(micro.vega.rows() * micro.vega.cols()),
                                &*r.vega.begin()));
    return r;
}

// main.h:575-696: the superbucket by bump and reprice, the validator of the
// AAD superbucket above: delta by a 1e-8 spot bump; each risk-view spread
// bumped by 1e-5 in turn, the local vols recalibrated (dupireCalib with the
// risk view, mcMdlDupire.h:337-388) and the product repriced on the GPU
inline SuperbucketResults dupireSuperbucketBump(const double spot, const double maxDt, const string& productId,
                                                const std::map<string, double>& notionals,
                                                const vector<double>& inclSpots, const double maxDs,
                                                const vector<Time>& inclTimes, const double maxDtVol,
                                                const vector<double>& strikes, const vector<Time>& mats,
                                                const double vol, const double jmpIntens, const double jmpAverage,
                                                const double jmpStd, const NumericalParam& num)
{
    SuperbucketResults results;
    auto params = dupireCalib(inclSpots, maxDs, inclTimes, maxDtVol, spot, vol, jmpIntens, jmpAverage, jmpStd);
    Dupire<double> model(spot, params.spots, params.times, params.lVols, maxDt);
    const Product<double>* product = getProduct<double>(productId);
    if (!product) throw std::runtime_error("dupireSuperbucketBump() : product not found");
    auto baseVals = value(model, *product, num);
    const vector<string>& allPayoffs = baseVals.identifiers;
    vector<double> vnots(allPayoffs.size(), 0.0);
    for (const auto& notional : notionals) {
        auto it = std::find(allPayoffs.begin(), allPayoffs.end(), notional.first);
        if (it == allPayoffs.end()) throw std::runtime_error("dupireSuperbucketBump() : payoff not found");
        vnots[std::distance(allPayoffs.begin(), it)] = notional.second;
    }
    auto book = [&](const ValueResults& v) {
        return std::inner_product(vnots.begin(), vnots.end(), v.values.begin(), 0.0);
    };
    results.value = book(baseVals);
    // delta
    Dupire<double> bumpedSpot(spot + 1.0e-08, params.spots, params.times, params.lVols, maxDt);
    results.delta = (book(value(bumpedSpot, *product, num)) - results.value) * 1.0e+08;
    // vega: one spread of the risk view at a time
    const cf_merton_ivs ivs{spot, vol, jmpIntens, jmpAverage, jmpStd};
    const size_t n = strikes.size(), m = mats.size();
    results.vega.resize(n, m);
    vector<double> spreads(n * m, 0.0);
    for (size_t i = 0; i < n; ++i)
        for (size_t j = 0; j < m; ++j) {
            spreads[i * m + j] += 1.0e-05;
            int nS = 0, nT = 0;
            cfCheck(cf_dupire_calib_view(&ivs, inclSpots.data(), int(inclSpots.size()), maxDs, inclTimes.data(),
                                         int(inclTimes.size()), maxDtVol, strikes.data(), int(n), mats.data(), int(m),
                                         spreads.data(), nullptr, nullptr, nullptr, &nS, &nT, 0));
            vector<double> bs(nS);
            vector<Time> bt(nT);
            matrix<double> blv(nS, nT);
            cfCheck(cf_dupire_calib_view(&ivs, inclSpots.data(), int(inclSpots.size()), maxDs, inclTimes.data(),
                                         int(inclTimes.size()), maxDtVol, strikes.data(), int(n), mats.data(), int(m),
                                         spreads.data(), bs.data(), bt.data(), &*blv.begin(), &nS, &nT, nS * nT));
            Dupire<double> bumped(spot, bs, bt, blv, maxDt);
            results.vega[i][j] = (book(value(bumped, *product, num)) - results.value) * 1.0e+05;
            spreads[i * m + j] -= 1.0e-05;
        }
    results.strikes = strikes;
    results.mats = mats;
    return results;
}

}  // namespace cfamd
