// Store-level C ABI (include/cfamd.h): the id-keyed entry points the
// reference exposes to its front end (xlExport.cpp:84-386, :546-772), built on
// the C++ mirror of store.h and main.h in include/compfinance_amd/.
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/cfamd.h"
#include "../../include/compfinance_amd/main.h"
#include "cf_internal.h"

namespace {

// Status code of a failed store call from the reference's error messages.
int fail_code(const std::string& msg)
{
    if (msg.find("not retrieve") != std::string::npos || msg.find("not found") != std::string::npos)
        return CF_ERR_NOT_FOUND;
    if (msg.find("only the Sobol") != std::string::npos || msg.find("not supported") != std::string::npos)
        return CF_ERR_UNSUPPORTED;
    return CF_ERR_ARG;
}

cfamd::NumericalParam to_num(const cf_num* n)
{
    if (!n) throw std::runtime_error("null numerical parameters");
    cfamd::NumericalParam num;
    num.parallel = n->parallel != 0;
    num.useSobol = n->use_sobol != 0;
    num.numPath = n->num_path;
    num.seed1 = n->seed1;
    num.seed2 = n->seed2;
    num.precision = n->precision;
    num.aadMode = n->aad_mode;
    if (n->first_path != 0) throw std::runtime_error("the store entry points always start at path 0");
    return num;
}

std::string id_of(const char* id)
{
    if (!id || !*id) throw std::runtime_error("empty store id");
    return std::string(id);
}

std::map<std::string, double> notionals_of(const char* const* labels, const double* notionals, int n)
{
    if (n < 0 || (n > 0 && (!labels || !notionals))) throw std::runtime_error("bad notionals");
    std::map<std::string, double> out;
    for (int i = 0; i < n; ++i) {
        if (!labels[i]) throw std::runtime_error("null payoff label");
        out[labels[i]] = notionals[i];
    }
    return out;
}

template <class F>
int store_call(F&& f)
{
    try {
        return f();
    } catch (const std::exception& e) {
        cfk::set_last_error(e.what());
        return fail_code(e.what());
    }
}

void copy_out(const std::vector<double>& v, double* out, int cap)
{
    if (out && cap > 0) std::memcpy(out, v.data(), sizeof(double) * std::min<size_t>(v.size(), size_t(cap)));
}

}  // namespace

extern "C" {

int cf_put_black_scholes(double spot, double vol, int q_spot, double rate, double div, const char* id)
{
    return store_call([&] {
        cfamd::putBlackScholes(spot, vol, q_spot != 0, rate, div, id_of(id));
        return CF_OK;
    });
}

int cf_put_dupire(double spot, const double* spots, int n_spots, const double* times, int n_times,
                  const double* vols, double max_dt, const char* id)
{
    return store_call([&] {
        if (n_spots < 1 || n_times < 1 || !spots || !times || !vols) throw std::runtime_error("empty local volatility grid");
        std::vector<double> s(spots, spots + n_spots), t(times, times + n_times);
        cfamd::matrix<double> v{size_t(n_spots), size_t(n_times)};
        std::copy(vols, vols + size_t(n_spots) * n_times, v.begin());
        // validate through the engine before storing
        cfamd::Dupire<double> probe(spot, s, t, v, max_dt);
        const cf_model d = probe.descriptor();
        cfamd::cfCheck(cf_num_params(&d));
        cfamd::putDupire(spot, s, t, v, max_dt, id_of(id));
        return CF_OK;
    });
}

int cf_put_european(double strike, double exercise_date, double settlement_date, const char* id)
{
    return store_call([&] {
        cfamd::putEuropean(strike, exercise_date, settlement_date > 0 ? settlement_date : exercise_date, id_of(id));
        return CF_OK;
    });
}

int cf_put_barrier(double strike, double barrier, double maturity, double monitor_freq, double smooth, int call_put,
                   const char* id)
{
    return store_call([&] {
        cfamd::putBarrier(strike, barrier, maturity, monitor_freq, smooth, call_put != 0, id_of(id));
        return CF_OK;
    });
}

int cf_put_europeans(const double* maturities, const double* strikes, int n, const char* id)
{
    return store_call([&] {
        if (n < 1 || !maturities || !strikes) throw std::runtime_error("no options");
        cfamd::putEuropeans(std::vector<double>(maturities, maturities + n), std::vector<double>(strikes, strikes + n),
                            id_of(id));
        return CF_OK;
    });
}

int cf_put_asian(double strike, double maturity, double fixing_freq, double snap_freq, int call_put, const char* id)
{
    return store_call([&] {
        cfamd::putAsian(strike, maturity, fixing_freq, snap_freq, call_put != 0, id_of(id));
        return CF_OK;
    });
}

int cf_put_book(const char* const* part_ids, int n, const char* id)
{
    return store_call([&] {
        if (n < 1 || !part_ids) throw std::runtime_error("empty book");
        std::vector<std::string> ids;
        for (int i = 0; i < n; ++i) ids.push_back(id_of(part_ids[i]));
        cfamd::putBook(ids, id_of(id));
        return CF_OK;
    });
}

int cf_value(const char* model_id, const char* product_id, const cf_num* num, double* values, int cap)
{
    return store_call([&] {
        auto r = cfamd::value(id_of(model_id), id_of(product_id), to_num(num));
        copy_out(r.values, values, cap);
        return int(r.values.size());
    });
}

int cf_aad_risk_one(const char* model_id, const char* product_id, const cf_num* num, const char* risk_payoff,
                    double* payoff_values, int cap_pay, double* risk_value, double* risks, int cap_risk)
{
    return store_call([&] {
        auto r = cfamd::AADriskOne(id_of(model_id), id_of(product_id), to_num(num), risk_payoff ? risk_payoff : "");
        copy_out(r.payoffValues, payoff_values, cap_pay);
        if (risk_value) *risk_value = r.riskPayoffValue;
        copy_out(r.risks, risks, cap_risk);
        return int(r.risks.size());
    });
}

int cf_aad_risk_multi(const char* model_id, const char* product_id, const cf_num* num, double* payoff_values,
                      int cap_pay, double* risks, int cap_risk)
{
    return store_call([&] {
        auto r = cfamd::AADriskMulti(id_of(model_id), id_of(product_id), to_num(num));
        copy_out(r.values, payoff_values, cap_pay);
        copy_out(std::vector<double>(r.risks.begin(), r.risks.end()), risks, cap_risk);
        return int(r.params.size());
    });
}

int cf_aad_risk_aggregate(const char* model_id, const char* product_id, const char* const* labels,
                          const double* notionals, int n, const cf_num* num, double* payoff_values, int cap_pay,
                          double* risk_value, double* risks, int cap_risk)
{
    return store_call([&] {
        auto r = cfamd::AADriskAggregate(id_of(model_id), id_of(product_id), notionals_of(labels, notionals, n),
                                         to_num(num));
        copy_out(r.payoffValues, payoff_values, cap_pay);
        if (risk_value) *risk_value = r.riskPayoffValue;
        copy_out(r.risks, risks, cap_risk);
        return int(r.risks.size());
    });
}

int cf_dupire_aad_risk(const char* model_id, const char* product_id, const char* const* labels,
                       const double* notionals, int n, const cf_num* num, double* value, double* delta, double* vega,
                       int cap_vega)
{
    return store_call([&] {
        auto r = cfamd::dupireAADRisk(id_of(model_id), id_of(product_id), notionals_of(labels, notionals, n),
                                      to_num(num));
        if (value) *value = r.value;
        if (delta) *delta = r.delta;
        std::vector<double> v(r.vega.begin(), r.vega.end());
        copy_out(v, vega, cap_vega);
        return int(v.size());
    });
}

void cf_store_clear(void) { cfamd::clearStores(); }

int cf_dupire_superbucket_bump(const cf_merton_ivs* ivs, double max_dt, const char* product_id,
                               const char* const* labels, const double* notionals, int n,
                               const double* incl_spots, int n_incl_spots, double max_ds,
                               const double* incl_times, int n_incl_times, double max_dt_vol,
                               const double* strikes, int n_strikes, const double* mats, int n_mats,
                               const cf_num* num, double* value, double* delta, double* vega)
{
    return store_call([&] {
        if (!ivs || !value || !delta || !vega) throw std::runtime_error("null argument");
        if (n_incl_spots < 1 || n_incl_times < 1 || n_strikes < 1 || n_mats < 1 || !incl_spots || !incl_times ||
            !strikes || !mats)
            throw std::runtime_error("empty grid input");
        auto r = cfamd::dupireSuperbucketBump(
            ivs->spot, max_dt, id_of(product_id), notionals_of(labels, notionals, n),
            std::vector<double>(incl_spots, incl_spots + n_incl_spots), max_ds,
            std::vector<double>(incl_times, incl_times + n_incl_times), max_dt_vol,
            std::vector<double>(strikes, strikes + n_strikes), std::vector<double>(mats, mats + n_mats), ivs->vol,
            ivs->jmp_intens, ivs->jmp_average, ivs->jmp_std, to_num(num));
        *value = r.value;
        *delta = r.delta;
        std::copy(r.vega.begin(), r.vega.end(), vega);
        return CF_OK;
    });
}

}  // extern "C"
