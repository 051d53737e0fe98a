// C++17 mirror of the reference model/product store (store.h:30-282):
// id-keyed registries holding a <double, Number> pair of each model and
// product, the put functions of the configurations' models and products, and
// getModel<T> / getProduct<T>.  Unlike the reference, the registries are
// guarded by a mutex so that several host threads can use them.
#pragma once

#include <mutex>
#include <unordered_map>
#include <utility>

#include "mcMdl.h"
#include "mcPrd.h"

namespace cfamd {

using ModelStore = std::unordered_map<string, std::pair<unique_ptr<Model<double>>, unique_ptr<Model<Number>>>>;
using ProductStore =
    std::unordered_map<string, std::pair<unique_ptr<Product<double>>, unique_ptr<Product<Number>>>>;

inline ModelStore modelStore;
inline ProductStore productStore;
inline std::recursive_mutex storeMutex;

inline void putBlackScholes(const double spot, const double vol, const bool qSpot, const double rate,
                            const double div, const string& store)
{
    auto mdl = std::make_unique<BlackScholes<double>>(spot, vol, qSpot, rate, div);
    auto riskMdl = std::make_unique<BlackScholes<Number>>(spot, vol, qSpot, rate, div);
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    modelStore[store] = std::make_pair(std::move(mdl), std::move(riskMdl));
}

inline void putDupire(const double spot, const vector<double>& spots, const vector<Time>& times,
                      const matrix<double>& vols, const double maxDt, const string& store)
{
    auto mdl = std::make_unique<Dupire<double>>(spot, spots, times, vols, maxDt);
    auto riskMdl = std::make_unique<Dupire<Number>>(spot, spots, times, vols, maxDt);
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    modelStore[store] = std::make_pair(std::move(mdl), std::move(riskMdl));
}

inline void putEuropean(const double strike, const Time exerciseDate, const Time settlementDate,
                        const string& store)
{
    auto prd = std::make_unique<European<double>>(strike, exerciseDate, settlementDate);
    auto riskPrd = std::make_unique<European<Number>>(strike, exerciseDate, settlementDate);
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    productStore[store] = std::make_pair(std::move(prd), std::move(riskPrd));
}

// store.h:144-163: non-positive smoothing becomes 1e-8
inline void putBarrier(const double strike, const double barrier, const Time maturity, const double monitorFreq,
                       const double smooth, const bool callPut, const string& store)
{
    const double smoothFactor = smooth <= 0 ? 1.0e-08 : smooth;
    auto prd = std::make_unique<UOC<double>>(strike, barrier, maturity, monitorFreq, smoothFactor, callPut);
    auto riskPrd = std::make_unique<UOC<Number>>(strike, barrier, maturity, monitorFreq, smoothFactor, callPut);
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    productStore[store] = std::make_pair(std::move(prd), std::move(riskPrd));
}

// store.h:184-205: maturities in increasing order, one strike per entry
inline void putEuropeans(const vector<Time>& maturities, const vector<double>& strikes, const string& store)
{
    std::map<Time, vector<double>> options;
    for (size_t i = 0; i < maturities.size(); ++i) options[maturities[i]].push_back(strikes[i]);
    auto prd = std::make_unique<Europeans<double>>(options);
    auto riskPrd = std::make_unique<Europeans<Number>>(options);
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    productStore[store] = std::make_pair(std::move(prd), std::move(riskPrd));
}

// Arithmetic-average option (no reference counterpart; see Asian<T>)
inline void putAsian(const double strike, const Time maturity, const Time fixingFreq, const Time snapFreq,
                     const bool callPut, const string& store)
{
    auto prd = std::make_unique<Asian<double>>(strike, maturity, fixingFreq, snapFreq, callPut);
    auto riskPrd = std::make_unique<Asian<Number>>(strike, maturity, fixingFreq, snapFreq, callPut);
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    productStore[store] = std::make_pair(std::move(prd), std::move(riskPrd));
}

// Book of stored products on shared paths (no reference counterpart; see
// Book<T>); throws if a part is not in the store.
inline void putBook(const vector<string>& parts, const string& store)
{
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    vector<const Product<double>*> pd;
    vector<const Product<Number>*> pn;
    for (const string& id : parts) {
        auto it = productStore.find(id);
        if (it == productStore.end()) throw std::runtime_error("Could not retrieve product " + id);
        pd.push_back(it->second.first.get());
        pn.push_back(it->second.second.get());
    }
    auto prd = std::make_unique<Book<double>>(pd);
    auto riskPrd = std::make_unique<Book<Number>>(pn);
    productStore[store] = std::make_pair(std::move(prd), std::move(riskPrd));
}

template <class T>
const Model<T>* getModel(const string& store);

template <>
inline const Model<double>* getModel(const string& store)
{
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    auto it = modelStore.find(store);
    return it == modelStore.end() ? nullptr : it->second.first.get();
}

template <>
inline const Model<Number>* getModel(const string& store)
{
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    auto it = modelStore.find(store);
    return it == modelStore.end() ? nullptr : it->second.second.get();
}

template <class T>
const Product<T>* getProduct(const string& store);

template <>
inline const Product<double>* getProduct(const string& store)
{
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    auto it = productStore.find(store);
    return it == productStore.end() ? nullptr : it->second.first.get();
}

template <>
inline const Product<Number>* getProduct(const string& store)
{
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    auto it = productStore.find(store);
    return it == productStore.end() ? nullptr : it->second.second.get();
}

inline const vector<string>* getPayoffLabels(const string& store)
{
    const Product<double>* p = getProduct<double>(store);
    return p ? &p->payoffLabels() : nullptr;
}

inline void clearStores()
{
    std::lock_guard<std::recursive_mutex> lk(storeMutex);
    modelStore.clear();
    productStore.clear();
}

}  // namespace cfamd
