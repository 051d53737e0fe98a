// C++17 mirror of the reference products on the MI355X engine: European<T>
// (mcPrd.h:29-126), UOC<T> (mcPrd.h:128-289) and Europeans<T>
// (mcPrd.h:291-400).  Constructors, timeline(), defline() and payoffLabels()
// keep the reference's signatures and results; the payoffs are evaluated on
// the device, fused into the path kernel.
#pragma once

#include <map>

#include "mcBase.h"

namespace cfamd {

template <class T>
class PrdBase : public Product<T> {
protected:
    vector<Time> myTimeline;
    vector<SampleDef> myDefline;
    mutable vector<string> myLabels;

    void loadTimeline()
    {
        const cf_product d = this->descriptor();
        const int n = cfCheck(cf_product_timeline(&d, nullptr, 0));
        myTimeline.resize(n);
        cf_product_timeline(&d, myTimeline.data(), n);
    }

public:
    const vector<Time>& timeline() const override { return myTimeline; }
    const vector<SampleDef>& defline() const override { return myDefline; }
    const vector<string>& payoffLabels() const override
    {
        if (myLabels.empty()) {
            const cf_product d = this->descriptor();
            const int n = cfCheck(cf_num_payoffs(&d));
            char buf[512];
            for (int k = 0; k < n; ++k) {
                cfCheck(cf_payoff_label(&d, k, buf, sizeof(buf)));
                myLabels.emplace_back(buf);
            }
        }
        return myLabels;
    }
};

// European call (mcPrd.h:44-86)
template <class T>
class European : public PrdBase<T> {
    double myStrike;
    Time myExerciseDate, mySettlementDate;

public:
    European(const double strike, const Time exerciseDate, const Time settlementDate)
        : myStrike(strike), myExerciseDate(exerciseDate), mySettlementDate(settlementDate)
    {
        this->loadTimeline();
        SampleDef s;
        s.numeraire = true;
        s.forwardMats.push_back({settlementDate});
        s.discountMats.push_back(settlementDate);
        this->myDefline = {s};
    }
    European(const double strike, const Time exerciseDate) : European(strike, exerciseDate, exerciseDate) {}
    unique_ptr<Product<T>> clone() const override { return std::make_unique<European<T>>(*this); }
    cf_product descriptor() const override
    {
        cf_product d{};
        d.kind = CF_PRODUCT_EUROPEAN;
        d.strike = myStrike;
        d.exercise = myExerciseDate;
        d.settlement = mySettlementDate;
        return d;
    }
};

// Up-and-out call or put with smoothed barrier (mcPrd.h:149-208)
template <class T>
class UOC : public PrdBase<T> {
    bool myCallPut;
    double myStrike, myBarrier;
    Time myMaturity, myMonitorFreq;
    double mySmooth;

public:
    UOC(const double strike, const double barrier, const Time maturity, const Time monitorFreq, const double smooth,
        const bool callPut = false)
        : myCallPut(callPut), myStrike(strike), myBarrier(barrier), myMaturity(maturity),
          myMonitorFreq(monitorFreq), mySmooth(smooth)
    {
        this->loadTimeline();
        for (size_t i = 0; i < this->myTimeline.size(); ++i) {
            SampleDef s;
            s.numeraire = false;
            s.forwardMats.push_back({this->myTimeline[i]});
            this->myDefline.push_back(s);
        }
        this->myDefline.back().numeraire = true;
    }
    unique_ptr<Product<T>> clone() const override { return std::make_unique<UOC<T>>(*this); }
    cf_product descriptor() const override
    {
        cf_product d{};
        d.kind = CF_PRODUCT_BARRIER;
        d.strike = myStrike;
        d.barrier = myBarrier;
        d.maturity = myMaturity;
        d.monitor_freq = myMonitorFreq;
        d.smooth = mySmooth;
        d.call_put = myCallPut ? 1 : 0;
        return d;
    }
};

// Calls on several maturities and strikes, maturity major (mcPrd.h:305-336)
template <class T>
class Europeans : public PrdBase<T> {
    vector<Time> myMaturities;
    vector<vector<double>> myStrikes;
    vector<double> myFlatMats, myFlatStrikes;

public:
    explicit Europeans(const std::map<Time, vector<double>>& options)
    {
        for (const auto& p : options) {
            myMaturities.push_back(p.first);
            myStrikes.push_back(p.second);
            for (double k : p.second) {
                myFlatMats.push_back(p.first);
                myFlatStrikes.push_back(k);
            }
            SampleDef s;
            s.numeraire = true;
            s.forwardMats.push_back({p.first});
            this->myDefline.push_back(s);
        }
        this->loadTimeline();
    }
    const vector<Time>& maturities() const { return myMaturities; }
    const vector<vector<double>>& strikes() const { return myStrikes; }
    unique_ptr<Product<T>> clone() const override { return std::make_unique<Europeans<T>>(*this); }
    cf_product descriptor() const override
    {
        cf_product d{};
        d.kind = CF_PRODUCT_EUROPEANS;
        d.n_options = int(myFlatMats.size());
        d.maturities = myFlatMats.data();
        d.strikes = myFlatStrikes.data();
        return d;
    }
};

}  // namespace cfamd
