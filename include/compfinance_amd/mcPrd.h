// C++17 mirror of the reference products on the MI355X engine: European<T>
// (mcPrd.h:29-126), UOC<T> (mcPrd.h:128-289) and Europeans<T>
// (mcPrd.h:291-400).  Constructors, timeline(), defline() and payoffLabels()
// keep the reference's signatures and results; the payoffs are evaluated on
// the device, fused into the path kernel.  Two products without a reference
// class, written to the same Product<T> interface (mcBase.h:122-152): the
// arithmetic-average Asian<T> and Book<T>, several products on shared paths.
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

// Arithmetic-average call or put on the spot: fixings at fixingFreq,
// 2 fixingFreq, ... up to the maturity, each moved to the nearest date of the
// grid 0, snapFreq, ..., maturity when snapFreq > 0; paid at maturity.
template <class T>
class Asian : public PrdBase<T> {
    double myStrike;
    Time myMaturity, myFixingFreq, mySnapFreq;
    bool myCallPut;

public:
    Asian(const double strike, const Time maturity, const Time fixingFreq, const Time snapFreq = 0.0,
          const bool callPut = false)
        : myStrike(strike), myMaturity(maturity), myFixingFreq(fixingFreq), mySnapFreq(snapFreq), myCallPut(callPut)
    {
        this->loadTimeline();
        for (size_t i = 0; i < this->myTimeline.size(); ++i) {
            SampleDef s;
            s.numeraire = i + 1 == this->myTimeline.size();
            s.forwardMats.push_back({this->myTimeline[i]});
            this->myDefline.push_back(s);
        }
    }
    unique_ptr<Product<T>> clone() const override { return std::make_unique<Asian<T>>(*this); }
    cf_product descriptor() const override
    {
        cf_product d{};
        d.kind = CF_PRODUCT_ASIAN;
        d.strike = myStrike;
        d.maturity = myMaturity;
        d.fixing_freq = myFixingFreq;
        d.snap_freq = mySnapFreq;
        d.call_put = myCallPut ? 1 : 0;
        return d;
    }
};

// Several products simulated on the same paths; payoffs concatenated in part
// order and labelled "<part index>:<part label>".  Parts: European settled at
// exercise, UOC, Europeans, Asian; at most 16 barriers and 4 fixing schedules.
template <class T>
class Book : public PrdBase<T> {
    vector<unique_ptr<Product<T>>> myParts;
    mutable vector<cf_product> myPartDesc;

public:
    explicit Book(const vector<const Product<T>*>& parts)
    {
        for (const Product<T>* p : parts) myParts.push_back(p->clone());
        this->loadTimeline();
        for (size_t i = 0; i < this->myTimeline.size(); ++i) {
            SampleDef s;
            s.numeraire = true;
            s.forwardMats.push_back({this->myTimeline[i]});
            this->myDefline.push_back(s);
        }
    }
    Book(const Book& o) : PrdBase<T>(o)
    {
        for (const auto& p : o.myParts) myParts.push_back(p->clone());
    }
    unique_ptr<Product<T>> clone() const override { return std::make_unique<Book<T>>(*this); }
    cf_product descriptor() const override
    {
        myPartDesc.clear();
        for (const auto& p : myParts) myPartDesc.push_back(p->descriptor());
        cf_product d{};
        d.kind = CF_PRODUCT_BOOK;
        d.n_parts = int(myPartDesc.size());
        d.parts = myPartDesc.data();
        return d;
    }
};

}  // namespace cfamd
