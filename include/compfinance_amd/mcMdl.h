// C++17 mirror of the reference models on the MI355X engine:
// matrix<T> (matrix.h:9-105), BlackScholes<T> (mcMdlBS.h:21-351) and
// Dupire<T> (mcMdlDupire.h:28-281).  Constructors, parameters() and
// parameterLabels() keep the reference's signatures and parameter order;
// the simulation itself runs on the device from descriptor().
#pragma once

#include <sstream>

#include "mcBase.h"

namespace cfamd {

// Row-major matrix (matrix.h:9-105)
template <class T>
class matrix {
    size_t myRows = 0, myCols = 0;
    vector<T> myVector;

public:
    matrix() = default;
    matrix(const size_t rows, const size_t cols) : myRows(rows), myCols(cols), myVector(rows * cols) {}
    template <class U>
    matrix(const matrix<U>& rhs) : myRows(rhs.rows()), myCols(rhs.cols()), myVector(rhs.begin(), rhs.end())
    {
    }
    void resize(const size_t rows, const size_t cols)
    {
        myRows = rows;
        myCols = cols;
        if (myVector.size() < rows * cols) myVector = vector<T>(rows * cols);
    }
    size_t rows() const { return myRows; }
    size_t cols() const { return myCols; }
    T* operator[](const size_t row) { return &myVector[row * myCols]; }
    const T* operator[](const size_t row) const { return &myVector[row * myCols]; }
    bool empty() const { return myVector.empty(); }
    typename vector<T>::iterator begin() { return myVector.begin(); }
    typename vector<T>::iterator end() { return myVector.begin() + myRows * myCols; }
    typename vector<T>::const_iterator begin() const { return myVector.begin(); }
    typename vector<T>::const_iterator end() const { return myVector.begin() + myRows * myCols; }
};

template <class T>
inline const vector<string>& labelsOf(const cf_model& d, vector<string>& cache)
{
    if (cache.empty()) {
        const int n = cfCheck(cf_num_params(&d));
        char buf[256];
        for (int k = 0; k < n; ++k) {
            cfCheck(cf_param_label(&d, k, buf, sizeof(buf)));
            cache.emplace_back(buf);
        }
    }
    return cache;
}

// BlackScholes<T> (mcMdlBS.h:21-351); parameters [spot, vol, rate, div]
template <class T>
class BlackScholes : public Model<T> {
    T mySpot, myRate, myDiv, myVol;
    const bool mySpotMeasure;
    vector<T*> myParameters;
    mutable vector<string> myLabels;

    void setParamPointers() { myParameters = {&mySpot, &myVol, &myRate, &myDiv}; }

public:
    template <class U>
    BlackScholes(const U spot, const U vol, const bool spotMeasure = false, const U rate = U(0.0),
                 const U div = U(0.0))
        : mySpot(spot), myRate(rate), myDiv(div), myVol(vol), mySpotMeasure(spotMeasure)
    {
        setParamPointers();
    }
    BlackScholes(const BlackScholes& rhs)
        : mySpot(rhs.mySpot), myRate(rhs.myRate), myDiv(rhs.myDiv), myVol(rhs.myVol),
          mySpotMeasure(rhs.mySpotMeasure), myLabels(rhs.myLabels)
    {
        setParamPointers();
    }
    T spot() const { return mySpot; }
    const T vol() const { return myVol; }
    const T rate() const { return myRate; }
    const T div() const { return myDiv; }
    const vector<T*>& parameters() override { return myParameters; }
    const vector<string>& parameterLabels() const override { return labelsOf<T>(descriptor(), myLabels); }
    unique_ptr<Model<T>> clone() const override { return std::make_unique<BlackScholes<T>>(*this); }
    cf_model descriptor() const override
    {
        cf_model d{};
        d.kind = CF_MODEL_BLACK_SCHOLES;
        d.spot = toDouble(mySpot);
        d.vol = toDouble(myVol);
        d.rate = toDouble(myRate);
        d.div = toDouble(myDiv);
        d.spot_measure = mySpotMeasure ? 1 : 0;
        return d;
    }
};

// Dupire<T> (mcMdlDupire.h:28-281); parameters [spot, vols spot major]
template <class T>
class Dupire : public Model<T> {
    T mySpot;
    const vector<double> mySpots;
    const vector<Time> myTimes;
    matrix<T> myVols;
    const Time myMaxDt;
    vector<T*> myParameters;
    mutable vector<string> myLabels;
    mutable vector<double> myVolValues;

    void setParamPointers()
    {
        myParameters.assign(1 + myVols.rows() * myVols.cols(), nullptr);
        myParameters[0] = &mySpot;
        std::transform(myVols.begin(), myVols.end(), std::next(myParameters.begin()), [](T& v) { return &v; });
    }

public:
    template <class U>
    Dupire(const U spot, const vector<double> spots, const vector<Time> times, const matrix<U> vols,
           const Time maxDt = 0.25)
        : mySpot(spot), mySpots(spots), myTimes(times), myVols(vols), myMaxDt(maxDt)
    {
        setParamPointers();
    }
    Dupire(const Dupire& rhs)
        : mySpot(rhs.mySpot), mySpots(rhs.mySpots), myTimes(rhs.myTimes), myVols(rhs.myVols),
          myMaxDt(rhs.myMaxDt), myLabels(rhs.myLabels)
    {
        setParamPointers();
    }
    T spot() const { return mySpot; }
    const vector<double>& spots() const { return mySpots; }
    const vector<Time>& times() const { return myTimes; }
    const matrix<T>& vols() const { return myVols; }
    const vector<T*>& parameters() override { return myParameters; }
    const vector<string>& parameterLabels() const override { return labelsOf<T>(descriptor(), myLabels); }
    unique_ptr<Model<T>> clone() const override { return std::make_unique<Dupire<T>>(*this); }
    cf_model descriptor() const override
    {
        myVolValues.resize(myVols.rows() * myVols.cols());
        std::transform(myVols.begin(), myVols.end(), myVolValues.begin(), [](const T& v) { return toDouble(v); });
        cf_model d{};
        d.kind = CF_MODEL_DUPIRE;
        d.spot = toDouble(mySpot);
        d.n_spots = int(mySpots.size());
        d.n_times = int(myTimes.size());
        d.spots = mySpots.data();
        d.times = myTimes.data();
        d.vols = myVolValues.data();
        d.max_dt = myMaxDt;
        return d;
    }
};

}  // namespace cfamd
