// Local-vol surface construction from an implied-vol surface (FP64, host)
// and the superbucket: risks to the implied-vol surface from the engine's
// local-vol risks.
//
// Reference semantics followed:
//   normal density / cdf   gaussians.h:13-40 (Zelen and Severo's cdf)
//   Black-Scholes, implied vol, Merton   analytics.h:8-91
//   implied surface, risk view, Dupire's formula   ivs.h:26-179
//   interpolation          interp.h:26-110 (smooth-step 2D for the risk view)
//   calibration            mcMdlDupire.h:283-388 (dupireCalibMaturity, dupireCalib)
//   superbucket            main.h:414-569
//
// The reference differentiates the calibration with its tape; here the
// derivative of each local vol to the risk-view spreads is written out by
// the chain rule over the same finite-difference formula, with the tape's
// conventions: d normalCdf = normalDens (AADExpr.h:434-446), the flat
// extrapolation of local vols outside 2.5 standard deviations copies the
// edge vol (one shared tape node), and the local-vol adjoints are *assigned*
// to those nodes (main.h:538-545), so a shared node keeps the last one.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "cf_host.h"

namespace cfh {
namespace {

constexpr double kTiny = 1.0e-08;   // EPS, gaussians.h:8

double normal_dens(double x)
{
    return x < -10.0 || 10.0 < x ? 0.0 : std::exp(-0.5 * x * x) / 2.506628274631;
}

double normal_cdf(double x)
{
    if (x < -10.0) return 0.0;
    if (x > 10.0) return 1.0;
    if (x < 0.0) return 1.0 - normal_cdf(-x);
    const double t = 1.0 / (1.0 + 0.2316419 * x);
    const double pol = t * (0.319381530 + t * (-0.356563782 + t * (1.781477937 + t * (-1.821255978 + t * 1.330274429))));
    return 1.0 - normal_dens(x) * pol;
}

// Black-Scholes call (zero rate) and its derivative to the vol with the
// tape's derivative of the cdf
struct CallV {
    double price, vega;
};

CallV black_scholes(double spot, double strike, double vol, double mat)
{
    const double sq = std::sqrt(mat);
    const double std_ = vol * sq;
    if (std_ <= kTiny) return {std::max(0.0, spot - strike), 0.0};
    const double lm = std::log(spot / strike);
    const double d2 = lm / std_ - 0.5 * std_;
    const double d1 = d2 + std_;
    const double dd2 = (-lm / (std_ * std_) - 0.5) * sq;   // d d2 / d vol
    const double dd1 = dd2 + sq;
    return {spot * normal_cdf(d1) - strike * normal_cdf(d2),
            spot * normal_dens(d1) * dd1 - strike * normal_dens(d2) * dd2};
}

double bs_price(double spot, double strike, double vol, double mat) { return black_scholes(spot, strike, vol, mat).price; }

// analytics.h:23-53: bracketing, bisection to 1e-12, then a secant step.
// The reference leaves the lower bracket's price unset until the bisection
// first moves the lower end; it is the price at the initial lower bound here.
double implied_vol(double spot, double strike, double prem, double mat)
{
    if (prem <= std::max(0.0, spot - strike) + kTiny) return 0.0;
    double u = 0.5;
    while (bs_price(spot, strike, u, mat) < prem) u *= 2;
    double l = 0.05;
    while (bs_price(spot, strike, l, mat) > prem) l /= 2;
    double pu = bs_price(spot, strike, u, mat);
    double pl = bs_price(spot, strike, l, mat);
    while (u - l > 1.e-12) {
        const double m = 0.5 * (u + l);
        const double p = bs_price(spot, strike, m, mat);
        if (p > prem) {
            u = m;
            pu = p;
        } else {
            l = m;
            pl = p;
        }
    }
    return l + (prem - pl) / (pu - pl) * (u - l);
}

// analytics.h:58-91: ten terms of the jump series
double merton(double spot, double strike, double vol, double mat, double intens, double meanJmp, double stdJmp)
{
    const double varJmp = stdJmp * stdJmp;
    const double mv2 = meanJmp + 0.5 * varJmp;
    const double comp = intens * (std::exp(mv2) - 1);
    const double var = vol * vol;
    const double intensT = intens * mat;
    unsigned fact = 1;
    double iT = 1.0, result = 0.0;
    for (unsigned n = 0; n < 10; ++n) {
        const double s = spot * std::exp(n * mv2 - comp * mat);
        const double v = std::sqrt(var + n * varJmp / mat);
        const double prob = std::exp(-intensT) * iT / fact;
        result += prob * bs_price(s, strike, v, mat);
        fact *= n + 1;
        iT *= intensT;
    }
    return result;
}

struct Merton {
    double spot, vol, intens, meanJmp, stdJmp;
    double implied(double strike, double mat) const
    {
        return implied_vol(spot, strike, merton(spot, strike, vol, mat, intens, meanJmp, stdJmp), mat);
    }
};

// Weights of the smooth-step interpolation of interp.h:26-63 on one axis:
// up to two (index, weight) pairs.
int step_weights(const vector<double>& xs, double x0, int idx[2], double w[2])
{
    const auto it = std::upper_bound(xs.begin(), xs.end(), x0);
    if (it == xs.end()) {
        idx[0] = int(xs.size()) - 1;
        w[0] = 1.0;
        return 1;
    }
    if (it == xs.begin()) {
        idx[0] = 0;
        w[0] = 1.0;
        return 1;
    }
    const int n = int(it - xs.begin()) - 1;
    const double t = (x0 - xs[n]) / (xs[n + 1] - xs[n]);
    const double s = t * t * (3.0 - 2 * t);
    idx[0] = n;
    w[0] = 1.0 - s;
    idx[1] = n + 1;
    w[1] = s;
    return 2;
}

// Smooth-step interpolation with flat extrapolation (interp.h:26-63), in the
// reference's order of operations.
double smooth_interp(const vector<double>& xs, const double* ys, double x0)
{
    const auto it = std::upper_bound(xs.begin(), xs.end(), x0);
    if (it == xs.end()) return ys[xs.size() - 1];
    if (it == xs.begin()) return ys[0];
    const size_t n = size_t(it - xs.begin()) - 1;
    const double t = (x0 - xs[n]) / (xs[n + 1] - xs[n]);
    return ys[n] + (ys[n + 1] - ys[n]) * t * t * (3.0 - 2 * t);
}

// Risk view grid (ivs.h:26-81): strikes x maturities of implied-vol spreads
// (strike major; empty: all 0).
struct View {
    vector<double> strikes, mats;
    vector<double> spreads;
    bool empty() const { return strikes.empty(); }
    // spread(strike, mat): interp2D<true> (interp.h:66-110), the smooth step
    // over maturities within two strike rows, then across them
    double spread(double strike, double mat) const
    {
        if (empty() || spreads.empty()) return 0.0;
        const size_t n = strikes.size(), m = mats.size();
        const auto it = std::upper_bound(strikes.begin(), strikes.end(), strike);
        const size_t n2 = size_t(it - strikes.begin());
        if (n2 == n) return smooth_interp(mats, &spreads[(n2 - 1) * m], mat);
        if (n2 == 0) return smooth_interp(mats, &spreads[0], mat);
        const size_t n1 = n2 - 1;
        const double z1 = smooth_interp(mats, &spreads[n1 * m], mat);
        const double z2 = smooth_interp(mats, &spreads[n2 * m], mat);
        const double t = (strike - strikes[n1]) / (strikes[n2] - strikes[n1]);
        return z1 + (z2 - z1) * t * t * (3.0 - 2 * t);
    }
    // sparse gradient of spread(strike, mat) (interp2D<true>, interp.h:66-110)
    void grad(double strike, double mat, std::map<int, double>& g, double scale) const
    {
        int ik[2], im[2];
        double wk[2], wm[2];
        const int nk = step_weights(strikes, strike, ik, wk);
        const int nm = step_weights(mats, mat, im, wm);
        for (int a = 0; a < nk; ++a)
            for (int b = 0; b < nm; ++b) g[ik[a] * int(mats.size()) + im[b]] += scale * wk[a] * wm[b];
    }
};

// Dupire's formula (ivs.h:119-138) with, when a view is given, the sparse
// derivative of the local vol to the view's spreads.
double local_vol(const Merton& ivs, double strike, double mat, const View* view, std::map<int, double>* d)
{
    struct Pt {
        double k, t;
    } pts[5] = {{strike, mat}, {strike, mat - 1.0e-04}, {strike, mat + 1.0e-04}, {strike - 1.0e-04, mat},
                {strike + 1.0e-04, mat}};
    // the implied vol of each point plus the view's spread there (ivs.h:100-110)
    CallV c[5];
    for (int p = 0; p < 5; ++p)
        c[p] = black_scholes(ivs.spot, pts[p].k,
                             ivs.implied(pts[p].k, pts[p].t) + (view ? view->spread(pts[p].k, pts[p].t) : 0.0),
                             pts[p].t);
    const double ct = (c[2].price - c[1].price) * 0.5e04;
    const double ckk = (c[3].price + c[4].price - 2.0 * c[0].price) * 1.0e08;
    const double r = 2.0 * ct / ckk;
    const double lv = std::sqrt(r) / strike;
    if (view && d && !view->empty()) {
        // d lv = (1 / strike) (0.5 / sqrt(r)) (2 d ct / ckk - 2 ct d ckk / ckk^2)
        const double a = (0.5 / std::sqrt(r)) / strike;
        const double dct = a * 2.0 / ckk;
        const double dckk = -a * 2.0 * ct / (ckk * ckk);
        const double dc[5] = {-2.0 * 1.0e08 * dckk, -0.5e04 * dct, 0.5e04 * dct, 1.0e08 * dckk, 1.0e08 * dckk};
        for (int p = 0; p < 5; ++p)
            if (dc[p] != 0.0 && c[p].vega != 0.0) view->grad(pts[p].k, pts[p].t, *d, dc[p] * c[p].vega);
    }
    return lv;
}

struct Calib {
    vector<double> spots, times;
    vector<double> lvols;                       // [spot][time]
    vector<int> node;                           // [spot][time] -> computed spot index of the shared value
    vector<std::map<int, double>> dlv;          // [spot][time] derivative to the view (computed nodes)
};

// dupireCalib (mcMdlDupire.h:337-388) with the per-maturity cut of
// dupireCalibMaturity (mcMdlDupire.h:286-333)
Calib calibrate(const Merton& ivs, const vector<double>& inclSpots, double maxDs, const vector<double>& inclTimes,
                double maxDt, const View* view, bool derivatives = true)
{
    Calib C;
    C.spots = fill_data(inclSpots, maxDs, 0.01, {});
    C.times = fill_data(inclTimes, maxDt, kOneHour, {maxDt});
    const int nS = int(C.spots.size()), nT = int(C.times.size());
    C.lvols.assign(size_t(nS) * nT, 0.0);
    C.node.assign(size_t(nS) * nT, 0);
    if (view && derivatives) C.dlv.resize(size_t(nS) * nT);
    for (int j = 0; j < nT; ++j) {
        const double mat = C.times[j];
        const double atm = black_scholes(ivs.spot, ivs.spot, ivs.implied(ivs.spot, mat), mat).price;
        const double sd = atm * 2.506628274631;
        int il = 0;
        while (il < nS && C.spots[il] < ivs.spot - 2.5 * sd) ++il;
        int ih = nS - 1;
        while (ih >= 0 && C.spots[ih] > ivs.spot + 2.5 * sd) --ih;
        for (int i = il; i <= ih; ++i) {
            const size_t e = size_t(i) * nT + j;
            C.lvols[e] = local_vol(ivs, C.spots[i], mat, view, C.dlv.empty() ? nullptr : &C.dlv[e]);
            C.node[e] = i;
        }
        for (int i = 0; i < il; ++i) {
            C.lvols[size_t(i) * nT + j] = C.lvols[size_t(il) * nT + j];
            C.node[size_t(i) * nT + j] = il;
        }
        for (int i = ih + 1; i < nS; ++i) {
            C.lvols[size_t(i) * nT + j] = C.lvols[size_t(ih) * nT + j];
            C.node[size_t(i) * nT + j] = ih;
        }
    }
    return C;
}

vector<double> vec(const double* p, int n)
{
    if (n < 1 || !p) fail(CF_ERR_ARG, "empty grid input");
    for (int i = 0; i < n; ++i) check_finite(p[i], "grid input");
    return vector<double>(p, p + n);
}

}  // namespace
}  // namespace cfh

using namespace cfh;

extern "C" {

int cf_dupire_calib(const cf_merton_ivs* ivs, const double* incl_spots, int n_incl_spots, double max_ds,
                    const double* incl_times, int n_incl_times, double max_dt, double* spots, double* times,
                    double* lvols, int* n_spots, int* n_times, int cap)
{
    return guarded([&] {
        if (!ivs || !n_spots || !n_times) fail(CF_ERR_ARG, "null argument");
        if (!(max_ds > 0.0) || !(max_dt > 0.0)) fail(CF_ERR_ARG, "grid spacing must be positive");
        const Merton m{ivs->spot, ivs->vol, ivs->jmp_intens, ivs->jmp_average, ivs->jmp_std};
        const Calib C = calibrate(m, vec(incl_spots, n_incl_spots), max_ds, vec(incl_times, n_incl_times), max_dt,
                                  nullptr);
        *n_spots = int(C.spots.size());
        *n_times = int(C.times.size());
        if (int(C.lvols.size()) <= cap) {
            if (spots) std::copy(C.spots.begin(), C.spots.end(), spots);
            if (times) std::copy(C.times.begin(), C.times.end(), times);
            if (lvols) std::copy(C.lvols.begin(), C.lvols.end(), lvols);
        }
        return CF_OK;
    });
}

int cf_dupire_calib_view(const cf_merton_ivs* ivs, const double* incl_spots, int n_incl_spots, double max_ds,
                         const double* incl_times, int n_incl_times, double max_dt, const double* strikes,
                         int n_strikes, const double* mats, int n_mats, const double* spreads, double* spots,
                         double* times, double* lvols, int* n_spots, int* n_times, int cap)
{
    return guarded([&] {
        if (!ivs || !n_spots || !n_times || !spreads) fail(CF_ERR_ARG, "null argument");
        if (!(max_ds > 0.0) || !(max_dt > 0.0)) fail(CF_ERR_ARG, "grid spacing must be positive");
        const Merton m{ivs->spot, ivs->vol, ivs->jmp_intens, ivs->jmp_average, ivs->jmp_std};
        View view{vec(strikes, n_strikes), vec(mats, n_mats), {}};
        view.spreads = vec(spreads, n_strikes * n_mats);
        const Calib C = calibrate(m, vec(incl_spots, n_incl_spots), max_ds, vec(incl_times, n_incl_times), max_dt,
                                  &view, false);
        *n_spots = int(C.spots.size());
        *n_times = int(C.times.size());
        if (int(C.lvols.size()) <= cap) {
            if (spots) std::copy(C.spots.begin(), C.spots.end(), spots);
            if (times) std::copy(C.times.begin(), C.times.end(), times);
            if (lvols) std::copy(C.lvols.begin(), C.lvols.end(), lvols);
        }
        return CF_OK;
    });
}

int cf_superbucket_vega(const cf_merton_ivs* ivs, const double* incl_spots, int n_incl_spots, double max_ds,
                        const double* incl_times, int n_incl_times, double max_dt_vol, const double* strikes,
                        int n_strikes, const double* mats, int n_mats, const double* microbucket, int n_micro,
                        double* vega)
{
    return guarded([&] {
        if (!ivs || !microbucket || !vega) fail(CF_ERR_ARG, "null argument");
        const Merton m{ivs->spot, ivs->vol, ivs->jmp_intens, ivs->jmp_average, ivs->jmp_std};
        const View view{vec(strikes, n_strikes), vec(mats, n_mats), {}};
        const Calib C = calibrate(m, vec(incl_spots, n_incl_spots), max_ds, vec(incl_times, n_incl_times),
                                  max_dt_vol, &view);
        const int nS = int(C.spots.size()), nT = int(C.times.size());
        if (n_micro != nS * nT) fail(CF_ERR_ARG, "microbucket size differs from the calibrated grid");
        // adjoints assigned to the local-vol nodes in spot-major order: a node
        // shared by flat-extrapolated vols keeps the last assignment
        vector<double> adj(size_t(nS) * nT, 0.0);
        for (int i = 0; i < nS; ++i)
            for (int j = 0; j < nT; ++j) adj[size_t(C.node[size_t(i) * nT + j]) * nT + j] = microbucket[size_t(i) * nT + j];
        std::fill(vega, vega + size_t(n_strikes) * n_mats, 0.0);
        for (size_t e = 0; e < adj.size(); ++e) {
            if (adj[e] == 0.0) continue;
            for (const auto& kv : C.dlv[e]) vega[kv.first] += adj[e] * kv.second;
        }
        return CF_OK;
    });
}

int cf_dupire_superbucket(const cf_merton_ivs* ivs, double max_dt, const char* product_id, const char* const* labels,
                          const double* notionals, int n, const double* incl_spots, int n_incl_spots, double max_ds,
                          const double* incl_times, int n_incl_times, double max_dt_vol, const double* strikes,
                          int n_strikes, const double* mats, int n_mats, const cf_num* num, double* value,
                          double* delta, double* vega)
{
    return guarded([&] {
        if (!ivs || !num || !value || !delta || !vega) fail(CF_ERR_ARG, "null argument");
        if (!product_id) fail(CF_ERR_ARG, "null product id");
        const Merton m{ivs->spot, ivs->vol, ivs->jmp_intens, ivs->jmp_average, ivs->jmp_std};
        const Calib C = calibrate(m, vec(incl_spots, n_incl_spots), max_ds, vec(incl_times, n_incl_times),
                                  max_dt_vol, nullptr);
        const int nS = int(C.spots.size()), nT = int(C.times.size());
        // microbucket on the device: dupireAADRisk on the calibrated model,
        // stored as "superbucket" (main.h:519-534)
        int rc = cf_put_dupire(ivs->spot, C.spots.data(), nS, C.times.data(), nT, C.lvols.data(), max_dt,
                               "superbucket");
        if (rc != CF_OK) fail(rc, g_err);
        vector<double> micro(size_t(nS) * nT);
        rc = cf_dupire_aad_risk("superbucket", product_id, labels, notionals, n, num, value, delta, micro.data(),
                                int(micro.size()));
        if (rc < 0) fail(rc, g_err);
        rc = cf_superbucket_vega(ivs, incl_spots, n_incl_spots, max_ds, incl_times, n_incl_times, max_dt_vol, strikes,
                                 n_strikes, mats, n_mats, micro.data(), int(micro.size()), vega);
        if (rc != CF_OK) fail(rc, g_err);
        return CF_OK;
    });
}

}  // extern "C"
