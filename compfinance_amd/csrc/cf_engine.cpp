// Host engine behind the descriptor-level C ABI (include/cfamd.h):
// validates the descriptors, builds the timelines and tables in FP64 exactly
// as the reference does, stages them on the device, launches the path kernel
// and the cross-block reductions, and maps adjoints back to model parameters.
//
// Reference steps restated here:
//   product timelines     mcPrd.h:44-81, :149-208, :305-336; store.h:153, :190-195
//   Dupire timeline/init  mcMdlDupire.h:165-217 with fillData utility.h:11-79
//   BS timeline/init      mcMdlBS.h:149-277
//   labels                mcPrd.h:66-80, :197-207, :325-335; mcMdlDupire.h:95-107;
//                         mcMdlBS.h:87-90
//   risk normalisation    mcBase.h:738-746 (sum of adjoints / nPath)
//   adjoint of the local-vol pre-interpolation (before the tape mark):
//                         mcMdlDupire.h:203-215, mcBase.h:717-735
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <list>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iomanip>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cfamd.h"
#include "cf_host.h"
#include "cf_internal.h"
#include "sobol_init.inc"

namespace cfh {

// Launch-shape overrides for tuning sweeps: CF_BLOCK_THREADS (64..1024, a
// multiple of 64) and CF_LAYOUT (full, lean).  Unset in normal use.
int forced_block_threads()   // 0: by path count
{
    static const int v = [] {
        const char* e = std::getenv("CF_BLOCK_THREADS");
        const int b = e ? std::atoi(e) : 0;
        return (b >= 64 && b <= 1024 && b % 64 == 0) ? b : 0;
    }();
    return v;
}

// CF_IVBAR_COPIES=1 keeps one local-vol adjoint grid per block (tuning).
bool ivbar_copies_allowed()
{
    static const bool v = [] {
        const char* e = std::getenv("CF_IVBAR_COPIES");
        return !(e && std::string(e) == "1");
    }();
    return v;
}

int forced_layout()   // -1 automatic, 0 FULL, 1 LEAN
{
    static const int v = [] {
        const char* e = std::getenv("CF_LAYOUT");
        if (!e) return -1;
        if (std::string(e) == "full") return 0;
        if (std::string(e) == "lean") return 1;
        return -1;
    }();
    return v;
}

// ------------------------------------------------------------------ errors

thread_local string g_err;
thread_local cf_kernel_stats g_stats{};


void fail(int code, const string& msg) { throw CfError(code, msg); }


void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) fail(CF_ERR_DEVICE, string(what) + ": " + hipGetErrorString(e));
}

string fmt2(double x)
{
    std::ostringstream o;
    o << std::fixed << std::setprecision(2) << x;
    return o.str();
}

// ------------------------------------------------------------------ timelines

// utility.h:11-79: union with the added points (equality within minDx), then
// supplementary points so that no gap exceeds maxDx.
vector<double> fill_data(const vector<double>& original, double maxDx, double minDx, const vector<double>& add)
{
    vector<double> seq;
    if (!add.empty()) {
        std::set_union(original.begin(), original.end(), add.begin(), add.end(), std::back_inserter(seq),
                       [minDx](double x, double y) { return x < y - minDx; });
    } else {
        seq = original;
    }
    vector<double> out;
    out.push_back(seq[0]);
    for (size_t k = 1; k < seq.size(); ++k) {
        const double cur = out.back(), nxt = seq[k];
        if (nxt - cur > maxDx) {
            const int pts = int((nxt - cur) / maxDx - kEps) + 1;
            const double spacing = (nxt - cur) / pts;
            double t = cur + spacing;
            while (t < nxt - minDx) {
                out.push_back(t);
                t += spacing;
            }
        }
        out.push_back(nxt);
    }
    return out;
}


Interp1 interp_weights(const double* xs, int n, double x)
{
    const double* it = std::upper_bound(xs, xs + n, x);
    if (it == xs + n) return {n - 1, n - 1, 1.0, 0.0};
    if (it == xs) return {0, 0, 1.0, 0.0};
    const int k = int(it - xs) - 1;
    const double t = (x - xs[k]) / (xs[k + 1] - xs[k]);
    return {k, k + 1, 1.0 - t, t};
}

double interp_value(const double* xs, int n, const double* ys, double x)
{
    const double* it = std::upper_bound(xs, xs + n, x);
    if (it == xs + n) return ys[n - 1];
    if (it == xs) return ys[0];
    const size_t k = size_t(it - xs) - 1;
    const double x1 = xs[k], x2 = xs[k + 1], y1 = ys[k], y2 = ys[k + 1];
    const double t = (x - x1) / (x2 - x1);
    return y1 + (y2 - y1) * t;
}

// ------------------------------------------------------------------ plans



void check_finite(double x, const char* what)
{
    if (!std::isfinite(x)) fail(CF_ERR_ARG, string(what) + " must be finite");
}

// Dates freq, 2 freq, ... accumulated as the barrier's monitoring dates
// (mcPrd.h:162-193) up to the maturity, which closes the list; today first
// when asked.
vector<double> periodic_dates(double maturity, double freq, bool with_today)
{
    vector<double> out;
    if (with_today) out.push_back(0.0);
    double t = 0.0 + freq;
    while (maturity - t > kOneHour) {
        out.push_back(t);
        t += freq;
    }
    out.push_back(maturity);
    return out;
}

// Asian fixing dates, moved to the nearest date of the snapping grid when
// snap > 0 (ties to the earlier date).  From the last fixing back, a fixing
// whose nearest date is not before the next fixing's takes the grid date just
// before it (a last monthly fixing a few days before maturity), so fixings
// stay distinct.
vector<double> asian_fixings(double maturity, double freq, double snap)
{
    vector<double> fix = periodic_dates(maturity, freq, false);
    if (snap > 0.0) {
        const vector<double> grid = periodic_dates(maturity, snap, true);
        for (size_t k = fix.size(); k-- > 0;) {
            double& t = fix[k];
            auto it = std::lower_bound(grid.begin(), grid.end(), t);
            if (it == grid.end()) t = grid.back();
            else if (it != grid.begin() && t - *(it - 1) <= *it - t) t = *(it - 1);
            else t = *it;
            if (k + 1 < fix.size() && !(t < fix[k + 1])) {
                auto nx = std::lower_bound(grid.begin(), grid.end(), fix[k + 1]);
                t = nx == grid.begin() ? 0.0 : *(nx - 1);
            }
            if (t <= 0.0) fail(CF_ERR_ARG, "Asian fixings closer than the snapping grid");
        }
    }
    return fix;
}

ProductPlan plan_product(const cf_product* p)
{
    if (!p) fail(CF_ERR_ARG, "null product");
    ProductPlan d;
    d.kind = p->kind;
    if (p->kind == CF_PRODUCT_EUROPEAN) {
        // mcPrd.h:44-81
        check_finite(p->strike, "strike");
        check_finite(p->exercise, "exercise date");
        if (p->exercise < 0.0) fail(CF_ERR_ARG, "exercise date must be non-negative");
        d.strike = p->strike;
        d.exercise = p->exercise;
        d.settlement = p->settlement > 0.0 ? p->settlement : p->exercise;
        d.timeline = {d.exercise};
        SampleDefH s;
        s.numeraire = true;
        s.has_fwd = true;
        s.fwd_mat = d.settlement;
        s.has_disc = true;
        s.disc_mat = d.settlement;
        d.defs = {s};
        d.n_pay = 1;
    } else if (p->kind == CF_PRODUCT_BARRIER) {
        // mcPrd.h:149-208; store.h:153
        check_finite(p->strike, "strike");
        check_finite(p->barrier, "barrier");
        check_finite(p->maturity, "maturity");
        check_finite(p->monitor_freq, "monitoring frequency");
        if (!(p->monitor_freq > 0.0)) fail(CF_ERR_ARG, "monitoring frequency must be positive");
        if (!(p->maturity > 0.0)) fail(CF_ERR_ARG, "maturity must be positive");
        if (p->maturity / p->monitor_freq > 20000.0) fail(CF_ERR_CAPACITY, "too many monitoring dates");
        d.strike = p->strike;
        d.barrier = p->barrier;
        d.smooth = p->smooth <= 0 ? kEps : p->smooth;
        d.put = p->call_put != 0;
        d.freq = p->monitor_freq;
        d.maturity = p->maturity;
        d.timeline = periodic_dates(p->maturity, p->monitor_freq, true);
        for (size_t i = 0; i < d.timeline.size(); ++i) {
            SampleDefH s;
            s.numeraire = false;
            s.has_fwd = true;
            s.fwd_mat = d.timeline[i];
            d.defs.push_back(s);
        }
        d.defs.back().numeraire = true;
        d.n_pay = 2;
    } else if (p->kind == CF_PRODUCT_EUROPEANS) {
        // store.h:184-205, mcPrd.h:305-336
        if (p->n_options <= 0 || !p->maturities || !p->strikes) fail(CF_ERR_ARG, "no options");
        std::map<double, vector<double>> opts;
        for (int i = 0; i < p->n_options; ++i) {
            check_finite(p->maturities[i], "maturity");
            check_finite(p->strikes[i], "strike");
            if (p->maturities[i] < 0.0) fail(CF_ERR_ARG, "maturities must be non-negative");
            opts[p->maturities[i]].push_back(p->strikes[i]);
        }
        for (auto& kv : opts) {
            d.mats.push_back(kv.first);
            d.strikes.push_back(kv.second);
            d.timeline.push_back(kv.first);
            SampleDefH s;
            s.numeraire = true;
            s.has_fwd = true;
            s.fwd_mat = kv.first;
            d.defs.push_back(s);
        }
        d.pay_first.push_back(0);
        for (auto& ks : d.strikes) {
            for (double k : ks) d.pay_strike.push_back(k);
            d.pay_first.push_back(int(d.pay_strike.size()));
        }
        d.n_pay = int(d.pay_strike.size());
    } else if (p->kind == CF_PRODUCT_ASIAN) {
        check_finite(p->strike, "strike");
        check_finite(p->maturity, "maturity");
        check_finite(p->fixing_freq, "fixing frequency");
        check_finite(p->snap_freq, "snapping frequency");
        if (!(p->maturity > 0.0)) fail(CF_ERR_ARG, "maturity must be positive");
        if (!(p->fixing_freq > 0.0)) fail(CF_ERR_ARG, "fixing frequency must be positive");
        if (p->snap_freq < 0.0) fail(CF_ERR_ARG, "snapping frequency must be non-negative");
        if (p->maturity / p->fixing_freq > 20000.0) fail(CF_ERR_CAPACITY, "too many fixing dates");
        if (p->snap_freq > 0.0 && p->maturity / p->snap_freq > 20000.0) fail(CF_ERR_CAPACITY, "snapping grid too fine");
        d.strike = p->strike;
        d.maturity = p->maturity;
        d.put = p->call_put != 0;
        d.fix_freq = p->fixing_freq;
        d.snap = p->snap_freq;
        d.timeline = asian_fixings(p->maturity, p->fixing_freq, p->snap_freq);
        for (size_t i = 0; i < d.timeline.size(); ++i) {
            SampleDefH s;
            s.numeraire = i + 1 == d.timeline.size();
            s.has_fwd = true;
            s.fwd_mat = d.timeline[i];
            d.defs.push_back(s);
        }
        d.n_pay = 1;
    } else if (p->kind == CF_PRODUCT_BOOK) {
        if (p->n_parts < 1 || !p->parts) fail(CF_ERR_ARG, "empty book");
        if (p->n_parts > 4096) fail(CF_ERR_CAPACITY, "more than 4096 products in a book");
        vector<double> all;
        for (int i = 0; i < p->n_parts; ++i) {
            const cf_product* q = p->parts + i;
            if (q->kind == CF_PRODUCT_BOOK) fail(CF_ERR_UNSUPPORTED, "a book cannot hold a book");
            ProductPlan part = plan_product(q);
            if (part.kind == CF_PRODUCT_EUROPEAN && part.settlement != part.exercise)
                fail(CF_ERR_UNSUPPORTED, "a book holds Europeans settled on their exercise date only");
            all.insert(all.end(), part.timeline.begin(), part.timeline.end());
            d.n_pay += part.n_pay;
            d.parts.push_back(std::move(part));
        }
        int n_bar = 0;
        vector<vector<double>> sched;
        for (const auto& q : d.parts) {
            if (q.kind == CF_PRODUCT_BARRIER) ++n_bar;
            if (q.kind == CF_PRODUCT_ASIAN && std::find(sched.begin(), sched.end(), q.timeline) == sched.end())
                sched.push_back(q.timeline);
        }
        if (n_bar > cfk::kMaxBookBarriers)
            fail(CF_ERR_CAPACITY, "a book holds at most " + std::to_string(cfk::kMaxBookBarriers) + " barriers");
        if (int(sched.size()) > cfk::kMaxBookGroups)
            fail(CF_ERR_CAPACITY,
                 "a book holds at most " + std::to_string(cfk::kMaxBookGroups) + " Asian fixing schedules");
        std::sort(all.begin(), all.end());
        all.erase(std::unique(all.begin(), all.end()), all.end());
        d.timeline = all;
        for (double t : all) {
            SampleDefH s;
            s.numeraire = true;
            s.has_fwd = true;
            s.fwd_mat = t;
            d.defs.push_back(s);
        }
    } else {
        fail(CF_ERR_UNSUPPORTED, "product kind not supported by the GPU engine");
    }
    return d;
}

vector<string> payoff_labels(const ProductPlan& d)
{
    vector<string> out;
    if (d.kind == CF_PRODUCT_ASIAN) {
        out.push_back(string(d.put ? "asian put " : "asian call ") + fmt2(d.maturity) + " " + fmt2(d.strike));
        return out;
    }
    if (d.kind == CF_PRODUCT_BOOK) {
        for (size_t i = 0; i < d.parts.size(); ++i)
            for (const string& l : payoff_labels(d.parts[i])) out.push_back(std::to_string(i) + ":" + l);
        return out;
    }
    if (d.kind == CF_PRODUCT_EUROPEAN) {
        if (d.settlement == d.exercise) out.push_back("call " + fmt2(d.strike) + " " + fmt2(d.exercise));
        else out.push_back("call " + fmt2(d.strike) + " " + fmt2(d.exercise) + " " + fmt2(d.settlement));
    } else if (d.kind == CF_PRODUCT_BARRIER) {
        const string vanilla = string(d.put ? "put " : "call ") + fmt2(d.maturity) + " " + fmt2(d.strike);
        out.push_back(vanilla + " up and out " + fmt2(d.barrier) + " monitoring freq " + fmt2(d.freq) + " smooth " +
                      fmt2(d.smooth));
        out.push_back(vanilla);
    } else {
        for (size_t i = 0; i < d.mats.size(); ++i)
            for (double k : d.strikes[i]) out.push_back("call " + fmt2(d.mats[i]) + " " + fmt2(k));
    }
    return out;
}

// Kernel product of a plan.
int kernel_product(const ProductPlan& d)
{
    switch (d.kind) {
    case CF_PRODUCT_EUROPEAN: return cfk::PRD_EUROPEAN;
    case CF_PRODUCT_BARRIER: return cfk::PRD_BARRIER;
    case CF_PRODUCT_EUROPEANS: return cfk::PRD_EUROPEANS;
    default: return cfk::PRD_BOOK;
    }
}


BookTables build_book(const ProductPlan& d, double width0)
{
    vector<const ProductPlan*> parts;
    if (d.kind == CF_PRODUCT_BOOK) {
        for (const auto& q : d.parts) parts.push_back(&q);
    } else {
        parts.push_back(&d);
    }
    const vector<double>& tl = d.timeline;
    const int ns = int(tl.size());
    auto at = [&](double t) {
        auto it = std::lower_bound(tl.begin(), tl.end(), t);
        if (it == tl.end() || *it != t) fail(CF_ERR_ARG, "book date not on its timeline");
        return int(it - tl.begin());
    };
    vector<vector<std::pair<double, int>>> van(ns);
    vector<uint32_t> bar_mask(ns, 0u), bar_last(ns, 0u), grp_mask(ns, 0u), grp_last(ns, 0u);
    vector<double> bar_par;
    vector<int> bar_col, bar_put;
    vector<vector<double>> sched;                       // fixing schedules
    vector<vector<std::tuple<double, int, int>>> members;   // per schedule: strike, column, put
    int col = 0;
    for (const ProductPlan* q : parts) {
        if (q->kind == CF_PRODUCT_EUROPEAN) {
            van[at(q->exercise)].push_back({q->strike, col++});
        } else if (q->kind == CF_PRODUCT_EUROPEANS) {
            for (size_t i = 0; i < q->mats.size(); ++i)
                for (double k : q->strikes[i]) van[at(q->mats[i])].push_back({k, col++});
        } else if (q->kind == CF_PRODUCT_BARRIER) {
            const int b = int(bar_col.size());
            if (b >= cfk::kMaxBookBarriers)
                fail(CF_ERR_CAPACITY, "a book holds at most " + std::to_string(cfk::kMaxBookBarriers) + " barriers");
            for (double t : q->timeline) bar_mask[at(t)] |= 1u << b;
            bar_last[at(q->timeline.back())] |= 1u << b;
            const double w = width0 * q->smooth;
            bar_par.insert(bar_par.end(), {q->strike, q->barrier + w, q->barrier - w, 1.0 / (2 * w)});
            bar_col.push_back(col);
            bar_put.push_back(q->put ? 1 : 0);
            col += 2;
        } else if (q->kind == CF_PRODUCT_ASIAN) {
            size_t g = 0;
            while (g < sched.size() && sched[g] != q->timeline) ++g;
            if (g == sched.size()) {
                if (int(g) >= cfk::kMaxBookGroups)
                    fail(CF_ERR_CAPACITY,
                         "a book holds at most " + std::to_string(cfk::kMaxBookGroups) + " Asian fixing schedules");
                sched.push_back(q->timeline);
                members.emplace_back();
            }
            members[g].push_back({q->strike, col++, q->put ? 1 : 0});
        } else {
            fail(CF_ERR_UNSUPPORTED, "product kind cannot be held in a book");
        }
    }
    for (size_t g = 0; g < sched.size(); ++g) {
        for (double t : sched[g]) grp_mask[at(t)] |= 1u << g;
        grp_last[at(sched[g].back())] |= 1u << g;
    }
    vector<int> van_first(1, 0), van_col, grp_first(1, 0), as_col, as_put;
    vector<double> van_strike, grp_inv, as_strike;
    for (const auto& v : van) {
        for (const auto& kc : v) {
            van_strike.push_back(kc.first);
            van_col.push_back(kc.second);
        }
        van_first.push_back(int(van_strike.size()));
    }
    for (size_t g = 0; g < sched.size(); ++g) {
        grp_inv.push_back(1.0 / double(sched[g].size()));
        for (const auto& m : members[g]) {
            as_strike.push_back(std::get<0>(m));
            as_col.push_back(std::get<1>(m));
            as_put.push_back(std::get<2>(m));
        }
        grp_first.push_back(int(as_strike.size()));
    }
    BookTables B;
    auto put = [&B](const void* p, size_t bytes) {
        // 16-byte aligned parts: the kernel reads a barrier's four parameters at once
        const size_t off = (B.blob.size() + 15) & ~size_t(15);
        B.blob.resize(off + bytes);
        if (bytes) std::memcpy(B.blob.data() + off, p, bytes);
        return int(off);
    };
    cfk::BookLayout& L = B.L;
    L.n_van = int(van_strike.size());
    L.n_bar = int(bar_col.size());
    L.n_grp = int(sched.size());
    L.n_asian = int(as_strike.size());
    L.van_strike = put(van_strike.data(), van_strike.size() * 8);
    L.van_col = put(van_col.data(), van_col.size() * 4);
    // a sample's row, one 16-byte read: barriers observing (bits 0-15) and
    // maturing (16-31), schedules fixing (bits 0-3) and ending (4-7), and its
    // vanillas [van_first[s], van_first[s + 1])
    vector<uint32_t> masks(size_t(ns) * 4);
    for (int k = 0; k < ns; ++k) {
        masks[4 * k] = bar_mask[k] | (bar_last[k] << 16);
        masks[4 * k + 1] = grp_mask[k] | (grp_last[k] << 4);
        masks[4 * k + 2] = uint32_t(van_first[k]);
        masks[4 * k + 3] = uint32_t(van_first[k + 1]);
    }
    L.masks = put(masks.data(), masks.size() * 4);
    L.bar_par = put(bar_par.data(), bar_par.size() * 8);
    const vector<float> bar_par_f(bar_par.begin(), bar_par.end());
    L.bar_par_f = put(bar_par_f.data(), bar_par_f.size() * 4);
    for (size_t b = 0; b < bar_col.size(); ++b) B.lo_min = std::min(B.lo_min, bar_par[4 * b + 2]);
    L.bar_col = put(bar_col.data(), bar_col.size() * 4);
    L.bar_put = put(bar_put.data(), bar_put.size() * 4);
    L.grp_inv = put(grp_inv.data(), grp_inv.size() * 8);
    const vector<float> grp_inv_f(grp_inv.begin(), grp_inv.end());
    L.grp_inv_f = put(grp_inv_f.data(), grp_inv_f.size() * 4);
    L.grp_first = put(grp_first.data(), grp_first.size() * 4);
    L.as_strike = put(as_strike.data(), as_strike.size() * 8);
    L.as_col = put(as_col.data(), as_col.size() * 4);
    L.as_put = put(as_put.data(), as_put.size() * 4);
    B.blob.resize((B.blob.size() + 15) & ~size_t(15));
    L.bytes = int(B.blob.size());
    return B;
}


void validate_model(const cf_model* m)
{
    if (!m) fail(CF_ERR_ARG, "null model");
    check_finite(m->spot, "spot");
    if (!(m->spot > 0.0)) fail(CF_ERR_ARG, "spot must be positive");
    if (m->kind == CF_MODEL_BLACK_SCHOLES) {
        check_finite(m->vol, "vol");
        check_finite(m->rate, "rate");
        check_finite(m->div, "div");
    } else if (m->kind == CF_MODEL_DUPIRE) {
        if (m->n_spots < 1 || m->n_times < 1 || !m->spots || !m->times || !m->vols)
            fail(CF_ERR_ARG, "empty local volatility grid");
        if (m->n_spots > 1024) fail(CF_ERR_CAPACITY, "more than 1024 local volatility spots");
        for (int i = 0; i < m->n_spots; ++i) {
            check_finite(m->spots[i], "local volatility spot");
            if (!(m->spots[i] > 0.0)) fail(CF_ERR_ARG, "local volatility spots must be positive");
            if (i > 0 && !(m->spots[i] > m->spots[i - 1])) fail(CF_ERR_ARG, "local volatility spots must increase");
        }
        for (int j = 0; j < m->n_times; ++j) {
            check_finite(m->times[j], "local volatility time");
            if (j > 0 && !(m->times[j] > m->times[j - 1])) fail(CF_ERR_ARG, "local volatility times must increase");
        }
        for (long k = 0; k < long(m->n_spots) * m->n_times; ++k) check_finite(m->vols[k], "local volatility");
        check_finite(m->max_dt, "maxDt");
        if (!(m->max_dt > 0.0)) fail(CF_ERR_ARG, "maxDt must be positive");
    } else {
        fail(CF_ERR_UNSUPPORTED, "model kind not supported by the GPU engine");
    }
}

ModelPlan plan_model(const cf_model* m, const ProductPlan& p)
{
    validate_model(m);
    ModelPlan d;
    d.kind = m->kind;
    d.spot = m->spot;
    const size_t ns = p.timeline.size();
    if (m->kind == CF_MODEL_DUPIRE) {
        // mcMdlDupire.h:165-193
        d.n_spots = m->n_spots;
        d.n_times = m->n_times;
        d.spots.assign(m->spots, m->spots + m->n_spots);
        d.times.assign(m->times, m->times + m->n_times);
        const double today = 0.0;
        d.timeline = fill_data(p.timeline, m->max_dt, kHalfDay, {today});
        d.n_steps = int(d.timeline.size()) - 1;
        d.step_sample.assign(d.timeline.size(), -1);
        int s = 0;
        for (size_t i = 0; i < d.timeline.size(); ++i)
            if (std::binary_search(p.timeline.begin(), p.timeline.end(), d.timeline[i])) d.step_sample[i] = s++;
        if (size_t(s) != ns) fail(CF_ERR_ARG, "product dates closer than half a day cannot be simulated");
        // the kernels take the last sample at the last timeline point
        if (ns == 0 || d.step_sample.back() != int(ns) - 1) fail(CF_ERR_ARG, "product has no sample at its last date");
        // mcMdlDupire.h:91-92, :195-217
        d.log_spots.resize(d.n_spots);
        for (int i = 0; i < d.n_spots; ++i) d.log_spots[i] = std::log(d.spots[i]);
        d.inv_dx.assign(d.n_spots, 0.0);
        for (int i = 0; i + 1 < d.n_spots; ++i) d.inv_dx[i] = 1.0 / (d.log_spots[i + 1] - d.log_spots[i]);
        d.iv.resize(size_t(d.n_steps) * d.n_spots);
        for (int i = 0; i < d.n_steps; ++i) {
            const double sqrtdt = std::sqrt(d.timeline[i + 1] - d.timeline[i]);
            for (int j = 0; j < d.n_spots; ++j)
                d.iv[size_t(i) * d.n_spots + j] =
                    sqrtdt * interp_value(m->times, m->n_times, m->vols + size_t(j) * m->n_times, d.timeline[i]);
        }
        d.ff.assign(ns, 1.0);
        d.disc.assign(ns, 1.0);
        d.num.assign(ns, 1.0);
        d.s_t0 = std::exp(std::log(m->spot));   // spot written at t = 0 (mcMdlDupire.h:245-253)
    } else {
        // mcMdlBS.h:149-195
        d.vol = m->vol;
        d.rate = m->rate;
        d.div = m->div;
        d.spot_measure = m->spot_measure != 0;
        const double today = 0.0;
        d.timeline.assign(1, today);
        for (double t : p.timeline)
            if (t > today) d.timeline.push_back(t);
        const bool todayOn = p.timeline[0] == today;
        d.n_steps = int(d.timeline.size()) - 1;
        d.step_sample.assign(d.timeline.size(), -1);
        int s = 0;
        if (todayOn) d.step_sample[0] = s++;
        for (int i = 1; i <= d.n_steps; ++i) d.step_sample[i] = s++;
        if (size_t(s) != ns) fail(CF_ERR_ARG, "product timeline not simulable in Black-Scholes");
        // the kernels take the last sample at the last timeline point
        if (ns == 0 || d.step_sample.back() != int(ns) - 1) fail(CF_ERR_ARG, "product has no sample at its last date");
        // mcMdlBS.h:197-277
        const double mu = d.rate - d.div;
        d.drift.resize(d.n_steps);
        d.stdv.resize(d.n_steps);
        for (int i = 0; i < d.n_steps; ++i) {
            const double dt = d.timeline[i + 1] - d.timeline[i];
            d.stdv[i] = d.vol * std::sqrt(dt);
            d.drift[i] = d.spot_measure ? (mu + 0.5 * d.vol * d.vol) * dt : (mu - 0.5 * d.vol * d.vol) * dt;
        }
        d.ff.assign(ns, 1.0);
        d.disc.assign(ns, 1.0);
        d.num.assign(ns, 1.0);
        d.bs_aux.assign(3 * ns + d.n_steps, 0.0);
        for (size_t i = 0; i < ns; ++i) {
            const SampleDefH& def = p.defs[i];
            const double t = p.timeline[i];
            if (def.numeraire) d.num[i] = d.spot_measure ? std::exp(d.div * t) / d.spot : std::exp(d.rate * t);
            if (def.has_disc) d.disc[i] = std::exp(-d.rate * (def.disc_mat - t));
            if (def.has_fwd) d.ff[i] = std::exp(mu * (def.fwd_mat - t));
            d.bs_aux[i] = t;
            d.bs_aux[ns + i] = def.has_fwd ? def.fwd_mat - t : 0.0;
            d.bs_aux[2 * ns + i] = def.has_disc ? def.disc_mat - t : 0.0;
        }
        for (int i = 0; i < d.n_steps; ++i) d.bs_aux[3 * ns + i] = d.timeline[i + 1] - d.timeline[i];
        d.s_t0 = m->spot * d.ff[0];   // forward of the first sample at t = 0
    }
    if (d.n_steps > cfk::kSobolDims)
        fail(CF_ERR_CAPACITY, "more simulation steps than Sobol dimensions (" + std::to_string(cfk::kSobolDims) + ")");
    return d;
}

vector<string> param_labels(const cf_model* m)
{
    vector<string> out;
    out.push_back("spot");
    if (m->kind == CF_MODEL_BLACK_SCHOLES) {
        out.push_back("vol");
        out.push_back("rate");
        out.push_back("div");
    } else {
        for (int i = 0; i < m->n_spots; ++i)
            for (int j = 0; j < m->n_times; ++j) out.push_back("lvol " + fmt2(m->spots[i]) + " " + fmt2(m->times[j]));
    }
    return out;
}

// ------------------------------------------------------------------ device






constexpr size_t kMaxPlans = 64;

std::mutex g_engines_mu;
std::map<int, std::unique_ptr<Engine>> g_engines;

Engine& engine()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_engines_mu);
    auto& slot = g_engines[dev];
    if (!slot) {
        auto e = std::make_unique<Engine>();
        e->device = dev;
        hip_check(hipDeviceGetAttribute(&e->num_cus, hipDeviceAttributeMultiprocessorCount, dev),
                  "hipDeviceGetAttribute");
        hip_check(hipDeviceGetAttribute(&e->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
                  "hipDeviceGetAttribute");
        hip_check(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking), "hipStreamCreate");
        for (auto& v : e->ev) hip_check(hipEventCreate(&v), "hipEventCreate");
        // direction numbers [32][kSobolDims], Joe-Kuo recurrence
        vector<uint32_t> dirs(32 * size_t(cfk::kSobolDims));
        for (int d = 0; d < cfk::kSobolDims; ++d) {
            const unsigned s = kSobolInit[d][0], a = kSobolInit[d][1];
            uint32_t v[32];
            for (unsigned k = 0; k < 32; ++k) {
                if (s == 0) {
                    v[k] = 1u << (31 - k);   // van der Corput dimension
                } else if (k < s) {
                    v[k] = kSobolInit[d][2 + k] << (31 - k);
                } else {
                    uint32_t x = v[k - s] ^ (v[k - s] >> s);
                    for (unsigned j = 1; j < s; ++j)
                        if ((a >> (s - 1 - j)) & 1u) x ^= v[k - j];
                    v[k] = x;
                }
            }
            for (int k = 0; k < 32; ++k) dirs[size_t(k) * cfk::kSobolDims + d] = v[k];
        }
        // into the library's module variable: creating the engine allocates no device memory
        const int rc = cfk::upload_sobol_dirs(dirs.data(), &e->d_dirs);
        if (rc != 0) hip_check(hipError_t(rc), "direction numbers");
        slot = std::move(e);
    }
    return *slot;
}


void validate_num(const cf_num* num)
{
    if (!num) fail(CF_ERR_ARG, "null numerical parameters");
    if (num->num_path < 1) fail(CF_ERR_ARG, "numPath must be positive");
    if (num->first_path < 0) fail(CF_ERR_ARG, "first path must be non-negative");
    if (int64_t(num->first_path) + num->num_path >= (int64_t(1) << 32) - 1)
        fail(CF_ERR_CAPACITY, "Sobol index beyond 2^32 - 1");
    if (num->precision != CF_FP64 && num->precision != CF_FP32) fail(CF_ERR_ARG, "unknown precision");
    if (num->aad_mode != CF_AAD_CHECKPOINT && num->aad_mode != CF_AAD_RECOMPUTE) fail(CF_ERR_ARG, "unknown AAD mode");
}

// Byte image of everything a prepared plan depends on.
// adj: cfk::AdjointKind
string plan_key(const cf_model* m, const cf_product* p, int precision, int adj, int block)
{
    string k;
    auto put = [&k](const void* d, size_t n) { k.append(static_cast<const char*>(d), n); };
    auto putd = [&](double x) { put(&x, sizeof x); };
    auto puti = [&](int x) { put(&x, sizeof x); };
    puti(precision);
    puti(adj);
    puti(block);
    puti(m->kind);
    putd(m->spot);
    if (m->kind == CF_MODEL_BLACK_SCHOLES) {
        putd(m->vol);
        putd(m->rate);
        putd(m->div);
        puti(m->spot_measure);
    } else {
        puti(m->n_spots);
        puti(m->n_times);
        putd(m->max_dt);
        put(m->spots, sizeof(double) * size_t(std::max(m->n_spots, 0)));
        put(m->times, sizeof(double) * size_t(std::max(m->n_times, 0)));
        put(m->vols, sizeof(double) * size_t(std::max(m->n_spots, 0)) * size_t(std::max(m->n_times, 0)));
    }
    std::function<void(const cf_product*)> put_product = [&](const cf_product* q) {
        puti(q->kind);
        if (q->kind == CF_PRODUCT_EUROPEANS) {
            puti(q->n_options);
            put(q->maturities, sizeof(double) * size_t(std::max(q->n_options, 0)));
            put(q->strikes, sizeof(double) * size_t(std::max(q->n_options, 0)));
        } else if (q->kind == CF_PRODUCT_BOOK) {
            puti(q->n_parts);
            for (int i = 0; i < q->n_parts; ++i) put_product(q->parts + i);
        } else {
            putd(q->strike);
            putd(q->exercise);
            putd(q->settlement);
            putd(q->barrier);
            putd(q->maturity);
            putd(q->monitor_freq);
            putd(q->smooth);
            puti(q->call_put);
            putd(q->fixing_freq);
            putd(q->snap_freq);
        }
    };
    put_product(p);
    return k;
}

// mrg32k3a jump matrices (mrg32k3a.h:264-381): for j = 0..31 the two
// component matrices raised to dim * 2^j, modulo m1 and m2, [j][component][9].
vector<uint64_t> mrg_jumps(int dim)
{
    using u128 = unsigned __int128;
    const uint64_t m[2] = {4294967087ull, 4294944443ull};
    const uint64_t base[2][9] = {{0, 1403580ull, 4294967087ull - 810728ull, 1, 0, 0, 0, 1, 0},
                                 {527612ull, 0, 4294944443ull - 1370589ull, 1, 0, 0, 0, 1, 0}};
    auto mul = [](const uint64_t* a, const uint64_t* b, uint64_t mod, uint64_t* r) {
        uint64_t t[9];
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k) {
                u128 acc = 0;
                for (int l = 0; l < 3; ++l) acc += u128(a[3 * i + l]) * b[3 * l + k];
                t[3 * i + k] = uint64_t(acc % mod);
            }
        std::memcpy(r, t, sizeof t);
    };
    vector<uint64_t> out(32 * 2 * 9);
    for (int c = 0; c < 2; ++c) {
        // A^dim by squaring
        uint64_t p[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, a[9];
        std::memcpy(a, base[c], sizeof a);
        for (unsigned e = unsigned(std::max(dim, 0)); e; e >>= 1) {
            if (e & 1) mul(p, a, m[c], p);
            mul(a, a, m[c], a);
        }
        for (int j = 0; j < 32; ++j) {
            std::memcpy(&out[size_t(2 * j + c) * 9], p, sizeof p);
            mul(p, p, m[c], p);
        }
    }
    return out;
}

// Plans, tables and layout of a model and product; FULL layout when it fits,
// LEAN otherwise.
int call_block(const Engine& E, const cf_num* num, bool fp64)
{
    if (const int f = forced_block_threads()) return f;
    // measured on MI355X (C3 grid, 2^12..2^19 paths): 512-thread blocks up to
    // 8 waves per compute unit cut the FP32 AAD kernel from 0.16 to 0.12-0.13
    // ms and FP32 pricing from 0.076 to 0.067; FP64 AAD runs 0.20-0.21 ms on
    // 256 threads up to 4 waves per compute unit, against 0.34-0.36 on 1024
    const int64_t waves = (std::max<int64_t>(num->num_path, 1) + cfk::kWave - 1) / cfk::kWave;
    const int64_t per_cu = (waves + std::max(E.num_cus, 1) - 1) / std::max(E.num_cus, 1);
    int w = fp64 ? 4 : 8;
    while (w < 16 && w < per_cu) w *= 2;
    return w * cfk::kWave;
}

std::shared_ptr<Prepared> prepare(Engine& E, const cf_model* mdl, const cf_product* prd, bool fp64, int adj,
                                  int block)
{
    const bool adjoint = adj != cfk::ADJ_NONE;
    auto P = std::make_shared<Prepared>();
    P->pd = plan_product(prd);
    P->md = plan_model(mdl, P->pd);
    const ProductPlan& pd = P->pd;
    const ModelPlan& md = P->md;
    const bool dup = md.kind == CF_MODEL_DUPIRE;

    vector<float> knots_f;
    cfk::KnotTable<double> kt64;
    cfk::KnotTable<float> kt32;
    if (dup) {
        if (fp64) {
            kt64 = cfk::build_knot_table(md.log_spots.data(), md.n_spots);
            P->bt = {int(kt64.entries.size()), kt64.fix, kt64.invh, kt64.c0};
        } else {
            // FP32 state is log(S / S0): knots shifted by log(S0) in FP64 first
            const double ls0 = std::log(md.spot);
            knots_f.resize(md.n_spots);
            for (int j = 0; j < md.n_spots; ++j) knots_f[j] = float(md.log_spots[j] - ls0);
            kt32 = cfk::build_knot_table(knots_f.data(), md.n_spots);
            P->bt = {int(kt32.entries.size()), kt32.fix, double(kt32.invh), double(kt32.c0)};
        }
    }
    const bool book = kernel_product(pd) == cfk::PRD_BOOK;
    if (book) P->book = build_book(pd, md.s_t0);
    P->cfg = cfk::KernelConfig{fp64 ? 0 : 1, md.kind, kernel_product(pd), adjoint, false, adj == cfk::ADJ_RECOMPUTE};
    P->block = std::min(block, cfk::book_block_limit(P->cfg.product, fp64 ? 8 : 4, adj));
    const int waves = P->block / cfk::kWave;
    const int forced = forced_layout();
    auto lds_bytes = [&](int copies) {
        return cfk::smem_layout(md.n_steps, md.n_spots, int(pd.timeline.size()), pd.n_pay, waves, fp64 ? 8 : 4, dup,
                                pd.kind == CF_PRODUCT_EUROPEANS, adjoint, P->cfg.lean, dup ? P->bt.nb : 0, copies)
            .total;
    };
    P->ivbar_copies = 1;
    // the FULL checkpoint kernels keep each state's bracket in a byte, and the
    // FULL kernels' knot search reads at most two knots of a bucket
    const int lean_from = forced == 1 || (dup && (md.n_spots > cfk::kMaxStoredKnots || P->bt.fix > 2)) ? 1 : 0;
    for (int lean = lean_from; lean < 2; ++lean) {
        P->cfg.lean = lean != 0;
        P->smem = lds_bytes(1);
        if (P->smem <= size_t(E.max_lds)) break;
    }
    if (P->smem > size_t(E.max_lds))
        fail(CF_ERR_CAPACITY, "tables need " + std::to_string(P->smem) + " bytes of LDS per block, device allows " +
                                  std::to_string(E.max_lds));
    int rc = cfk::path_kernel_occupancy(P->cfg, P->block, P->smem, &P->per_cu);
    if (rc != 0) hip_check(hipError_t(rc), "occupancy");
    if (P->per_cu < 1) fail(CF_ERR_CAPACITY, "path kernel does not fit on a compute unit");
    // a second local-vol adjoint grid (odd lanes) when it costs no occupancy
    if (adjoint && dup && ivbar_copies_allowed()) {
        const size_t smem2 = lds_bytes(2);
        int per_cu2 = 0;
        if (smem2 <= size_t(E.max_lds)) {
            rc = cfk::path_kernel_occupancy(P->cfg, P->block, smem2, &per_cu2);
            if (rc != 0) hip_check(hipError_t(rc), "occupancy");
            if (per_cu2 >= P->per_cu) {
                P->smem = smem2;
                P->ivbar_copies = 2;
            }
        }
    }

    Staging st;
    P->o_ss = st.add(md.step_sample);
    if (dup) {
        // interval tables: interval c in [0, m] starts at knot c - 1 (the
        // first knot for c = 0) with inverse width 0 outside the knots; rows
        // padded with their first and last value (cfk::lv_cols)
        const int m = md.n_spots, mp = cfk::lv_cols(m);
        auto tables = [&](auto real, const vector<double>& knots_r) {
            using T = decltype(real);
            vector<T> kx(size_t(m) + 1), ih(size_t(m) + 1, T(0)), v(size_t(md.n_steps) * mp);
            kx[0] = T(knots_r[0]);
            for (int j = 0; j < m; ++j) kx[size_t(j) + 1] = T(knots_r[size_t(j)]);
            for (int j = 0; j + 1 < m; ++j) ih[size_t(j) + 1] = T(md.inv_dx[size_t(j)]);
            for (int i = 0; i < md.n_steps; ++i) {
                const double* r = &md.iv[size_t(i) * m];
                T* d = &v[size_t(i) * mp];
                d[0] = T(r[0]);
                for (int j = 0; j < m; ++j) d[j + 1] = T(r[j]);
                d[m + 1] = T(r[m - 1]);
            }
            P->o_knots = st.add(kx);
            P->o_invdx = st.add(ih);
            P->o_iv = st.add(v);
            // every node at least the recovery threshold: so is every
            // interpolated vol (a convex combination of two nodes), and the
            // reverse sweep never draws a normal again
            const T vmin = T(sizeof(T) == 8 ? cfk::kRecoverVolF64 : cfk::kRecoverVolF32);
            P->iv_recover = std::all_of(v.begin(), v.end(), [vmin](T y) { return std::isfinite(double(y)) && y >= vmin; });
        };
        if (fp64) {
            tables(double(), md.log_spots);
        } else {
            const vector<double> kf(knots_f.begin(), knots_f.end());   // exact: floats widen
            tables(float(), kf);
        }
    } else if (fp64) {
        P->o_drift = st.add(md.drift);
        P->o_std = st.add(md.stdv);
    } else {
        vector<float> a(md.drift.begin(), md.drift.end()), b(md.stdv.begin(), md.stdv.end());
        P->o_drift = st.add(a);
        P->o_std = st.add(b);
    }
    if (dup) {
        // bucket entries: the bucket's knot count and its next two knots; NaN
        // past the last knot, which no state (not even +inf) is >= to
        if (fp64) P->o_bucket = st.add(kt64.entries);
        else P->o_bucket = st.add(kt32.entries);
    }
    P->o_ff = st.add(md.ff);
    P->o_disc = st.add(md.disc);
    P->o_num = st.add(md.num);
    P->o_aux = st.add(md.bs_aux);
    if (pd.kind == CF_PRODUCT_EUROPEANS) {
        P->o_pf = st.add(pd.pay_first);
        P->o_pk = st.add(pd.pay_strike);
    }
    if (book) P->o_book = st.add(P->book.blob);
    {
        const vector<uint64_t> jump = mrg_jumps(md.n_steps);
        P->o_mrg = st.add(jump);
    }
    P->staged = std::move(st.bytes);

    if (dup && adjoint) {
        P->sqrtdt.resize(md.n_steps);
        P->tw.resize(md.n_steps);
        for (int i = 0; i < md.n_steps; ++i) {
            P->sqrtdt[i] = std::sqrt(md.timeline[i + 1] - md.timeline[i]);
            P->tw[i] = interp_weights(md.times.data(), md.n_times, md.timeline[i]);
        }
    }
    return P;
}

std::shared_ptr<Prepared> lookup(Engine& E, const cf_model* mdl, const cf_product* prd, bool fp64, int adj,
                                 bool with_tables, int block)
{
    if (!mdl) fail(CF_ERR_ARG, "null model");
    if (!prd) fail(CF_ERR_ARG, "null product");
    const string key = plan_key(mdl, prd, fp64 ? CF_FP64 : CF_FP32, adj, block);
    std::lock_guard<std::mutex> lk(E.plan_mu);
    std::shared_ptr<Prepared> P;
    for (auto it = E.plans.begin(); it != E.plans.end(); ++it) {
        if (it->first == key) {
            E.plans.splice(E.plans.begin(), E.plans, it);
            P = E.plans.front().second;
            break;
        }
    }
    if (!P) {
        P = prepare(E, mdl, prd, fp64, adj, block);
        E.plans.emplace_front(key, P);
        if (E.plans.size() > kMaxPlans) E.plans.pop_back();
    }
    if (with_tables && !P->tables.p) {
        P->tables.ensure(P->staged.size());
        hip_check(hipMemcpy(P->tables.p, P->staged.data(), P->staged.size(), hipMemcpyHostToDevice), "table upload");
    }
    return P;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

CallShape shape_call(const Engine& E, const Prepared& PR, const cf_num* num, bool adjoint)
{
    CallShape S;
    S.first = num->first_path;
    S.n_path = num->num_path;
    S.adjoint = adjoint;
    S.fp64 = PR.cfg.precision == 0;
    S.waves = PR.block / cfk::kWave;
    const int64_t q0 = (S.first + 1) >> 6, q1 = (S.first + S.n_path) >> 6;
    S.merge = cfk::merge_tail_blocks(S.first, S.n_path);
    const int64_t blocks = q1 - q0 + 1 - (S.merge ? 1 : 0);
    const int64_t want = (blocks + S.waves - 1) / S.waves;
    S.grid = int(std::min<int64_t>(want, int64_t(PR.per_cu) * E.num_cus));
    S.slots = S.grid * S.waves;
    S.npay = PR.pd.n_pay;
    S.stride = cfk::out_stride(S.npay);
    const bool dup = PR.md.kind == CF_MODEL_DUPIRE;
    S.ivn = adjoint && dup ? size_t(PR.md.n_steps) * PR.md.n_spots : 0;
    S.ck_rows = adjoint ? cfk::ckpt_rows(PR.md.n_steps, dup, PR.cfg.recompute, S.fp64 ? 8 : 4) : 0;
    S.n_steps = PR.md.n_steps;
    S.mrg = !num->use_sobol && PR.md.n_steps > 0;
    // the reference converts the int seeds to unsigned (main.h:55-57); states
    // are equal modulo m, so the seeds are reduced
    S.seed1 = uint64_t(unsigned(num->seed1)) % 4294967087ull;
    S.seed2 = uint64_t(unsigned(num->seed2)) % 4294944443ull;
    return S;
}

static size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

size_t CallBytes::total() const
{
    return al256(block_out) + al256(reduced) + al256(ivbar_blocks) + al256(ivbar) + al256(weights) + al256(ckpt);
}

CallBytes call_bytes(const CallShape& S)
{
    CallBytes b{};
    b.block_out = size_t(S.grid) * S.stride * 8;
    b.reduced = size_t(S.stride) * 8;
    b.ivbar_blocks = size_t(S.grid) * S.ivn * 8;
    b.ivbar = S.ivn * 8;
    b.weights = S.adjoint ? size_t(S.npay) * 8 : 0;
    b.ckpt = size_t(S.slots) * S.ck_rows * cfk::kWave * (S.fp64 ? 8 : 4);
    return b;
}

cfk::LaunchParams launch_params(const Engine& E, const Prepared& PR, const CallShape& S, const CallBuffers& B)
{
    const ProductPlan& pd = PR.pd;
    const ModelPlan& md = PR.md;
    const bool dup = md.kind == CF_MODEL_DUPIRE;
    const unsigned char* dt = B.tables;
    cfk::LaunchParams P;
    std::memset(&P, 0, sizeof P);   // padding too: sequences compare byte for byte
    P.first_path = S.first;
    P.n_path = S.n_path;
    P.merge_tail = S.merge ? 1 : 0;
    P.n_steps = md.n_steps;
    P.dirs = E.d_dirs;
    P.n_spots = md.n_spots;
    P.knots = dt + PR.o_knots;
    P.inv_dx = dt + PR.o_invdx;
    P.iv = dt + PR.o_iv;
    P.bucket = dt + PR.o_bucket;
    P.n_buckets = dup ? PR.bt.nb : 0;
    P.bucket_fix = PR.bt.fix;
    P.bucket_invh = PR.bt.invh;
    P.bucket_c0 = PR.bt.c0;
    P.drift = dt + PR.o_drift;
    P.stdv = dt + PR.o_std;
    P.spot = md.spot;
    P.x0 = S.fp64 ? std::log(md.spot) : 0.0;
    P.spot_measure = md.spot_measure ? 1 : 0;
    P.bs_vol = md.vol;
    P.bs_aux = reinterpret_cast<const double*>(dt + PR.o_aux);
    P.n_samples = int(pd.timeline.size());
    P.step_sample = reinterpret_cast<const int*>(dt + PR.o_ss);
    P.ff = reinterpret_cast<const double*>(dt + PR.o_ff);
    P.disc = reinterpret_cast<const double*>(dt + PR.o_disc);
    P.num = reinterpret_cast<const double*>(dt + PR.o_num);
    P.n_pay = pd.n_pay;
    P.strike = pd.strike;
    if (pd.kind == CF_PRODUCT_BARRIER) {
        // mcPrd.h:247-250: width from the spot written at t = 0, no derivative
        const double smooth = md.s_t0 * pd.smooth;
        P.two_smooth = 2 * smooth;
        P.inv_two_smooth = 1.0 / P.two_smooth;
        P.bar_smooth = pd.barrier + smooth;
        P.minus_smooth = pd.barrier - smooth;
        // the state is log S in FP64 and log(S / S0) in FP32; a state at or
        // below bar_xlo gives a spot under B - smooth after exp's rounding
        P.bar_xlo = -HUGE_VAL;
        if (P.minus_smooth > 0) {
            const double xl = S.fp64 ? std::log(P.minus_smooth) : std::log(P.minus_smooth / md.spot);
            P.bar_xlo = xl - (S.fp64 ? 1e-10 : 1e-4) * (1 + std::fabs(xl));
        }
    }
    P.put = pd.put ? 1 : 0;
    P.pay_first = reinterpret_cast<const int*>(dt + PR.o_pf);
    P.pay_strike = reinterpret_cast<const double*>(dt + PR.o_pk);
    P.book = dt + PR.o_book;
    P.bk = PR.book.L;
    P.book_lo = PR.book.lo_min;

    P.block_out = B.block_out;
    P.per_path = B.per_path;
    P.weights = B.weights;
    P.ckpt = B.ckpt;
    P.ivbar_blocks = B.ivbar_blocks;
    P.per_path_agg = B.per_path_agg;
    if (S.mrg) {
        P.mrg_jump = reinterpret_cast<const uint64_t*>(dt + PR.o_mrg);
        P.mrg_seed1 = uint32_t(S.seed1);
        P.mrg_seed2 = uint32_t(S.seed2);
    }
    P.ivbar_copies = PR.ivbar_copies;
    P.iv_recover = PR.iv_recover ? 1 : 0;
    P.rev_point = md.n_steps;
    return P;
}

void make_sequence(const Prepared& PR, const CallShape& S, const CallBuffers& B, const cfk::LaunchParams& P,
                   Sequence* q)
{
    std::memset(q, 0, sizeof *q);
    q->cfg = PR.cfg;
    q->cfg.mrg = S.mrg;
    std::memcpy(&q->P, &P, sizeof P);
    q->grid = S.grid;
    q->block = PR.block;
    q->smem = PR.smem;
    q->block_out = B.block_out;
    q->stride = S.stride;
    q->reduced = B.reduced;
    if (S.ivn) {
        q->ivn = int64_t(S.ivn);
        q->ivbar_blocks = B.ivbar_blocks;
        q->ivbar = B.ivbar;
    }
}

void ready_sequence(const Sequence& q)
{
    // the plan's kernel was set up at planning; the mrg32k3a and
    // itemized-risk variants of it are set up on first use (cached by the
    // launcher)
    if (!q.cfg.mrg && !q.cfg.item) return;
    int per_cu = 0;
    const int rc = cfk::path_kernel_occupancy(q.cfg, q.block, q.smem, &per_cu);
    if (rc != 0) hip_check(hipError_t(rc), "occupancy");
}

void enqueue_sequence(const Sequence& q, hipStream_t stream, hipEvent_t ev_kernel0, hipEvent_t ev_kernel1)
{
    if (ev_kernel0) hip_check(hipEventRecord(ev_kernel0, stream), "hipEventRecord");
    int rc = cfk::launch_path_kernel(q.cfg, q.P, q.grid, q.block, q.smem, stream);
    if (rc != 0) hip_check(hipError_t(rc), "path kernel launch");
    if (ev_kernel1) hip_check(hipEventRecord(ev_kernel1, stream), "hipEventRecord");
    rc = cfk::launch_reduce_rows(q.block_out, q.stride, q.reduced, q.ivbar_blocks, int(q.ivn), q.ivbar, q.grid,
                                 stream);
    if (rc != 0) hip_check(hipError_t(rc), "reduce launch");
}

void enqueue_call(const Prepared& PR, const CallShape& S, const CallBuffers& B, const cfk::LaunchParams& P,
                  hipStream_t stream, hipEvent_t ev_kernel0, hipEvent_t ev_kernel1)
{
    Sequence q;
    make_sequence(PR, S, B, P, &q);
    ready_sequence(q);
    enqueue_sequence(q, stream, ev_kernel0, ev_kernel1);
}

void collect_results(const Prepared& PR, const CallShape& S, const double* red, const double* ivbar, const Call& c)
{
    const int npay = S.npay;
    const double np = double(S.n_path);
    for (int k = 0; k < npay; ++k) {
        const double mean = red[k] / np;
        if (c.means) c.means[k] = mean;
        if (c.std_errs) {
            const double var = std::max(0.0, red[npay + k] / np - mean * mean);
            c.std_errs[k] = std::sqrt(var / np);
        }
    }
    if (!S.adjoint) return;
    if (c.agg_mean) c.agg_mean[0] = red[2 * npay] / np;
    const double* padj = red + 2 * npay + 2;
    const ModelPlan& md = PR.md;
    if (md.kind == CF_MODEL_DUPIRE) {
        const int nS = md.n_spots, nT = md.n_times;
        double* r = c.risks;
        std::fill(r, r + 1 + size_t(nS) * nT, 0.0);
        r[0] = padj[0] / np;
        // adjoint of interpVols[i][j] = sqrt(dt_i) * interp_t(vols[j], t_i)
        for (int i = 0; i < md.n_steps; ++i) {
            const Interp1& ip = PR.tw[i];
            for (int j = 0; j < nS; ++j) {
                const double a = ivbar[size_t(i) * nS + j] * PR.sqrtdt[i];
                if (a == 0.0) continue;
                r[1 + size_t(j) * nT + ip.j0] += a * ip.w0;
                if (ip.w1 != 0.0) r[1 + size_t(j) * nT + ip.j1] += a * ip.w1;
            }
        }
        for (size_t k = 1; k < 1 + size_t(nS) * nT; ++k) r[k] /= np;
    } else {
        for (int k = 0; k < 4; ++k) c.risks[k] = padj[k] / np;
    }
}

int adjoint_kind(const Call& c)
{
    if (!c.adjoint) return cfk::ADJ_NONE;
    return c.num->aad_mode == CF_AAD_RECOMPUTE ? cfk::ADJ_RECOMPUTE : cfk::ADJ_CHECKPOINT;
}

// Synchronous call on the engine's own stream and buffers.
int run(const cf_model* mdl, const cf_product* prd, Call& c)
{
    const auto t_start = std::chrono::steady_clock::now();
    validate_num(c.num);
    if (c.adjoint && !c.weights) fail(CF_ERR_ARG, "null aggregation weights");
    const bool fp64 = c.num->precision == CF_FP64;

    Engine& E = engine();
    std::lock_guard<std::mutex> lk(E.sync_mu);
    hip_check(hipSetDevice(E.device), "hipSetDevice");
    const std::shared_ptr<Prepared> PP = lookup(E, mdl, prd, fp64, adjoint_kind(c), true, call_block(E, c.num, fp64));
    const Prepared& PR = *PP;
    const CallShape S = shape_call(E, PR, c.num, c.adjoint);
    const CallBytes nb = call_bytes(S);

    double* h_in = static_cast<double*>(E.pin_in.ensure(size_t(S.npay) * 8));
    double* h_out = static_cast<double*>(E.pin_out.ensure((size_t(S.stride) + S.ivn) * 8));
    CallBuffers B;
    B.tables = static_cast<const unsigned char*>(PR.tables.p);
    B.block_out = static_cast<double*>(E.block_out.ensure(nb.block_out));
    // the reduced sums and the local-vol adjoints side by side, written by the
    // reduction straight into mapped host memory
    B.reduced = static_cast<double*>(E.pin_out.dev);
    if (S.ivn) {
        B.ivbar_blocks = static_cast<double*>(E.ivbar_blocks.ensure(nb.ivbar_blocks));
        B.ivbar = B.reduced + S.stride;
    }
    if (c.adjoint) {
        B.weights = static_cast<double*>(E.weights.ensure(nb.weights));
        B.ckpt = E.ckpt.ensure(nb.ckpt);
    }
    if (c.per_path) B.per_path = static_cast<double*>(E.per_path.ensure(size_t(S.n_path) * S.npay * 8));
    if (c.per_path_agg) B.per_path_agg = static_cast<double*>(E.per_path_agg.ensure(size_t(S.n_path) * 8));

    hip_check(hipEventRecord(E.ev[0], E.stream), "hipEventRecord");
    if (c.adjoint) {
        // the device copy of the weights is reused while the call's weights are
        // the last ones uploaded to the same buffer (only this path writes it)
        const size_t wb = size_t(S.npay) * 8;
        const bool same = E.weights_on_device == B.weights && E.weights_last.size() == size_t(S.npay) &&
                          std::memcmp(E.weights_last.data(), c.weights, wb) == 0;
        if (!same) {
            std::memcpy(h_in, c.weights, wb);
            hip_check(hipMemcpyAsync(B.weights, h_in, wb, hipMemcpyHostToDevice, E.stream), "upload");
            E.weights_last.assign(c.weights, c.weights + S.npay);
            E.weights_on_device = B.weights;
        }
    }
    const cfk::LaunchParams P = launch_params(E, PR, S, B);
    const double setup_ms = ms_since(t_start);
    enqueue_call(PR, S, B, P, E.stream, E.ev[1], E.ev[2]);
    hip_check(hipEventRecord(E.ev[3], E.stream), "hipEventRecord");

    if (B.per_path)
        hip_check(hipMemcpyAsync(c.per_path, B.per_path, size_t(S.n_path) * S.npay * 8, hipMemcpyDeviceToHost,
                                 E.stream), "download");
    if (B.per_path_agg)
        hip_check(hipMemcpyAsync(c.per_path_agg, B.per_path_agg, size_t(S.n_path) * 8, hipMemcpyDeviceToHost,
                                 E.stream), "download");
    const double queued_ms = ms_since(t_start);
    hip_check(hipStreamSynchronize(E.stream), "kernel execution");
    const double waited_ms = ms_since(t_start);

    float t_main = 0, t_red = 0, t_all = 0;
    hip_check(hipEventElapsedTime(&t_main, E.ev[1], E.ev[2]), "event");
    hip_check(hipEventElapsedTime(&t_red, E.ev[2], E.ev[3]), "event");
    hip_check(hipEventElapsedTime(&t_all, E.ev[0], E.ev[3]), "event");

    collect_results(PR, S, h_out, h_out + S.stride, c);

    g_stats.main_kernel_ms = t_main;
    g_stats.reduce_ms = t_red;
    g_stats.total_ms = t_all;
    g_stats.grid_blocks = S.grid;
    g_stats.block_threads = PR.block;
    g_stats.lds_bytes = int(PR.smem);
    g_stats.path_steps = S.n_path * int64_t(PR.md.n_steps);
    g_stats.setup_ms = setup_ms;
    g_stats.host_ms = ms_since(t_start);
    g_stats.lean_layout = PR.cfg.lean ? 1 : 0;
    g_stats.enqueue_ms = queued_ms - setup_ms;
    g_stats.wait_ms = waited_ms - queued_ms;
    return CF_OK;
}

}  // namespace cfh

using namespace cfh;

void cfk::set_last_error(const char* msg) { g_err = msg; }

// ------------------------------------------------------------------ C ABI

extern "C" {

const char* cf_last_error(void) { return g_err.c_str(); }

const char* cf_version(void) { return "compfinance-amd 0.1 (gfx950)"; }

int cf_sobol_points(int64_t first_path, int64_t n_paths, int dim, int precision, int layout, uint32_t* words,
                    void* normals)
{
    return guarded([&] {
        if (first_path < 0 || n_paths < 1 || first_path + n_paths > int64_t(0xffffffff))
            fail(CF_ERR_ARG, "Sobol path range out of range");
        if (dim < 1 || dim > cfk::kSobolDims) fail(CF_ERR_ARG, "Sobol dimension out of range");
        if (precision != CF_FP64 && precision != CF_FP32) fail(CF_ERR_ARG, "unknown precision");
        if (layout < -1 || layout > 2) fail(CF_ERR_ARG, "unknown layout");
        Engine& E = engine();
        const int rb = precision == CF_FP64 ? 8 : 4;
        if (layout == -1) layout = cfk::sobol_points_smem(rb, false, dim, 256) > size_t(E.max_lds) ? 1 : 0;
        const bool lean = layout == 1;
        if (cfk::sobol_points_smem(rb, lean, dim, 256) > size_t(E.max_lds))
            fail(CF_ERR_CAPACITY, "Sobol tables do not fit the block's shared memory");
        const size_t n = size_t(n_paths) * size_t(dim);
        std::lock_guard<std::mutex> lk(E.sync_mu);
        struct Owned : DevBuf {
            ~Owned() { if (p) (void)hipFree(p); }
        } dw, dn;
        if (words) dw.ensure(n * 4);
        if (normals) dn.ensure(n * size_t(rb));
        const int rc = cfk::launch_sobol_points(rb, layout, E.d_dirs, dim, first_path, n_paths,
                                                static_cast<uint32_t*>(dw.p), dn.p, E.stream);
        if (rc != 0) hip_check(hipError_t(rc), "Sobol points launch");
        if (words) hip_check(hipMemcpyAsync(words, dw.p, n * 4, hipMemcpyDeviceToHost, E.stream), "copy");
        if (normals) hip_check(hipMemcpyAsync(normals, dn.p, n * size_t(rb), hipMemcpyDeviceToHost, E.stream), "copy");
        hip_check(hipStreamSynchronize(E.stream), "Sobol points");
        return CF_OK;
    });
}

int cf_device_synchronize(void)
{
    return guarded([&] {
        hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return CF_OK;
    });
}

int cf_last_kernel_stats(cf_kernel_stats* out)
{
    if (!out) {
        g_err = "null stats";
        return CF_ERR_ARG;
    }
    *out = g_stats;
    return CF_OK;
}

int cf_num_payoffs(const cf_product* prd)
{
    return guarded([&] { return plan_product(prd).n_pay; });
}

int cf_num_params(const cf_model* mdl)
{
    return guarded([&] {
        validate_model(mdl);
        return mdl->kind == CF_MODEL_BLACK_SCHOLES ? 4 : 1 + mdl->n_spots * mdl->n_times;
    });
}

static int copy_label(const string& s, char* buf, size_t cap)
{
    if (!buf || cap == 0) fail(CF_ERR_ARG, "null label buffer");
    const size_t n = std::min(s.size(), cap - 1);
    std::memcpy(buf, s.data(), n);
    buf[n] = 0;
    return int(s.size());
}

int cf_payoff_label(const cf_product* prd, int k, char* buf, size_t cap)
{
    return guarded([&] {
        const auto labels = payoff_labels(plan_product(prd));
        if (k < 0 || k >= int(labels.size())) fail(CF_ERR_ARG, "payoff index out of range");
        return copy_label(labels[k], buf, cap);
    });
}

int cf_param_label(const cf_model* mdl, int k, char* buf, size_t cap)
{
    return guarded([&] {
        validate_model(mdl);
        const auto labels = param_labels(mdl);
        if (k < 0 || k >= int(labels.size())) fail(CF_ERR_ARG, "parameter index out of range");
        return copy_label(labels[k], buf, cap);
    });
}

int cf_product_timeline(const cf_product* prd, double* times, int cap)
{
    return guarded([&] {
        const auto p = plan_product(prd);
        if (cap > 0 && !times) fail(CF_ERR_ARG, "null times");
        for (int i = 0; i < int(p.timeline.size()) && i < cap; ++i) times[i] = p.timeline[i];
        return int(p.timeline.size());
    });
}

int cf_sim_timeline(const cf_model* mdl, const cf_product* prd, double* times, int* common, int cap)
{
    return guarded([&] {
        const auto p = plan_product(prd);
        const auto m = plan_model(mdl, p);
        for (int i = 0; i < int(m.timeline.size()) && i < cap; ++i) {
            if (times) times[i] = m.timeline[i];
            if (common) common[i] = m.step_sample[i] >= 0 ? 1 : 0;
        }
        return int(m.timeline.size());
    });
}

int cf_mc_simul(const cf_model* mdl, const cf_product* prd, const cf_num* num, double* means, double* std_errs,
                double* per_path)
{
    return guarded([&] {
        Call c;
        c.num = num;
        c.means = means;
        c.std_errs = std_errs;
        c.per_path = per_path;
        if (!means) fail(CF_ERR_ARG, "null means");
        return run(mdl, prd, c);
    });
}

int cf_mc_simul_aad(const cf_model* mdl, const cf_product* prd, const cf_num* num, const double* weights,
                    double* payoff_means, double* agg_mean, double* risks, double* per_path_payoffs,
                    double* per_path_agg)
{
    return guarded([&] {
        if (!risks) fail(CF_ERR_ARG, "null risks");
        Call c;
        c.adjoint = true;
        c.num = num;
        c.weights = weights;
        c.means = payoff_means;
        c.agg_mean = agg_mean;
        c.risks = risks;
        c.per_path = per_path_payoffs;
        c.per_path_agg = per_path_agg;
        return run(mdl, prd, c);
    });
}

namespace cfh {

// Time of the last observation of each payoff column, in column order: a
// European of a Europeans product at its maturity, any other payoff at its
// product's last date (its value depends on no later state).
static void column_last_times(const ProductPlan& d, vector<double>& out)
{
    if (d.kind == CF_PRODUCT_EUROPEANS) {
        for (size_t s = 0; s + 1 < d.pay_first.size(); ++s)
            for (int c = d.pay_first[s]; c < d.pay_first[s + 1]; ++c) out.push_back(d.timeline[s]);
    } else if (d.kind == CF_PRODUCT_BOOK) {
        for (const ProductPlan& q : d.parts) column_last_times(q, out);
    } else {
        for (int c = 0; c < d.n_pay; ++c) out.push_back(d.timeline.back());
    }
}

// The simulation timeline point of each payoff column's last observation
// (n_steps where it is not found, which sweeps every step).
static vector<int> payoff_last_points(const Prepared& PR)
{
    vector<double> t;
    column_last_times(PR.pd, t);
    const vector<double>& tl = PR.pd.timeline;
    const vector<int>& ss = PR.md.step_sample;
    vector<int> pts(t.size(), PR.md.n_steps);
    for (size_t c = 0; c < t.size(); ++c) {
        const auto it = std::lower_bound(tl.begin(), tl.end(), t[c]);
        if (it == tl.end() || *it != t[c]) continue;
        const int s = int(it - tl.begin());
        for (int ip = 0; ip < int(ss.size()); ++ip)
            if (ss[size_t(ip)] == s) pts[c] = std::min(ip, PR.md.n_steps);
    }
    return pts;
}

// Itemized risk over shared paths (the reference's single sweep with vector
// adjoints, mcBase.h:858-989): one forward-only launch keeps every block of 64
// paths' checkpoints and end states, then one reverse-only launch per payoff
// (one-hot weights) reads them.  Returns false, doing nothing, when the kept
// checkpoints do not fit the device's free memory.
bool run_multi(const cf_model* mdl, const cf_product* prd, const cf_num* num, double* payoff_means, double* risks)
{
    const bool fp64 = num->precision == CF_FP64;
    Engine& E = engine();
    std::lock_guard<std::mutex> lk(E.sync_mu);
    hip_check(hipSetDevice(E.device), "hipSetDevice");
    const std::shared_ptr<Prepared> PP =
        lookup(E, mdl, prd, fp64, cfk::ADJ_CHECKPOINT, true, call_block(E, num, fp64));
    const Prepared& PR = *PP;
    const CallShape S = shape_call(E, PR, num, true);
    const CallBytes nb = call_bytes(S);
    const int64_t q0 = (S.first + 1) >> 6, q1 = (S.first + S.n_path) >> 6;
    const size_t blocks64 = size_t(q1 - q0 + 1 - (S.merge ? 1 : 0));
    const size_t rb = fp64 ? 8 : 4;
    const size_t ck_bytes = blocks64 * size_t(S.ck_rows) * cfk::kWave * rb;
    const size_t st_bytes = blocks64 * size_t(cfk::state_rows(PR.cfg.product == cfk::PRD_BOOK)) * cfk::kWave * rb;
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    if (ck_bytes + st_bytes + (size_t(256) << 20) > free_b) return false;

    struct Owned : DevBuf {
        ~Owned() { if (p) (void)hipFree(p); }
    } ckb, stb;
    ckb.ensure(ck_bytes);
    stb.ensure(st_bytes);
    CallBuffers B;
    B.tables = static_cast<const unsigned char*>(PR.tables.p);
    B.block_out = static_cast<double*>(E.block_out.ensure(nb.block_out));
    B.reduced = static_cast<double*>(E.reduced.ensure((size_t(S.stride) + S.ivn) * 8));
    if (S.ivn) {
        B.ivbar_blocks = static_cast<double*>(E.ivbar_blocks.ensure(nb.ivbar_blocks));
        B.ivbar = B.reduced + S.stride;
    }
    B.weights = static_cast<double*>(E.weights.ensure(nb.weights));
    E.weights_on_device = nullptr;   // this call rewrites the weights buffer
    B.ckpt = ckb.p;
    cfk::LaunchParams P = launch_params(E, PR, S, B);
    P.state = stb.p;

    const int npay = S.npay;
    const int npar = PR.md.kind == CF_MODEL_DUPIRE ? 1 + PR.md.n_spots * PR.md.n_times : 4;
    vector<double> w(static_cast<size_t>(npay), 0.0), red(static_cast<size_t>(S.stride) + S.ivn),
        col(static_cast<size_t>(npar)), agg(1);
    // each payoff's reverse starts at its last observation
    const vector<int> last_pt = payoff_last_points(PR);
    auto launch = [&](int mode, int k) {
        std::fill(w.begin(), w.end(), 0.0);
        w[size_t(k)] = 1.0;
        hip_check(hipMemcpyAsync(B.weights, w.data(), size_t(npay) * 8, hipMemcpyHostToDevice, E.stream), "upload");
        P.ck_mode = mode;
        P.rev_point = mode == 2 ? last_pt[size_t(k)] : PR.md.n_steps;
        Sequence q;
        make_sequence(PR, S, B, P, &q);
        q.cfg.item = true;
        ready_sequence(q);
        enqueue_sequence(q, E.stream, nullptr, nullptr);
        hip_check(hipMemcpyAsync(red.data(), B.reduced, red.size() * 8, hipMemcpyDeviceToHost, E.stream), "download");
        hip_check(hipStreamSynchronize(E.stream), "kernel execution");
    };
    // the forward: payoff means (no adjoint results yet)
    launch(1, 0);
    {
        CallShape S_fwd = S;
        S_fwd.adjoint = false;
        Call c;
        c.num = num;
        c.means = payoff_means;
        collect_results(PR, S_fwd, red.data(), red.data() + S.stride, c);
    }
    // one reverse per payoff
    for (int k = 0; k < npay; ++k) {
        launch(2, k);
        Call c;
        c.adjoint = true;
        c.num = num;
        c.agg_mean = agg.data();
        c.risks = col.data();
        collect_results(PR, S, red.data(), red.data() + S.stride, c);
        for (int i = 0; i < npar; ++i) risks[size_t(i) * npay + k] = col[size_t(i)];
    }
    return true;
}

}  // namespace cfh

int cf_mc_simul_aad_multi(const cf_model* mdl, const cf_product* prd, const cf_num* num, double* payoff_means,
                          double* risks)
{
    return guarded([&] {
        if (!risks) fail(CF_ERR_ARG, "null risks");
        validate_num(num);
        validate_model(mdl);
        if (num->aad_mode == CF_AAD_CHECKPOINT && run_multi(mdl, prd, num, payoff_means, risks))
            return mdl->kind == CF_MODEL_BLACK_SCHOLES ? 4 : 1 + mdl->n_spots * mdl->n_times;
        const int npay = plan_product(prd).n_pay;
        validate_model(mdl);
        const int npar = mdl->kind == CF_MODEL_BLACK_SCHOLES ? 4 : 1 + mdl->n_spots * mdl->n_times;
        // one adjoint sweep per payoff over the same paths (one-hot weights)
        vector<double> w(size_t(npay), 0.0), col(static_cast<size_t>(npar), 0.0), agg(1, 0.0);
        for (int k = 0; k < npay; ++k) {
            std::fill(w.begin(), w.end(), 0.0);
            w[k] = 1.0;
            Call c;
            c.adjoint = true;
            c.num = num;
            c.weights = w.data();
            c.means = k == 0 ? payoff_means : nullptr;
            c.agg_mean = agg.data();
            c.risks = col.data();
            run(mdl, prd, c);
            for (int i = 0; i < npar; ++i) risks[size_t(i) * npay + k] = col[i];
        }
        return npar;
    });
}

int64_t cf_aad_footprint(const cf_model* mdl, const cf_product* prd, const cf_num* num)
{
    int64_t bytes = 0;
    const int rc = guarded([&] {
        validate_num(num);
        Engine& E = engine();
        const bool fp64 = num->precision == CF_FP64;
        const int adj = num->aad_mode == CF_AAD_RECOMPUTE ? cfk::ADJ_RECOMPUTE : cfk::ADJ_CHECKPOINT;
        const std::shared_ptr<Prepared> P = lookup(E, mdl, prd, fp64, adj, false, call_block(E, num, fp64));
        bytes = int64_t(call_bytes(shape_call(E, *P, num, true)).total() + P->staged.size());
        return CF_OK;
    });
    return rc == CF_OK ? bytes : int64_t(rc);
}

}  // extern "C"
