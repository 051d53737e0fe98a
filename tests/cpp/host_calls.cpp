// Host-only paths of the library's C ABI (include/cfamd.h), for a build of
// the host sources under AddressSanitizer and UndefinedBehaviorSanitizer:
// labels, timelines and counts of every model and product kind, planning and
// validation errors, the store (put, value errors, clear), the FP64 local-vol
// calibration and the superbucket chain, and the entry points that need a
// device failing cleanly without one.  Every call's status is checked; the
// program prints "ok" and exits 0, or names the first unexpected result.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cfamd.h"

static int g_bad = 0;
#define EXPECT(c)                                                               \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::printf("FAILED %s:%d %s (%s)\n", __FILE__, __LINE__, #c,       \
                        cf_last_error());                                       \
            ++g_bad;                                                            \
        }                                                                       \
    } while (0)

static cf_model dupire(std::vector<double>& spots, std::vector<double>& times, std::vector<double>& vols, int ns,
                       int nt)
{
    spots.resize(ns);
    times.resize(nt);
    vols.resize(size_t(ns) * nt);
    for (int i = 0; i < ns; ++i) spots[i] = 50.0 + 150.0 * i / (ns > 1 ? ns - 1 : 1);
    for (int j = 0; j < nt; ++j) times[j] = (j + 1) / 12.0;
    for (int i = 0; i < ns; ++i)
        for (int j = 0; j < nt; ++j) vols[size_t(i) * nt + j] = 0.15 + 0.01 * i / ns + 0.05 * j / nt;
    cf_model m{};
    m.kind = CF_MODEL_DUPIRE;
    m.spot = 100.0;
    m.n_spots = ns;
    m.n_times = nt;
    m.spots = spots.data();
    m.times = times.data();
    m.vols = vols.data();
    m.max_dt = 1.0 / 52;
    return m;
}

static void describe(const cf_model* m, const cf_product* p)
{
    const int npay = cf_num_payoffs(p), npar = cf_num_params(m);
    EXPECT(npay > 0);
    EXPECT(npar > 0);
    char buf[128];
    for (int k = 0; k < npay; ++k) EXPECT(cf_payoff_label(p, k, buf, sizeof buf) > 0);
    for (int k = 0; k < npar; k += 7) EXPECT(cf_param_label(m, k, buf, sizeof buf) > 0);
    EXPECT(cf_payoff_label(p, npay, buf, sizeof buf) < 0);          // out of range
    cf_payoff_label(p, 0, buf, 2);                                   // short buffer
    std::vector<double> t(2048);
    std::vector<int> common(2048);
    const int nt = cf_product_timeline(p, t.data(), int(t.size()));
    EXPECT(nt > 0);
    const int ns = cf_sim_timeline(m, p, t.data(), common.data(), int(t.size()));
    EXPECT(ns >= nt);
    cf_sim_timeline(m, p, t.data(), common.data(), 1);               // short buffer
}

int main()
{
    std::vector<double> spots, times, vols;
    cf_model lv = dupire(spots, times, vols, 30, 36);
    cf_model bs{};
    bs.kind = CF_MODEL_BLACK_SCHOLES;
    bs.spot = 100.0;
    bs.vol = 0.2;
    bs.rate = 0.03;
    bs.div = 0.01;

    cf_product eu{};
    eu.kind = CF_PRODUCT_EUROPEAN;
    eu.strike = 100.0;
    eu.exercise = 1.0;
    eu.settlement = 1.0;
    cf_product uoc{};
    uoc.kind = CF_PRODUCT_BARRIER;
    uoc.strike = 120.0;
    uoc.barrier = 150.0;
    uoc.maturity = 3.0;
    uoc.monitor_freq = 1.0 / 52;
    uoc.smooth = 0.05;
    double mats[3] = {1.0, 2.0, 2.0}, ks[3] = {90.0, 100.0, 110.0};
    cf_product eus{};
    eus.kind = CF_PRODUCT_EUROPEANS;
    eus.n_options = 3;
    eus.maturities = mats;
    eus.strikes = ks;
    cf_product asian{};
    asian.kind = CF_PRODUCT_ASIAN;
    asian.strike = 100.0;
    asian.maturity = 3.0;
    asian.fixing_freq = 1.0 / 12;
    asian.snap_freq = 1.0 / 52;
    cf_product parts[4] = {eu, uoc, eus, asian};
    cf_product book{};
    book.kind = CF_PRODUCT_BOOK;
    book.n_parts = 4;
    book.parts = parts;

    for (const cf_model* m : {&lv, &bs})
        for (const cf_product* p : {&eu, &uoc, &eus, &asian, &book}) describe(m, p);
    std::vector<double> s60, t60, v60;
    cf_model lv60 = dupire(s60, t60, v60, 60, 60);
    describe(&lv60, &uoc);

    // validation errors
    cf_model bad = lv;
    bad.n_spots = 0;
    EXPECT(cf_num_params(&bad) < 0);
    cf_product badp = uoc;
    badp.maturity = -1.0;
    EXPECT(cf_num_payoffs(&badp) < 0 || cf_product_timeline(&badp, nullptr, 0) < 0);
    EXPECT(cf_num_payoffs(nullptr) < 0);
    cf_product badb = book;
    badb.n_parts = 0;
    EXPECT(cf_num_payoffs(&badb) < 0);

    // the store
    EXPECT(cf_put_black_scholes(100.0, 0.2, 0, 0.0, 0.0, "bs") == 0);
    EXPECT(cf_put_dupire(100.0, spots.data(), 30, times.data(), 36, vols.data(), 1.0 / 52, "lv") == 0);
    EXPECT(cf_put_european(100.0, 1.0, 1.0, "eu") == 0);
    EXPECT(cf_put_barrier(120.0, 150.0, 3.0, 1.0 / 52, 0.05, 0, "uoc") == 0);
    EXPECT(cf_put_europeans(mats, ks, 3, "eus") == 0);
    EXPECT(cf_put_asian(100.0, 3.0, 1.0 / 12, 1.0 / 52, 0, "asian") == 0);
    const char* ids[] = {"eu", "uoc", "eus", "asian"};
    EXPECT(cf_put_book(ids, 4, "book") == 0);
    const char* missing[] = {"eu", "nope"};
    EXPECT(cf_put_book(missing, 2, "book2") < 0);
    cf_num num{};
    num.parallel = 1;
    num.use_sobol = 1;
    num.num_path = 1024;
    double vals[64];
    EXPECT(cf_value("nope", "eu", &num, vals, 64) < 0);               // unknown model
    EXPECT(cf_value("bs", "nope", &num, vals, 64) < 0);               // unknown product
    num.num_path = 0;
    EXPECT(cf_value("bs", "eu", &num, vals, 64) < 0);                 // bad path count
    cf_store_clear();
    EXPECT(cf_value("bs", "eu", &num, vals, 64) < 0);                 // cleared

    // FP64 calibration and the superbucket chain (host)
    cf_merton_ivs ivs{100.0, 0.15, 0.05, -0.15, 0.10};
    double incl_s[2] = {80.0, 120.0}, incl_t[2] = {0.5, 1.0};
    int ns = 0, nt = 0;
    EXPECT(cf_dupire_calib(&ivs, incl_s, 2, 5.0, incl_t, 2, 0.25, nullptr, nullptr, nullptr, &ns, &nt, 0) == 0);
    EXPECT(ns > 0 && nt > 0);
    std::vector<double> cs(ns), ct(nt), clv(size_t(ns) * nt);
    EXPECT(cf_dupire_calib(&ivs, incl_s, 2, 5.0, incl_t, 2, 0.25, cs.data(), ct.data(), clv.data(), &ns, &nt,
                           ns * nt) == 0);
    for (double v : clv) EXPECT(v > 0.0 && v < 2.0);
    std::vector<double> micro(size_t(ns) * nt, 1e-3), vega(6);
    double sk[3] = {90.0, 100.0, 110.0}, sm[2] = {0.5, 1.0};
    EXPECT(cf_superbucket_vega(&ivs, incl_s, 2, 5.0, incl_t, 2, 0.25, sk, 3, sm, 2, micro.data(), ns * nt,
                               vega.data()) == 0);
    EXPECT(cf_superbucket_vega(&ivs, incl_s, 2, 5.0, incl_t, 2, 0.25, sk, 3, sm, 2, micro.data(), ns * nt - 1,
                               vega.data()) < 0);                     // wrong microbucket size

    // entry points that need a device fail cleanly without one (or run with one)
    num.num_path = 1024;
    std::vector<double> means(8), risks(1200);
    double agg = 0.0, w[2] = {1.0, 0.0};
    cf_mc_simul_aad(&lv, &uoc, &num, w, means.data(), &agg, risks.data(), nullptr, nullptr);
    EXPECT(cf_mc_simul(&lv, &uoc, nullptr, means.data(), nullptr, nullptr) < 0);   // null parameters
    cf_server* srv = nullptr;
    cf_server_opts opts{4, int64_t(1) << 26, 1, 1};
    if (cf_server_create(&opts, &srv) == 0) EXPECT(cf_server_destroy(srv) == 0);
    opts.slots = 0;
    EXPECT(cf_server_create(&opts, &srv) < 0);                        // bad options

    if (g_bad) return 1;
    std::printf("ok\n");
    return 0;
}
