// A reference-style caller compiled against the C++ mirror headers and
// libcfamd.so: the store, the main.h entry points, the simulation templates,
// the calibration, bump and reprice.  `host` exercises what runs without a GPU (labels,
// timelines, calibration, errors); `device` runs the simulations.  Prints one
// JSON object for the test to compare.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "compfinance_amd/main.h"

using namespace cfamd;

static void printVec(const char* name, const std::vector<double>& v, bool last = false)
{
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char** argv)
{
    const std::string mode = argc > 1 ? argv[1] : "host";
    std::vector<double> spots, times;
    for (int i = 0; i < 30; ++i) spots.push_back(50.0 + 150.0 * i / 29);
    for (int j = 0; j < 36; ++j) times.push_back((j + 1) / 12.0);
    matrix<double> vols(spots.size(), times.size());
    for (size_t i = 0; i < spots.size(); ++i)
        for (size_t j = 0; j < times.size(); ++j) {
            const double l = std::log(spots[i] / 100.0);
            vols[i][j] = 0.15 + 0.10 * l * l + 0.05 * std::exp(-times[j]);
        }
    putDupire(100.0, spots, times, vols, 1.0 / 52, "lv");
    putBarrier(120.0, 150.0, 3.0, 1.0 / 52, 0.05, false, "uoc");
    putEuropeans({3.0, 3.0, 2.0}, {100.0, 120.0, 110.0}, "eus");
    putAsian(100.0, 3.0, 1.0 / 12, 1.0 / 52, false, "asian");
    putBook({"eus", "uoc", "asian"}, "book");

    std::printf("{\n");
    const auto& labels = *getPayoffLabels("book");
    std::printf("\"book_labels\": [");
    for (size_t i = 0; i < labels.size(); ++i) std::printf("%s\"%s\"", i ? ", " : "", labels[i].c_str());
    std::printf("],\n");
    printVec("book_timeline", getProduct<double>("book")->timeline());
    auto cal = dupireCalib({50.0, 100.0, 200.0}, 5.0, {0.25, 1.0, 3.0}, 0.1, 100.0, 0.15, 0.05, -0.15, 0.1);
    printVec("calib_lvols", std::vector<double>(cal.lVols.begin(), cal.lVols.end()));
    std::string err;
    try {
        value("nope", "uoc", NumericalParam{});
    } catch (const std::exception& e) {
        err = e.what();
    }
    std::printf("\"missing_model_error\": \"%s\"", err.c_str());

    if (mode == "device") {
        NumericalParam num;
        num.numPath = 1 << 14;
        auto v = value("lv", "uoc", num);
        std::printf(",\n");
        printVec("uoc_values", v.values);
        auto r = AADriskOne("lv", "uoc", num);
        printVec("uoc_risks", r.risks);
        auto a = AADriskAggregate("lv", "book", {{labels[0], 1.0}, {labels[3], -0.5}, {labels[5], 2.0}}, num);
        printVec("book_risks", a.risks);
        // bump and reprice (main.h:316-359): the risks of the barrier payoff
        auto br = bumpRisk("lv", "uoc", num);
        std::vector<double> bump0;
        for (size_t i = 0; i < br.risks.rows(); ++i) bump0.push_back(br.risks[i][0]);
        printVec("bump_risks", bump0);
        // the template level, as mcBase.h callers use it
        const Model<Number>* mdl = getModel<Number>("lv");
        const Product<Number>* prd = getProduct<Number>("uoc");
        Sobol rng;
        auto sim = mcParallelSimulAAD(*prd, *mdl, rng, 1 << 14,
                                      [](const vector<Number>& v) { return v[0]; });
        printVec("template_risks", sim.risks, true);
    } else {
        std::printf("\n");
    }
    std::printf("}\n");
    return 0;
}
