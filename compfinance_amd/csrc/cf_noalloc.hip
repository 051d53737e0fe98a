// cf_mc_simul_aad_noalloc: the AAD simulation of cf_mc_simul_aad for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  Every device buffer of the
// call (the plan's tables, the weights, per-block partial sums, local-vol
// adjoint grids, checkpoints, reduced sums, optional per-path outputs) is
// carved from one module array, `scratch`, of kNoAllocScratchBytes per device.
// No hipMalloc, hipFree, pinned allocation or allocating library routine is
// called: uploads and downloads go straight between the caller's host memory
// and the array.
//
// The array is per device and per process and keeps the last call's contents.
// Nothing in it has to start at zero: every kernel writes each partial sum and
// adjoint grid it later reads.  Calls of one process on one device take the
// engine's lock and run one after the other; without that lock two calls on
// one device would race on the array.
#include <hip/hip_runtime.h>

#include <cstring>
#include <list>
#include <map>
#include <mutex>

#include "../../include/cfamd.h"
#include "cf_host.h"
#include "cf_internal.h"

namespace {

constexpr size_t kNoAllocScratchBytes = size_t(CF_NOALLOC_SCRATCH_MIB) << 20;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Address of this device's copy of `scratch`, looked up once per device with
// the device current.  The host-side name `scratch` is not a device address.
std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch_ptr;

unsigned char* scratch_on(int dev)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_ptr.find(dev);
    if (it != g_scratch_ptr.end()) return it->second;
    void* p = nullptr;
    cfh::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    return static_cast<unsigned char*>(p);
}

// Host plans of this entry point: kept apart from the engine's cache, whose
// plans own device tables (evicting one frees them).  These own none.
constexpr size_t kMaxNoAllocPlans = 16;
std::mutex g_plans_mu;
std::list<std::pair<std::string, std::shared_ptr<cfh::Prepared>>> g_plans;

std::shared_ptr<cfh::Prepared> host_plan(cfh::Engine& E, const cf_model* mdl, const cf_product* prd, bool fp64,
                                         int adj, int block)
{
    if (!mdl) cfh::fail(CF_ERR_ARG, "null model");
    if (!prd) cfh::fail(CF_ERR_ARG, "null product");
    const std::string key = cfh::plan_key(mdl, prd, fp64 ? CF_FP64 : CF_FP32, adj, block);
    std::lock_guard<std::mutex> lk(g_plans_mu);
    for (auto it = g_plans.begin(); it != g_plans.end(); ++it) {
        if (it->first == key) {
            g_plans.splice(g_plans.begin(), g_plans, it);
            return g_plans.front().second;
        }
    }
    auto P = cfh::prepare(E, mdl, prd, fp64, adj, block);
    g_plans.emplace_front(key, P);
    if (g_plans.size() > kMaxNoAllocPlans) g_plans.pop_back();
    return P;
}

size_t al256(size_t x) { return (x + 255) & ~size_t(255); }

int run_noalloc(const cf_model* mdl, const cf_product* prd, cfh::Call& c)
{
    using namespace cfh;
    validate_num(c.num);
    if (!c.weights) fail(CF_ERR_ARG, "null aggregation weights");
    const bool fp64 = c.num->precision == CF_FP64;

    Engine& E = engine();
    std::lock_guard<std::mutex> lk(E.sync_mu);
    hip_check(hipSetDevice(E.device), "hipSetDevice");
    unsigned char* base = scratch_on(E.device);
    const std::shared_ptr<Prepared> PP = host_plan(E, mdl, prd, fp64, adjoint_kind(c), call_block(E, c.num, fp64));
    const Prepared& PR = *PP;
    const CallShape S = shape_call(E, PR, c.num, true);
    const CallBytes nb = call_bytes(S);

    // carve the call's buffers from the array
    size_t off = 0;
    auto take = [&](size_t bytes) -> unsigned char* {
        unsigned char* p = base + off;
        off += al256(bytes);
        return p;
    };
    CallBuffers B;
    unsigned char* tables = take(PR.staged.size());
    B.tables = tables;
    B.weights = reinterpret_cast<double*>(take(nb.weights));
    B.block_out = reinterpret_cast<double*>(take(nb.block_out));
    const size_t red_bytes = (size_t(S.stride) + S.ivn) * 8;
    B.reduced = reinterpret_cast<double*>(take(red_bytes));
    if (S.ivn) {
        B.ivbar_blocks = reinterpret_cast<double*>(take(nb.ivbar_blocks));
        B.ivbar = B.reduced + S.stride;
    }
    B.ckpt = take(nb.ckpt);
    if (c.per_path) B.per_path = reinterpret_cast<double*>(take(size_t(S.n_path) * S.npay * 8));
    if (c.per_path_agg) B.per_path_agg = reinterpret_cast<double*>(take(size_t(S.n_path) * 8));
    if (off > kNoAllocScratchBytes)
        fail(CF_ERR_CAPACITY, "the call needs " + std::to_string(off) + " bytes of device scratch; the "
                              "no-allocation entry point holds " + std::to_string(kNoAllocScratchBytes) +
                              " (fewer paths, the recompute adjoint, or cf_mc_simul_aad)");

    // tables and weights straight from host memory, the call's work, the
    // reduced sums back into host memory
    std::vector<double> red(red_bytes / 8);
    hip_check(hipMemcpyAsync(tables, PR.staged.data(), PR.staged.size(), hipMemcpyHostToDevice, E.stream),
              "upload");
    hip_check(hipMemcpyAsync(B.weights, c.weights, size_t(S.npay) * 8, hipMemcpyHostToDevice, E.stream), "upload");
    const cfk::LaunchParams P = launch_params(E, PR, S, B);
    enqueue_call(PR, S, B, P, E.stream, nullptr, nullptr);
    hip_check(hipMemcpyAsync(red.data(), B.reduced, red_bytes, hipMemcpyDeviceToHost, E.stream), "download");
    if (B.per_path)
        hip_check(hipMemcpyAsync(c.per_path, B.per_path, size_t(S.n_path) * S.npay * 8, hipMemcpyDeviceToHost,
                                 E.stream), "download");
    if (B.per_path_agg)
        hip_check(hipMemcpyAsync(c.per_path_agg, B.per_path_agg, size_t(S.n_path) * 8, hipMemcpyDeviceToHost,
                                 E.stream), "download");
    hip_check(hipStreamSynchronize(E.stream), "kernel execution");
    collect_results(PR, S, red.data(), red.data() + S.stride, c);
    // the engine's own weights buffer is untouched; its repeat check stays valid
    return CF_OK;
}

}  // namespace

int cf_mc_simul_aad_noalloc(const cf_model* mdl, const cf_product* prd, const cf_num* num, const double* weights,
                            double* payoff_means, double* agg_mean, double* risks, double* per_path_payoffs,
                            double* per_path_agg)
{
    return cfh::guarded([&] {
        if (!risks) cfh::fail(CF_ERR_ARG, "null risks");
        cfh::Call c;
        c.adjoint = true;
        c.num = num;
        c.weights = weights;
        c.means = payoff_means;
        c.agg_mean = agg_mean;
        c.risks = risks;
        c.per_path = per_path_payoffs;
        c.per_path_agg = per_path_agg;
        return run_noalloc(mdl, prd, c);
    });
}
