// Host-side engine types shared by the synchronous entry points
// (cf_engine.cpp), the store (cf_store.cpp) and the request server
// (cf_server.cpp): plans of products and models, prepared plans with their
// device tables, the per-device engine state, and the pieces of one
// simulation call (geometry, device buffers, launch, results).
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cfamd.h"
#include "cf_internal.h"

namespace cfh {

using std::size_t;
using std::string;
using std::vector;

constexpr double kHalfDay = 0.00136986301369863;   // mcMdlDupire.h:26
constexpr double kOneHour = 0.000114469;           // mcPrd.h:26
constexpr double kEps = 1.0e-08;                   // utility.h:5, store.h:153

// ------------------------------------------------------------------ errors

extern thread_local string g_err;
extern thread_local cf_kernel_stats g_stats;

struct CfError : std::runtime_error {
    int code;
    CfError(int c, const string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const string& msg);
void hip_check(hipError_t e, const char* what);

// Runs f, mapping exceptions to a status code and the thread's last error.
template <class F>
int guarded(F&& f)
{
    try {
        return f();
    } catch (const CfError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return CF_ERR_CAPACITY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CF_ERR_ARG;
    }
}

// ------------------------------------------------------------------ plans

// linear interpolation with flat extrapolation (interp.h:26-63) and its
// two-node weights
struct Interp1 {
    int j0, j1;
    double w0, w1;
};

struct SampleDefH {
    bool numeraire = true;
    bool has_fwd = false, has_disc = false;
    double fwd_mat = 0.0, disc_mat = 0.0;
};

struct ProductPlan {
    int kind = 0;
    vector<double> timeline;
    vector<SampleDefH> defs;
    double strike = 0.0, barrier = 0.0, smooth = 0.0, freq = 0.0, maturity = 0.0;
    double exercise = 0.0, settlement = 0.0;
    bool put = false;
    vector<double> mats;              // Europeans: distinct maturities, increasing
    vector<vector<double>> strikes;   // Europeans: strikes per maturity, input order
    vector<int> pay_first;
    vector<double> pay_strike;
    int n_pay = 0;
    // Asian: fixing dates (its timeline), fixing and snapping periods
    double fix_freq = 0.0, snap = 0.0;
    // Book: the parts, in payoff-column order
    vector<ProductPlan> parts;
};

// Book tables (an Asian alone is a book of one part), laid out in one blob.
// width0: the spot written at t = 0, which scales every barrier's smoothing
// width (mcPrd.h:247-250).
struct BookTables {
    vector<unsigned char> blob;
    cfk::BookLayout L{};
    double lo_min = HUGE_VAL;      // the lowest B - w of the barriers
};

struct ModelPlan {
    int kind = 0;
    double spot = 0.0;
    vector<double> timeline;
    vector<int> step_sample;   // [n_steps + 1]
    int n_steps = 0;
    // Dupire
    int n_spots = 0, n_times = 0;
    vector<double> spots, times, log_spots, inv_dx, iv;   // iv [n_steps][n_spots]
    // Black-Scholes
    double vol = 0.0, rate = 0.0, div = 0.0;
    bool spot_measure = false;
    vector<double> drift, stdv;
    vector<double> ff, disc, num;   // per product sample
    vector<double> bs_aux;
    double s_t0 = 0.0;              // first sample's forward (barrier smoothing width)
};

// Bucket table of the kernels' upper_bound (KnotSearch in cf_path_kernel.h),
// built with the same real type and the same two rounded operations the
// kernel uses to map a state to its bucket.
// Shape of the knot search table (cfk::build_knot_table) handed to the kernel.
struct BucketTable {
    int nb = 1, fix = 1;
    double invh = 0.0, c0 = 0.0;   // values of the kernel's real type
};

// ------------------------------------------------------------------ device

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    void* ensure(size_t bytes)
    {
        if (bytes <= cap && p) return p;
        if (p) hip_check(hipFree(p), "hipFree");
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes, 256);
        hip_check(hipMalloc(&p, want), "hipMalloc");
        cap = want;
        return p;
    }
};

// Grow-only pinned host buffer (uploads and downloads without staging copies).
// Pinned host memory, coherent and mapped: kernels write results straight
// into it (dev is its device address), so a call ends with its last kernel
// rather than with a copy.
struct PinnedBuf {
    void* p = nullptr;
    void* dev = nullptr;
    size_t cap = 0;
    void* ensure(size_t bytes)
    {
        if (bytes <= cap && p) return p;
        if (p) hip_check(hipHostFree(p), "hipHostFree");
        p = dev = nullptr;
        cap = 0;
        const size_t want = std::max<size_t>(bytes, 4096);
        hip_check(hipHostMalloc(&p, want, hipHostMallocMapped | hipHostMallocCoherent), "hipHostMalloc");
        cap = want;
        hip_check(hipHostGetDevicePointer(&dev, p, 0), "hipHostGetDevicePointer");
        return p;
    }
};

// Host staging of a prepared plan's tables, one upload.
struct Staging {
    vector<unsigned char> bytes;
    template <class T>
    size_t add(const T* data, size_t n)
    {
        const size_t off = (bytes.size() + 255) & ~size_t(255);
        bytes.resize(off + n * sizeof(T));
        if (n) std::memcpy(bytes.data() + off, data, n * sizeof(T));
        return off;
    }
    template <class T>
    size_t add(const vector<T>& v) { return add(v.data(), v.size()); }
};

// A model and product prepared for one precision and one kernel: the plans,
// the knot search table, the shared-memory layout and the tables resident in
// device memory.  Prepared once per distinct (model, product, precision,
// adjoint) and reused by later calls.
struct Prepared {
    ProductPlan pd;
    ModelPlan md;
    BucketTable bt;
    cfk::KernelConfig cfg{};
    size_t smem = 0;
    int per_cu = 0;
    int block = 0;
    int ivbar_copies = 1;   // LDS local-vol adjoint grids per block
    bool iv_recover = false;   // Dupire: every local-vol table entry at least kRecoverVolF64 / F32
    BookTables book;
    vector<unsigned char> staged;   // the tables as uploaded
    DevBuf tables;                  // device copy for the synchronous calls
    size_t o_ss = 0, o_knots = 0, o_invdx = 0, o_iv = 0, o_drift = 0, o_std = 0, o_bucket = 0;
    size_t o_ff = 0, o_disc = 0, o_num = 0, o_aux = 0, o_pf = 0, o_pk = 0, o_book = 0;
    size_t o_mrg = 0;   // mrg32k3a jump matrices [32][2][9] for dim = n_steps
    // adjoint of the local-vol pre-interpolation: sqrt(dt_i) and the time
    // interpolation weights of each step
    vector<double> sqrtdt;
    vector<Interp1> tw;
    ~Prepared()
    {
        if (tables.p) (void)hipFree(tables.p);
    }
};

// Per-device state: direction numbers, the plan cache, and the stream and
// grow-only buffers of the synchronous entry points.
struct Engine {
    int device = -1;
    int num_cus = 0;
    int max_lds = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    uint32_t* d_dirs = nullptr;
    DevBuf block_out, ivbar_blocks, reduced, per_path, per_path_agg, ckpt, weights;
    PinnedBuf pin_in, pin_out;
    // the aggregation weights last uploaded to `weights` (skipped when a call repeats them)
    vector<double> weights_last;
    const void* weights_on_device = nullptr;
    std::list<std::pair<string, std::shared_ptr<Prepared>>> plans;   // most recent first
    std::mutex plan_mu;   // the plan cache
    std::mutex sync_mu;   // the synchronous entry points' stream and buffers
};

Engine& engine();

// ------------------------------------------------------------------ one call

struct Call {
    bool adjoint = false;
    const cf_num* num = nullptr;
    const double* weights = nullptr;
    // outputs
    double* means = nullptr;
    double* std_errs = nullptr;
    double* per_path = nullptr;
    double* agg_mean = nullptr;
    double* risks = nullptr;
    double* per_path_agg = nullptr;
};

void validate_num(const cf_num* num);
void check_finite(double x, const char* what);
void validate_model(const cf_model* m);
// utility.h:11-79
vector<double> fill_data(const vector<double>& original, double maxDx, double minDx, const vector<double>& add);
ProductPlan plan_product(const cf_product* p);
ModelPlan plan_model(const cf_model* m, const ProductPlan& p);
vector<string> payoff_labels(const ProductPlan& d);

// adj: cfk::AdjointKind.  The host plan of a model and product (tables
// staged on the host, none on the device) and its cache key.
// block: threads per block of the path kernel (call_block), capped for books.
std::shared_ptr<Prepared> prepare(Engine& E, const cf_model* mdl, const cf_product* prd, bool fp64, int adj,
                                  int block);
string plan_key(const cf_model* m, const cf_product* p, int precision, int adj, int block);
int adjoint_kind(const Call& c);

// Threads per block of the path kernel for a call of num->num_path paths: the
// fewest waves per block that still give every compute unit its share of the
// paths' waves, at least 8 (FP32) or 4 (FP64) and at most 16.  Fewer waves on
// a compute unit run each path's dependent steps faster, so a call too small
// to fill the device finishes sooner; from 16 waves per compute unit on, the
// 1024-thread blocks are the fastest.
int call_block(const Engine& E, const cf_num* num, bool fp64);

// adj: cfk::AdjointKind.  Host plan from the engine's cache (prepared on a
// miss); with_tables also makes its tables resident on the device.
std::shared_ptr<Prepared> lookup(Engine& E, const cf_model* mdl, const cf_product* prd, bool fp64, int adj,
                                 bool with_tables, int block);

// Launch geometry of one call: a persistent grid of at most the resident
// blocks of the device, enough waves for the 64-path Sobol blocks.
struct CallShape {
    int64_t first = 0, n_path = 0;
    bool mrg = false;               // mrg32k3a normals (generated in the path kernel)
    uint64_t seed1 = 0, seed2 = 0;  // reduced modulo m1, m2
    int grid = 0, waves = 0, slots = 0, stride = 0, npay = 0, ck_rows = 0, n_steps = 0;
    size_t ivn = 0;                 // Dupire adjoint: steps x spots
    bool adjoint = false, fp64 = false;
    bool merge = false;             // the last block of 64 Sobol indices runs in the first block's lanes
};
CallShape shape_call(const Engine& E, const Prepared& PR, const cf_num* num, bool adjoint);

// Device buffers of one call.
struct CallBuffers {
    const unsigned char* tables = nullptr;
    double* block_out = nullptr;    // [grid][stride]
    double* reduced = nullptr;      // [stride]
    double* ivbar_blocks = nullptr; // [grid][ivn]
    double* ivbar = nullptr;        // [ivn]
    double* weights = nullptr;      // [npay]
    void* ckpt = nullptr;           // [slots][ck_rows][64] reals
    double* per_path = nullptr;     // [n_path][npay] or null
    double* per_path_agg = nullptr; // [n_path] or null
};

// Bytes of each device buffer (0 when unused) and of all of them, each part
// aligned to 256 bytes, tables excluded.
struct CallBytes {
    size_t block_out, reduced, ivbar_blocks, ivbar, weights, ckpt;
    size_t total() const;
};
CallBytes call_bytes(const CallShape& S);

cfk::LaunchParams launch_params(const Engine& E, const Prepared& PR, const CallShape& S, const CallBuffers& B);

// The device work of one call as plain values: the path kernel and the
// cross-block reductions.  Enqueueing
// reads nothing else, so two calls whose sequences are equal byte for byte
// enqueue the same work (the server replays a captured graph on that
// equality).  Built zero-filled, padding included.
struct Sequence {
    cfk::KernelConfig cfg;
    cfk::LaunchParams P;
    int grid, block;
    size_t smem;
    double* block_out;             // [grid][stride]
    int stride;
    double* reduced;               // [stride]
    int64_t ivn;
    double* ivbar_blocks;          // [grid][ivn]
    double* ivbar;                 // [ivn]
};
void make_sequence(const Prepared& PR, const CallShape& S, const CallBuffers& B, const cfk::LaunchParams& P,
                   Sequence* q);
// Host-side set-up of the sequence's kernels (LDS limits); not a stream
// operation, so it runs before a capture.
void ready_sequence(const Sequence& q);
// Enqueues the sequence on `stream`; ev_kernel brackets the path kernel
// (either may be null).
void enqueue_sequence(const Sequence& q, hipStream_t stream, hipEvent_t ev_kernel0, hipEvent_t ev_kernel1);

// make_sequence + ready_sequence + enqueue_sequence.
void enqueue_call(const Prepared& PR, const CallShape& S, const CallBuffers& B, const cfk::LaunchParams& P,
                  hipStream_t stream, hipEvent_t ev_kernel0, hipEvent_t ev_kernel1);

// Means, standard errors, aggregate and parameter risks from the reduced
// sums (red: [stride], ivbar: [ivn] local-vol adjoints).
void collect_results(const Prepared& PR, const CallShape& S, const double* red, const double* ivbar,
                     const Call& c);

double ms_since(std::chrono::steady_clock::time_point t0);

}  // namespace cfh
