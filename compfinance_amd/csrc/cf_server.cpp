// Request server: continuous batching of pricing and risk requests on one
// GPU.  No counterpart in the reference, whose entry points are synchronous
// and not re-entrant (AAD.cpp:22-26, store.h:35-36); each request runs the
// same path kernels and reductions as mcParallelSimul / mcParallelSimulAAD
// (mcBase.h:314-400, :564-753) through the shared call pieces of cf_host.h.
//
// Organisation
//   * submit (any host thread): deep copy of the request, plan on the host
//     (timeline, tables, launch geometry; validation errors surface here),
//     device footprint; the request joins a FIFO queue;
//   * one dispatcher thread per server: admits queued requests while a slot
//     (a HIP stream with its pinned staging) is free and the request's device
//     block fits in the server's arena (one allocation made at creation, so
//     admission never calls hipMalloc); an AAD request whose checkpoint
//     variant does not fit runs the recompute variant when that fits;
//     requests that arrive while others run start at once, and the kernels
//     of different slots overlap on the device;
//   * steady state: when a slot sees the same launch sequence a third time
//     (the exact bytes of every operation's parameters: same plan, shape,
//     arena block and staging), it captures the uploads, kernels and result
//     copies in a HIP graph; later requests with that sequence replay the
//     graph in one launch.  Requests seen once or twice launch directly, so a
//     mix of distinct requests, or of pairs, pays no instantiation;
//   * completion: a host function queued on the slot's stream after the
//     result copies signals the dispatcher, which forms the results, frees
//     the block and the slot and admits more work.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <deque>
#include <map>
#include <thread>

#include "cf_host.h"

namespace cfh {
namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator of 256-byte aligned blocks inside one device
// allocation, with coalescing of free neighbours.
class Arena {
  public:
    void init(size_t bytes)
    {
        cap_ = (bytes + 255) & ~size_t(255);
        hip_check(hipMalloc(&base_, cap_), "hipMalloc (server arena)");
        free_[0] = cap_;
    }
    ~Arena()
    {
        if (base_) (void)hipFree(base_);
    }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = std::max<size_t>((bytes + 255) & ~size_t(255), 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[*off + bytes] = rest;
            live_[*off] = bytes;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off)
    {
        auto it = live_.find(off);
        if (it == live_.end()) return;
        size_t b = it->second;
        used_ -= b;
        live_.erase(it);
        size_t o = off;
        auto nx = free_.lower_bound(o);
        if (nx != free_.end() && nx->first == o + b) {
            b += nx->second;
            free_.erase(nx);
        }
        auto pv = free_.lower_bound(o);
        if (pv != free_.begin()) {
            --pv;
            if (pv->first + pv->second == o) {
                o = pv->first;
                b += pv->second;
                free_.erase(pv);
            }
        }
        free_[o] = b;
    }
    size_t size_of(size_t off) const
    {
        auto it = live_.find(off);
        return it == live_.end() ? 0 : it->second;
    }
    unsigned char* base() const { return static_cast<unsigned char*>(base_); }
    size_t capacity() const { return cap_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

  private:
    void* base_ = nullptr;
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_, live_;
};

// A product descriptor with its arrays and parts owned.
struct OwnedProduct {
    cf_product d{};
    vector<double> mats, strikes;
    vector<OwnedProduct> parts;
    vector<cf_product> part_desc;

    void copy(const cf_product& p)
    {
        d = p;
        if (p.kind == CF_PRODUCT_EUROPEANS && p.n_options > 0 && p.maturities && p.strikes) {
            mats.assign(p.maturities, p.maturities + p.n_options);
            strikes.assign(p.strikes, p.strikes + p.n_options);
        }
        if (p.kind == CF_PRODUCT_BOOK && p.n_parts > 0 && p.parts) {
            parts.resize(size_t(p.n_parts));
            for (int i = 0; i < p.n_parts; ++i) parts[i].copy(p.parts[i]);
        }
        relink();
    }
    void relink()
    {
        d.maturities = mats.empty() ? nullptr : mats.data();
        d.strikes = strikes.empty() ? nullptr : strikes.data();
        part_desc.clear();
        for (auto& q : parts) {
            q.relink();
            part_desc.push_back(q.d);
        }
        d.parts = part_desc.empty() ? nullptr : part_desc.data();
    }
};

struct Request {
    int64_t id = 0;
    cf_model model{};
    vector<double> spots, times, vols;
    OwnedProduct prd;
    cf_num num{};
    bool aad = false;
    vector<double> weights;
    // plans: requested variant and, for a checkpoint AAD, the recompute one
    std::shared_ptr<Prepared> plan, plan_rc;
    CallShape shape, shape_rc;
    size_t bytes = 0, bytes_rc = 0;
    bool use_rc = false;
    bool waited = false;   // a cf_server_wait holds it
    int waits = 0;   // admission passes that found no room
    // results
    Call call;
    vector<double> means, std_errs, risks;
    double agg = 0.0;
    int status = CF_OK;
    string error;
    bool done = false;
    Clock::time_point t_submit, t_planned, t_start, t_launched, t_done;
    float kernel_ms = 0.f, device_ms = 0.f;
    size_t off = 0;       // arena block (alone)
    size_t out_off = 0;   // first result in the slot's mapped memory (doubles)
    int packed = 1;       // requests in its launch
};

struct Server;

// An instantiated graph of one launch sequence and the bytes that name it.
struct GraphEntry {
    vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
};

// Instantiated graphs kept per slot; the least recently used goes first.
constexpr size_t kGraphsPerSlot = 8;
// Sequences launched directly that a slot remembers with their count, to
// capture at the kCaptureAt-th sighting.
constexpr size_t kSeenPerSlot = 32;
constexpr int kCaptureAt = 3;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;   // around the path kernel
    hipEvent_t ev_a = nullptr, ev_b = nullptr; // around the whole sequence
    PinnedBuf pin_in, pin_out;
    vector<Request*> batch;   // the requests of its launch (empty: idle)
    size_t off = 0;           // their arena block
    Clock::time_point t_end;
    Server* srv = nullptr;
    int index = 0;
    vector<GraphEntry> graphs;   // used only by the dispatcher while the slot is idle
    std::deque<std::pair<vector<unsigned char>, int>> seen;   // key, direct launches
    uint64_t tick = 0;
};

// Requests passed over for lack of room at most this many times; then the
// queue waits for the oldest one.
constexpr int kMaxWaits = 8;

struct Server {
    Engine* E = nullptr;
    int device = 0;
    bool auto_recompute = true;
    bool graphs = true;
    int max_pack = cfk::kMaxPack;
    Arena arena;
    vector<std::unique_ptr<Slot>> slots;
    std::mutex mu;
    std::condition_variable cv, cv_done;
    std::deque<Request*> queue;
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    vector<int> finished;   // slots whose work ended (host functions)
    int running = 0;
    bool stop = false;
    int64_t next_id = 1;
    cf_server_stats st{};
    Clock::time_point t0 = Clock::now();
    std::thread thread;
};

void on_slot_done(void* user)
{
    Slot* s = static_cast<Slot*>(user);
    Server* S = s->srv;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        s->t_end = Clock::now();
        S->finished.push_back(s->index);
    }
    S->cv.notify_all();
}

// Device bytes of a request's block: tables then the call buffers.
size_t block_bytes(const Prepared& PR, const CallShape& S)
{
    return ((PR.staged.size() + 255) & ~size_t(255)) + call_bytes(S).total();
}

// Host-to-device copies of a request before its sequence (the results come
// back through the slot's mapped host memory, written by the reduction).
struct Copies {
    void* tables_dst;
    const void* tables_src;
    size_t tables_bytes;
    void* weights_dst;
    const void* weights_src;
    size_t weights_bytes;
};

void enqueue_request(const Sequence& q, const Copies& c, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1)
{
    hip_check(hipMemcpyAsync(c.tables_dst, c.tables_src, c.tables_bytes, hipMemcpyHostToDevice, st), "upload");
    if (c.weights_bytes)
        hip_check(hipMemcpyAsync(c.weights_dst, c.weights_src, c.weights_bytes, hipMemcpyHostToDevice, st), "upload");
    enqueue_sequence(q, st, ev0, ev1);
}

// Appends the bytes of x to a graph key.
template <typename T>
void key_add(vector<unsigned char>& key, const T& x)
{
    const unsigned char* b = reinterpret_cast<const unsigned char*>(&x);
    key.insert(key.end(), b, b + sizeof x);
}

// The slot's graph of the launch sequence named by `key` (the bytes of every
// operation's parameters): found by its bytes; else, on the kCaptureAt-th
// sighting of the key, captured from `enqueue` (the same calls as a direct
// launch) and instantiated (*captured set); else null: launch directly.
template <typename Enqueue>
hipGraphExec_t slot_graph(Slot& s, vector<unsigned char> key, const Enqueue& enqueue, bool* captured)
{
    ++s.tick;
    *captured = false;
    for (GraphEntry& g : s.graphs)
        if (g.key == key) {
            g.last_use = s.tick;
            return g.exec;
        }
    auto seen = std::find_if(s.seen.begin(), s.seen.end(), [&](const auto& e) { return e.first == key; });
    if (seen == s.seen.end()) {
        s.seen.emplace_back(std::move(key), 1);
        if (s.seen.size() > kSeenPerSlot) s.seen.pop_front();
        return nullptr;
    }
    if (++seen->second < kCaptureAt) return nullptr;
    s.seen.erase(seen);
    hipGraph_t graph = nullptr;
    hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    try {
        enqueue();
    } catch (...) {
        if (hipStreamEndCapture(s.stream, &graph) == hipSuccess && graph) (void)hipGraphDestroy(graph);
        throw;
    }
    hip_check(hipStreamEndCapture(s.stream, &graph), "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    hip_check(e, "hipGraphInstantiate");
    if (s.graphs.size() >= kGraphsPerSlot) {
        // the slot is idle, so none of its graphs is executing
        auto lru = std::min_element(s.graphs.begin(), s.graphs.end(),
                                    [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
        (void)hipGraphExecDestroy(lru->exec);
        s.graphs.erase(lru);
    }
    s.graphs.push_back(GraphEntry{std::move(key), exec, s.tick});
    *captured = true;
    return exec;
}

// Launches request r in slot s from its arena block (dispatcher thread, no
// server lock held).
void launch(Server& S, Slot& s, Request& r)
{
    const Prepared& PR = r.use_rc ? *r.plan_rc : *r.plan;
    const CallShape& SH = r.use_rc ? r.shape_rc : r.shape;
    const CallBytes nb = call_bytes(SH);
    unsigned char* p = S.arena.base() + r.off;
    auto take = [&p](size_t bytes) {
        unsigned char* q = bytes ? p : nullptr;
        p += (bytes + 255) & ~size_t(255);
        return q;
    };
    // the reduction writes the sums and the local-vol adjoints straight into
    // the slot's mapped host memory (their arena space stays reserved)
    s.pin_out.ensure((size_t(SH.stride) + SH.ivn) * 8);
    double* hout_dev = static_cast<double*>(s.pin_out.dev);
    CallBuffers B;
    unsigned char* tables = take(PR.staged.size());
    B.tables = tables;
    B.block_out = reinterpret_cast<double*>(take(nb.block_out));
    (void)take(nb.reduced);
    B.reduced = hout_dev;
    B.ivbar_blocks = reinterpret_cast<double*>(take(nb.ivbar_blocks));
    (void)take(nb.ivbar);
    B.ivbar = SH.ivn ? hout_dev + SH.stride : nullptr;
    B.weights = reinterpret_cast<double*>(take(nb.weights));
    B.ckpt = take(nb.ckpt);
    // host staging: tables then weights, one upload each
    const size_t tb = PR.staged.size();
    unsigned char* hin = static_cast<unsigned char*>(s.pin_in.ensure(tb + nb.weights + 8));
    std::memcpy(hin, PR.staged.data(), tb);
    if (r.aad) std::memcpy(hin + tb, r.weights.data(), nb.weights);
    Copies c;
    std::memset(&c, 0, sizeof c);
    c.tables_dst = tables;
    c.tables_src = hin;
    c.tables_bytes = tb;
    if (r.aad) {
        c.weights_dst = B.weights;
        c.weights_src = hin + tb;
        c.weights_bytes = nb.weights;
    }
    const cfk::LaunchParams P = launch_params(*S.E, PR, SH, B);
    Sequence q;
    make_sequence(PR, SH, B, P, &q);
    ready_sequence(q);
    bool captured = false;
    hipGraphExec_t exec = nullptr;
    if (S.graphs) {
        vector<unsigned char> key;
        key_add(key, q);
        key_add(key, c);
        exec = slot_graph(s, std::move(key), [&] { enqueue_request(q, c, s.stream, nullptr, nullptr); }, &captured);
    }
    hip_check(hipEventRecord(s.ev_a, s.stream), "hipEventRecord");
    if (exec) {
        {
            std::lock_guard<std::mutex> lk(S.mu);
            ++(captured ? S.st.graph_captures : S.st.graph_replays);
        }
        hip_check(hipEventRecord(s.ev0, s.stream), "hipEventRecord");
        hip_check(hipGraphLaunch(exec, s.stream), "hipGraphLaunch");
        hip_check(hipEventRecord(s.ev1, s.stream), "hipEventRecord");
    } else {
        enqueue_request(q, c, s.stream, s.ev0, s.ev1);
    }
    hip_check(hipEventRecord(s.ev_b, s.stream), "hipEventRecord");
    hip_check(hipLaunchHostFunc(s.stream, on_slot_done, &s), "hipLaunchHostFunc");
    r.t_launched = Clock::now();
}

// A request that may share a launch: Sobol normals, its primary plan, and
// too few blocks to fill the device alone.
bool packable(const Server& S, const Request& r)
{
    return !r.use_rc && !r.shape.mrg && !r.plan->cfg.item && r.shape.grid < S.E->num_cus * r.plan->per_cu;
}

// Requests b can join a launch of a: the same path kernel and block size.
bool pack_compatible(const Request& a, const Request& b)
{
    const cfk::KernelConfig &x = a.plan->cfg, &y = b.plan->cfg;
    return x.precision == y.precision && x.model == y.model && x.product == y.product && x.adjoint == y.adjoint &&
           x.lean == y.lean && x.recompute == y.recompute && a.plan->block == b.plan->block && a.aad == b.aad;
}

// Launches the requests of a pack in slot s from the slot's arena block, as
// one packed path-kernel launch and one packed reduction (dispatcher thread,
// no lock held).  Block layout: every request's tables and weights first
// (one upload from the same layout in pinned memory), then each request's
// call buffers.  Each request runs on its own blocks with the launch
// parameters, grid and LDS layout it has alone, so its results are those of
// its launch alone.
void launch_pack(Server& S, Slot& s)
{
    const int n = int(s.batch.size());
    unsigned char* base = S.arena.base() + s.off;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    size_t staged = 0, outs = 0;
    for (Request* r : s.batch) {
        staged += al(r->plan->staged.size()) + al(call_bytes(r->shape).weights);
        outs += (size_t(r->shape.stride) + r->shape.ivn + 7) & ~size_t(7);
    }
    unsigned char* hin = static_cast<unsigned char*>(s.pin_in.ensure(staged));
    s.pin_out.ensure(outs * 8);
    double* hout_dev = static_cast<double*>(s.pin_out.dev);

    cfk::PackArgs A;
    cfk::ReducePack RP;
    std::memset(&A, 0, sizeof A);   // padding too: keys compare byte for byte
    std::memset(&RP, 0, sizeof RP);
    A.n = RP.n = n;
    size_t a_off = 0, b_off = staged, o_off = 0, smem = 0;
    for (int i = 0; i < n; ++i) {
        Request& r = *s.batch[size_t(i)];
        const Prepared& PR = *r.plan;
        const CallShape& SH = r.shape;
        const CallBytes nb = call_bytes(SH);
        CallBuffers B;
        B.tables = base + a_off;
        std::memcpy(hin + a_off, PR.staged.data(), PR.staged.size());
        a_off += al(PR.staged.size());
        if (nb.weights) {
            B.weights = reinterpret_cast<double*>(base + a_off);
            std::memcpy(hin + a_off, r.weights.data(), nb.weights);
            a_off += al(nb.weights);
        }
        auto take = [&](size_t bytes) {
            unsigned char* q = bytes ? base + b_off : nullptr;
            b_off += al(bytes);
            return q;
        };
        B.block_out = reinterpret_cast<double*>(take(nb.block_out));
        (void)take(nb.reduced);
        B.ivbar_blocks = reinterpret_cast<double*>(take(nb.ivbar_blocks));
        (void)take(nb.ivbar);
        B.ckpt = take(nb.ckpt);
        r.out_off = o_off;
        B.reduced = hout_dev + o_off;
        B.ivbar = SH.ivn ? B.reduced + SH.stride : nullptr;
        o_off += (size_t(SH.stride) + SH.ivn + 7) & ~size_t(7);
        A.p[i] = launch_params(*S.E, PR, SH, B);
        A.first[i + 1] = A.first[i] + SH.grid;
        smem = std::max(smem, PR.smem);
        RP.in0[i] = B.block_out;
        RP.cols0[i] = SH.stride;
        RP.out0[i] = B.reduced;
        RP.in1[i] = B.ivbar_blocks;
        RP.cols1[i] = int(SH.ivn);
        RP.out1[i] = B.ivbar;
        RP.rows[i] = SH.grid;
        RP.first[i + 1] = RP.first[i] + cfk::reduce_blocks(SH.stride) + cfk::reduce_blocks(int(SH.ivn));
        r.packed = n;
    }
    if (b_off > S.arena.size_of(s.off)) fail(CF_ERR_CAPACITY, "packed block exceeds its arena allocation");
    cfk::KernelConfig cfg = s.batch[0]->plan->cfg;
    cfg.packed = true;
    const int block = s.batch[0]->plan->block;
    int per_cu = 0;
    int rc = cfk::path_kernel_occupancy(cfg, block, smem, &per_cu);   // sets the LDS limit once per size
    if (rc != 0) hip_check(hipError_t(rc), "occupancy");
    if (per_cu < 1) fail(CF_ERR_CAPACITY, "packed kernel does not fit on a compute unit");

    auto enqueue = [&](hipEvent_t e0, hipEvent_t e1) {
        hip_check(hipMemcpyAsync(base, hin, staged, hipMemcpyHostToDevice, s.stream), "upload");
        if (e0) hip_check(hipEventRecord(e0, s.stream), "hipEventRecord");
        int e = cfk::launch_path_kernel_packed(cfg, A, block, smem, s.stream);
        if (e != 0) hip_check(hipError_t(e), "packed path kernel launch");
        if (e1) hip_check(hipEventRecord(e1, s.stream), "hipEventRecord");
        e = cfk::launch_reduce_packed(RP, s.stream);
        if (e != 0) hip_check(hipError_t(e), "packed reduce launch");
    };
    bool captured = false;
    hipGraphExec_t exec = nullptr;
    if (S.graphs) {
        vector<unsigned char> key;
        key_add(key, A);
        key_add(key, RP);
        key_add(key, cfg);
        key_add(key, block);
        key_add(key, smem);
        key_add(key, staged);
        key_add(key, hin);
        exec = slot_graph(s, std::move(key), [&] { enqueue(nullptr, nullptr); }, &captured);
    }
    {
        std::lock_guard<std::mutex> lk(S.mu);
        ++S.st.packed_launches;
        S.st.packed_requests += n;
        if (exec) ++(captured ? S.st.graph_captures : S.st.graph_replays);
    }
    hip_check(hipEventRecord(s.ev_a, s.stream), "hipEventRecord");
    if (exec) {
        hip_check(hipEventRecord(s.ev0, s.stream), "hipEventRecord");
        hip_check(hipGraphLaunch(exec, s.stream), "hipGraphLaunch");
        hip_check(hipEventRecord(s.ev1, s.stream), "hipEventRecord");
    } else {
        enqueue(s.ev0, s.ev1);
    }
    hip_check(hipEventRecord(s.ev_b, s.stream), "hipEventRecord");
    hip_check(hipLaunchHostFunc(s.stream, on_slot_done, &s), "hipLaunchHostFunc");
    const Clock::time_point t = Clock::now();
    for (Request* r : s.batch) r->t_launched = t;
}

// Results of request r that ran in slot s (dispatcher thread, no lock).
void finish(Server& S, Slot& s, Request& r)
{
    const Prepared& PR = r.use_rc ? *r.plan_rc : *r.plan;
    const CallShape& SH = r.use_rc ? r.shape_rc : r.shape;
    const double* hout = static_cast<const double*>(s.pin_out.p) + r.out_off;
    r.means.assign(size_t(SH.npay), 0.0);
    r.std_errs.assign(size_t(SH.npay), 0.0);
    Call c;
    c.adjoint = r.aad;
    c.means = r.means.data();
    c.std_errs = r.std_errs.data();
    c.agg_mean = &r.agg;
    if (r.aad) {
        r.risks.assign(size_t(PR.md.kind == CF_MODEL_DUPIRE ? 1 + PR.md.n_spots * PR.md.n_times : 4), 0.0);
        c.risks = r.risks.data();
    }
    collect_results(PR, SH, hout, hout + SH.stride, c);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s.ev0, s.ev1) == hipSuccess) r.kernel_ms = ms;
    if (hipEventElapsedTime(&ms, s.ev_a, s.ev_b) == hipSuccess) r.device_ms = ms;
    (void)S;
}

void dispatcher(Server* Sp)
{
    Server& S = *Sp;
    (void)hipSetDevice(S.device);
    std::unique_lock<std::mutex> lk(S.mu);
    for (;;) {
        // ---- completions
        while (!S.finished.empty()) {
            const int si = S.finished.back();
            S.finished.pop_back();
            Slot& s = *S.slots[si];
            const vector<Request*> batch = s.batch;
            const Clock::time_point t_end = s.t_end;
            lk.unlock();
            for (Request* r : batch) {
                if (r->status != CF_OK) continue;
                try {
                    finish(S, s, *r);
                } catch (const CfError& e) {
                    r->status = e.code;
                    r->error = e.what();
                } catch (const std::exception& e) {
                    r->status = CF_ERR_DEVICE;
                    r->error = e.what();
                }
            }
            lk.lock();
            for (Request* r : batch) {
                r->t_done = t_end;
                r->done = true;
                --S.running;
                ++S.st.completed;
                if (r->status != CF_OK) ++S.st.failed;
            }
            S.arena.release(s.off);
            s.batch.clear();
            S.cv_done.notify_all();
        }
        // ---- admissions: free slots, oldest requests first; a request that
        // does not fit is passed over at most kMaxWaits times
        vector<Slot*> start;
        for (auto& sp : S.slots) {
            if (!sp->batch.empty() || S.queue.empty()) continue;
            Request* pick = nullptr;
            for (auto it = S.queue.begin(); it != S.queue.end(); ++it) {
                Request* r = *it;
                size_t off = 0;
                bool ok = S.arena.alloc(r->bytes, &off);
                if (!ok && r->plan_rc && S.auto_recompute && S.arena.alloc(r->bytes_rc, &off)) {
                    ok = true;
                    r->use_rc = true;
                    ++S.st.recompute_switches;
                }
                if (ok) {
                    r->off = off;
                    pick = r;
                    S.queue.erase(it);
                    break;
                }
                ++S.st.memory_waits;
                if (++r->waits > kMaxWaits) break;
            }
            if (!pick) break;
            vector<Request*> batch{pick};
            size_t off = pick->off;
            // queued requests of the same kernel that fit beside it on the
            // device join its launch, in one arena block
            if (S.max_pack > 1 && packable(S, *pick)) {
                int cap = S.E->num_cus * pick->plan->per_cu, grid = pick->shape.grid;
                size_t bytes = pick->bytes;
                for (Request* r : S.queue) {
                    if (int(batch.size()) >= S.max_pack) break;
                    if (!packable(S, *r) || !pack_compatible(*pick, *r)) continue;
                    const int cap_r = std::min(cap, S.E->num_cus * r->plan->per_cu);
                    if (grid + r->shape.grid > cap_r) continue;
                    cap = cap_r;
                    grid += r->shape.grid;
                    bytes += r->bytes;
                    batch.push_back(r);
                }
                if (batch.size() > 1) {
                    S.arena.release(pick->off);
                    if (S.arena.alloc(bytes, &off)) {
                        for (size_t i = 1; i < batch.size(); ++i)
                            S.queue.erase(std::find(S.queue.begin(), S.queue.end(), batch[i]));
                    } else {
                        // no room for the pack: the first request alone, in the
                        // room its own block had a moment ago
                        if (!S.arena.alloc(pick->bytes, &pick->off)) {
                            S.queue.push_front(pick);
                            break;
                        }
                        off = pick->off;
                        batch.resize(1);
                    }
                }
            }
            for (Request* r : batch) {
                if (S.running > 0) ++S.st.admitted_busy;
                ++S.running;
                r->t_start = Clock::now();
            }
            S.st.peak_running = std::max<int64_t>(S.st.peak_running, S.running);
            S.st.peak_arena_bytes = std::max<int64_t>(S.st.peak_arena_bytes, int64_t(S.arena.used()));
            sp->batch = std::move(batch);
            sp->off = off;
            start.push_back(sp.get());
        }
        if (!start.empty()) {
            lk.unlock();
            for (Slot* sl : start) {
                try {
                    if (sl->batch.size() == 1) launch(S, *sl, *sl->batch[0]);
                    else launch_pack(S, *sl);
                } catch (const std::exception& e) {
                    // nothing of these requests is queued after the failing
                    // call (the completion signal is queued last); complete
                    // them through the same path with the error
                    (void)hipStreamSynchronize(sl->stream);
                    std::lock_guard<std::mutex> g(S.mu);
                    const CfError* ce = dynamic_cast<const CfError*>(&e);
                    for (Request* r : sl->batch) {
                        r->status = ce ? ce->code : CF_ERR_DEVICE;
                        r->error = e.what();
                    }
                    sl->t_end = Clock::now();
                    S.finished.push_back(sl->index);
                }
            }
            lk.lock();
            continue;
        }
        if (S.stop && S.queue.empty() && S.running == 0 && S.finished.empty()) break;
        S.cv.wait(lk);
    }
}

Request* find(Server& S, int64_t id)
{
    auto it = S.reqs.find(id);
    if (it == S.reqs.end()) fail(CF_ERR_NOT_FOUND, "unknown request id");
    return it->second.get();
}

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace
}  // namespace cfh

using namespace cfh;

extern "C" {

int cf_server_create(const cf_server_opts* opts, cf_server** out)
{
    return guarded([&] {
        if (!out) fail(CF_ERR_ARG, "null server handle");
        const int n_slots = opts ? opts->slots : 4;
        const int64_t budget = opts ? opts->hbm_budget : (int64_t(16) << 30);
        if (n_slots < 1 || n_slots > 16) fail(CF_ERR_ARG, "slots must be in 1..16");
        if (budget < (int64_t(1) << 20)) fail(CF_ERR_ARG, "device budget below 1 MiB");
        auto S = std::make_unique<Server>();
        S->E = &engine();
        S->device = S->E->device;
        S->auto_recompute = opts ? opts->auto_recompute != 0 : true;
        S->graphs = opts ? opts->graphs != 0 : true;
        const int mp = opts ? opts->max_pack : 0;
        if (mp < 0 || mp > cfk::kMaxPack) fail(CF_ERR_ARG, "max_pack must be in 0..8");
        S->max_pack = mp == 0 ? cfk::kMaxPack : mp;
        hip_check(hipSetDevice(S->device), "hipSetDevice");
        S->arena.init(size_t(budget));
        for (int i = 0; i < n_slots; ++i) {
            auto s = std::make_unique<Slot>();
            hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreate(&s->ev0), "hipEventCreate");
            hip_check(hipEventCreate(&s->ev1), "hipEventCreate");
            hip_check(hipEventCreate(&s->ev_a), "hipEventCreate");
            hip_check(hipEventCreate(&s->ev_b), "hipEventCreate");
            s->srv = S.get();
            s->index = i;
            s->pin_in.ensure(size_t(1) << 20);   // staging sized for the common requests up front
            s->pin_out.ensure(size_t(1) << 20);
            S->slots.push_back(std::move(s));
        }
        S->st.arena_bytes = int64_t(S->arena.capacity());
        S->thread = std::thread(dispatcher, S.get());
        *out = reinterpret_cast<cf_server*>(S.release());
        return CF_OK;
    });
}

int cf_server_submit(cf_server* srv, const cf_request* req, int64_t* id)
{
    return guarded([&] {
        if (!srv || !req || !id) fail(CF_ERR_ARG, "null argument");
        Server& S = *reinterpret_cast<Server*>(srv);
        auto r = std::make_unique<Request>();
        r->t_submit = Clock::now();
        validate_num(&req->num);
        r->num = req->num;
        r->aad = req->aad != 0;
        r->model = req->model;
        if (req->model.kind == CF_MODEL_DUPIRE && req->model.spots && req->model.times && req->model.vols &&
            req->model.n_spots > 0 && req->model.n_times > 0) {
            const size_t ns = size_t(req->model.n_spots), nt = size_t(req->model.n_times);
            r->spots.assign(req->model.spots, req->model.spots + ns);
            r->times.assign(req->model.times, req->model.times + nt);
            r->vols.assign(req->model.vols, req->model.vols + ns * nt);
            r->model.spots = r->spots.data();
            r->model.times = r->times.data();
            r->model.vols = r->vols.data();
        }
        r->prd.copy(req->product);
        const bool fp64 = r->num.precision == CF_FP64;
        int adj = cfk::ADJ_NONE;
        if (r->aad) adj = r->num.aad_mode == CF_AAD_RECOMPUTE ? cfk::ADJ_RECOMPUTE : cfk::ADJ_CHECKPOINT;
        const int block = call_block(*S.E, &r->num, fp64);
        r->plan = lookup(*S.E, &r->model, &r->prd.d, fp64, adj, false, block);
        r->shape = shape_call(*S.E, *r->plan, &r->num, r->aad);
        r->bytes = block_bytes(*r->plan, r->shape);
        if (r->aad) {
            const int npay = r->plan->pd.n_pay;
            r->weights.assign(size_t(npay), 0.0);
            if (req->weights) std::copy(req->weights, req->weights + npay, r->weights.begin());
            else r->weights[0] = 1.0;
        }
        if (adj == cfk::ADJ_CHECKPOINT) {
            r->plan_rc = lookup(*S.E, &r->model, &r->prd.d, fp64, cfk::ADJ_RECOMPUTE, false, block);
            r->shape_rc = shape_call(*S.E, *r->plan_rc, &r->num, true);
            r->bytes_rc = block_bytes(*r->plan_rc, r->shape_rc);
        }
        const size_t least = r->plan_rc && S.auto_recompute ? std::min(r->bytes, r->bytes_rc) : r->bytes;
        if (least > S.arena.capacity())
            fail(CF_ERR_CAPACITY, "request needs " + std::to_string(least) + " device bytes, server holds " +
                                      std::to_string(S.arena.capacity()));
        r->t_planned = Clock::now();
        {
            std::lock_guard<std::mutex> lk(S.mu);
            if (S.stop) fail(CF_ERR_ARG, "server is shutting down");
            r->id = S.next_id++;
            *id = r->id;
            S.queue.push_back(r.get());
            S.reqs[r->id] = std::move(r);
            ++S.st.submitted;
        }
        S.cv.notify_all();
        return CF_OK;
    });
}

int cf_server_poll(cf_server* srv, int64_t id)
{
    return guarded([&] {
        if (!srv) fail(CF_ERR_ARG, "null server handle");
        Server& S = *reinterpret_cast<Server*>(srv);
        std::lock_guard<std::mutex> lk(S.mu);
        return find(S, id)->done ? 1 : 0;
    });
}

int cf_server_wait(cf_server* srv, int64_t id, cf_result* out)
{
    return guarded([&] {
        if (!srv || !out) fail(CF_ERR_ARG, "null argument");
        Server& S = *reinterpret_cast<Server*>(srv);
        std::unique_lock<std::mutex> lk(S.mu);
        Request* r = find(S, id);
        // one waiter per request: the waiter releases it
        if (r->waited) fail(CF_ERR_ARG, "request is already being waited on");
        r->waited = true;
        S.cv_done.wait(lk, [r] { return r->done; });
        std::unique_ptr<Request> owned = std::move(S.reqs[id]);
        S.reqs.erase(id);
        lk.unlock();
        out->status = r->status;
        std::snprintf(out->error, sizeof out->error, "%s", r->error.c_str());
        out->n_pay = int(r->means.size());
        out->n_param = int(r->risks.size());
        if (out->means) std::copy_n(r->means.begin(), std::min(out->n_pay, std::max(out->cap_pay, 0)), out->means);
        if (out->std_errs)
            std::copy_n(r->std_errs.begin(), std::min(out->n_pay, std::max(out->cap_pay, 0)), out->std_errs);
        if (out->risks)
            std::copy_n(r->risks.begin(), std::min(out->n_param, std::max(out->cap_risk, 0)), out->risks);
        out->agg_mean = r->agg;
        out->queue_ms = ms_between(r->t_submit, r->t_start);
        out->latency_ms = ms_between(r->t_submit, r->t_done);
        out->kernel_ms = r->kernel_ms;
        out->submit_ms = ms_between(S.t0, r->t_submit);
        out->device_bytes = int64_t(r->use_rc ? r->bytes_rc : r->bytes);
        out->recomputed = r->use_rc ? 1 : 0;
        out->packed = r->packed;
        out->plan_ms = ms_between(r->t_submit, r->t_planned);
        out->launch_ms = r->t_launched > r->t_start ? ms_between(r->t_start, r->t_launched) : 0.0;
        out->device_ms = r->device_ms;
        if (r->status != CF_OK) {
            g_err = r->error;
            return r->status;
        }
        return CF_OK;
    });
}

int cf_server_get_stats(cf_server* srv, cf_server_stats* out)
{
    return guarded([&] {
        if (!srv || !out) fail(CF_ERR_ARG, "null argument");
        Server& S = *reinterpret_cast<Server*>(srv);
        std::lock_guard<std::mutex> lk(S.mu);
        *out = S.st;
        out->arena_bytes = int64_t(S.arena.capacity());
        out->queued = int64_t(S.queue.size());
        out->running = S.running;
        return CF_OK;
    });
}

int cf_server_destroy(cf_server* srv)
{
    return guarded([&] {
        if (!srv) return CF_OK;
        std::unique_ptr<Server> S(reinterpret_cast<Server*>(srv));
        {
            std::lock_guard<std::mutex> lk(S->mu);
            S->stop = true;
        }
        S->cv.notify_all();
        if (S->thread.joinable()) S->thread.join();
        for (auto& s : S->slots) {
            (void)hipStreamSynchronize(s->stream);
            for (GraphEntry& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
            s->graphs.clear();
            (void)hipEventDestroy(s->ev0);
            (void)hipEventDestroy(s->ev1);
            (void)hipEventDestroy(s->ev_a);
            (void)hipEventDestroy(s->ev_b);
            (void)hipStreamDestroy(s->stream);
            if (s->pin_in.p) (void)hipHostFree(s->pin_in.p);
            if (s->pin_out.p) (void)hipHostFree(s->pin_out.p);
            s->pin_in.p = s->pin_out.p = nullptr;
        }
        return CF_OK;
    });
}

}  // extern "C"
