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
//   * steady state: when a slot sees the same launch sequence a second time
//     (the exact bytes of every operation's parameters: same plan, shape,
//     arena block and staging), it captures the uploads, kernels and result
//     copies in a HIP graph; later requests with that sequence replay the
//     graph in one launch.  One-off requests launch directly, so a mix of
//     distinct requests pays no instantiation;
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
    Clock::time_point t_submit, t_start, t_done;
    float kernel_ms = 0.f;
    size_t off = 0;
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
// Sequences launched directly that a slot remembers, to capture on a repeat.
constexpr size_t kSeenPerSlot = 32;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    PinnedBuf pin_in, pin_out;
    Request* req = nullptr;
    Clock::time_point t_end;
    Server* srv = nullptr;
    int index = 0;
    vector<GraphEntry> graphs;   // used only by the dispatcher while the slot is idle
    std::deque<vector<unsigned char>> seen;
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

// The slot's graph of (q, c): found by its bytes; else, on the second
// sighting of the sequence, captured from the same enqueue as a direct
// launch and instantiated (*captured set); else null: launch directly.
hipGraphExec_t slot_graph(Slot& s, const Sequence& q, const Copies& c, bool* captured)
{
    vector<unsigned char> key(sizeof q + sizeof c);
    std::memcpy(key.data(), &q, sizeof q);
    std::memcpy(key.data() + sizeof q, &c, sizeof c);
    ++s.tick;
    *captured = false;
    for (GraphEntry& g : s.graphs)
        if (g.key == key) {
            g.last_use = s.tick;
            return g.exec;
        }
    auto seen = std::find(s.seen.begin(), s.seen.end(), key);
    if (seen == s.seen.end()) {
        s.seen.push_back(std::move(key));
        if (s.seen.size() > kSeenPerSlot) s.seen.pop_front();
        return nullptr;
    }
    s.seen.erase(seen);
    hipGraph_t graph = nullptr;
    hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    try {
        enqueue_request(q, c, s.stream, nullptr, nullptr);
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
    B.normals = take(nb.normals);
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
    hipGraphExec_t exec = S.graphs ? slot_graph(s, q, c, &captured) : nullptr;
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
    hip_check(hipLaunchHostFunc(s.stream, on_slot_done, &s), "hipLaunchHostFunc");
}

// Results of the request that ran in slot s (dispatcher thread, no lock).
void finish(Server& S, Slot& s, Request& r)
{
    const Prepared& PR = r.use_rc ? *r.plan_rc : *r.plan;
    const CallShape& SH = r.use_rc ? r.shape_rc : r.shape;
    const double* hout = static_cast<const double*>(s.pin_out.p);
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
            Request* r = s.req;
            const Clock::time_point t_end = s.t_end;
            lk.unlock();
            if (r->status == CF_OK) {
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
            const int code = r->status;
            r->t_done = t_end;
            r->done = true;
            S.arena.release(r->off);
            s.req = nullptr;
            --S.running;
            ++S.st.completed;
            if (code != CF_OK) ++S.st.failed;
            S.cv_done.notify_all();
        }
        // ---- admissions: free slots, oldest requests first; a request that
        // does not fit is passed over at most kMaxWaits times
        vector<std::pair<Slot*, Request*>> start;
        for (auto& sp : S.slots) {
            if (sp->req || S.queue.empty()) continue;
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
            if (S.running > 0) ++S.st.admitted_busy;
            ++S.running;
            S.st.peak_running = std::max<int64_t>(S.st.peak_running, S.running);
            S.st.peak_arena_bytes = std::max<int64_t>(S.st.peak_arena_bytes, int64_t(S.arena.used()));
            sp->req = pick;
            pick->t_start = Clock::now();
            start.push_back({sp.get(), pick});
        }
        if (!start.empty()) {
            lk.unlock();
            for (auto& sr : start) {
                try {
                    launch(S, *sr.first, *sr.second);
                } catch (const std::exception& e) {
                    // nothing of this request is queued after the failing call
                    // (the completion signal is queued last); complete it
                    // through the same path with the error
                    (void)hipStreamSynchronize(sr.first->stream);
                    std::lock_guard<std::mutex> g(S.mu);
                    const CfError* ce = dynamic_cast<const CfError*>(&e);
                    sr.second->status = ce ? ce->code : CF_ERR_DEVICE;
                    sr.second->error = e.what();
                    sr.first->t_end = Clock::now();
                    S.finished.push_back(sr.first->index);
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
        hip_check(hipSetDevice(S->device), "hipSetDevice");
        S->arena.init(size_t(budget));
        for (int i = 0; i < n_slots; ++i) {
            auto s = std::make_unique<Slot>();
            hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
            hip_check(hipEventCreate(&s->ev0), "hipEventCreate");
            hip_check(hipEventCreate(&s->ev1), "hipEventCreate");
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
            (void)hipStreamDestroy(s->stream);
            if (s->pin_in.p) (void)hipHostFree(s->pin_in.p);
            if (s->pin_out.p) (void)hipHostFree(s->pin_out.p);
            s->pin_in.p = s->pin_out.p = nullptr;
        }
        return CF_OK;
    });
}

}  // extern "C"
