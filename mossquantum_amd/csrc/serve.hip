// Request server: continuous batching of energy + adjoint-gradient requests.
//
// State-vector requests join the batched adjoint engine (mossq_batch.h) at any
// iteration: admission allocates their psi/lambda vectors from a size-class
// cache under the server's device-memory budget, uploads their matrices through
// a pinned staging ring, and starts them at the current iteration; a request
// retires when its reverse pass ends, its results are copied back
// asynchronously and its memory returns to the cache.  States stay resident on
// the device for the whole request.  Density-matrix requests, state-vector
// programs with gates wider than two qubits, and state-vector requests of at
// least kFusedMinQubits qubits run on worker threads, each with its own
// context (stream), so their kernels overlap the engine's and each other's on
// the device.  The fused state-vector requests have workers of their own, so
// they never queue behind the longer density-matrix work.  The large
// state-vector requests use the fused adjoint (LDS-tile sweeps forward and,
// with gradient probes, in reverse); density-matrix requests of 6 or more
// qubits use the fused density adjoint (chunk plans cut at the parameterized
// steps).  Workers build these plans off the submitting thread.
//
// New capability: the reference evaluates one objective at a time
// (src/differentiation/objectives.jl:183-218).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "mossq_batch.h"

namespace mossq {

constexpr int kMaxActive = 4096;     // engine items resident at once
constexpr int kMaxGrads = 1024;      // gradient slots per engine item
constexpr int kInflight = 8;         // engine iterations queued ahead of the device
constexpr u64 kRingBytes = 64ull << 20;
constexpr int kWorkers = 4;          // worker threads (contexts) beside the engine
constexpr int kFusedWorkers = 2;     // of which serve the fused state-vector queue
constexpr int kFusedMinQubits = 14;  // state-vector requests this wide take the fused adjoint

struct SRequest {
    i64 ticket = 0;
    int n = 0, is_density = 0;
    std::vector<mossq_op> ops;
    std::vector<c128> pool;
    const Obs *obs = nullptr;
    Program prog;  // host program (steps + pool)
    u64 bytes = 0;  // reserved against the budget
    // engine state
    bool engine = false;
    bool fused = false;  // worker request run with the fused adjoint
    int slot = -1;  // engine item slot (values / grads arena)
    int t0 = 0, nfwd = 0, nchunk = 1;
    u64 vec_bytes = 0, pool_bytes = 0, ops_bytes = 0;
    c128 *vec = nullptr, *d_pool = nullptr;
    BOp *d_ops = nullptr;
    hipEvent_t done_ev = nullptr;
    // results
    std::atomic<int> state{0};  // 0 queued, 1 running, 2 results pending copy, 3 done, -1 failed
    double value = 0;
    std::vector<double> grads;
    std::string error;
};

// Pinned host staging for asynchronous uploads: a ring whose regions are
// reused only after the copies reading them have completed.
struct PinnedRing {
    char *buf = nullptr;
    u64 cap = 0, head = 0;
    struct Pending {
        u64 lo, hi;
        hipEvent_t ev;
    };
    std::deque<Pending> pending;
    std::vector<hipEvent_t> spare;
    int init(u64 bytes) {
        cap = bytes;
        MOSSQ_HIP(hipHostMalloc((void **)&buf, bytes, hipHostMallocDefault));
        return MOSSQ_OK;
    }
    int reserve(u64 bytes, char **p, u64 *lo) {
        MOSSQ_CHECK(bytes <= cap, MOSSQ_EOOM, "request too large for the staging ring");
        bytes = (bytes + 255) & ~u64(255);
        if (head + bytes > cap) head = 0;
        const u64 a = head, b = head + bytes;
        while (!pending.empty()) {
            const Pending &q = pending.front();
            bool overlap = q.lo < b && a < q.hi;
            if (!overlap && hipEventQuery(q.ev) != hipSuccess) break;
            MOSSQ_HIP(hipEventSynchronize(q.ev));
            spare.push_back(q.ev);
            pending.pop_front();
        }
        for (const Pending &q : pending)
            if (q.lo < b && a < q.hi) MOSSQ_HIP(hipEventSynchronize(q.ev));
        *p = buf + a;
        *lo = a;
        head = b;
        return MOSSQ_OK;
    }
    int mark(hipStream_t s, u64 lo, u64 hi) {
        hipEvent_t ev;
        if (!spare.empty()) {
            ev = spare.back();
            spare.pop_back();
        } else {
            MOSSQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        }
        MOSSQ_HIP(hipEventRecord(ev, s));
        pending.push_back({lo, hi, ev});
        return MOSSQ_OK;
    }
    void destroy() {
        for (auto &q : pending) (void)hipEventDestroy(q.ev);
        for (auto &e : spare) (void)hipEventDestroy(e);
        if (buf) (void)hipHostFree(buf);
    }
};

struct Server {
    Ctx *ctx = nullptr;
    // The requests' device memory: the engine's vectors and tables come from
    // ectx (a context on ctx's stream), the workers' from their own contexts;
    // all of them, cached blocks included, count against `mem` (limit =
    // budget).
    Ctx *ectx = nullptr;
    std::vector<Ctx *> wctxs;  // worker contexts
    MemGroup mem;
    u64 budget = 0;
    std::atomic<u64> reserved{0};
    std::mutex mu;
    std::deque<SRequest *> queue;                     // waiting for admission
    std::unordered_map<i64, SRequest *> requests;     // all live requests
    std::vector<SRequest *> active;                   // engine items
    i64 next_ticket = 1;
    int t = 0;
    PinnedRing ring;
    // engine tables
    BItemDev *d_items[2] = {nullptr, nullptr};
    BItemDev *h_items[2] = {nullptr, nullptr};
    hipEvent_t table_ev[2] = {nullptr, nullptr};
    int cur_table = 0;
    bool table_dirty = true;
    int nchunks = 0;
    c128 *part = nullptr;
    u64 part_cap = 0;
    double *d_values = nullptr, *d_grads = nullptr;
    double *h_results = nullptr;  // pinned: per slot 1 + kMaxGrads doubles
    std::vector<int> free_slots;
    hipEvent_t iter_ev[kInflight] = {};
    // retiring engine items (results copy in flight)
    std::vector<SRequest *> draining;
    std::vector<i64> completed;  // finished tickets not yet reported
    // workers
    std::vector<std::thread> workers;
    std::deque<SRequest *> wqueue;  // density-matrix and wide-gate requests
    std::deque<SRequest *> fqueue;  // fused state-vector requests
    std::condition_variable wcv;
    bool stop = false;
    int worker_busy = 0;
    u64 finished = 0;                  // requests finished so far (under mu)
    std::condition_variable done_cv;   // notified when one finishes
};

static int build_request_program(SRequest *r) {
    r->prog = Program();
    return program_build_host(&r->prog, r->n, r->is_density, int(r->ops.size()), r->ops.data(),
                              (const double *)r->pool.data(), r->pool.size());
}

// One worker request on worker context wc: upload, plan, adjoint, release.
static int run_worker_request(Ctx *wc, SRequest *r) {
    int st = program_upload(wc, &r->prog);
    if (st == MOSSQ_OK && r->fused) st = fused_plan_build(&r->prog);
    if (st == MOSSQ_OK && r->is_density) st = density_adjoint_plan_build(&r->prog);
    r->grads.assign(std::max(r->prog.nderiv, 1), 0.0);
    if (st == MOSSQ_OK)
        st = r->is_density ? adjoint_density(wc, &r->prog, r->obs, nullptr, 0, nullptr, &r->value, r->grads.data())
                           : adjoint_pure(wc, &r->prog, r->obs, nullptr, nullptr, &r->value, r->grads.data());
    fused_plan_free(&r->prog);
    if (r->prog.d_pool) dev_free(wc, r->prog.d_pool);
    r->prog.d_pool = nullptr;
    return st;
}

static void worker_main(Server *s, int wi) {
    Ctx *wc = s->wctxs[wi];
    (void)hipSetDevice(wc->device);
    std::deque<SRequest *> &q = wi < kFusedWorkers ? s->fqueue : s->wqueue;
    while (true) {
        SRequest *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->wcv.wait(lk, [&] { return s->stop || !q.empty(); });
            if (s->stop && q.empty()) return;
            r = q.front();
            q.pop_front();
            s->worker_busy++;
        }
        int st = run_worker_request(wc, r);
        // Admission counts each request's estimated bytes; when the blocks of
        // the requests running beside this one leave too little room (the
        // estimate fell short), wait for one of them to finish and run again.
        while (st == MOSSQ_EOOM) {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop || s->reserved.load() <= r->bytes) break;  // nothing else holds memory
            const u64 seen = s->finished;
            s->done_cv.wait_for(lk, std::chrono::milliseconds(50), [&] { return s->stop || s->finished != seen; });
            lk.unlock();
            st = run_worker_request(wc, r);
        }
        r->grads.resize(r->prog.nderiv);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (st != MOSSQ_OK) r->error = mossq_last_error();
            r->state = st == MOSSQ_OK ? 3 : -1;
            s->reserved -= r->bytes;
            s->worker_busy--;
            s->finished++;
            s->completed.push_back(r->ticket);
        }
        s->done_cv.notify_all();
    }
}

// Size of the block the block cache hands out for `bytes`.
static u64 block_bytes(u64 bytes) {
    u64 r = 4096;
    while (r < bytes) r <<= 1;
    return r;
}

// The engine's blocks for a request (vectors, matrix pool, one micro-op per
// instruction), as the cache rounds them.
static u64 engine_bytes(const SRequest *r) {
    return block_bytes(2 * (u64(1) << r->n) * sizeof(c128)) +
           block_bytes(std::max<size_t>(r->prog.h_pool.size(), 1) * sizeof(c128)) +
           block_bytes(std::max<size_t>(r->prog.steps.size(), 1) * sizeof(BOp));
}

static int ensure_part(Server *s, int need) {
    if (u64(need) <= s->part_cap) return MOSSQ_OK;
    u64 cap = std::max<u64>(u64(need) * 2, 1 << 16);
    c128 *p = nullptr;
    MOSSQ_TRY(dev_alloc(s->ctx, cap * sizeof(c128), (void **)&p));
    if (s->part) dev_free(s->ctx, s->part);  // ordered after queued kernels
    s->part = p;
    s->part_cap = cap;
    return MOSSQ_OK;
}

// Admit one engine request at iteration s->t (caller holds no lock on device work).
static int admit_engine(Server *s, SRequest *r) {
    std::vector<BOp> ops;
    int nd = 0;
    MOSSQ_TRY(batch_lower(&r->prog, ops, &nd));
    MOSSQ_CHECK(nd <= kMaxGrads, MOSSQ_EUNSUPPORTED, "request has too many parameters for the engine");
    MOSSQ_CHECK(!s->free_slots.empty(), MOSSQ_EOOM, "engine is full");
    Ctx *c = s->ctx;
    r->vec_bytes = 2 * (u64(1) << r->n) * sizeof(c128);
    r->pool_bytes = std::max<size_t>(r->prog.h_pool.size(), 1) * sizeof(c128);
    r->ops_bytes = std::max<size_t>(ops.size(), 1) * sizeof(BOp);
    MOSSQ_TRY(cached_alloc(s->ectx, r->vec_bytes, (void **)&r->vec));
    MOSSQ_TRY(cached_alloc(s->ectx, r->pool_bytes, (void **)&r->d_pool));
    MOSSQ_TRY(cached_alloc(s->ectx, r->ops_bytes, (void **)&r->d_ops));
    char *stage = nullptr;
    u64 lo = 0;
    const u64 up = r->pool_bytes + r->ops_bytes;
    MOSSQ_TRY(s->ring.reserve(up, &stage, &lo));
    memcpy(stage, r->prog.h_pool.data(), r->prog.h_pool.size() * sizeof(c128));
    memcpy(stage + r->pool_bytes, ops.data(), ops.size() * sizeof(BOp));
    MOSSQ_HIP(hipMemcpyAsync(r->d_pool, stage, r->pool_bytes, hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipMemcpyAsync(r->d_ops, stage + r->pool_bytes, r->ops_bytes, hipMemcpyHostToDevice, c->stream));
    MOSSQ_TRY(s->ring.mark(c->stream, lo, lo + up));
    MOSSQ_TRY(fill_basis0(c, r->vec, u64(1) << r->n));
    r->slot = s->free_slots.back();
    s->free_slots.pop_back();
    r->nfwd = int(ops.size());
    r->nchunk = batch_chunks(r->n);
    r->t0 = s->t;
    r->engine = true;
    r->state = 1;
    s->active.push_back(r);
    s->table_dirty = true;
    return MOSSQ_OK;
}

static int rebuild_table(Server *s) {
    const int k = s->cur_table ^ 1;
    if (s->table_ev[k]) MOSSQ_HIP(hipEventSynchronize(s->table_ev[k]));
    int ch = 0;
    BItemDev *h = s->h_items[k];
    for (size_t i = 0; i < s->active.size(); ++i) {
        SRequest *r = s->active[i];
        BItemDev &d = h[i];
        const u64 len = u64(1) << r->n;
        d.psi = r->vec;
        d.lam = r->vec + len;
        d.ops = r->d_ops;
        d.pool = r->d_pool;
        d.groups = r->obs->d_groups;
        d.mats = r->obs->d_mats;
        d.seed = r->obs->d_mats_seed;
        d.n = r->n;
        d.nfwd = r->nfwd;
        d.G = r->obs->G;
        d.t0 = r->t0;
        d.chunk0 = ch;
        d.nchunk = r->nchunk;
        d.gslot0 = r->slot * kMaxGrads;
        d.vslot = r->slot;
        d.value_only = 0;
        d.preset = 0;
        d.seed_twice = r->obs->seed_twice ? 1 : 0;
        ch += r->nchunk;
    }
    s->nchunks = ch;
    MOSSQ_TRY(ensure_part(s, ch));
    if (!s->active.empty())
        MOSSQ_HIP(hipMemcpyAsync(s->d_items[k], h, s->active.size() * sizeof(BItemDev), hipMemcpyHostToDevice,
                                 s->ctx->stream));
    if (!s->table_ev[k]) MOSSQ_HIP(hipEventCreateWithFlags(&s->table_ev[k], hipEventDisableTiming));
    MOSSQ_HIP(hipEventRecord(s->table_ev[k], s->ctx->stream));
    s->cur_table = k;
    s->table_dirty = false;
    return MOSSQ_OK;
}

// Retire engine items whose last micro-op ran in iteration t-1.
static int retire(Server *s) {
    Ctx *c = s->ctx;
    std::vector<SRequest *> keep;
    for (SRequest *r : s->active) {
        if (s->t - 1 < r->t0 + 2 * r->nfwd) {
            keep.push_back(r);
            continue;
        }
        const int nd = r->prog.nderiv;
        double *dst = s->h_results + size_t(r->slot) * (1 + kMaxGrads);
        MOSSQ_HIP(hipMemcpyAsync(dst, s->d_values + r->slot, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        if (nd)
            MOSSQ_HIP(hipMemcpyAsync(dst + 1, s->d_grads + size_t(r->slot) * kMaxGrads, nd * sizeof(double),
                                     hipMemcpyDeviceToHost, c->stream));
        if (!r->done_ev) MOSSQ_HIP(hipEventCreateWithFlags(&r->done_ev, hipEventDisableTiming));
        MOSSQ_HIP(hipEventRecord(r->done_ev, c->stream));
        // stream order keeps later users of these blocks behind this request's kernels
        MOSSQ_TRY(cached_free(s->ectx, r->vec));
        MOSSQ_TRY(cached_free(s->ectx, r->d_pool));
        MOSSQ_TRY(cached_free(s->ectx, r->d_ops));
        r->vec = r->d_pool = nullptr;
        r->d_ops = nullptr;
        r->state = 2;
        s->draining.push_back(r);
        s->table_dirty = true;
    }
    s->active.swap(keep);
    return MOSSQ_OK;
}

// Results whose copies have landed: move them out of pinned memory, free the slot.
static void collect(Server *s) {
    std::vector<SRequest *> keep;
    for (SRequest *r : s->draining) {
        if (hipEventQuery(r->done_ev) != hipSuccess) {
            keep.push_back(r);
            continue;
        }
        const double *src = s->h_results + size_t(r->slot) * (1 + kMaxGrads);
        r->value = src[0];
        r->grads.assign(src + 1, src + 1 + r->prog.nderiv);
        s->free_slots.push_back(r->slot);
        r->slot = -1;
        s->reserved -= r->bytes;
        r->state = 3;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->finished++;
            s->completed.push_back(r->ticket);
        }
        s->done_cv.notify_all();
    }
    s->draining.swap(keep);
}

static int server_step(Server *s, int *active, int *queued) {
    Ctx *c = s->ctx;
    collect(s);
    MOSSQ_TRY(retire(s));
    // admission (FIFO): engine requests start at this iteration; others go to the worker
    while (true) {
        SRequest *r = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (s->queue.empty()) break;
            r = s->queue.front();
            if (s->reserved + r->bytes > s->budget) break;
            if (r->engine && s->free_slots.empty()) break;
            s->queue.pop_front();
            s->reserved += r->bytes;
            if (!r->engine) {
                r->state = 1;
                (r->fused ? s->fqueue : s->wqueue).push_back(r);
                s->wcv.notify_all();
                continue;
            }
        }
        int st = admit_engine(s, r);
        if (st != MOSSQ_OK) {
            r->error = mossq_last_error();
            r->state = -1;
            s->reserved -= r->bytes;
            std::lock_guard<std::mutex> lk(s->mu);
            s->completed.push_back(r->ticket);
        }
    }
    if (s->table_dirty) MOSSQ_TRY(rebuild_table(s));
    if (!s->active.empty()) {
        hipEvent_t &ev = s->iter_ev[s->t % kInflight];
        if (ev)
            MOSSQ_HIP(hipEventSynchronize(ev));
        else
            MOSSQ_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        MOSSQ_TRY(batch_iteration(c, s->d_items[s->cur_table], int(s->active.size()), s->nchunks, s->t, s->part,
                                  s->d_values, s->d_grads));
        MOSSQ_HIP(hipEventRecord(ev, c->stream));
        s->t++;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (active)
        *active = int(s->active.size() + s->draining.size()) + int(s->wqueue.size() + s->fqueue.size()) +
                  s->worker_busy;
    if (queued) *queued = int(s->queue.size());
    return MOSSQ_OK;
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_server_create(mossq_ctx *ctx, uint64_t budget_bytes, mossq_server **out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(out && budget_bytes > 0, MOSSQ_EINVAL, "bad server arguments");
    Server *s = new Server();
    s->ctx = C(ctx);
    s->budget = budget_bytes;
    s->mem.limit = budget_bytes;
    int st = ctx_borrow(s->ctx, &s->ectx);
    if (st == MOSSQ_OK) {
        s->ectx->cache_allocs = true;
        s->ectx->group = &s->mem;
        s->mem.members.push_back(s->ectx);
    }
    for (int i = 0; i < kWorkers && st == MOSSQ_OK; ++i) {
        mossq_ctx *w = nullptr;
        st = mossq_ctx_create(s->ctx->device, 0, &w);
        if (st == MOSSQ_OK) {
            C(w)->cache_allocs = true;  // no device-wide waits when a request's buffers are released
            C(w)->group = &s->mem;
            s->mem.members.push_back(C(w));
            s->wctxs.push_back(C(w));
        }
    }
    if (st == MOSSQ_OK) st = s->ring.init(kRingBytes);
    for (int k = 0; st == MOSSQ_OK && k < 2; ++k) {
        st = dev_alloc(s->ctx, kMaxActive * sizeof(BItemDev), (void **)&s->d_items[k]);
        if (st == MOSSQ_OK && hipHostMalloc((void **)&s->h_items[k], kMaxActive * sizeof(BItemDev)) != hipSuccess)
            st = fail(MOSSQ_EOOM, "pinned item table");
    }
    if (st == MOSSQ_OK) st = dev_alloc(s->ctx, kMaxActive * sizeof(double), (void **)&s->d_values);
    if (st == MOSSQ_OK) st = dev_alloc(s->ctx, size_t(kMaxActive) * kMaxGrads * sizeof(double), (void **)&s->d_grads);
    if (st == MOSSQ_OK &&
        hipHostMalloc((void **)&s->h_results, size_t(kMaxActive) * (1 + kMaxGrads) * sizeof(double)) != hipSuccess)
        st = fail(MOSSQ_EOOM, "pinned result arena");
    if (st != MOSSQ_OK) {
        mossq_server_destroy(reinterpret_cast<mossq_server *>(s));
        return st;
    }
    for (int i = kMaxActive - 1; i >= 0; --i) s->free_slots.push_back(i);
    for (int i = 0; i < kWorkers; ++i) s->workers.push_back(std::thread(worker_main, s, i));
    *out = reinterpret_cast<mossq_server *>(s);
    return MOSSQ_OK;
}

int mossq_server_destroy(mossq_server *srv) {
    if (!srv) return MOSSQ_OK;
    Server *s = reinterpret_cast<Server *>(srv);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->wcv.notify_all();
    s->done_cv.notify_all();
    for (auto &w : s->workers)
        if (w.joinable()) w.join();
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto &kv : s->requests) {
        SRequest *r = kv.second;
        if (r->done_ev) (void)hipEventDestroy(r->done_ev);
        delete r;
    }
    for (int k = 0; k < 2; ++k) {
        if (s->d_items[k]) dev_free(s->ctx, s->d_items[k]);
        if (s->h_items[k]) (void)hipHostFree(s->h_items[k]);
        if (s->table_ev[k]) (void)hipEventDestroy(s->table_ev[k]);
    }
    for (auto &e : s->iter_ev)
        if (e) (void)hipEventDestroy(e);
    if (s->part) dev_free(s->ctx, s->part);
    if (s->d_values) dev_free(s->ctx, s->d_values);
    if (s->d_grads) dev_free(s->ctx, s->d_grads);
    if (s->h_results) (void)hipHostFree(s->h_results);
    s->ring.destroy();
    for (Ctx *w : s->wctxs) mossq_ctx_destroy(reinterpret_cast<mossq_ctx *>(w));
    if (s->ectx) mossq_ctx_destroy(reinterpret_cast<mossq_ctx *>(s->ectx));
    delete s;
    return MOSSQ_OK;
}

int mossq_server_submit(mossq_server *srv, int n, int is_density, int nops, const mossq_op *ops, const double *pool,
                        uint64_t pool_len, const mossq_obs *obs, int64_t *ticket) {
    MOSSQ_CHECK(srv && obs && ticket, MOSSQ_EINVAL, "null argument");
    Server *s = reinterpret_cast<Server *>(srv);
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    MOSSQ_CHECK(o->ctx == s->ctx, MOSSQ_EINVAL, "observable belongs to another context");
    MOSSQ_CHECK(o->n == n, MOSSQ_EINVAL, "observable and request qubit counts differ");
    SRequest *r = new SRequest();
    r->n = n;
    r->is_density = is_density ? 1 : 0;
    r->ops.assign(ops, ops + nops);
    r->pool.assign((const c128 *)pool, (const c128 *)pool + pool_len);
    r->obs = o;
    int st = build_request_program(r);
    if (st != MOSSQ_OK) {
        delete r;
        return st;
    }
    bool engine_ok = !r->is_density;
    if (engine_ok)
        for (const Step &x : r->prog.steps) engine_ok &= (!x.split && x.kb <= 2);
    engine_ok = engine_ok && r->prog.nderiv <= kMaxGrads;
    r->fused = !r->is_density && n >= kFusedMinQubits;
    r->engine = engine_ok && !r->fused;
    // Estimated device bytes, as the block cache rounds them: the program's
    // pool, the plan tables (bounded by a few times the pool), the states and
    // the partial sums.
    const u64 pool_blocks = block_bytes(r->prog.h_pool.size() * sizeof(c128)) * 9 + (4ull << 20);
    if (r->engine) {
        r->bytes = engine_bytes(r);
    } else if (r->fused) {
        // two states, gradient partials (one per wave per tile, or per pair-kernel block), plans
        const u64 row = std::max<u64>(u64(pair_parts(n)), u64(4) << (n - 12));
        const u64 parts = std::max<u64>(u64(std::max(r->prog.nderiv, 1)) * row, u64(o->G) * 512);
        r->bytes = 2 * block_bytes((u64(1) << n) * sizeof(c128)) + block_bytes(parts * sizeof(c128)) + pool_blocks;
    } else if (r->is_density) {
        // (the step-by-step layout's states; the fused chunk plans keep fewer)
        r->bytes = adjoint_density_bytes(&r->prog, o) + pool_blocks;
    } else {
        r->bytes = 2 * block_bytes(2 * (u64(1) << n) * sizeof(c128)) + pool_blocks;
    }
    if (r->bytes > s->budget) {
        delete r;
        return fail(MOSSQ_EOOM, "request needs more device memory than the server budget");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    s->requests[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
    return MOSSQ_OK;
}

int mossq_server_step(mossq_server *srv, int *active, int *queued) {
    MOSSQ_CHECK(srv, MOSSQ_EINVAL, "null server");
    Server *s = reinterpret_cast<Server *>(srv);
    MOSSQ_HIP(hipSetDevice(s->ctx->device));
    return server_step(s, active, queued);
}

int mossq_server_completed(mossq_server *srv, int64_t *tickets, int max, int *count) {
    MOSSQ_CHECK(srv && count && (tickets || max == 0), MOSSQ_EINVAL, "null argument");
    Server *s = reinterpret_cast<Server *>(srv);
    collect(s);
    std::lock_guard<std::mutex> lk(s->mu);
    int k = std::min<int>(max, int(s->completed.size()));
    for (int i = 0; i < k; ++i) tickets[i] = s->completed[i];
    s->completed.erase(s->completed.begin(), s->completed.begin() + k);
    *count = k;
    return MOSSQ_OK;
}

int mossq_server_memory(mossq_server *srv, uint64_t *in_use, uint64_t *peak) {
    MOSSQ_CHECK(srv, MOSSQ_EINVAL, "null server");
    Server *s = reinterpret_cast<Server *>(srv);
    std::lock_guard<std::mutex> lk(s->mem.mu);
    if (in_use) *in_use = s->mem.used;
    if (peak) *peak = s->mem.peak;
    return MOSSQ_OK;
}

int mossq_server_poll(mossq_server *srv, int64_t ticket, int *done, double *value, double *grads, int max_grads) {
    MOSSQ_CHECK(srv && done, MOSSQ_EINVAL, "null argument");
    Server *s = reinterpret_cast<Server *>(srv);
    SRequest *r = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->requests.find(ticket);
        MOSSQ_CHECK(it != s->requests.end(), MOSSQ_EINVAL, "unknown or already collected ticket");
        r = it->second;
    }
    if (r->state == 2) collect(s);
    int st = r->state;
    if (st == -1) {
        *done = 1;
        std::string e = r->error;
        std::lock_guard<std::mutex> lk(s->mu);
        s->requests.erase(ticket);
        delete r;
        return fail(MOSSQ_EDEVICE, "request failed: " + e);
    }
    if (st != 3) {
        *done = 0;
        return MOSSQ_OK;
    }
    *done = 1;
    if (value) *value = r->value;
    if (grads) {
        MOSSQ_CHECK(max_grads >= int(r->grads.size()), MOSSQ_EINVAL, "gradient buffer too small");
        std::copy(r->grads.begin(), r->grads.end(), grads);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->requests.erase(ticket);
    if (r->done_ev) (void)hipEventDestroy(r->done_ev);
    delete r;
    return MOSSQ_OK;
}

}  // extern "C"
