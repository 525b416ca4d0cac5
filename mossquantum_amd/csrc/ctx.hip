// Context, device buffers, copies, errors and events of libmossq_hip.so.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "mossq_internal.h"

namespace mossq {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

Ctx *C(mossq_ctx *ctx) { return reinterpret_cast<Ctx *>(ctx); }

int ctx_check(mossq_ctx *ctx) {
    MOSSQ_CHECK(ctx != nullptr, MOSSQ_EINVAL, "null context");
    MOSSQ_HIP(hipSetDevice(C(ctx)->device));
    return MOSSQ_OK;
}

// Returns `bytes` to the context's memory group (if any).
static void group_release(Ctx *c, u64 bytes) {
    if (!c->group || !bytes) return;
    std::lock_guard<std::mutex> lk(c->group->mu);
    c->group->used -= bytes;
}

// Takes `bytes` from the group; past its limit, every member's cache is
// released first.
static int group_reserve(Ctx *c, u64 bytes) {
    MemGroup *g = c->group;
    for (int pass = 0; pass < 2; ++pass) {
        {
            std::lock_guard<std::mutex> lk(g->mu);
            if (g->used + bytes <= g->limit) {
                g->used += bytes;
                g->peak = std::max(g->peak, g->used);
                return MOSSQ_OK;
            }
        }
        if (pass == 0)
            for (Ctx *m : g->members) MOSSQ_TRY(cache_flush(m));
    }
    return fail(MOSSQ_EOOM, "memory budget of the server exceeded: " + std::to_string(g->used) + " + " +
                                std::to_string(bytes) + " > " + std::to_string(g->limit));
}

// Erases p from the allocation list (caller holds mem_mu); returns its size.
static u64 forget_alloc(Ctx *c, void *p) {
    auto it = std::find_if(c->allocs.begin(), c->allocs.end(), [&](auto &a) { return a.first == p; });
    if (it == c->allocs.end()) return 0;
    const u64 b = it->second;
    c->in_use -= b;
    c->allocs.erase(it);
    return b;
}

int cache_flush(Ctx *c) {
    u64 freed = 0;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        if (!c->cached) return MOSSQ_OK;
        // cached blocks may still be read by work queued on either stream
        MOSSQ_HIP(hipStreamSynchronize(c->stream));
        if (c->aux) MOSSQ_HIP(hipStreamSynchronize(c->aux));
        for (auto &kv : c->cache)
            for (void *p : kv.second) {
                (void)hipFree(p);
                freed += forget_alloc(c, p);
            }
        c->cache.clear();
        c->in_cache.clear();
        c->cached = 0;
    }
    group_release(c, freed);
    return MOSSQ_OK;
}

static int raw_alloc(Ctx *c, u64 bytes, void **p) {
    if (c->budget && c->in_use + bytes > c->budget) MOSSQ_TRY(cache_flush(c));
    if (c->budget && c->in_use + bytes > c->budget)
        return fail(MOSSQ_EOOM, "device memory budget exceeded: " + std::to_string(c->in_use) + " + " +
                                    std::to_string(bytes) + " > " + std::to_string(c->budget));
    if (c->group) MOSSQ_TRY(group_reserve(c, bytes));
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && c->cached) {
        (void)hipGetLastError();
        const int st = cache_flush(c);
        if (st != MOSSQ_OK) {
            group_release(c, bytes);
            return st;
        }
        e = hipMalloc(p, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        group_release(c, bytes);
        return fail(MOSSQ_EOOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(c->mem_mu);
    c->allocs.emplace_back(*p, bytes);
    c->in_use += bytes;
    return MOSSQ_OK;
}

int cached_alloc(Ctx *c, u64 bytes, void **p) {
    u64 r = 4096;
    while (r < bytes) r <<= 1;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        auto it = c->cache.find(r);
        if (it != c->cache.end() && !it->second.empty()) {
            *p = it->second.back();
            it->second.pop_back();
            c->in_cache.erase(*p);
            c->cached -= r;
            return MOSSQ_OK;
        }
    }
    return raw_alloc(c, r, p);
}

int cached_free(Ctx *c, void *p) {
    if (!p) return MOSSQ_OK;
    u64 freed = 0;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        auto it = std::find_if(c->allocs.begin(), c->allocs.end(), [&](auto &a) { return a.first == p; });
        MOSSQ_CHECK(it != c->allocs.end(), MOSSQ_EINVAL, "pointer was not allocated by this context");
        MOSSQ_CHECK(!c->in_cache.count(p), MOSSQ_EINVAL, "device block released twice");
        if (c->cached + it->second <= kAllocCacheCap) {
            // kept for reuse on this stream, which orders it after the queued work
            c->cache[it->second].push_back(p);
            c->in_cache.insert(p);
            c->cached += it->second;
            return MOSSQ_OK;
        }
        MOSSQ_HIP(hipStreamSynchronize(c->stream));
        if (c->aux) MOSSQ_HIP(hipStreamSynchronize(c->aux));
        MOSSQ_HIP(hipFree(p));
        freed = forget_alloc(c, p);
    }
    group_release(c, freed);
    return MOSSQ_OK;
}

int dev_alloc(Ctx *c, u64 bytes, void **p) {
    if (bytes == 0) bytes = 16;
    if (c->cache_allocs) return cached_alloc(c, bytes, p);
    return raw_alloc(c, bytes, p);
}

int dev_free(Ctx *c, void *p) {
    if (!p) return MOSSQ_OK;
    if (c->cache_allocs) return cached_free(c, p);
    u64 freed = 0;
    {
        std::lock_guard<std::mutex> lk(c->mem_mu);
        auto it = std::find_if(c->allocs.begin(), c->allocs.end(), [&](auto &a) { return a.first == p; });
        MOSSQ_CHECK(it != c->allocs.end() && !c->in_cache.count(p), MOSSQ_EINVAL,
                    "pointer was not allocated by this context");
        // Memory may still be read by queued kernels: order the free after them.
        MOSSQ_HIP(hipStreamSynchronize(c->stream));
        if (c->aux) MOSSQ_HIP(hipStreamSynchronize(c->aux));
        MOSSQ_HIP(hipFree(p));
        freed = forget_alloc(c, p);
    }
    group_release(c, freed);
    return MOSSQ_OK;
}

int ctx_borrow(Ctx *base, Ctx **out) {
    Ctx *c = new Ctx();
    c->device = base->device;
    c->num_cu = base->num_cu;
    c->stream = base->stream;
    c->own_stream = false;
    *out = c;
    return MOSSQ_OK;
}

int ctx_aux(Ctx *c) {
    if (c->aux) return MOSSQ_OK;
    hipStream_t st = nullptr;
    MOSSQ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));  // default priority
    for (hipEvent_t &ev : c->aux_ev) {
        const hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            for (hipEvent_t &x : c->aux_ev)
                if (x) (void)hipEventDestroy(x), x = nullptr;
            (void)hipStreamDestroy(st);
            return fail(MOSSQ_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    c->aux = st;
    return MOSSQ_OK;
}

int get_scratch(Ctx *c, u64 bytes, void **p) {
    if (c->scratch_bytes < bytes) {
        MOSSQ_TRY(dev_free(c, c->scratch));
        c->scratch = nullptr;
        c->scratch_bytes = 0;
        u64 want = std::max<u64>(bytes, 1 << 20);
        MOSSQ_TRY(dev_alloc(c, want, &c->scratch));
        c->scratch_bytes = want;
    }
    *p = c->scratch;
    return MOSSQ_OK;
}

int stage_matrix(Ctx *c, const c128 *host, u64 count, const c128 **dev) {
    u64 bytes = count * sizeof(c128);
    if (c->mat_stage_bytes < bytes) {
        MOSSQ_TRY(dev_free(c, c->mat_stage));
        c->mat_stage = nullptr;
        c->mat_stage_bytes = 0;
        u64 want = std::max<u64>(bytes, 64 << 10);
        MOSSQ_TRY(dev_alloc(c, want, &c->mat_stage));
        c->mat_stage_bytes = want;
    }
    // Stream order keeps the copy behind every kernel that reads the previous
    // matrix; the wait makes sure `host` (caller memory) has been consumed.
    MOSSQ_HIP(hipMemcpyAsync(c->mat_stage, host, bytes, hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    *dev = reinterpret_cast<const c128 *>(c->mat_stage);
    return MOSSQ_OK;
}

int get_pinned(Ctx *c, u64 bytes, double **p) {
    if (c->host_pinned_bytes < bytes) {
        if (c->host_pinned) MOSSQ_HIP(hipHostFree(c->host_pinned));
        c->host_pinned = nullptr;
        u64 want = std::max<u64>(bytes, 64 << 10);
        MOSSQ_HIP(hipHostMalloc((void **)&c->host_pinned, want, hipHostMallocDefault));
        c->host_pinned_bytes = want;
    }
    *p = c->host_pinned;
    return MOSSQ_OK;
}

int get_update_pinned(Ctx *c, u64 bytes, unsigned char **p) {
    if (c->upd_pinned_bytes < bytes) {
        if (c->upd_pinned) MOSSQ_HIP(hipHostFree(c->upd_pinned));
        c->upd_pinned = nullptr;
        c->upd_pinned_bytes = 0;
        const u64 want = std::max<u64>(bytes + bytes / 4, 1 << 20);
        MOSSQ_HIP(hipHostMalloc((void **)&c->upd_pinned, want, hipHostMallocDefault));
        c->upd_pinned_bytes = want;
    }
    *p = c->upd_pinned;
    return MOSSQ_OK;
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_version(void) { return 1; }

const char *mossq_last_error(void) { return g_last_error.c_str(); }

int mossq_device_count(int *count) {
    MOSSQ_CHECK(count, MOSSQ_EINVAL, "null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return MOSSQ_OK;
}

int mossq_ctx_create(int device, uint64_t mem_budget_bytes, mossq_ctx **out) {
    MOSSQ_CHECK(out, MOSSQ_EINVAL, "null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(MOSSQ_EDEVICE, "no HIP device available (libmossq_hip has no CPU fallback)");
    }
    MOSSQ_CHECK(device >= 0 && device < n, MOSSQ_EINVAL, "device index out of range");
    MOSSQ_HIP(hipSetDevice(device));
    Ctx *c = new Ctx();
    c->device = device;
    c->budget = mem_budget_bytes;
    e = hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess || c->num_cu <= 0) {
        delete c;
        return fail(MOSSQ_EDEVICE, "hipDeviceGetAttribute(MultiprocessorCount) failed");
    }
    // Default priority, non-blocking w.r.t. the legacy null stream.
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(MOSSQ_EDEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < kNumEvents; ++i) {
        e = hipEventCreate(&c->events[i]);
        if (e != hipSuccess) {
            mossq_ctx_destroy(reinterpret_cast<mossq_ctx *>(c));
            return fail(MOSSQ_EDEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    // the allocation-free adjoint's scratch address on this device
    const int st = noalloc_scratch_init(device);
    if (st != MOSSQ_OK) {
        const std::string msg = mossq_last_error();
        mossq_ctx_destroy(reinterpret_cast<mossq_ctx *>(c));
        return fail(st, msg);
    }
    *out = reinterpret_cast<mossq_ctx *>(c);
    return MOSSQ_OK;
}

int mossq_ctx_destroy(mossq_ctx *ctx) {
    if (!ctx) return MOSSQ_OK;
    Ctx *c = C(ctx);
    MOSSQ_HIP(hipSetDevice(c->device));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    if (c->aux) {
        MOSSQ_HIP(hipStreamSynchronize(c->aux));
        for (hipEvent_t &ev : c->aux_ev)
            if (ev) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(c->aux);
    }
    u64 freed = 0;
    for (auto &a : c->allocs) {
        (void)hipFree(a.first);
        freed += a.second;
    }
    c->allocs.clear();
    group_release(c, freed);
    if (c->host_pinned) (void)hipHostFree(c->host_pinned);
    if (c->upd_pinned) (void)hipHostFree(c->upd_pinned);
    for (int i = 0; i < kNumEvents; ++i)
        if (c->events[i]) (void)hipEventDestroy(c->events[i]);
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MOSSQ_OK;
}

int mossq_sync(mossq_ctx *ctx) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_HIP(hipStreamSynchronize(C(ctx)->stream));
    return MOSSQ_OK;
}

int mossq_mem_in_use(mossq_ctx *ctx, uint64_t *bytes) {
    MOSSQ_CHECK(ctx && bytes, MOSSQ_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(C(ctx)->mem_mu);
    *bytes = C(ctx)->in_use;
    return MOSSQ_OK;
}

int mossq_alloc(mossq_ctx *ctx, uint64_t bytes, void **dptr) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(dptr, MOSSQ_EINVAL, "null output");
    return dev_alloc(C(ctx), bytes, dptr);
}

int mossq_free(mossq_ctx *ctx, void *dptr) {
    MOSSQ_TRY(ctx_check(ctx));
    return dev_free(C(ctx), dptr);
}

int mossq_memset_zero(mossq_ctx *ctx, void *dptr, uint64_t bytes) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(dptr || bytes == 0, MOSSQ_EINVAL, "null buffer");
    MOSSQ_HIP(hipMemsetAsync(dptr, 0, bytes, C(ctx)->stream));
    return MOSSQ_OK;
}

int mossq_copy_h2d(mossq_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK((dst && src) || bytes == 0, MOSSQ_EINVAL, "null buffer");
    MOSSQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, C(ctx)->stream));
    MOSSQ_HIP(hipStreamSynchronize(C(ctx)->stream));
    return MOSSQ_OK;
}

int mossq_copy_d2h(mossq_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK((dst && src) || bytes == 0, MOSSQ_EINVAL, "null buffer");
    MOSSQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C(ctx)->stream));
    MOSSQ_HIP(hipStreamSynchronize(C(ctx)->stream));
    return MOSSQ_OK;
}

int mossq_copy_d2d(mossq_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK((dst && src) || bytes == 0, MOSSQ_EINVAL, "null buffer");
    MOSSQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, C(ctx)->stream));
    return MOSSQ_OK;
}

int mossq_event_record(mossq_ctx *ctx, int slot) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(slot >= 0 && slot < kNumEvents, MOSSQ_EINVAL, "event slot out of range");
    MOSSQ_HIP(hipEventRecord(C(ctx)->events[slot], C(ctx)->stream));
    return MOSSQ_OK;
}

int mossq_event_elapsed_ms(mossq_ctx *ctx, int slot0, int slot1, float *ms) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(slot0 >= 0 && slot0 < kNumEvents && slot1 >= 0 && slot1 < kNumEvents && ms,
                MOSSQ_EINVAL, "bad event arguments");
    MOSSQ_HIP(hipEventSynchronize(C(ctx)->events[slot1]));
    MOSSQ_HIP(hipEventElapsedTime(ms, C(ctx)->events[slot0], C(ctx)->events[slot1]));
    return MOSSQ_OK;
}

}  // extern "C"
