// Internal declarations shared by the translation units of libmossq_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <map>
#include <unordered_set>
#include <vector>

#include "../../include/mossq_hip.h"

namespace mossq {

using u64 = uint64_t;
using i64 = int64_t;

// ---------------------------------------------------------------- errors ----
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define MOSSQ_HIP(call)                                                              \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return ::mossq::fail(MOSSQ_EDEVICE, std::string(#call) + ": " +          \
                                                    hipGetErrorString(e_));          \
    } while (0)

#define MOSSQ_CHECK(cond, code, msg)                                                 \
    do {                                                                             \
        if (!(cond)) return ::mossq::fail((code), (msg));                            \
    } while (0)

#define MOSSQ_TRY(expr)                                                              \
    do {                                                                             \
        int s_ = (expr);                                                             \
        if (s_ != MOSSQ_OK) return s_;                                               \
    } while (0)

// --------------------------------------------------------------- complex ----
// complex128 as double2 (re, im); 16-byte aligned -> one dwordx4 access.
struct __align__(16) c128 {
    double re, im;
};

__host__ __device__ __forceinline__ c128 mk(double r, double i) { return c128{r, i}; }
__host__ __device__ __forceinline__ c128 cadd(c128 a, c128 b) { return c128{a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ c128 csub(c128 a, c128 b) { return c128{a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ c128 cmul(c128 a, c128 b) {
    return c128{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
// acc + a*b
__host__ __device__ __forceinline__ c128 cfma(c128 a, c128 b, c128 acc) {
    return c128{fma(a.re, b.re, fma(-a.im, b.im, acc.re)), fma(a.re, b.im, fma(a.im, b.re, acc.im))};
}
// acc + conj(a)*b
__host__ __device__ __forceinline__ c128 cfmaconj(c128 a, c128 b, c128 acc) {
    return c128{fma(a.re, b.re, fma(a.im, b.im, acc.re)), fma(a.re, b.im, fma(-a.im, b.re, acc.im))};
}
__host__ __device__ __forceinline__ c128 cconj(c128 a) { return c128{a.re, -a.im}; }

// Streaming accesses (each amplitude is read and written once per sweep):
// nontemporal hints keep the state from displacing cached data and measured
// ~5% faster one-qubit sweeps on MI355X.
__device__ __forceinline__ c128 load_nt(const c128 *p) {
    return c128{__builtin_nontemporal_load(&p->re), __builtin_nontemporal_load(&p->im)};
}
__device__ __forceinline__ void store_nt(c128 *p, c128 x) {
    __builtin_nontemporal_store(x.re, &p->re);
    __builtin_nontemporal_store(x.im, &p->im);
}


// ------------------------------------------------------ wave reductions ----
// Lane exchanges of a double without LDS: DPP moves inside 16-lane rows and
// gfx950's permlane swaps across rows, two 32-bit halves each.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, int(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, int(v >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(unsigned)lo | ((long long)hi << 32));
}

// Sum over the 64 lanes of a wave, the same in every lane for a given lane
// position and run (the order is fixed): rotations by 8, 4, 2, 1 inside each
// row (DPP row_ror), then the rows pairwise (v_permlane16_swap) and the
// halves (v_permlane32_swap).  Every step is a VALU op, where a shuffle
// through ds_bpermute waits on LDS each time.
__device__ __forceinline__ double wave_sum_f64(double x) {
    x += dpp_f64<0x128>(x);  // row_ror:8
    x += dpp_f64<0x124>(x);  // row_ror:4
    x += dpp_f64<0x122>(x);  // row_ror:2
    x += dpp_f64<0x121>(x);  // row_ror:1
    {
        const long long v = __double_as_longlong(x);
        const auto a = __builtin_amdgcn_permlane16_swap(int(v), int(v), false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(int(v >> 32), int(v >> 32), false, false);
        // a[0]: even rows keep x, odd rows take the even row below; a[1] the other way
        x = __longlong_as_double((long long)(unsigned)a[0] | ((long long)b[0] << 32)) +
            __longlong_as_double((long long)(unsigned)a[1] | ((long long)b[1] << 32));
    }
    {
        const long long v = __double_as_longlong(x);
        const auto a = __builtin_amdgcn_permlane32_swap(int(v), int(v), false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(int(v >> 32), int(v >> 32), false, false);
        x = __longlong_as_double((long long)(unsigned)a[0] | ((long long)b[0] << 32)) +
            __longlong_as_double((long long)(unsigned)a[1] | ((long long)b[1] << 32));
    }
    return x;
}

// ----------------------------------------------------------------- bits ----
// Insert a zero bit at position q of x (bits >= q move up by one).
__host__ __device__ __forceinline__ u64 insert0(u64 x, int q) {
    u64 lo = x & ((u64(1) << q) - 1);
    return ((x >> q) << (q + 1)) | lo;
}
// Insert zeros at k increasing positions.
template <int K>
__host__ __device__ __forceinline__ u64 insert0s(u64 x, const int *pos) {
#pragma unroll
    for (int j = 0; j < K; ++j) x = insert0(x, pos[j]);
    return x;
}
__host__ __device__ __forceinline__ u64 insert0s_dyn(u64 x, const int *pos, int k) {
    for (int j = 0; j < k; ++j) x = insert0(x, pos[j]);
    return x;
}

// ------------------------------------------------------------- context ----
constexpr int kNumEvents = 16;

struct Ctx;

// A device-memory limit shared by several contexts (the request server's
// engine and worker contexts).  Every raw allocation of a member counts
// against it, cached blocks included; an allocation that would pass the limit
// first releases the cached blocks of every member.
struct MemGroup {
    std::mutex mu;
    u64 limit = 0, used = 0, peak = 0;
    std::vector<Ctx *> members;  // fixed before any member allocates
};

struct Ctx {
    int device = 0;
    int num_cu = 0;  // compute units of the device
    hipStream_t stream = nullptr;
    u64 budget = 0;  // 0 = unlimited
    u64 in_use = 0;
    std::vector<std::pair<void *, u64>> allocs;
    hipEvent_t events[kNumEvents] = {};
    // reusable device scratch (grown on demand, owned by the context)
    void *scratch = nullptr;
    u64 scratch_bytes = 0;
    void *mat_stage = nullptr;  // device staging for per-call matrices
    u64 mat_stage_bytes = 0;
    double *host_pinned = nullptr;  // pinned host staging for results
    u64 host_pinned_bytes = 0;
    unsigned char *upd_pinned = nullptr;  // pinned host staging of batched parameter updates
    u64 upd_pinned_bytes = 0;
    // second in-order stream (default priority) for work that overlaps the
    // main stream's, with two ordering events; created on first use
    hipStream_t aux = nullptr;
    hipEvent_t aux_ev[2] = {};
    // Block cache: cached_alloc rounds sizes up to powers of two and
    // cached_free keeps blocks (up to kAllocCacheCap bytes) for reuse instead
    // of hipFree, which waits for the whole device.  Blocks are reused on this
    // context's stream only, after the work queued on it; a budget or hipMalloc
    // failure releases them first.  The library's per-call scratch (Allocs)
    // always goes through it; with cache_allocs every dev_alloc does.
    bool cache_allocs = false;
    std::map<u64, std::vector<void *>> cache;
    std::unordered_set<void *> in_cache;  // blocks now held by the cache
    u64 cached = 0;
    // allocs, in_use, cache, in_cache and cached change under mem_mu: a
    // MemGroup member may release another member's cache from its own thread
    std::mutex mem_mu;
    MemGroup *group = nullptr;
    bool own_stream = true;  // false: `stream` is borrowed from another context
};
constexpr u64 kAllocCacheCap = u64(16) << 30;

// The context's second stream and its events (created on first use).
int ctx_aux(Ctx *c);
// A context without a stream of its own: allocations are its own (its cache,
// its group), launches and copies go to base's stream.
int ctx_borrow(Ctx *base, Ctx **out);
// Releases every cached block of c (after the work queued on its streams).
int cache_flush(Ctx *c);
// Address of the allocation-free adjoint's device array on `device` (which
// must be current), taken once per device at context creation (program.hip).
int noalloc_scratch_init(int device);

int ctx_check(mossq_ctx *ctx);
Ctx *C(mossq_ctx *ctx);
// Device allocation that respects the context's memory budget.
int dev_alloc(Ctx *c, u64 bytes, void **p);
int dev_free(Ctx *c, void *p);
// Scratch of at least `bytes`, valid until the next scratch request.
int get_scratch(Ctx *c, u64 bytes, void **p);
// Upload `count` complex numbers to the context's matrix staging area.
int stage_matrix(Ctx *c, const c128 *host, u64 count, const c128 **dev);
int get_pinned(Ctx *c, u64 bytes, double **p);
// Pinned host staging of at least `bytes` for batched parameter updates
// (its previous contents consumed: every such update waits for its copy).
int get_update_pinned(Ctx *c, u64 bytes, unsigned char **p);

// A device table that a kernel only reads, as a constant-memory pointer: a
// pointer read from memory is otherwise generic and loads through the flat
// path into vector registers, where constant loads of uniform addresses are
// scalar.
template <class T>
__device__ __forceinline__ const T *const_table(const T *p) {
    typedef __attribute__((address_space(4))) const T cT;
    return (const T *)(cT *)reinterpret_cast<uintptr_t>(p);
}

int cached_alloc(Ctx *c, u64 bytes, void **p);
int cached_free(Ctx *c, void *p);

// Orders the context's main stream after its second stream when the scope
// ends (declared after an Allocs, it runs first, so blocks read by work on the
// second stream return to the cache only behind that work).
struct AuxJoin {
    Ctx *c;
    bool active = false;
    explicit AuxJoin(Ctx *cc) : c(cc) {}
    ~AuxJoin() {
        if (active && c->aux) {
            (void)hipEventRecord(c->aux_ev[1], c->aux);
            (void)hipStreamWaitEvent(c->stream, c->aux_ev[1], 0);
        }
    }
};

// Scoped per-call scratch: everything obtained through get() goes back to the
// context's block cache on exit.
struct Allocs {
    Ctx *c;
    std::vector<void *> ptrs;
    explicit Allocs(Ctx *cc) : c(cc) {}
    int get(u64 bytes, void **p) {
        int st = cached_alloc(c, bytes == 0 ? 16 : bytes, p);
        if (st == MOSSQ_OK) ptrs.push_back(*p);
        return st;
    }
    ~Allocs() {
        for (void *p : ptrs) cached_free(c, p);
    }
};

// --------------------------------------------------------- gate launches ----
// Classification of a local matrix, used to pick a kernel.
struct MatInfo {
    bool monomial = false;  // exactly one nonzero per row and column
    int nchanged = 0;       // rows with a non-identity action (monomial only)
    uint8_t rows[32];       // changed rows
    uint8_t src[32];        // source column of each changed row
    c128 val[32];           // matrix entry of each changed row
};
MatInfo classify(const c128 *m, int dim);

// Apply a dense matrix on `k` (sorted) bits of a 2^nbits vector; dispatches to the
// monomial kernel when `info.monomial`.  `dmat` is the device copy of the matrix.
int launch_dense(Ctx *c, c128 *v, int nbits, int k, const int *bits, const c128 *dmat,
                 const MatInfo &info);
// Reverse-pass fused kernel: applies matrix V on x and (unless y_fixed) on y,
// and accumulates per-workgroup partials of sum_b x_b^dagger D y_b into part
// (if D != nullptr; pair_parts(nbits) partials).  k <= 4.
int launch_dense_pair(Ctx *c, c128 *x, c128 *y, int nbits, int k, const int *bits,
                      const c128 *dV, const c128 *dD, c128 *part, bool y_fixed);
int pair_parts(int nbits);
// Host-matrix application (monomial detection, staging) on `k` sorted bits.
int apply_host_matrix(Ctx *c, c128 *v, int nbits, int k, const int *bits, const c128 *m);
int fill_basis0(Ctx *c, c128 *v, u64 len);
int outer_product(Ctx *c, c128 *rho, const c128 *psi, int n);
// kron(conj(U), U): the density-matrix action of a k-qubit gate on bits (pos, pos+n).
void dm_superop_of_gate(const c128 *U, int k, std::vector<c128> &out);
void conj_transpose(const c128 *m, int dim, std::vector<c128> &out);

// Reduction helpers (deterministic two-stage).
int reduce_partials(Ctx *c, const c128 *part, int nparts, c128 *dev_out);

}  // namespace mossq
