// Programs (lowered circuits) and the adjoint gradient drivers.
//
// A program owns a device matrix pool holding, per instruction, the forward
// matrix, the reverse (uncompute) matrix and the gradient matrix, all already
// expressed on the bits of the stored vector:
//   state vector   gate   : U, U^dagger, D = (dU)^dagger U      on the k qubit bits
//   density matrix gate   : conj(U)(x)U, conj(U^dag)(x)U^dag,
//                           D^dagger(x)I + I(x)D^T              on bits (pos, pos+n)
//   density matrix channel: S, S^dagger                         on bits (pos, pos+n)
// so every launch is one sweep of a dense or monomial block kernel.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_set>

#include "mossq_program.h"

namespace mossq {

constexpr int kB = 256;

__global__ void k_fill_identity(c128 *__restrict__ v, int n) {
    const u64 N = u64(1) << n, len = N * N;
    for (u64 t = u64(blockIdx.x) * kB + threadIdx.x; t < len; t += u64(gridDim.x) * kB) {
        u64 i = t & (N - 1), j = t >> n;
        v[t] = mk(i == j ? 1.0 : 0.0, 0.0);
    }
}

__global__ __launch_bounds__(kB) void k_reduce_rows_real(const c128 *__restrict__ part, int nparts,
                                                         c128 *__restrict__ out) {
    __shared__ c128 red[kB / 64];
    const c128 *p = part + size_t(blockIdx.x) * nparts;
    c128 a = mk(0, 0);
    for (int j = threadIdx.x; j < nparts; j += kB) a = cadd(a, p[j]);
    for (int o = 32; o > 0; o >>= 1) {
        a.re += __shfl_xor(a.re, o, 64);
        a.im += __shfl_xor(a.im, o, 64);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        c128 s = red[0];
        for (int w = 1; w < kB / 64; ++w) s = cadd(s, red[w]);
        out[blockIdx.x] = s;
    }
}

static int reduce_rows(Ctx *c, const c128 *part, int nrows, int nparts, c128 *out) {
    if (nrows <= 0) return MOSSQ_OK;
    hipLaunchKernelGGL(k_reduce_rows_real, dim3(nrows), dim3(kB), 0, c->stream, part, nparts, out);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

static int push(std::vector<c128> &pool, const c128 *m, size_t count) {
    int off = int(pool.size());
    pool.insert(pool.end(), m, m + count);
    return off;
}

// D^dagger (x) I + I (x) D^T in the local order (row bits low, column bits high).
static void dm_grad_matrix(const c128 *D, int k, std::vector<c128> &out) {
    int d = 1 << k, L = d * d;
    out.assign(size_t(L) * L, mk(0, 0));
    for (int a = 0; a < d; ++a)
        for (int cc = 0; cc < d; ++cc)
            for (int b = 0; b < d; ++b)
                for (int e = 0; e < d; ++e) {
                    c128 v = mk(0, 0);
                    if (cc == e) v = cadd(v, cconj(D[b + d * a]));
                    if (a == b) v = cadd(v, D[e + d * cc]);
                    out[size_t(a + d * cc) + size_t(L) * (b + d * e)] = v;
                }
}

static int apply_step_fwd(Ctx *c, const Program *p, const Step &s, c128 *v) {
    if (s.split) {
        MOSSQ_TRY(launch_dense(c, v, p->nbits, s.k, s.bits, p->d_pool + s.fwd, s.ifwd));
        return launch_dense(c, v, p->nbits, s.k, s.bits + s.k, p->d_pool + s.fwd_cols, s.cols->fwd);
    }
    return launch_dense(c, v, p->nbits, s.kb, s.bits, p->d_pool + s.fwd, s.ifwd);
}

static int apply_step_rev(Ctx *c, const Program *p, const Step &s, c128 *v) {
    if (s.split) {
        MOSSQ_TRY(launch_dense(c, v, p->nbits, s.k, s.bits, p->d_pool + s.rev, s.irev));
        return launch_dense(c, v, p->nbits, s.k, s.bits + s.k, p->d_pool + s.rev_cols, s.cols->rev);
    }
    return launch_dense(c, v, p->nbits, s.kb, s.bits, p->d_pool + s.rev, s.irev);
}

int apply_step(Ctx *c, const Program *p, int i, c128 *v, bool reverse) {
    return reverse ? apply_step_rev(c, p, p->steps[i], v) : apply_step_fwd(c, p, p->steps[i], v);
}

// One reverse step of the pure-state adjoint on (y, l): U^dagger on both and,
// for a parameterized step, the partials of Re<y|D|l> into row grad_index.
int adjoint_rev_step(Ctx *c, const Program *p, int i, c128 *y, c128 *l, c128 *part, int row) {
    const Step &st = p->steps[i];
    if (st.grad >= 0)
        return launch_dense_pair(c, y, l, p->n, st.kb, st.bits, p->d_pool + st.rev, p->d_pool + st.grad,
                                 part + size_t(st.grad_index) * row, false);
    if (st.irev.monomial || st.kb > 4) {
        MOSSQ_TRY(apply_step_rev(c, p, st, y));
        return apply_step_rev(c, p, st, l);
    }
    return launch_dense_pair(c, y, l, p->n, st.kb, st.bits, p->d_pool + st.rev, nullptr, nullptr, false);
}

int adjoint_part_row(const Program *p) {
    return std::max(pair_parts(p->n), fused_adjoint_parts(p));
}

int run_steps(Ctx *c, const Program *p, c128 *state, int lo, int hi) {
    for (int i = lo; i < hi; ++i) MOSSQ_TRY(apply_step_fwd(c, p, p->steps[i], state));
    return MOSSQ_OK;
}

// ------------------------------------------------------------- pure adjoint ----
int adjoint_pure_enqueue(Ctx *c, const Program *p, const Obs *o, const c128 *in, c128 *out_state,
                         const AdjointScratch &s) {
    const u64 len = u64(1) << p->n;
    if (in)
        MOSSQ_HIP(hipMemcpyAsync(s.a, in, len * sizeof(c128), hipMemcpyDeviceToDevice, c->stream));
    else
        MOSSQ_TRY(fill_basis0(c, s.a, len));
    // forward y = C x (reference circuit * x, differentiation/circuit.jl:83)
    if (p->fused)
        MOSSQ_TRY(fused_run(c, p, s.a));
    else
        MOSSQ_TRY(run_steps(c, p, s.a, 0, int(p->steps.size())));
    if (out_state)
        MOSSQ_HIP(hipMemcpyAsync(out_state, s.a, len * sizeof(c128), hipMemcpyDeviceToDevice, c->stream));
    // value <H> and seed (H + H^dagger) y in one pass (differentiation/expectation.jl:32-35)
    MOSSQ_TRY(observe_dev(c, o, s.a, s.a, s.b, s.value_dev, s.part));
    // reverse: per instruction  lambda <- U^dag lambda, y <- U^dag y,
    // g = Re <y|D|lambda>  (differentiation/circuit.jl:136-144); fused into
    // LDS-tile sweeps when the program has an adjoint plan.
    const int P = adjoint_part_row(p);
    if (p->fused_rev) {
        MOSSQ_TRY(zero_fill(c, s.part, u64(std::max(p->nderiv, 1)) * P));
        MOSSQ_TRY(fused_adjoint_run(c, p, s.a, s.b, s.part, P));
    } else {
        for (int i = int(p->steps.size()) - 1; i >= 0; --i) MOSSQ_TRY(adjoint_rev_step(c, p, i, s.a, s.b, s.part, P));
    }
    return reduce_rows(c, s.part, p->nderiv, P, s.grads_dev);
}

// value = Re of the sum of nvalues device partial values; gradients copied out.
static int copy_results(Ctx *c, const Program *p, int nvalues, const c128 *value_dev, const c128 *grads_dev,
                        double *value, double *grads) {
    std::vector<c128> hv(nvalues), hg(p->nderiv);
    MOSSQ_HIP(hipMemcpyAsync(hv.data(), value_dev, sizeof(c128) * nvalues, hipMemcpyDeviceToHost, c->stream));
    if (p->nderiv)
        MOSSQ_HIP(hipMemcpyAsync(hg.data(), grads_dev, sizeof(c128) * p->nderiv, hipMemcpyDeviceToHost,
                                 c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    double v = 0;
    for (auto &x : hv) v += x.re;  // objective value = real(expectation)
    if (value) *value = v;
    if (grads)
        for (int j = 0; j < p->nderiv; ++j) grads[j] = hg[j].re;
    return MOSSQ_OK;
}

// ------------------------------------------------- allocation-free adjoint ----
// mossq_adjoint_noalloc keeps its scratch (the forward state, lambda, the
// gradient partials, the value and the gradient sums) in this file-scope
// device array instead of allocating: 64 MiB per device, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (its
// name is a host-side stand-in): set-up takes each device's address with
// hipGetSymbolAddress, once per device with that device current, and the
// entry point uses that address.  Every call on a device shares the array:
// calls that overlap on one device race.
constexpr u64 kNoAllocBytes = u64(64) << 20;
static __device__ __align__(256) c128 scratch[kNoAllocBytes / sizeof(c128)];
constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static c128 *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch_init(int device) {
    MOSSQ_CHECK(device >= 0 && device < kMaxDevices, MOSSQ_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return MOSSQ_OK;
    int cur = -1;
    MOSSQ_HIP(hipGetDevice(&cur));
    MOSSQ_CHECK(cur == device, MOSSQ_EINVAL, "scratch set-up needs its device current");
    void *ptr = nullptr;
    MOSSQ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    MOSSQ_CHECK(ptr != nullptr, MOSSQ_EDEVICE, "no device address for the allocation-free scratch");
    g_noalloc_ptr[device] = static_cast<c128 *>(ptr);
    return MOSSQ_OK;
}

// Carves the pure adjoint's buffers for program p out of the device's array.
static int noalloc_layout(const Ctx *c, const Program *p, AdjointScratch *s) {
    c128 *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        base = c->device >= 0 && c->device < kMaxDevices ? g_noalloc_ptr[c->device] : nullptr;
    }
    MOSSQ_CHECK(base, MOSSQ_EDEVICE, "allocation-free scratch was not set up for this device");
    const u64 len = u64(1) << p->n;
    const u64 nparts = std::max<u64>(u64(kObsParts), u64(std::max(p->nderiv, 1)) * adjoint_part_row(p));
    const u64 align = 256 / sizeof(c128);
    auto up = [&](u64 x) { return (x + align - 1) / align * align; };
    const u64 need = up(len) * 2 + up(nparts) + up(1) + up(u64(std::max(p->nderiv, 1)));
    MOSSQ_CHECK(need * sizeof(c128) <= kNoAllocBytes, MOSSQ_EINVAL,
                "the allocation-free adjoint needs " + std::to_string(need * sizeof(c128)) +
                    " bytes of scratch; it holds " + std::to_string(kNoAllocBytes) +
                    " (state vectors up to 20 qubits); use mossq_adjoint");
    u64 off = 0;
    s->a = base + off;
    off += up(len);
    s->b = base + off;
    off += up(len);
    s->part = base + off;
    off += up(nparts);
    s->value_dev = base + off;
    off += up(1);
    s->grads_dev = base + off;
    return MOSSQ_OK;
}

int adjoint_pure(Ctx *c, const Program *p, const Obs *o, const c128 *in, c128 *out_state, double *value,
                        double *grads) {
    const u64 len = u64(1) << p->n;
    Allocs al(c);
    AdjointScratch s;
    const int P = adjoint_part_row(p);
    u64 nparts = std::max<u64>(u64(kObsParts), u64(std::max(p->nderiv, 1)) * P);
    MOSSQ_TRY(al.get(len * sizeof(c128), (void **)&s.a));
    MOSSQ_TRY(al.get(len * sizeof(c128), (void **)&s.b));
    MOSSQ_TRY(al.get(nparts * sizeof(c128), (void **)&s.part));
    MOSSQ_TRY(al.get(sizeof(c128), (void **)&s.value_dev));
    MOSSQ_TRY(al.get(u64(std::max(p->nderiv, 1)) * sizeof(c128), (void **)&s.grads_dev));
    MOSSQ_TRY(adjoint_pure_enqueue(c, p, o, in, out_state, s));
    return copy_results(c, p, 1, s.value_dev, s.grads_dev, value, grads);
}

// ---------------------------------------------------------- density adjoint ----
// Checkpoints at segment starts (every s instructions); the reverse pass
// recomputes each segment once from its checkpoint and keeps the states
// before each of its instructions, so channels never need an inverse and
// memory is (N/s + s + 2) states instead of one per channel
// (the reference checkpoints before every channel, differentiation/circuit.jl:90-98).
u64 adjoint_density_bytes(const Program *p, const Obs *o) {
    const u64 bytes = (u64(1) << p->nbits) * sizeof(c128);
    const int N = int(p->steps.size());
    int seg = std::max(1, int(std::ceil(std::sqrt(double(std::max(N, 1))))));
    int nseg = std::max(1, (N + seg - 1) / seg);
    if (p->dm_adj) {
        // fused chunks: nseg checkpoints and seg - 1 recompute buffers
        density_adjoint_layout(p, &seg, &nseg);
        seg -= 1;
    }
    const int row = p->dm_adj ? density_adjoint_parts(p) : pair_parts(p->nbits);
    u64 parts = std::max<u64>(u64(o->G) * 64, u64(std::max(p->nderiv, 1)) * row);
    return bytes * u64(2 + nseg + seg) + (parts + o->G + p->nderiv + 1) * sizeof(c128);
}

int program_upload(Ctx *c, Program *p) {
    p->ctx = c;
    MOSSQ_TRY(cached_alloc(c, std::max<size_t>(p->h_pool.size(), 1) * sizeof(c128), (void **)&p->d_pool));
    if (!p->h_pool.empty())
        MOSSQ_HIP(hipMemcpyAsync(p->d_pool, p->h_pool.data(), p->h_pool.size() * sizeof(c128),
                                 hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    return MOSSQ_OK;
}

// The fused variant: forward and recompute run as chunk plans cut at the
// parameterized steps, so only the states before parameterized steps are kept
// and every chunk is one or a few LDS-tile sweeps (tile.hip).
static int adjoint_density_fused(Ctx *c, const Program *p, const Obs *o, const c128 *in, int in_is_density,
                                 c128 *out_state, double *value, double *grads) {
    const int n = p->n, nb = p->nbits;
    const u64 len = u64(1) << nb, bytes = len * sizeof(c128);
    int seg = 0, nseg = 0;
    density_adjoint_layout(p, &seg, &nseg);
    Allocs al(c);
    c128 *delta = nullptr, *work = nullptr;
    std::vector<c128 *> ckpt(nseg), stored(std::max(seg - 1, 1));
    MOSSQ_TRY(al.get(bytes, (void **)&delta));
    MOSSQ_TRY(al.get(bytes, (void **)&work));
    for (auto &q : ckpt) MOSSQ_TRY(al.get(bytes, (void **)&q));
    for (int i = 0; i + 1 < seg; ++i) MOSSQ_TRY(al.get(bytes, (void **)&stored[i]));
    const int P = density_adjoint_parts(p);
    c128 *part = nullptr, *value_dev = nullptr, *grads_dev = nullptr;
    u64 nparts = std::max<u64>(u64(o->G) * 64, u64(std::max(p->nderiv, 1)) * P);
    MOSSQ_TRY(al.get(nparts * sizeof(c128), (void **)&part));
    MOSSQ_TRY(al.get(u64(o->G) * sizeof(c128), (void **)&value_dev));
    MOSSQ_TRY(al.get(u64(std::max(p->nderiv, 1)) * sizeof(c128), (void **)&grads_dev));

    if (!in)
        MOSSQ_TRY(fill_basis0(c, work, len));
    else if (in_is_density)
        MOSSQ_HIP(hipMemcpyAsync(work, in, bytes, hipMemcpyDeviceToDevice, c->stream));
    else
        MOSSQ_TRY(outer_product(c, work, in, n));
    MOSSQ_TRY(density_adjoint_forward(c, p, work, ckpt.data()));
    if (out_state) MOSSQ_HIP(hipMemcpyAsync(out_state, work, bytes, hipMemcpyDeviceToDevice, c->stream));
    MOSSQ_TRY(dm_expect_groups_dev(c, o, work, value_dev, part));
    // seed: z * H^dagger as a dense density matrix (differentiation/expectation.jl:69-72)
    hipLaunchKernelGGL(k_fill_identity, dim3(4096), dim3(kB), 0, c->stream, work, n);
    MOSSQ_HIP(hipGetLastError());
    MOSSQ_TRY(apply_op_dev(c, o, work, delta, nb, OpMats::Adjoint));
    // the paired launches write pair_parts entries of their rows, the tile
    // probes whole rows; every other entry must read as zero
    MOSSQ_TRY(zero_fill(c, part, u64(std::max(p->nderiv, 1)) * P));
    std::vector<const c128 *> tab;
    density_adjoint_probe_table(p, ckpt.data(), stored.data(), tab);
    const c128 **d_tab = nullptr;
    MOSSQ_TRY(al.get(tab.size() * sizeof(c128 *), (void **)&d_tab));
    MOSSQ_HIP(hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(c128 *), hipMemcpyHostToDevice, c->stream));
    MOSSQ_TRY(density_adjoint_reverse(c, p, delta, ckpt.data(), stored.data(), d_tab, part, P));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));  // tab (host) is read by the copy above
    MOSSQ_TRY(reduce_rows(c, part, p->nderiv, P, grads_dev));
    return copy_results(c, p, o->G, value_dev, grads_dev, value, grads);
}

int adjoint_density(Ctx *c, const Program *p, const Obs *o, const c128 *in, int in_is_density,
                           c128 *out_state, double *value, double *grads) {
    if (p->dm_adj) return adjoint_density_fused(c, p, o, in, in_is_density, out_state, value, grads);
    const int n = p->n, nb = p->nbits;
    const u64 len = u64(1) << nb, bytes = len * sizeof(c128);
    const int N = int(p->steps.size());
    int seg = std::max(1, int(std::ceil(std::sqrt(double(std::max(N, 1))))));
    int nseg = std::max(1, (N + seg - 1) / seg);
    Allocs al(c);
    c128 *delta = nullptr, *work = nullptr;
    std::vector<c128 *> ckpt(nseg), stored(seg);
    MOSSQ_TRY(al.get(bytes, (void **)&delta));
    MOSSQ_TRY(al.get(bytes, (void **)&work));
    for (auto &q : ckpt) MOSSQ_TRY(al.get(bytes, (void **)&q));
    for (auto &q : stored) MOSSQ_TRY(al.get(bytes, (void **)&q));
    const int P = pair_parts(nb);
    c128 *part = nullptr, *value_dev = nullptr, *grads_dev = nullptr;
    u64 nparts = std::max<u64>(u64(o->G) * 64, u64(std::max(p->nderiv, 1)) * P);
    MOSSQ_TRY(al.get(nparts * sizeof(c128), (void **)&part));
    MOSSQ_TRY(al.get(u64(o->G) * sizeof(c128), (void **)&value_dev));
    MOSSQ_TRY(al.get(u64(std::max(p->nderiv, 1)) * sizeof(c128), (void **)&grads_dev));

    if (!in)
        MOSSQ_TRY(fill_basis0(c, work, len));
    else if (in_is_density)
        MOSSQ_HIP(hipMemcpyAsync(work, in, bytes, hipMemcpyDeviceToDevice, c->stream));
    else
        MOSSQ_TRY(outer_product(c, work, in, n));
    for (int i = 0; i < N; ++i) {
        if (i % seg == 0)
            MOSSQ_HIP(hipMemcpyAsync(ckpt[i / seg], work, bytes, hipMemcpyDeviceToDevice, c->stream));
        MOSSQ_TRY(apply_step_fwd(c, p, p->steps[i], work));
    }
    if (out_state) MOSSQ_HIP(hipMemcpyAsync(out_state, work, bytes, hipMemcpyDeviceToDevice, c->stream));
    MOSSQ_TRY(dm_expect_groups_dev(c, o, work, value_dev, part));
    // seed: z * H^dagger as a dense density matrix (differentiation/expectation.jl:69-72)
    hipLaunchKernelGGL(k_fill_identity, dim3(4096), dim3(kB), 0, c->stream, work, n);
    MOSSQ_HIP(hipGetLastError());
    MOSSQ_TRY(apply_op_dev(c, o, work, delta, nb, OpMats::Adjoint));
    for (int sg = nseg - 1; sg >= 0; --sg) {
        int lo = sg * seg, hi = std::min(N, lo + seg);
        MOSSQ_HIP(hipMemcpyAsync(stored[0], ckpt[sg], bytes, hipMemcpyDeviceToDevice, c->stream));
        for (int i = lo; i + 1 < hi; ++i) {
            MOSSQ_HIP(hipMemcpyAsync(stored[i - lo + 1], stored[i - lo], bytes, hipMemcpyDeviceToDevice, c->stream));
            MOSSQ_TRY(apply_step_fwd(c, p, p->steps[i], stored[i - lo + 1]));
        }
        for (int i = hi - 1; i >= lo; --i) {
            const Step &st = p->steps[i];
            if (st.grad >= 0) {
                // Delta <- U^dag Delta U; rho_before = U^dag rho U is the stored state;
                // g = Re[<Delta, D^dag rho> + <Delta, rho D>] (differentiation/circuit.jl:155-172)
                MOSSQ_TRY(launch_dense_pair(c, delta, stored[i - lo], nb, st.kb, st.bits, p->d_pool + st.rev,
                                            p->d_pool + st.grad, part + size_t(st.grad_index) * P, true));
            } else {
                // channel: Delta <- S^dagger Delta (differentiation/circuit.jl:107-108); gate: U^dag Delta U
                MOSSQ_TRY(apply_step_rev(c, p, st, delta));
            }
        }
    }
    MOSSQ_TRY(reduce_rows(c, part, p->nderiv, P, grads_dev));
    return copy_results(c, p, o->G, value_dev, grads_dev, value, grads);
}

static int check_mat_range(u64 off, u64 count, u64 pool_len) {
    MOSSQ_CHECK(off + count <= pool_len, MOSSQ_EINVAL, "matrix offset outside the pool");
    return MOSSQ_OK;
}

// Validate the instruction records and build the host side of a program
// (steps and matrix pool); no device work.
int program_build_host(Program *p, int n, int is_density, int nops, const mossq_op *ops, const double *pool,
                       uint64_t pool_len) {
    MOSSQ_CHECK((ops || nops == 0) && (pool || pool_len == 0) && nops >= 0, MOSSQ_EINVAL, "bad program arguments");
    MOSSQ_CHECK(n >= 1 && n <= (is_density ? 20 : 40), MOSSQ_EINVAL, "qubit count out of range");
    const c128 *P = (const c128 *)pool;
    p->n = n;
    p->is_density = is_density ? 1 : 0;
    p->nbits = is_density ? 2 * n : n;
    auto bad = [&](int code, const std::string &m) { return fail(code, m); };
    p->steps.reserve(size_t(nops));
    for (int i = 0; i < nops; ++i) {
        const mossq_op &op = ops[i];
        const int k = op.k;
        if (op.kind != MOSSQ_OP_GATE && op.kind != MOSSQ_OP_CHANNEL) return bad(MOSSQ_EINVAL, "unknown op kind");
        if (op.kind == MOSSQ_OP_CHANNEL && !is_density)
            return bad(MOSSQ_EINVAL, "channels act on density matrices: promote the state before the first channel");
        int kmax = op.kind == MOSSQ_OP_GATE ? 5 : 2;
        if (k < 1 || k > kmax || k > n) return bad(MOSSQ_EUNSUPPORTED, "instruction size not supported");
        for (int j = 0; j < k; ++j)
            if (op.pos[j] < 0 || op.pos[j] >= n || (j && op.pos[j] <= op.pos[j - 1]))
                return bad(MOSSQ_EINVAL, "positions must be increasing and in range");
        const int d = 1 << k;
        Step st;
        st.k = k;
        if (op.kind == MOSSQ_OP_CHANNEL) {
            const int L = d * d;
            if (check_mat_range(op.mat, u64(L) * L, pool_len) != MOSSQ_OK) return bad(MOSSQ_EINVAL, "bad matrix offset");
            st.channel = true;
            p->has_channel = true;
            st.kb = 2 * k;
            for (int j = 0; j < k; ++j) {
                st.bits[j] = op.pos[j];
                st.bits[j + k] = op.pos[j] + n;
            }
            std::vector<c128> Sd;
            conj_transpose(P + op.mat, L, Sd);
            st.fwd = push(p->h_pool, P + op.mat, size_t(L) * L);
            st.ifwd = classify(P + op.mat, L);
            st.rev = push(p->h_pool, Sd.data(), Sd.size());
            st.irev = classify(Sd.data(), L);
            if (op.dmat >= 0) return bad(MOSSQ_EUNSUPPORTED, "channels carry no parameters");
            p->steps.push_back(st);
            continue;
        }
        if (check_mat_range(op.mat, u64(d) * d, pool_len) != MOSSQ_OK) return bad(MOSSQ_EINVAL, "bad matrix offset");
        if (op.dmat >= 0 && check_mat_range(op.dmat, u64(d) * d, pool_len) != MOSSQ_OK)
            return bad(MOSSQ_EINVAL, "bad derivative offset");
        const c128 *U = P + op.mat;
        std::vector<c128> Ud;
        conj_transpose(U, d, Ud);
        if (!is_density) {
            st.kb = k;
            for (int j = 0; j < k; ++j) st.bits[j] = op.pos[j];
            st.fwd = push(p->h_pool, U, size_t(d) * d);
            st.ifwd = classify(U, d);
            st.rev = push(p->h_pool, Ud.data(), Ud.size());
            st.irev = classify(Ud.data(), d);
            if (op.dmat >= 0) {
                if (k > 4) return bad(MOSSQ_EUNSUPPORTED, "parameterized gates wider than 4 qubits");
                st.grad = push(p->h_pool, P + op.dmat, size_t(d) * d);
                st.grad_index = p->nderiv++;
            }
        } else if (k <= 2) {
            st.kb = 2 * k;
            for (int j = 0; j < k; ++j) {
                st.bits[j] = op.pos[j];
                st.bits[j + k] = op.pos[j] + n;
            }
            std::vector<c128> Sf, Sr;
            dm_superop_of_gate(U, k, Sf);
            dm_superop_of_gate(Ud.data(), k, Sr);
            const int L = d * d;
            st.fwd = push(p->h_pool, Sf.data(), Sf.size());
            st.ifwd = classify(Sf.data(), L);
            st.rev = push(p->h_pool, Sr.data(), Sr.size());
            st.irev = classify(Sr.data(), L);
            if (k == 2) st.u_src = push(p->h_pool, U, size_t(d) * d);
            if (op.dmat >= 0) {
                std::vector<c128> G;
                dm_grad_matrix(P + op.dmat, k, G);
                st.grad = push(p->h_pool, G.data(), G.size());
                st.grad_index = p->nderiv++;
            }
        } else {
            if (op.dmat >= 0)
                return bad(MOSSQ_EUNSUPPORTED, "parameterized density-matrix gates wider than 2 qubits");
            st.split = true;
            st.kb = 2 * k;
            for (int j = 0; j < k; ++j) {
                st.bits[j] = op.pos[j];
                st.bits[j + k] = op.pos[j] + n;
            }
            std::vector<c128> Uc(size_t(d) * d), Udc(size_t(d) * d);
            for (size_t q = 0; q < Uc.size(); ++q) {
                Uc[q] = cconj(U[q]);
                Udc[q] = cconj(Ud[q]);
            }
            st.fwd = push(p->h_pool, U, size_t(d) * d);
            st.ifwd = classify(U, d);
            auto cols = std::make_shared<SplitInfo>();
            st.fwd_cols = push(p->h_pool, Uc.data(), Uc.size());
            cols->fwd = classify(Uc.data(), d);
            st.rev = push(p->h_pool, Ud.data(), Ud.size());
            st.irev = classify(Ud.data(), d);
            st.rev_cols = push(p->h_pool, Udc.data(), Udc.size());
            cols->rev = classify(Udc.data(), d);
            st.cols = cols;
        }
        p->steps.push_back(st);
    }
    return MOSSQ_OK;
}

// One parameter update, prepared on the host (program_update_prepare), then
// committed: the new steps and pool and every plan's new values on the host,
// with the device copies listed in `copies`.
struct ProgramUpdate {
    Program q;       // the program rebuilt from the new records
    PlanUpdate pad;  // the 12-qubit form's plans
    PlanUpdate main;
    std::vector<DevCopy> copies;
};

// A table segment of a batched update: `bytes` at `src` in the staging block
// go to device address `dst`.
struct ScatterSeg {
    u64 src, bytes, dst, unused;
};

// Copies every segment from the staging block to its table (32-bit words;
// each workgroup takes whole segments).
__global__ void __launch_bounds__(256) k_scatter(const unsigned char *stage, const ScatterSeg *segs, u64 nseg) {
    for (u64 s = blockIdx.x; s < nseg; s += gridDim.x) {
        const ScatterSeg g = segs[s];
        const uint32_t *src = reinterpret_cast<const uint32_t *>(stage + g.src);
        uint32_t *dst = reinterpret_cast<uint32_t *>(g.dst);
        const u64 words = g.bytes / 4;
        for (u64 k = threadIdx.x; k < words; k += blockDim.x) dst[k] = src[k];
    }
}

// f(i) for i in [0, m) on nt threads (this one included).
template <class F>
static void run_on_threads(int nt, int m, F f) {
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < m; i = next++) f(i);
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

static bool same_steps(const Program &a, const Program &b, std::vector<char> &moved) {
    moved.assign(a.steps.size(), 0);
    if (a.steps.size() != b.steps.size() || a.h_pool.size() != b.h_pool.size() || a.nderiv != b.nderiv) return false;
    bool m = false;
    auto same_info = [&](const MatInfo &x, const MatInfo &y) {
        if (x.monomial != y.monomial || x.nchanged != y.nchanged) return false;
        for (int j = 0; x.monomial && j < x.nchanged; ++j) {
            if (x.rows[j] != y.rows[j] || x.src[j] != y.src[j]) return false;
            if (x.val[j].re != y.val[j].re || x.val[j].im != y.val[j].im) m = true;
        }
        return true;
    };
    for (size_t i = 0; i < a.steps.size(); ++i) {
        const Step &x = a.steps[i], &y = b.steps[i];
        m = false;
        if (x.channel != y.channel || x.k != y.k || x.kb != y.kb || memcmp(x.bits, y.bits, sizeof(x.bits)) != 0 ||
            x.fwd != y.fwd || x.rev != y.rev || x.grad != y.grad || x.grad_index != y.grad_index ||
            x.u_src != y.u_src || x.split != y.split || x.fwd_cols != y.fwd_cols || x.rev_cols != y.rev_cols)
            return false;
        if (!same_info(x.ifwd, y.ifwd) || !same_info(x.irev, y.irev))
            return false;
        if (x.split && (!x.cols || !y.cols || !same_info(x.cols->fwd, y.cols->fwd) ||
                        !same_info(x.cols->rev, y.cols->rev)))
            return false;
        moved[i] = m ? 1 : 0;
    }
    return true;
}

// Host half of a parameter update of p (no device calls; updates of
// different programs may be prepared on different threads): the program
// rebuilt from the new records, its plans and those of its 12-qubit form
// rebuilt and checked, and the device copies listed.  Refusals leave p as it
// was.
static int program_update_prepare(Program *p, int nops, const mossq_op *ops, const double *pool, uint64_t pool_len,
                                  ProgramUpdate *u) {
    Program &q = u->q;
    MOSSQ_TRY(program_build_host(&q, p->n, p->is_density, nops, ops, pool, pool_len));
    std::vector<char> moved;
    MOSSQ_CHECK(same_steps(*p, q, moved), MOSSQ_EINVAL, "the records describe another program");
    // A program graph recorded the entries of the monomial steps its run
    // launches on their own (kernel arguments); those must not move under it.
    for (size_t i = 0; i < moved.size() && p->live_step_graphs > 0; ++i)
        MOSSQ_CHECK(!(moved[i] && fused_run_launches_step(p, int(i))), MOSSQ_EINVAL,
                    "a program graph launches a step whose entries would change: destroy it first");
    if (p->pad) {
        Program qp;
        qp.ctx = p->ctx;
        qp.n = qp.nbits = p->pad->n;
        qp.nderiv = q.nderiv;
        qp.steps = q.steps;
        qp.h_pool = q.h_pool;
        MOSSQ_TRY(fused_plan_update_prepare(p->pad, &qp, &u->pad));
        fused_plan_update_copies(u->pad, u->copies);
    }
    MOSSQ_TRY(fused_plan_update_prepare(p, &q, &u->main));
    fused_plan_update_copies(u->main, u->copies);
    // (the 12-qubit form shares d_pool)
    u->copies.push_back({p->d_pool, q.h_pool.data(), q.h_pool.size() * sizeof(c128)});
    return MOSSQ_OK;
}

// The new values into p's host tables (and its 12-qubit form's); the listed
// copies write them to the device.
static void program_update_commit(Program *p, ProgramUpdate *u) {
    if (p->pad) {
        fused_plan_update_commit(&u->pad);
        p->pad->steps = u->q.steps;
        p->pad->h_pool = u->q.h_pool;
    }
    fused_plan_update_commit(&u->main);
    p->steps.swap(u->q.steps);
    p->h_pool.swap(u->q.h_pool);
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_program_create(mossq_ctx *ctx, int n, int is_density, int nops, const mossq_op *ops, const double *pool,
                         uint64_t pool_len, int fuse, mossq_program **out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(out, MOSSQ_EINVAL, "null output");
    Program *p = new Program();
    p->ctx = C(ctx);
    int s = program_build_host(p, n, is_density, nops, ops, pool, pool_len);
    if (s != MOSSQ_OK) {
        delete p;
        return s;
    }
    Ctx *c = p->ctx;
    s = cached_alloc(c, std::max<size_t>(p->h_pool.size(), 1) * sizeof(c128), (void **)&p->d_pool);
    if (s != MOSSQ_OK) {
        delete p;
        return s;
    }
    if (!p->h_pool.empty())
        MOSSQ_HIP(hipMemcpyAsync(p->d_pool, p->h_pool.data(), p->h_pool.size() * sizeof(c128),
                                 hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    if (fuse) {
        s = fused_plan_build(p);
        if (s != MOSSQ_OK) {
            mossq_program_destroy(reinterpret_cast<mossq_program *>(p));
            return s;
        }
    }
    *out = reinterpret_cast<mossq_program *>(p);
    return MOSSQ_OK;
}

// Same steps up to matrix values: kinds, bits, pool offsets and, for
// monomial matrices, the pattern of changed rows.  moved[i] reports whether
// monomial step i's entries differ.
int mossq_program_update(mossq_ctx *ctx, mossq_program *prog, int nops, const mossq_op *ops, const double *pool,
                         uint64_t pool_len) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(prog, MOSSQ_EINVAL, "null argument");
    Program *p = reinterpret_cast<Program *>(prog);
    Ctx *c = C(ctx);
    MOSSQ_CHECK(p->ctx == c, MOSSQ_EINVAL, "program belongs to another context");
    ProgramUpdate u;
    MOSSQ_TRY(program_update_prepare(p, nops, ops, pool, pool_len, &u));
    program_update_commit(p, &u);
    for (const DevCopy &cp : u.copies)
        if (cp.bytes) MOSSQ_HIP(hipMemcpyAsync(cp.dst, cp.src, cp.bytes, hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    return MOSSQ_OK;
}

int mossq_program_update_many(mossq_ctx *ctx, int m, mossq_program *const *progs, const int *nops,
                              const mossq_op *const *ops, const double *const *pools, const uint64_t *pool_lens,
                              int threads, int *status) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(m >= 0 && (m == 0 || (progs && nops && ops && pools && pool_lens)), MOSSQ_EINVAL, "bad arguments");
    for (int i = 0; status && i < m; ++i) status[i] = MOSSQ_EINVAL;
    Ctx *c = C(ctx);
    {
        std::unordered_set<const void *> seen;
        for (int i = 0; i < m; ++i) {
            MOSSQ_CHECK(progs[i], MOSSQ_EINVAL, "null program");
            MOSSQ_CHECK(reinterpret_cast<Program *>(progs[i])->ctx == c, MOSSQ_EINVAL,
                        "program belongs to another context");
            MOSSQ_CHECK(seen.insert(progs[i]).second, MOSSQ_EINVAL, "a program appears twice");
        }
    }
    if (m == 0) return MOSSQ_OK;
    std::vector<std::unique_ptr<ProgramUpdate>> ups(m);
    std::vector<int> st(m, MOSSQ_OK);
    std::vector<std::string> msg(m);
    const int nt = std::max(1, std::min({threads > 0 ? threads : int(std::thread::hardware_concurrency()), 64, m}));
    // host phase on nt threads: prepare every update
    run_on_threads(nt, m, [&](int i) {
        ups[i].reset(new ProgramUpdate());
        st[i] = program_update_prepare(reinterpret_cast<Program *>(progs[i]), nops[i], ops[i], pools[i],
                                       pool_lens[i], ups[i].get());
        if (st[i] != MOSSQ_OK) {
            msg[i] = mossq_last_error();
            ups[i].reset();
        }
    });
    // one staging layout: every copy's bytes at 16-byte aligned offsets,
    // then the segment table
    std::vector<u64> first(m + 1, 0);  // first segment of each update
    u64 data = 0;
    for (int i = 0; i < m; ++i) {
        first[i + 1] = first[i];
        if (!ups[i]) continue;
        for (const DevCopy &cp : ups[i]->copies) {
            if (!cp.bytes) continue;
            MOSSQ_CHECK(cp.bytes % 4 == 0, MOSSQ_EINVAL, "table sizes are whole 32-bit words");
            data += (cp.bytes + 15) & ~u64(15);
            ++first[i + 1];
        }
    }
    const u64 nseg = first[m];
    const u64 total = data + nseg * sizeof(ScatterSeg);
    if (nseg) {
        unsigned char *host = nullptr, *dev = nullptr;
        // an earlier call that returned on an error may have left its copy
        // from the pinned staging in flight
        MOSSQ_HIP(hipStreamSynchronize(c->stream));
        MOSSQ_TRY(get_update_pinned(c, total, &host));
        Allocs al(c);
        MOSSQ_TRY(al.get(total, (void **)&dev));
        ScatterSeg *segs = reinterpret_cast<ScatterSeg *>(host + data);
        // offsets of each update's first byte
        std::vector<u64> at(m + 1, 0);
        for (int i = 0; i < m; ++i) {
            at[i + 1] = at[i];
            if (ups[i])
                for (const DevCopy &cp : ups[i]->copies) at[i + 1] += cp.bytes ? (cp.bytes + 15) & ~u64(15) : 0;
        }
        // host phase on nt threads: stage the values, fill the segments, commit
        run_on_threads(nt, m, [&](int i) {
            if (!ups[i]) return;
            u64 o = at[i], sg = first[i];
            for (const DevCopy &cp : ups[i]->copies) {
                if (!cp.bytes) continue;
                memcpy(host + o, cp.src, cp.bytes);
                segs[sg++] = ScatterSeg{o, cp.bytes, reinterpret_cast<u64>(cp.dst), 0};
                o += (cp.bytes + 15) & ~u64(15);
            }
            program_update_commit(reinterpret_cast<Program *>(progs[i]), ups[i].get());
        });
        MOSSQ_HIP(hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, c->stream));
        const unsigned grid = unsigned(std::min<u64>(nseg, 8192));
        hipLaunchKernelGGL(k_scatter, dim3(grid), dim3(256), 0, c->stream, dev,
                           reinterpret_cast<const ScatterSeg *>(dev + data), nseg);
        MOSSQ_HIP(hipGetLastError());
        // the staging block returns to the cache behind the scatter; the wait
        // also keeps the pinned staging from being rewritten under the copy
        MOSSQ_HIP(hipStreamSynchronize(c->stream));
    }
    for (int i = 0; status && i < m; ++i) status[i] = st[i];
    for (int i = 0; i < m; ++i)
        if (st[i] != MOSSQ_OK) return fail(st[i], "program " + std::to_string(i) + ": " + msg[i]);
    return MOSSQ_OK;
}

int mossq_program_destroy(mossq_program *prog) {
    if (!prog) return MOSSQ_OK;
    Program *p = reinterpret_cast<Program *>(prog);
    MOSSQ_CHECK(p->live_graphs == 0, MOSSQ_EINVAL, "a graph captured from the program is still alive");
    (void)hipSetDevice(p->ctx->device);
    fused_plan_free(p);
    int s = cached_free(p->ctx, p->d_pool);
    delete p;
    return s;
}

int mossq_program_sweeps(const mossq_program *prog, int *sweeps) {
    MOSSQ_CHECK(prog && sweeps, MOSSQ_EINVAL, "null argument");
    const Program *p = reinterpret_cast<const Program *>(prog);
    *sweeps = p->fused ? fused_plan_sweeps(p) : int(p->steps.size());
    return MOSSQ_OK;
}

int mossq_program_run(mossq_ctx *ctx, const mossq_program *prog, void *state) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(prog && state, MOSSQ_EINVAL, "null argument");
    const Program *p = reinterpret_cast<const Program *>(prog);
    Ctx *c = C(ctx);
    MOSSQ_CHECK(p->ctx == c, MOSSQ_EINVAL, "program belongs to another context");
    if (p->fused) return fused_run(c, p, (c128 *)state);
    return run_steps(c, p, (c128 *)state, 0, int(p->steps.size()));
}

int mossq_program_autotune(mossq_ctx *ctx, mossq_program *prog, int reps, int *variants, double *best_ms) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(prog && variants && best_ms && reps >= 1, MOSSQ_EINVAL, "bad autotune arguments");
    Program *p = reinterpret_cast<Program *>(prog);
    MOSSQ_CHECK(p->ctx == C(ctx), MOSSQ_EINVAL, "program belongs to another context");
    return fused_plan_autotune(p, reps, variants, best_ms);
}


int mossq_program_capture(mossq_ctx *ctx, const mossq_program *prog, void *state, mossq_graph **out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(prog && state && out, MOSSQ_EINVAL, "null argument");
    const Program *p = reinterpret_cast<const Program *>(prog);
    Ctx *c = C(ctx);
    MOSSQ_CHECK(p->ctx == c, MOSSQ_EINVAL, "program belongs to another context");
    // The run path launches kernels only (no allocation, copy or memset), so
    // the captured graph holds kernel nodes alone.
    MOSSQ_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int s = p->fused ? fused_run(c, p, (c128 *)state) : run_steps(c, p, (c128 *)state, 0, int(p->steps.size()));
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (s != MOSSQ_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        if (s != MOSSQ_OK) return s;
        MOSSQ_HIP(e);
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        MOSSQ_HIP(ei);
    }
    auto *gr = new mossq_graph();
    gr->ctx = c;
    gr->prog = p;
    p->live_graphs++;
    p->live_step_graphs++;
    gr->graph = g;
    gr->exec = ex;
    *out = gr;
    return MOSSQ_OK;
}

int mossq_graph_launch(mossq_ctx *ctx, mossq_graph *graph) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(graph, MOSSQ_EINVAL, "null argument");
    Ctx *c = C(ctx);
    MOSSQ_CHECK(graph->ctx == c, MOSSQ_EINVAL, "graph belongs to another context");
    MOSSQ_HIP(hipGraphLaunch(graph->exec, c->stream));
    return MOSSQ_OK;
}

int mossq_graph_destroy(mossq_graph *graph) {
    if (!graph) return MOSSQ_OK;
    (void)hipSetDevice(graph->ctx->device);
    (void)hipStreamSynchronize(graph->ctx->stream);
    if (graph->exec) (void)hipGraphExecDestroy(graph->exec);
    if (graph->graph) (void)hipGraphDestroy(graph->graph);
    if (graph->prog) {
        Program *p = const_cast<Program *>(graph->prog);
        p->live_step_graphs--;
        if (--p->live_graphs == 0) fused_plan_release_retired(p);
    }
    for (const Program *cp : graph->progs) {
        Program *p = const_cast<Program *>(cp);
        if (--p->live_graphs == 0) fused_plan_release_retired(p);
    }
    if (graph->batch) batch_run_free(graph->batch);
    delete graph;
    return MOSSQ_OK;
}

int mossq_adjoint(mossq_ctx *ctx, const mossq_program *prog, const mossq_obs *obs, const void *in,
                  int in_is_density, void *out_state, double *value, double *grads) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(prog && obs, MOSSQ_EINVAL, "null argument");
    const Program *p = reinterpret_cast<const Program *>(prog);
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    Ctx *c = C(ctx);
    MOSSQ_CHECK(p->ctx == c && o->ctx == c, MOSSQ_EINVAL, "program/observable belong to another context");
    MOSSQ_CHECK(o->n == p->n, MOSSQ_EINVAL, "observable and program qubit counts differ");
    if (!p->is_density) {
        MOSSQ_CHECK(!in_is_density, MOSSQ_EINVAL, "a state-vector program needs a state-vector input");
        return adjoint_pure(c, p, o, (const c128 *)in, (c128 *)out_state, value, grads);
    }
    return adjoint_density(c, p, o, (const c128 *)in, in_is_density, (c128 *)out_state, value, grads);
}

int mossq_adjoint_noalloc(mossq_ctx *ctx, const mossq_program *prog, const mossq_obs *obs, const void *in,
                          void *out_state, double *value, double *grads) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(prog && obs, MOSSQ_EINVAL, "null argument");
    const Program *p = reinterpret_cast<const Program *>(prog);
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    Ctx *c = C(ctx);
    MOSSQ_CHECK(p->ctx == c && o->ctx == c, MOSSQ_EINVAL, "program/observable belong to another context");
    MOSSQ_CHECK(o->n == p->n, MOSSQ_EINVAL, "observable and program qubit counts differ");
    MOSSQ_CHECK(!p->is_density, MOSSQ_EUNSUPPORTED, "the allocation-free adjoint takes state-vector programs");
    AdjointScratch s;
    MOSSQ_TRY(noalloc_layout(c, p, &s));
    MOSSQ_TRY(adjoint_pure_enqueue(c, p, o, (const c128 *)in, (c128 *)out_state, s));
    return copy_results(c, p, 1, s.value_dev, s.grads_dev, value, grads);
}

}  // extern "C"
