// Batched adjoint gradients over independent programs with uneven qubit counts
// (the engine of mossq_batch_adjoint and of the request server).
#include <algorithm>

#include "mossq_batch.h"

namespace mossq {

constexpr int kBT = 256;

__device__ __forceinline__ c128 bwave_sum(c128 a) {
    for (int o = 32; o > 0; o >>= 1) {
        a.re += __shfl_xor(a.re, o, 64);
        a.im += __shfl_xor(a.im, o, 64);
    }
    return a;
}

// Deterministic block sum; valid in thread 0.
__device__ __forceinline__ c128 bblock_sum(c128 a, c128 *red) {
    a = bwave_sum(a);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
    __syncthreads();
    c128 s = mk(0, 0);
    if (threadIdx.x == 0)
        for (int w = 0; w < kBT / 64; ++w) s = cadd(s, red[w]);
    return s;
}

// One chunk of a gate micro-op: forward (psi <- U psi) or reverse
// (psi, lambda <- U^dagger, gradient partial Re <psi|D|lambda>).
template <int K>
__device__ __forceinline__ void bgate_chunk(const BItemDev &it, const BOp &op, bool fwd, int c, c128 *sm, c128 *red,
                                            c128 *__restrict__ part) {
    constexpr int D = 1 << K;
    u64 off[D];
    off[0] = 0;
    off[1] = u64(1) << op.b0;
    if (K == 2) {
        off[2] = u64(1) << op.b1;
        off[3] = off[1] | off[2];
    }
    const bool grad = !fwd && op.grad >= 0;
    if (threadIdx.x < D * D) {
        sm[threadIdx.x] = it.pool[(fwd ? op.fwd : op.rev) + threadIdx.x];
        if (grad) sm[16 + threadIdx.x] = it.pool[op.grad + threadIdx.x];
    }
    __syncthreads();
    c128 *psi = it.psi, *lam = it.lam;
    const u64 nblk = u64(1) << (it.n - K), per = (nblk + it.nchunk - 1) / u64(it.nchunk);
    const u64 b_lo = u64(c) * per, b_hi = min(nblk, b_lo + per);
    c128 acc = mk(0, 0);
    for (u64 b = b_lo + threadIdx.x; b < b_hi; b += kBT) {
        u64 base = insert0(b, op.b0);
        if (K == 2) base = insert0(base, op.b1);
        c128 x[D], y[D];
#pragma unroll
        for (int q = 0; q < D; ++q) x[q] = load_nt(psi + base + off[q]);
#pragma unroll
        for (int r = 0; r < D; ++r) {
            c128 a = mk(0, 0);
#pragma unroll
            for (int q = 0; q < D; ++q) a = cfma(sm[r + D * q], x[q], a);
            y[r] = a;
        }
#pragma unroll
        for (int r = 0; r < D; ++r) store_nt(psi + base + off[r], y[r]);
        if (!fwd) {
            c128 l[D], lo2[D];
#pragma unroll
            for (int q = 0; q < D; ++q) l[q] = load_nt(lam + base + off[q]);
#pragma unroll
            for (int r = 0; r < D; ++r) {
                c128 a = mk(0, 0);
#pragma unroll
                for (int q = 0; q < D; ++q) a = cfma(sm[r + D * q], l[q], a);
                lo2[r] = a;
            }
#pragma unroll
            for (int r = 0; r < D; ++r) store_nt(lam + base + off[r], lo2[r]);
            if (grad) {
                // g = Re <psi|D|lambda> on the updated vectors (differentiation/circuit.jl:136-144)
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    c128 dl = mk(0, 0);
#pragma unroll
                    for (int q = 0; q < D; ++q) dl = cfma(sm[16 + r + D * q], lo2[q], dl);
                    acc = cfmaconj(y[r], dl, acc);
                }
            }
        }
    }
    if (grad) {
        c128 s = bblock_sum(acc, red);
        if (threadIdx.x == 0) part[it.chunk0 + c] = s;
    }
}

// Observe step on rows [i0, i1) of one item: E partial = sum_i conj(psi_i)
// (H psi)_i and lambda = (H + H^dagger) psi.  Only the nonzero flip patterns
// of each group load; the diagonal terms of all groups share psi[i].  The
// group records are read as uniform (scalar) loads; the matrices come from
// LDS when they fit, else from the observable's device tables.
constexpr int kObsMatsLds = 512;

__device__ __forceinline__ c128 observe_rows(const BItemDev &it, u64 i0, u64 i1, const ObsGroup *groups,
                                             const c128 *mats, const c128 *seed) {
    const c128 *__restrict__ psi = it.psi;
    c128 acc = mk(0, 0);
    if (it.seed_twice || it.value_only) {
        // one sum: the seed is exactly 2 (M psi)_i (scaling by 2 is exact, so
        // this equals the two-sum form bit for bit)
        u64 i = i0 + threadIdx.x;
        for (; i + kBT < i1; i += 2 * kBT) {  // two rows at a time: their loads overlap
            const u64 ib = i + kBT;
            c128 hpa = mk(0, 0), dha = mk(0, 0), hpb = mk(0, 0), dhb = mk(0, 0);
            obs_row_pair(i, ib, psi, it.G, groups, mats, hpa, dha, hpb, dhb);
            const c128 xa = psi[i], xb = psi[ib];
            hpa = cfma(dha, xa, hpa);
            hpb = cfma(dhb, xb, hpb);
            if (!it.value_only) {
                it.lam[i] = mk(2 * hpa.re, 2 * hpa.im);
                it.lam[ib] = mk(2 * hpb.re, 2 * hpb.im);
            }
            acc = cfmaconj(xa, hpa, acc);
            acc = cfmaconj(xb, hpb, acc);
        }
        for (; i < i1; i += kBT) {
            c128 hp = mk(0, 0), dh = mk(0, 0), u0 = mk(0, 0), u1 = mk(0, 0);
            obs_row<false>(i, psi, it.G, groups, mats, mats, hp, dh, u0, u1);
            const c128 x0 = psi[i];
            hp = cfma(dh, x0, hp);
            if (!it.value_only) it.lam[i] = mk(2 * hp.re, 2 * hp.im);
            acc = cfmaconj(x0, hp, acc);
        }
        return acc;
    }
    for (u64 i = i0 + threadIdx.x; i < i1; i += kBT) {
        c128 hp = mk(0, 0), sd = mk(0, 0), dh = mk(0, 0), ds = mk(0, 0);
        obs_row<true>(i, psi, it.G, groups, mats, seed, hp, dh, sd, ds);
        const c128 x0 = psi[i];
        hp = cfma(dh, x0, hp);
        if (!it.value_only) it.lam[i] = cfma(ds, x0, sd);
        acc = cfmaconj(x0, hp, acc);
    }
    return acc;
}

// Observe step on chunk c of item it, its partial energy into part.
__device__ __forceinline__ void observe_chunk(const BItemDev &it, int c, c128 *red, c128 *__restrict__ part) {
    __shared__ c128 s_mats[kObsMatsLds], s_seed[kObsMatsLds];
    const u64 len = u64(1) << it.n, per = len / u64(it.nchunk);
    const ObsGroup last = it.groups[it.G - 1];
    const int nm = last.moff + (1 << (2 * last.k));
    c128 acc;
    if (nm <= kObsMatsLds) {
        for (int j = threadIdx.x; j < nm; j += kBT) {
            s_mats[j] = it.mats[j];
            if (!it.seed_twice && !it.value_only) s_seed[j] = it.seed[j];
        }
        __syncthreads();
        acc = observe_rows(it, u64(c) * per, u64(c + 1) * per, const_table(it.groups), s_mats, s_seed);
    } else {
        acc = observe_rows(it, u64(c) * per, u64(c + 1) * per, const_table(it.groups), const_table(it.mats),
                           const_table(it.seed));
    }
    c128 s = bblock_sum(acc, red);
    if (threadIdx.x == 0) part[it.chunk0 + c] = s;
}

// Workgroups are dealt to the 8 XCDs round robin (workgroup b runs on XCD
// b mod 8).  The chunk of workgroup b: XCD x takes a contiguous range of the
// nb chunks (q or q + 1 of them), so an item's chunks, whose partner rows are
// each other's, share one XCD's L2 instead of each XCD fetching them.
__device__ __forceinline__ int xcd_chunk(int b, int nb) {
    const int x = b & 7, k = b >> 3, q = nb >> 3, r = nb & 7;
    return x * q + (x < r ? x : r) + k;
}

// The observe step alone, for a table whose items are all at it (the tiled
// items of a batch): without the gate micro-ops' registers it runs at a
// higher occupancy, which hides more of the partner loads.
__global__ __launch_bounds__(kBT) void k_bobserve(const BItemDev *__restrict__ items, int nitems,
                                                  c128 *__restrict__ part) {
    __shared__ c128 red[kBT / 64];
    const int cb = xcd_chunk(int(blockIdx.x), int(gridDim.x));
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (items[mid].chunk0 <= cb)
            lo = mid;
        else
            hi = mid - 1;
    }
    observe_chunk(items[lo], cb - items[lo].chunk0, red, part);
}

// One engine iteration for every active item: a workgroup owns one chunk of
// one item and performs that item's current micro-op on it.
__global__ __launch_bounds__(kBT, 2) void k_biter(const BItemDev *__restrict__ items, int nitems, int t,
                                               c128 *__restrict__ part) {
    __shared__ c128 red[kBT / 64];
    __shared__ c128 sm[32];
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (items[mid].chunk0 <= int(blockIdx.x))
            lo = mid;
        else
            hi = mid - 1;
    }
    const BItemDev &it = items[lo];
    const int n = it.n, nfwd = it.nfwd;
    const int c = int(blockIdx.x) - it.chunk0;
    const int pc = t - it.t0;
    if (pc < 0 || pc > (it.value_only ? nfwd : 2 * nfwd)) return;
    c128 *psi = it.psi, *lam = it.lam;
    if (pc == nfwd) {
        observe_chunk(it, c, red, part);
        return;
    }
    const bool fwd = pc < nfwd;
    const BOp op = it.ops[fwd ? pc : 2 * nfwd - pc];
    if (op.kind == 1)
        bgate_chunk<1>(it, op, fwd, c, sm, red, part);
    else
        bgate_chunk<2>(it, op, fwd, c, sm, red, part);
}

// Per-item reductions of this iteration's partials (fixed order).
__global__ __launch_bounds__(kBT) void k_breduce(const BItemDev *__restrict__ items, int t,
                                                 const c128 *__restrict__ part, double *__restrict__ values,
                                                 double *__restrict__ grads) {
    __shared__ c128 red[kBT / 64];
    const BItemDev &it = items[blockIdx.x];
    const int pc = t - it.t0;
    int slot = -1;
    bool is_value = false;
    if (pc == it.nfwd) {
        is_value = true;
    } else if (!it.value_only && pc > it.nfwd && pc <= 2 * it.nfwd) {
        const BOp op = it.ops[2 * it.nfwd - pc];
        if (op.grad >= 0) slot = it.gslot0 + op.gslot;
    }
    if (!is_value && slot < 0) return;
    c128 a = mk(0, 0);
    for (int j = threadIdx.x; j < it.nchunk; j += kBT) a = cadd(a, part[it.chunk0 + j]);
    c128 s = bblock_sum(a, red);
    if (threadIdx.x == 0) {
        if (is_value)
            values[it.vslot] = s.re;
        else
            grads[slot] = s.re;
    }
}

int batch_chunks(int n) {
    u64 pairs = u64(1) << (n - 1);
    u64 c = pairs / kBatchPairsPerChunk;
    return int(std::max<u64>(c, 1));
}

int batch_lower(const Program *p, std::vector<BOp> &ops, int *nderiv) {
    MOSSQ_CHECK(!p->is_density, MOSSQ_EUNSUPPORTED, "the batched engine runs state-vector programs");
    ops.clear();
    for (const Step &st : p->steps) {
        MOSSQ_CHECK(!st.split && st.kb <= 2, MOSSQ_EUNSUPPORTED, "batched engine: gates on at most two qubits");
        BOp o{};
        o.kind = st.kb;
        o.b0 = st.bits[0];
        o.b1 = st.kb == 2 ? st.bits[1] : 0;
        o.fwd = st.fwd;
        o.rev = st.rev;
        o.grad = st.grad;
        o.gslot = st.grad_index;
        ops.push_back(o);
    }
    *nderiv = p->nderiv;
    return MOSSQ_OK;
}

int batch_iteration_on(hipStream_t st, const BItemDev *items, int nitems, int nchunks_total, int t, c128 *part,
                       double *values, double *grads) {
    if (nitems == 0) return MOSSQ_OK;
    hipLaunchKernelGGL(k_biter, dim3(nchunks_total), dim3(kBT), 0, st, items, nitems, t, part);
    MOSSQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_breduce, dim3(nitems), dim3(kBT), 0, st, items, t, part, values, grads);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

int batch_iteration(Ctx *c, const BItemDev *items, int nitems, int nchunks_total, int t, c128 *part, double *values,
                    double *grads) {
    return batch_iteration_on(c->stream, items, nitems, nchunks_total, t, part, values, grads);
}

// Initial states |0...0> for the items of a table whose psi is not preset; a
// workgroup clears one chunk of one item.
__global__ __launch_bounds__(kBT) void k_binit(const BItemDev *__restrict__ items, int nitems) {
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (items[mid].chunk0 <= int(blockIdx.x))
            lo = mid;
        else
            hi = mid - 1;
    }
    const BItemDev &it = items[lo];
    if (it.preset) return;
    const u64 len = u64(1) << it.n, per = len / u64(it.nchunk);
    const u64 c = u64(blockIdx.x) - u64(it.chunk0);
    for (u64 i = c * per + threadIdx.x; i < (c + 1) * per; i += kBT) it.psi[i] = mk(i == 0 ? 1.0 : 0.0, 0.0);
}

// Gradient rows of the items whose reverse pass ran as batched fused sweeps:
// one workgroup sums one parameter's row of partials in a fixed order.
struct BRows {
    const c128 *part;
    int row, count, gslot0, p0;
};

__global__ __launch_bounds__(kBT) void k_brows(const BRows *__restrict__ rows, int nrows, double *__restrict__ grads) {
    __shared__ c128 red[kBT / 64];
    int lo = 0, hi = nrows - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (rows[mid].p0 <= int(blockIdx.x))
            lo = mid;
        else
            hi = mid - 1;
    }
    const BRows r = rows[lo];
    const int j = int(blockIdx.x) - r.p0;
    const c128 *src = r.part + size_t(j) * r.row;
    c128 a = mk(0, 0);
    for (int k = threadIdx.x; k < r.row; k += kBT) a = cadd(a, src[k]);
    const c128 s = bblock_sum(a, red);
    if (threadIdx.x == 0) grads[r.gslot0 + j] = s.re;
}

// Everything one batch needs on the device, built on the host once: the
// item tables, the launch tables of the tiled sweeps, and the scratch
// vectors.  An engine call is prepare, enqueue, finish; a captured batch graph
// keeps its BatchRun (and the memory) until the graph is destroyed.
struct BatchRun {
    Ctx *c;
    int B = 0;
    bool value_only = false;
    std::vector<const Program *> progs;
    std::vector<int64_t> grad_offset;
    Allocs al;
    c128 *vec = nullptr, *part = nullptr, *part2 = nullptr, *rpart = nullptr;
    BOp *d_ops = nullptr;
    BItemDev *d_items = nullptr;
    BRows *d_rows = nullptr;
    double *d_vals = nullptr, *d_grads = nullptr;
    int nchunks[2] = {0, 0}, ntab[2] = {0, 0}, tmax = 0, np = 0, nrows = 0;
    i64 dtot = 0, rev_parts = 0;
    BatchSweeps fwd, rev;
    std::vector<std::pair<c128 *, const void *>> presets;  // (psi, init) copies of each call
    std::vector<u64> preset_len;
    // inputs of the reverse launch tables when their build is deferred until the
    // forward launches are enqueued (rev_deferred)
    bool rev_deferred = false;
    std::vector<const Program *> tprogs;
    std::vector<u64> tx, tl, tpart;
    std::vector<int> trow;
    u64 vec_len = 0;
    explicit BatchRun(Ctx *cc) : c(cc), al(cc) {}
};

// Batch of B items.  Items whose programs carry all-sweep fused plans (12 or
// more qubits) run their forward and reverse passes as batched LDS-tile sweeps
// (launch s: sweep s of every such item) around one observe launch; the others
// run the gate-by-gate engine micro-program.
// value_only: forward and observe steps only (no lambda vectors, no gradients);
// init[b], when given, is item b's initial state (else |0...0>).
static int batch_prepare(BatchRun &R, int B, const Program *const *progs_in, const Obs *const *obs,
                         const int64_t *grad_offset, bool value_only, const void *const *init,
                         bool defer_rev = false) {
    Ctx *c = R.c;
    R.B = B;
    R.value_only = value_only;
    // A program of fewer than 12 qubits with a 12-qubit form (Program::pad)
    // runs as that form's tile sweeps, from |0...0> (an item with its own
    // initial state keeps the gate-by-gate engine: its vector has 2^n entries).
    R.progs.assign(progs_in, progs_in + B);
    for (int b = 0; b < B; ++b) {
        const Program *pd = fused_pad(progs_in[b]);
        if (pd && !(init && init[b]) && fused_plan_batchable(pd, !value_only)) R.progs[b] = pd;
    }
    const Program *const *progs = R.progs.data();
    R.grad_offset.assign(B, 0);
    if (!value_only) R.grad_offset.assign(grad_offset, grad_offset + B);
    std::vector<BItemHost> items(B);
    std::vector<BOp> all_ops;
    std::vector<int> op_off(B), nder(B);
    std::vector<char> tiled(B);
    std::vector<int> order;  // tiled items first, then stepped ones
    i64 total_amps = 0;
    for (int b = 0; b < B; ++b) {
        tiled[b] = fused_plan_batchable(progs[b], !value_only);
        if (tiled[b]) {
            items[b].ops.clear();
            nder[b] = progs[b]->nderiv;
            if (!value_only) R.rev_parts += i64(std::max(progs[b]->nderiv, 1)) * fused_adjoint_parts(progs[b]);
        } else {
            MOSSQ_TRY(batch_lower(progs[b], items[b].ops, &nder[b]));
        }
        op_off[b] = int(all_ops.size());
        all_ops.insert(all_ops.end(), items[b].ops.begin(), items[b].ops.end());
        items[b].nfwd = int(items[b].ops.size());
        items[b].nchunk = batch_chunks(progs[b]->n);
        items[b].gslot0 = value_only ? 0 : int(grad_offset[b]);
        total_amps += i64(1) << progs[b]->n;
        if (!value_only) R.dtot = std::max<i64>(R.dtot, grad_offset[b] + nder[b]);
        const int k = tiled[b] ? 0 : 1;
        R.nchunks[k] += items[b].nchunk;
        ++R.ntab[k];
        if (!tiled[b]) R.tmax = std::max(R.tmax, (value_only ? 1 : 2) * items[b].nfwd);
    }
    for (int k = 0; k < 2; ++k)
        for (int b = 0; b < B; ++b)
            if ((tiled[b] != 0) == (k == 0)) order.push_back(b);
    Allocs &al = R.al;
    MOSSQ_TRY(al.get(u64(total_amps) * (value_only ? 1 : 2) * sizeof(c128), (void **)&R.vec));
    MOSSQ_TRY(al.get(u64(std::max(R.nchunks[0], 1)) * sizeof(c128), (void **)&R.part));
    MOSSQ_TRY(al.get(u64(std::max(R.nchunks[1], 1)) * sizeof(c128), (void **)&R.part2));
    MOSSQ_TRY(al.get(std::max<size_t>(all_ops.size(), 1) * sizeof(BOp), (void **)&R.d_ops));
    MOSSQ_TRY(al.get(u64(B) * sizeof(BItemDev), (void **)&R.d_items));
    MOSSQ_TRY(al.get(u64(B) * sizeof(double), (void **)&R.d_vals));
    MOSSQ_TRY(al.get(u64(std::max<i64>(R.dtot, 1)) * sizeof(double), (void **)&R.d_grads));
    if (R.rev_parts) MOSSQ_TRY(al.get(u64(R.rev_parts) * sizeof(c128), (void **)&R.rpart));
    c128 *vec = R.vec;
    std::vector<BItemDev> hd(B);
    std::vector<const Program *> tprogs;
    std::vector<u64> tx, tl, tpart;
    std::vector<int> trow;
    std::vector<BRows> rows;
    i64 vo = 0, ro = 0;
    int ch[2] = {0, 0};
    for (int pos = 0; pos < B; ++pos) {
        const int b = order[pos];
        const int k = tiled[b] ? 0 : 1;
        const i64 len = i64(1) << progs[b]->n;
        BItemDev &d = hd[pos];
        d.psi = vec + vo;
        d.lam = value_only ? nullptr : vec + vo + len;
        vo += (value_only ? 1 : 2) * len;
        d.ops = R.d_ops + op_off[b];
        d.pool = progs[b]->d_pool;
        d.groups = obs[b]->d_groups;
        d.mats = obs[b]->d_mats;
        d.seed = obs[b]->d_mats_seed;
        d.n = progs[b]->n;
        d.nfwd = items[b].nfwd;
        d.G = obs[b]->G;
        d.t0 = 0;
        d.chunk0 = ch[k];
        d.nchunk = items[b].nchunk;
        d.gslot0 = items[b].gslot0;
        d.vslot = b;
        d.value_only = value_only ? 1 : 0;
        d.preset = (init && init[b]) ? 1 : 0;
        d.seed_twice = obs[b]->seed_twice ? 1 : 0;
        ch[k] += d.nchunk;
        if (d.preset) {
            R.presets.push_back({d.psi, init[b]});
            R.preset_len.push_back(u64(len));
        }
        if (tiled[b]) {
            tprogs.push_back(progs[b]);
            tx.push_back(u64(d.psi - vec));
            tl.push_back(value_only ? 0 : u64(d.lam - vec));
            const int row = value_only ? 0 : fused_adjoint_parts(progs[b]);
            tpart.push_back(u64(ro));
            trow.push_back(row);
            if (!value_only && nder[b] > 0) {
                rows.push_back(BRows{R.rpart + ro, row, nder[b], d.gslot0, R.np});
                R.np += nder[b];
            }
            ro += i64(std::max(nder[b], 1)) * row;
        }
    }
    const int nt = int(tprogs.size());
    if (nt) {
        const u64 vec_len = u64(total_amps) * (value_only ? 1 : 2);
        MOSSQ_TRY(batch_sweeps_build(c, al, nt, tprogs.data(), vec, vec_len, tx.data(), nullptr, nullptr, 0, nullptr,
                                     nullptr, false, &R.fwd));
        if (!value_only && defer_rev) {
            R.rev_deferred = true;
            R.tprogs.swap(tprogs);
            R.tx.swap(tx);
            R.tl.swap(tl);
            R.tpart.swap(tpart);
            R.trow.swap(trow);
            R.vec_len = vec_len;
        } else if (!value_only) {
            MOSSQ_TRY(batch_sweeps_build(c, al, nt, tprogs.data(), vec, vec_len, tx.data(), tl.data(), R.rpart,
                                         u64(R.rev_parts), tpart.data(), trow.data(), true, &R.rev));
        }
    }
    R.nrows = int(rows.size());
    if (!rows.empty()) {
        MOSSQ_TRY(al.get(rows.size() * sizeof(BRows), (void **)&R.d_rows));
        MOSSQ_HIP(hipMemcpyAsync(R.d_rows, rows.data(), rows.size() * sizeof(BRows), hipMemcpyHostToDevice,
                                 c->stream));
    }
    if (!all_ops.empty())
        MOSSQ_HIP(hipMemcpyAsync(R.d_ops, all_ops.data(), all_ops.size() * sizeof(BOp), hipMemcpyHostToDevice,
                                 c->stream));
    MOSSQ_HIP(hipMemcpyAsync(R.d_items, hd.data(), hd.size() * sizeof(BItemDev), hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));  // the host tables go out of scope
    return MOSSQ_OK;
}

// The device work of one engine call on the context stream: initial states,
// tiled sweeps and observe, gate-by-gate items on the second stream (joined
// before return).  Kernel launches, event waits and device-to-device copies of
// the initial states only, so a stream capture of it replays as a whole.
static int batch_enqueue(BatchRun &R) {
    Ctx *c = R.c;
    for (size_t i = 0; i < R.presets.size(); ++i)
        MOSSQ_HIP(hipMemcpyAsync(R.presets[i].first, R.presets[i].second, R.preset_len[i] * sizeof(c128),
                                 hipMemcpyDeviceToDevice, c->stream));
    const BItemDev *tab[2] = {R.d_items, R.d_items + R.ntab[0]};
    for (int k = 0; k < 2; ++k)
        if (R.ntab[k]) {
            hipLaunchKernelGGL(k_binit, dim3(R.nchunks[k]), dim3(kBT), 0, c->stream, tab[k], R.ntab[k]);
            MOSSQ_HIP(hipGetLastError());
        }
    // With both kinds of items, the gate-by-gate items (many small launches)
    // run on the second stream beside the tiled sweeps; they share no memory
    // but the result arrays, where their slots are disjoint.
    const bool overlap = R.ntab[0] && R.ntab[1];
    hipStream_t s2 = c->stream;
    AuxJoin join(c);  // on every return, before the scratch can go back to the cache
    if (overlap) {
        MOSSQ_TRY(ctx_aux(c));
        MOSSQ_HIP(hipEventRecord(c->aux_ev[0], c->stream));
        MOSSQ_HIP(hipStreamWaitEvent(c->aux, c->aux_ev[0], 0));
        s2 = c->aux;
        join.active = true;
    }
    // The tiled sweeps (a few long launches) are enqueued before the
    // gate-by-gate items' many short ones, so the device starts on them while
    // the host is still enqueueing the second stream's launches.
    if (R.ntab[0]) {
        MOSSQ_TRY(batch_sweeps_run(c, R.fwd, false));
        // observe (the tiled items have no engine gate steps: t = 0 is their observe step)
        hipLaunchKernelGGL(k_bobserve, dim3(R.nchunks[0]), dim3(kBT), 0, c->stream, tab[0], R.ntab[0], R.part);
        MOSSQ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_breduce, dim3(R.ntab[0]), dim3(kBT), 0, c->stream, tab[0], 0, R.part, R.d_vals,
                           R.d_grads);
        MOSSQ_HIP(hipGetLastError());
    }
    if (R.ntab[0] && !R.value_only) {
        if (R.rev_deferred) {
            // built now, while the device runs the forward launches; the table
            // copies queue behind them on the stream
            MOSSQ_TRY(batch_sweeps_build(c, R.al, int(R.tprogs.size()), R.tprogs.data(), R.vec, R.vec_len,
                                         R.tx.data(), R.tl.data(), R.rpart, u64(R.rev_parts), R.tpart.data(),
                                         R.trow.data(), true, &R.rev, false));
            R.rev_deferred = false;
        }
        MOSSQ_TRY(zero_fill(c, R.rpart, u64(R.rev_parts)));
        MOSSQ_TRY(batch_sweeps_run(c, R.rev, true));
        if (R.np) {
            hipLaunchKernelGGL(k_brows, dim3(R.np), dim3(kBT), 0, c->stream, R.d_rows, R.nrows, R.d_grads);
            MOSSQ_HIP(hipGetLastError());
        }
    }
    if (R.ntab[1])
        for (int t = 0; t <= R.tmax; ++t)
            MOSSQ_TRY(batch_iteration_on(s2, tab[1], R.ntab[1], R.nchunks[1], t, R.part2, R.d_vals, R.d_grads));
    if (overlap) {
        join.active = false;
        MOSSQ_HIP(hipEventRecord(c->aux_ev[1], c->aux));
        MOSSQ_HIP(hipStreamWaitEvent(c->stream, c->aux_ev[1], 0));
    }
    return MOSSQ_OK;
}

// Values and gradients to the host (one synchronisation).
static int batch_finish(BatchRun &R, double *values, double *grads) {
    Ctx *c = R.c;
    std::vector<double> hv(R.B), hg(std::max<i64>(R.dtot, 1));
    MOSSQ_HIP(hipMemcpyAsync(hv.data(), R.d_vals, R.B * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipMemcpyAsync(hg.data(), R.d_grads, hg.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    for (int b = 0; b < R.B; ++b) {
        values[b] = hv[b];
        if (!R.value_only)
            for (int j = 0; j < R.progs[b]->nderiv; ++j) grads[R.grad_offset[b] + j] = hg[R.grad_offset[b] + j];
    }
    return MOSSQ_OK;
}

static int batch_adjoint_engine(Ctx *c, int B, const Program *const *progs, const Obs *const *obs,
                                const int64_t *grad_offset, double *values, double *grads, bool value_only,
                                const void *const *init) {
    BatchRun R(c);
    MOSSQ_TRY(batch_prepare(R, B, progs, obs, grad_offset, value_only, init, true));
    MOSSQ_TRY(batch_enqueue(R));
    return batch_finish(R, values, grads);
}

void batch_run_free(BatchRun *r) { delete r; }

static int batch_items_check(Ctx *c, int B, mossq_program *const *progs, mossq_obs *const *obs,
                             std::vector<const Program *> &P, std::vector<const Obs *> &O) {
    P.resize(B);
    O.resize(B);
    for (int b = 0; b < B; ++b) {
        P[b] = reinterpret_cast<const Program *>(progs[b]);
        O[b] = reinterpret_cast<const Obs *>(obs[b]);
        MOSSQ_CHECK(P[b] && O[b], MOSSQ_EINVAL, "null program or observable in batch");
        MOSSQ_CHECK(P[b]->ctx == c && O[b]->ctx == c, MOSSQ_EINVAL, "batch item belongs to another context");
        MOSSQ_CHECK(!P[b]->is_density, MOSSQ_EUNSUPPORTED, "the batched engine takes state-vector programs");
        MOSSQ_CHECK(O[b]->n == P[b]->n, MOSSQ_EINVAL, "observable and program qubit counts differ");
    }
    return MOSSQ_OK;
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_batch_adjoint(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                        const int64_t *grad_offset, double *values, double *grads) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(B >= 0 && (B == 0 || (progs && obs && grad_offset && values && grads)), MOSSQ_EINVAL,
                "bad batch arguments");
    Ctx *c = C(ctx);
    std::vector<const Program *> P;
    std::vector<const Obs *> O;
    MOSSQ_TRY(batch_items_check(c, B, progs, obs, P, O));
    for (int b = 0; b < B; ++b) MOSSQ_CHECK(grad_offset[b] >= 0, MOSSQ_EINVAL, "negative gradient offset");
    if (B == 0) return MOSSQ_OK;
    return batch_adjoint_engine(c, B, P.data(), O.data(), grad_offset, values, grads, false, nullptr);
}

int mossq_batch_adjoint_init(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                             const void *const *init, const int64_t *grad_offset, double *values, double *grads) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(B >= 0 && (B == 0 || (progs && obs && grad_offset && values && grads)), MOSSQ_EINVAL,
                "bad batch arguments");
    Ctx *c = C(ctx);
    std::vector<const Program *> P;
    std::vector<const Obs *> O;
    MOSSQ_TRY(batch_items_check(c, B, progs, obs, P, O));
    for (int b = 0; b < B; ++b) MOSSQ_CHECK(grad_offset[b] >= 0, MOSSQ_EINVAL, "negative gradient offset");
    if (B == 0) return MOSSQ_OK;
    return batch_adjoint_engine(c, B, P.data(), O.data(), grad_offset, values, grads, false, init);
}

int mossq_batch_capture(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                        const int64_t *grad_offset, mossq_graph **out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(B >= 1 && progs && obs && grad_offset && out, MOSSQ_EINVAL, "bad batch arguments");
    Ctx *c = C(ctx);
    std::vector<const Program *> P;
    std::vector<const Obs *> O;
    MOSSQ_TRY(batch_items_check(c, B, progs, obs, P, O));
    for (int b = 0; b < B; ++b) MOSSQ_CHECK(grad_offset[b] >= 0, MOSSQ_EINVAL, "negative gradient offset");
    auto *R = new BatchRun(c);
    int s = batch_prepare(*R, B, P.data(), O.data(), grad_offset, false, nullptr);
    if (s == MOSSQ_OK) s = ctx_aux(c);  // created outside the capture
    if (s != MOSSQ_OK) {
        delete R;
        return s;
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        s = batch_enqueue(*R);
        e = hipStreamEndCapture(c->stream, &g);
    }
    hipGraphExec_t ex = nullptr;
    if (s == MOSSQ_OK && e == hipSuccess) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (s != MOSSQ_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        delete R;
        if (s != MOSSQ_OK) return s;
        return fail(MOSSQ_EDEVICE, std::string("batch capture: ") + hipGetErrorString(e));
    }
    auto *gr = new mossq_graph();
    gr->ctx = c;
    gr->batch = R;
    for (const Program *p : P)
        if (std::find(gr->progs.begin(), gr->progs.end(), p) == gr->progs.end()) gr->progs.push_back(p);
    for (const Program *p : gr->progs) p->live_graphs++;
    gr->graph = g;
    gr->exec = ex;
    *out = gr;
    return MOSSQ_OK;
}

int mossq_batch_graph_run(mossq_ctx *ctx, mossq_graph *graph, double *values, double *grads) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(graph && graph->batch && values && grads, MOSSQ_EINVAL, "bad batch graph arguments");
    Ctx *c = C(ctx);
    MOSSQ_CHECK(graph->ctx == c, MOSSQ_EINVAL, "graph belongs to another context");
    MOSSQ_HIP(hipGraphLaunch(graph->exec, c->stream));
    return batch_finish(*graph->batch, values, grads);
}

int mossq_batch_expectation(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                            const void *const *init, double *values) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(B >= 0 && (B == 0 || (progs && obs && values)), MOSSQ_EINVAL, "bad batch arguments");
    Ctx *c = C(ctx);
    std::vector<const Program *> P;
    std::vector<const Obs *> O;
    MOSSQ_TRY(batch_items_check(c, B, progs, obs, P, O));
    if (B == 0) return MOSSQ_OK;
    std::vector<int64_t> zero(B, 0);
    return batch_adjoint_engine(c, B, P.data(), O.data(), zero.data(), values, nullptr, true, init);
}

}  // extern "C"
