// Observables (QubitsOperator grouped by site tuple), expectations, operator
// action, dot products, traces and readout marginals.
//
// All reductions are deterministic: each workgroup writes a partial sum to a
// fixed slot and a single-workgroup kernel adds the partials in a fixed order,
// so repeated calls return bit-identical values (the reference's threaded
// parallel_sum adds unordered atomics, utils/parallel_for.jl:47-80).
#include "mossq_obs.h"

namespace mossq {

constexpr int kB = 256;
constexpr int kMaxParts = 512;  // partial sums per group / per reduction

static inline unsigned blocks_for(u64 work, int items, unsigned cap) {
    u64 g = (work + u64(kB) * items - 1) / (u64(kB) * items);
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

__device__ __forceinline__ c128 wave_sum(c128 a) {
    for (int o = 32; o > 0; o >>= 1) {
        a.re += __shfl_xor(a.re, o, 64);
        a.im += __shfl_xor(a.im, o, 64);
    }
    return a;
}

// Block reduction of one c128 per thread; result valid in thread 0.
__device__ __forceinline__ c128 block_sum(c128 a) {
    __shared__ c128 red[kB / 64];
    a = wave_sum(a);
    int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = a;
    __syncthreads();
    c128 s = mk(0, 0);
    if (threadIdx.x == 0)
        for (int i = 0; i < kB / 64; ++i) s = cadd(s, red[i]);
    return s;
}

// out[g] = sum_{j < nparts} part[g * stride + j], one workgroup per g.
__global__ __launch_bounds__(kB) void k_reduce_rows(const c128 *__restrict__ part, int nparts, int stride,
                                                    c128 *__restrict__ out) {
    const c128 *p = part + size_t(blockIdx.x) * stride;
    c128 a = mk(0, 0);
    for (int j = threadIdx.x; j < nparts; j += kB) a = cadd(a, p[j]);
    c128 s = block_sum(a);
    if (threadIdx.x == 0) out[blockIdx.x] = s;
}

int reduce_partials(Ctx *c, const c128 *part, int nparts, c128 *dev_out) {
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kB), 0, c->stream, part, nparts, nparts, dev_out);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// ------------------------------------------------------------ expectation ----
// One pass over the state for every group of the observable.  Row i's action
// (A psi)_i = sum_g sum_x A_g[r, r ^ x] psi[i ^ flip_g(x)] (r = local index of
// i on g's sites) loads only the flip patterns in xmask, and every group's
// diagonal term shares the one load of psi_i.  Each workgroup writes the
// partial sum of conj(phi_i) (M psi)_i over its rows; with OUT = 1 it also
// stores out_i = (W psi)_i (W = M + M^dagger: the adjoint seed), and with
// OUT = 2 (every W entry exactly 2 M: a Hermitian observable) out_i =
// 2 (M psi)_i from the one sum, equal to the W form bit for bit.  Group
// records are read as uniform (scalar) loads; the matrices, gathered at each
// row's local index, are staged in LDS when they fit.
// (The reference sums one pass per site group, observables/expectation.jl:12-33.)
constexpr int kObsLdsMats = 512;

template <int OUT>
__global__ __launch_bounds__(kB) void k_observe(const c128 *__restrict__ phi, const c128 *__restrict__ psi, u64 len,
                                                int G, const ObsGroup *__restrict__ groups,
                                                const c128 *__restrict__ mats, const c128 *__restrict__ wmats,
                                                c128 *__restrict__ out, c128 *__restrict__ part) {
    constexpr bool WRITE = OUT == 1;
    __shared__ c128 s_m[kObsLdsMats], s_w[WRITE ? kObsLdsMats : 1];
    const ObsGroup last = groups[G - 1];
    const int nm = last.moff + (1 << (2 * last.k));
    const bool staged = nm <= kObsLdsMats;
    if (staged) {
        for (int j = threadIdx.x; j < nm; j += kB) {
            s_m[j] = mats[j];
            if (WRITE) s_w[j] = wmats[j];
        }
        __syncthreads();
    }
    const c128 *M = staged ? s_m : mats;
    const c128 *W = staged ? s_w : wmats;
    // blockIdx.y: the column of a batch (consecutive vectors of len entries)
    phi += u64(blockIdx.y) * len;
    psi += u64(blockIdx.y) * len;
    if (OUT) out += u64(blockIdx.y) * len;
    c128 acc = mk(0, 0);
    for (u64 i = u64(blockIdx.x) * kB + threadIdx.x; i < len; i += u64(gridDim.x) * kB) {
        c128 hp = mk(0, 0), dh = mk(0, 0), wp = mk(0, 0), dw = mk(0, 0);
        obs_row<WRITE>(i, psi, G, groups, M, W, hp, dh, wp, dw);
        const c128 x0 = psi[i];
        hp = cfma(dh, x0, hp);
        acc = cfmaconj(phi[i], hp, acc);
        if (OUT == 1) out[i] = cfma(dw, x0, wp);
        if (OUT == 2) out[i] = mk(2 * hp.re, 2 * hp.im);
    }
    const c128 s = block_sum(acc);
    if (threadIdx.x == 0) part[size_t(blockIdx.y) * gridDim.x + blockIdx.x] = s;
}

int observe_dev(Ctx *c, const Obs *o, const c128 *phi, const c128 *psi, c128 *seed, c128 *dev_out, c128 *part) {
    const u64 len = u64(1) << o->n;
    const unsigned nb = blocks_for(len, 4, kObsParts);
    if (seed && o->seed_twice)
        hipLaunchKernelGGL(k_observe<2>, dim3(nb), dim3(kB), 0, c->stream, phi, psi, len, o->G, o->d_groups,
                           o->d_mats, o->d_mats, seed, part);
    else if (seed)
        hipLaunchKernelGGL(k_observe<1>, dim3(nb), dim3(kB), 0, c->stream, phi, psi, len, o->G, o->d_groups,
                           o->d_mats, o->d_mats_seed, seed, part);
    else
        hipLaunchKernelGGL(k_observe<0>, dim3(nb), dim3(kB), 0, c->stream, phi, psi, len, o->G, o->d_groups,
                           o->d_mats, o->d_mats, nullptr, part);
    MOSSQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kB), 0, c->stream, part, int(nb), int(nb), dev_out);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

int expect_total_host(Ctx *c, const Obs *o, const c128 *phi, const c128 *psi, double *out) {
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (u64(kObsParts) + 1) * sizeof(c128), &scr));
    c128 *part = (c128 *)scr;
    c128 *res = part + kObsParts;
    MOSSQ_TRY(observe_dev(c, o, phi, psi, nullptr, res, part));
    c128 h;
    MOSSQ_HIP(hipMemcpyAsync(&h, res, sizeof(c128), hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    out[0] = h.re;
    out[1] = h.im;
    return MOSSQ_OK;
}

// ---------------------------------------------------------- operator action ----
// out[i] = sum_g sum_c M_g[r_g(i), c] psi[i with the sites of g set to c],
// summed over the nonzero flip patterns c = r ^ x of each group.
// nbits may exceed the observable's n (left action on density-matrix rows).
__global__ __launch_bounds__(kB) void k_apply_op(c128 *__restrict__ out, const c128 *__restrict__ psi, u64 len,
                                                 int G, const ObsGroup *__restrict__ groups,
                                                 const c128 *__restrict__ mats) {
    for (u64 i = u64(blockIdx.x) * kB + threadIdx.x; i < len; i += u64(gridDim.x) * kB) {
        c128 acc = mk(0, 0), dg = mk(0, 0), unused0 = mk(0, 0), unused1 = mk(0, 0);
        obs_row<false>(i, psi, G, groups, mats, mats, acc, dg, unused0, unused1);
        out[i] = cfma(dg, psi[i], acc);
    }
}

int apply_op_dev(Ctx *c, const Obs *o, const c128 *psi, c128 *out, int nbits, OpMats which) {
    u64 len = u64(1) << nbits;
    const c128 *m = which == OpMats::Plain ? o->d_mats : which == OpMats::Seed ? o->d_mats_seed : o->d_mats_adj;
    hipLaunchKernelGGL(k_apply_op, dim3(blocks_for(len, 2, 1u << 20)), dim3(kB), 0, c->stream, out, psi, len,
                       o->G, o->d_groups, m);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// ---------------------------------------------------- density expectation ----
// sum_g sum_b sum_{r,c} M_g[r,c] rho[b|c, b|r]
__global__ __launch_bounds__(kB) void k_dm_expect_groups(const c128 *__restrict__ rho, int n,
                                                         const ObsGroup *__restrict__ groups,
                                                         const c128 *__restrict__ mats, c128 *__restrict__ part,
                                                         int nparts) {
    const ObsGroup g = groups[blockIdx.y];
    const int D = 1 << g.k;
    const u64 N = u64(1) << n;
    const u64 nbase = u64(1) << (n - g.k);
    rho += u64(blockIdx.z) * N * N;  // the column of a batch
    c128 acc = mk(0, 0);
    for (u64 t = u64(blockIdx.x) * kB + threadIdx.x; t < nbase; t += u64(gridDim.x) * kB) {
        u64 b = insert0s_dyn(t, g.sites, g.k);
        for (int r = 0; r < D; ++r) {
            u64 orr = 0;
            for (int j = 0; j < g.k; ++j)
                if (r & (1 << j)) orr |= u64(1) << g.sites[j];
            for (int cc = 0; cc < D; ++cc) {
                u64 oc = 0;
                for (int j = 0; j < g.k; ++j)
                    if (cc & (1 << j)) oc |= u64(1) << g.sites[j];
                acc = cfma(mats[g.moff + r + D * cc], rho[(b | oc) + N * (b | orr)], acc);
            }
        }
    }
    c128 s = block_sum(acc);
    if (threadIdx.x == 0) part[(size_t(blockIdx.z) * gridDim.y + blockIdx.y) * nparts + blockIdx.x] = s;
}

// ------------------------------------------------------------ dot / trace ----
__global__ __launch_bounds__(kB) void k_dot(const c128 *__restrict__ a, const c128 *__restrict__ b, u64 len,
                                            u64 stride_a, c128 *__restrict__ part) {
    c128 acc = mk(0, 0);
    for (u64 i = u64(blockIdx.x) * kB + threadIdx.x; i < len; i += u64(gridDim.x) * kB)
        acc = cfmaconj(a[i * stride_a], b[i * stride_a], acc);
    c128 s = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ __launch_bounds__(kB) void k_trace(const c128 *__restrict__ rho, u64 N, c128 *__restrict__ part) {
    c128 acc = mk(0, 0);
    for (u64 i = u64(blockIdx.x) * kB + threadIdx.x; i < N; i += u64(gridDim.x) * kB)
        acc = cadd(acc, rho[i + N * i]);
    c128 s = block_sum(acc);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

static int finish_scalar(Ctx *c, c128 *part, int nb, double *out) {
    c128 *res = part + kMaxParts;
    MOSSQ_TRY(reduce_partials(c, part, nb, res));
    c128 h;
    MOSSQ_HIP(hipMemcpyAsync(&h, res, sizeof(c128), hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    out[0] = h.re;
    out[1] = h.im;
    return MOSSQ_OK;
}

// ----------------------------------------------------------- probabilities ----
// part[chunk * nbins + m] = sum of |amp|^2 (or Re rho_ii) over the chunk's
// members of readout bin m.
__global__ __launch_bounds__(kB) void k_marginal(const c128 *__restrict__ s, int n, int is_dm, int k,
                                                 ObsSites q, ObsSites rest, u64 nchunks, u64 chunk_len,
                                                 double *__restrict__ part) {
    const u64 nbins = u64(1) << k;
    const u64 total = nbins * nchunks;
    const u64 N = u64(1) << n;
    for (u64 t = u64(blockIdx.x) * kB + threadIdx.x; t < total; t += u64(gridDim.x) * kB) {
        u64 m = t % nbins, ch = t / nbins;
        u64 hi = 0;
        for (int j = 0; j < k; ++j)
            if (m & (u64(1) << j)) hi |= u64(1) << q.s[j];
        double acc = 0;
        for (u64 r = ch * chunk_len; r < (ch + 1) * chunk_len; ++r) {
            u64 lo = 0;
            for (int j = 0; j < n - k; ++j)
                if (r & (u64(1) << j)) lo |= u64(1) << rest.s[j];
            u64 i = hi | lo;
            if (is_dm) {
                acc += s[i + N * i].re;
            } else {
                c128 a = s[i];
                acc += a.re * a.re + a.im * a.im;
            }
        }
        part[ch * nbins + m] = acc;
    }
}

__global__ __launch_bounds__(kB) void k_marginal_reduce(const double *__restrict__ part, u64 nbins, u64 nchunks,
                                                        double *__restrict__ out) {
    for (u64 m = u64(blockIdx.x) * kB + threadIdx.x; m < nbins; m += u64(gridDim.x) * kB) {
        double acc = 0;
        for (u64 ch = 0; ch < nchunks; ++ch) acc += part[ch * nbins + m];
        out[m] = acc;
    }
}

// Per-block bounds of the diagonal of rho: part[3 b + 0] = max |Im rho_ii|,
// part[3 b + 1] = min Re rho_ii, part[3 b + 2] = 1 if every entry seen is
// finite (else 0).
__global__ __launch_bounds__(kB) void k_diag_bounds(const c128 *__restrict__ rho, int n, double *__restrict__ part) {
    __shared__ double red[3][kB / 64];
    const u64 N = u64(1) << n;
    double mi = 0, mr = INFINITY, fin = 1;
    for (u64 i = u64(blockIdx.x) * kB + threadIdx.x; i < N; i += u64(gridDim.x) * kB) {
        const c128 d = rho[i + N * i];
        if (!isfinite(d.re) || !isfinite(d.im)) fin = 0;
        mi = fmax(mi, fabs(d.im));
        mr = fmin(mr, d.re);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mi = fmax(mi, __shfl_xor(mi, o, 64));
        mr = fmin(mr, __shfl_xor(mr, o, 64));
        fin = fmin(fin, __shfl_xor(fin, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = mi;
        red[1][threadIdx.x >> 6] = mr;
        red[2][threadIdx.x >> 6] = fin;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kB / 64; ++w) {
            mi = fmax(mi, red[0][w]);
            mr = fmin(mr, red[1][w]);
            fin = fmin(fin, red[2][w]);
        }
        part[3 * blockIdx.x + 0] = mi;
        part[3 * blockIdx.x + 1] = mr;
        part[3 * blockIdx.x + 2] = fin;
    }
}

int dm_expect_groups_dev(Ctx *c, const Obs *o, const c128 *rho, c128 *dev_out, c128 *part) {
    unsigned nb = blocks_for(u64(1) << o->n, 1, 64);
    hipLaunchKernelGGL(k_dm_expect_groups, dim3(nb, o->G), dim3(kB), 0, c->stream, rho, o->n, o->d_groups,
                       o->d_mats, part, int(nb));
    MOSSQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_rows, dim3(o->G), dim3(kB), 0, c->stream, part, int(nb), int(nb), dev_out);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_obs_create(mossq_ctx *ctx, int n, int ngroups, const int32_t *group_k, const int32_t *sites,
                     const double *mats, mossq_obs **out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(out && n >= 1 && n <= 40 && ngroups >= 1 && ngroups <= 4096 && group_k && mats, MOSSQ_EINVAL,
                "bad observable arguments");
    Obs *o = new Obs();
    o->ctx = C(ctx);
    o->n = n;
    o->G = ngroups;
    int so = 0, mo = 0;
    const c128 *M = (const c128 *)mats;
    for (int g = 0; g < ngroups; ++g) {
        ObsGroup gr{};
        gr.k = group_k[g];
        if (gr.k < 0 || gr.k > kMaxObsSites || gr.k > n) {
            delete o;
            return fail(MOSSQ_EUNSUPPORTED, "observable group size must be 0..5 and <= n");
        }
        for (int j = 0; j < gr.k; ++j) {
            gr.sites[j] = sites[so + j];
            bool ok = gr.sites[j] >= 0 && gr.sites[j] < n && (j == 0 || gr.sites[j] > gr.sites[j - 1]);
            if (!ok) {
                delete o;
                return fail(MOSSQ_EINVAL, "observable sites must be increasing and < n");
            }
        }
        gr.rshift = gr.k == 0 ? 0 : gr.sites[0];
        for (int j = 1; j < gr.k; ++j)
            if (gr.sites[j] != gr.sites[0] + j) gr.rshift = -1;
        so += gr.k;
        int D = 1 << gr.k;
        gr.moff = mo;
        for (int r = 0; r < D; ++r)
            for (int x = 0; x < D; ++x) {
                const c128 v = M[mo + r + D * (r ^ x)];
                if (v.re != 0.0 || v.im != 0.0) gr.xmask |= 1u << x;
            }
        for (int x = 1; x < D; ++x)
            if ((gr.xmask >> x) & 1u) {
                if (gr.nf >= kObsFlips) {
                    gr.nf = -1;
                    break;
                }
                gr.fx[gr.nf] = x;
                gr.fm[gr.nf] = group_flip(gr, x);
                ++gr.nf;
            }
        for (int i = 0; i < D * D; ++i) o->h_mats.push_back(M[mo + i]);
        // Adjoint seed matrix M + M^dagger, column-major: the seed of a real
        // objective is (conj(z) H + z H^dagger) psi at z = 1
        // (reference differentiation/expectation.jl:32-35).
        for (int cc = 0; cc < D; ++cc)
            for (int r = 0; r < D; ++r) {
                o->h_seed.push_back(cadd(M[mo + r + D * cc], cconj(M[mo + cc + D * r])));
                o->h_adj.push_back(cconj(M[mo + cc + D * r]));
            }
        mo += D * D;
        o->h_groups.push_back(gr);
    }
    o->seed_twice = true;
    for (size_t j = 0; j < o->h_seed.size(); ++j)
        if (o->h_seed[j].re != 2 * o->h_mats[j].re || o->h_seed[j].im != 2 * o->h_mats[j].im) o->seed_twice = false;
    Ctx *c = o->ctx;
    int st = cached_alloc(c, sizeof(ObsGroup) * ngroups, (void **)&o->d_groups);
    if (st == MOSSQ_OK) st = cached_alloc(c, sizeof(c128) * o->h_mats.size(), (void **)&o->d_mats);
    if (st == MOSSQ_OK) st = cached_alloc(c, sizeof(c128) * o->h_seed.size(), (void **)&o->d_mats_seed);
    if (st == MOSSQ_OK) st = cached_alloc(c, sizeof(c128) * o->h_adj.size(), (void **)&o->d_mats_adj);
    if (st != MOSSQ_OK) {
        mossq_obs_destroy(reinterpret_cast<mossq_obs *>(o));
        return st;
    }
    MOSSQ_HIP(hipMemcpyAsync(o->d_groups, o->h_groups.data(), sizeof(ObsGroup) * ngroups, hipMemcpyHostToDevice,
                             c->stream));
    MOSSQ_HIP(hipMemcpyAsync(o->d_mats, o->h_mats.data(), sizeof(c128) * o->h_mats.size(), hipMemcpyHostToDevice,
                             c->stream));
    MOSSQ_HIP(hipMemcpyAsync(o->d_mats_seed, o->h_seed.data(), sizeof(c128) * o->h_seed.size(),
                             hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipMemcpyAsync(o->d_mats_adj, o->h_adj.data(), sizeof(c128) * o->h_adj.size(),
                             hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    *out = reinterpret_cast<mossq_obs *>(o);
    return MOSSQ_OK;
}

int mossq_obs_destroy(mossq_obs *obs) {
    if (!obs) return MOSSQ_OK;
    Obs *o = reinterpret_cast<Obs *>(obs);
    Ctx *c = o->ctx;
    (void)hipSetDevice(c->device);
    int st = MOSSQ_OK;
    if (o->d_groups) st |= cached_free(c, o->d_groups);
    if (o->d_mats) st |= cached_free(c, o->d_mats);
    if (o->d_mats_seed) st |= cached_free(c, o->d_mats_seed);
    if (o->d_mats_adj) st |= cached_free(c, o->d_mats_adj);
    delete o;
    return st ? MOSSQ_EDEVICE : MOSSQ_OK;
}

int mossq_sv_expectation(mossq_ctx *ctx, const mossq_obs *obs, const void *psi, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && psi && out, MOSSQ_EINVAL, "null argument");
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    return expect_total_host(C(ctx), o, (const c128 *)psi, (const c128 *)psi, out);
}

int mossq_sv_cross_expectation(mossq_ctx *ctx, const mossq_obs *obs, const void *phi, const void *psi,
                               double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && phi && psi && out, MOSSQ_EINVAL, "null argument");
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    return expect_total_host(C(ctx), o, (const c128 *)phi, (const c128 *)psi, out);
}

int mossq_sv_apply_operator(mossq_ctx *ctx, const mossq_obs *obs, const void *psi, void *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && psi && out && psi != out, MOSSQ_EINVAL, "bad buffers (output must not alias input)");
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    return apply_op_dev(C(ctx), o, (const c128 *)psi, (c128 *)out, o->n, OpMats::Plain);
}

int mossq_dm_apply_operator(mossq_ctx *ctx, const mossq_obs *obs, const void *rho, void *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && rho && out && rho != out, MOSSQ_EINVAL, "bad buffers (output must not alias input)");
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    MOSSQ_CHECK(o->n <= 20, MOSSQ_EINVAL, "density matrix too large");
    return apply_op_dev(C(ctx), o, (const c128 *)rho, (c128 *)out, 2 * o->n, OpMats::Plain);
}

int mossq_dm_expectation(mossq_ctx *ctx, const mossq_obs *obs, const void *rho, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && rho && out, MOSSQ_EINVAL, "null argument");
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    MOSSQ_CHECK(o->n <= 20, MOSSQ_EINVAL, "density matrix too large");
    Ctx *c = C(ctx);
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (u64(o->G) * 64 + o->G) * sizeof(c128), &scr));
    c128 *part = (c128 *)scr, *res = part + size_t(o->G) * 64;
    MOSSQ_TRY(dm_expect_groups_dev(c, o, (const c128 *)rho, res, part));
    std::vector<c128> h(o->G);
    MOSSQ_HIP(hipMemcpyAsync(h.data(), res, sizeof(c128) * o->G, hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    double re = 0, im = 0;
    for (auto &v : h) {
        re += v.re;
        im += v.im;
    }
    out[0] = re;
    out[1] = im;
    return MOSSQ_OK;
}

int mossq_sv_expectation_batch(mossq_ctx *ctx, const mossq_obs *obs, const void *psi, int B, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && (psi || B == 0) && (out || B == 0) && B >= 0 && B <= 65535, MOSSQ_EINVAL,
                "bad batch expectation arguments");
    if (B == 0) return MOSSQ_OK;
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    Ctx *c = C(ctx);
    const u64 len = u64(1) << o->n;
    // partial sums: nb per column, the batch's total bounded by 4 kObsParts
    const unsigned nb = std::max(1u, std::min(blocks_for(len, 4, kObsParts), unsigned(4 * kObsParts / B)));
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (u64(B) * nb + B) * sizeof(c128), &scr));
    c128 *part = (c128 *)scr, *res = part + size_t(B) * nb;
    const c128 *v = (const c128 *)psi;
    hipLaunchKernelGGL(k_observe<0>, dim3(nb, B), dim3(kB), 0, c->stream, v, v, len, o->G, o->d_groups,
                       o->d_mats, o->d_mats, nullptr, part);
    MOSSQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_rows, dim3(B), dim3(kB), 0, c->stream, part, int(nb), int(nb), res);
    MOSSQ_HIP(hipGetLastError());
    std::vector<c128> h(B);
    MOSSQ_HIP(hipMemcpyAsync(h.data(), res, sizeof(c128) * B, hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; ++b) {
        out[2 * b] = h[b].re;
        out[2 * b + 1] = h[b].im;
    }
    return MOSSQ_OK;
}

int mossq_dm_expectation_batch(mossq_ctx *ctx, const mossq_obs *obs, const void *rho, int B, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(obs && (rho || B == 0) && (out || B == 0) && B >= 0 && B <= 65535, MOSSQ_EINVAL,
                "bad batch expectation arguments");
    if (B == 0) return MOSSQ_OK;
    const Obs *o = reinterpret_cast<const Obs *>(obs);
    MOSSQ_CHECK(o->n <= 20, MOSSQ_EINVAL, "density matrix too large");
    Ctx *c = C(ctx);
    const unsigned nb = blocks_for(u64(1) << o->n, 1, 64);
    const u64 rows = u64(B) * o->G;
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (rows * nb + rows) * sizeof(c128), &scr));
    c128 *part = (c128 *)scr, *res = part + rows * nb;
    hipLaunchKernelGGL(k_dm_expect_groups, dim3(nb, o->G, B), dim3(kB), 0, c->stream, (const c128 *)rho, o->n,
                       o->d_groups, o->d_mats, part, int(nb));
    MOSSQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_rows, dim3(unsigned(rows)), dim3(kB), 0, c->stream, part, int(nb), int(nb), res);
    MOSSQ_HIP(hipGetLastError());
    std::vector<c128> h(rows);
    MOSSQ_HIP(hipMemcpyAsync(h.data(), res, sizeof(c128) * rows, hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    for (int b = 0; b < B; ++b) {
        double re = 0, im = 0;
        for (int g = 0; g < o->G; ++g) {
            re += h[size_t(b) * o->G + g].re;
            im += h[size_t(b) * o->G + g].im;
        }
        out[2 * b] = re;
        out[2 * b + 1] = im;
    }
    return MOSSQ_OK;
}

int mossq_dot(mossq_ctx *ctx, const void *a, const void *b, uint64_t len, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(a && b && out, MOSSQ_EINVAL, "null argument");
    Ctx *c = C(ctx);
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (kMaxParts + 1) * sizeof(c128), &scr));
    c128 *part = (c128 *)scr;
    unsigned nb = blocks_for(len, 4, kMaxParts);
    hipLaunchKernelGGL(k_dot, dim3(nb), dim3(kB), 0, c->stream, (const c128 *)a, (const c128 *)b, len, u64(1),
                       part);
    MOSSQ_HIP(hipGetLastError());
    return finish_scalar(c, part, int(nb), out);
}

int mossq_dm_trace(mossq_ctx *ctx, const void *rho, int n, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(rho && out && n >= 0 && n <= 20, MOSSQ_EINVAL, "bad argument");
    Ctx *c = C(ctx);
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (kMaxParts + 1) * sizeof(c128), &scr));
    c128 *part = (c128 *)scr;
    u64 N = u64(1) << n;
    unsigned nb = blocks_for(N, 1, kMaxParts);
    hipLaunchKernelGGL(k_trace, dim3(nb), dim3(kB), 0, c->stream, (const c128 *)rho, N, part);
    MOSSQ_HIP(hipGetLastError());
    return finish_scalar(c, part, int(nb), out);
}

int mossq_probabilities(mossq_ctx *ctx, const void *state, int n, int is_density, int k, const int32_t *qubits,
                        double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(state && out && qubits && n >= 1 && n <= 40 && k >= 1 && k <= n && k <= 24, MOSSQ_EINVAL,
                "bad probabilities arguments");
    MOSSQ_CHECK(!is_density || n <= 20, MOSSQ_EINVAL, "density matrix too large");
    ObsSites q{}, rest{};
    bool used[64] = {false};
    for (int j = 0; j < k; ++j) {
        MOSSQ_CHECK(qubits[j] >= 0 && qubits[j] < n && !used[qubits[j]], MOSSQ_EINVAL,
                    "readout qubits must be distinct and in range");
        used[qubits[j]] = true;
        q.s[j] = qubits[j];
    }
    int r = 0;
    for (int b = 0; b < n; ++b)
        if (!used[b]) rest.s[r++] = b;
    Ctx *c = C(ctx);
    u64 nbins = u64(1) << k, nrest = u64(1) << (n - k);
    // enough (bin, chunk) work items to fill the device, chunks of equal length
    u64 nchunks = 1;
    while (nbins * nchunks < (u64(1) << 16) && nchunks < nrest) nchunks <<= 1;
    u64 chunk_len = nrest / nchunks;
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, (nbins * nchunks + nbins) * sizeof(double), &scr));
    double *part = (double *)scr, *res = part + nbins * nchunks;
    hipLaunchKernelGGL(k_marginal, dim3(blocks_for(nbins * nchunks, 1, 1u << 16)), dim3(kB), 0, c->stream,
                       (const c128 *)state, n, is_density, k, q, rest, nchunks, chunk_len, part);
    MOSSQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_marginal_reduce, dim3(blocks_for(nbins, 1, 4096)), dim3(kB), 0, c->stream, part, nbins,
                       nchunks, res);
    MOSSQ_HIP(hipGetLastError());
    MOSSQ_HIP(hipMemcpyAsync(out, res, nbins * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    return MOSSQ_OK;
}

int mossq_dm_diag_bounds(mossq_ctx *ctx, const void *rho, int n, double *out) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(rho && out && n >= 1 && n <= 20, MOSSQ_EINVAL, "bad diagonal-bounds arguments");
    Ctx *c = C(ctx);
    const unsigned nb = blocks_for(u64(1) << n, 1, 1024);
    void *scr = nullptr;
    MOSSQ_TRY(get_scratch(c, u64(3) * nb * sizeof(double), &scr));
    hipLaunchKernelGGL(k_diag_bounds, dim3(nb), dim3(kB), 0, c->stream, (const c128 *)rho, n, (double *)scr);
    MOSSQ_HIP(hipGetLastError());
    std::vector<double> h(size_t(3) * nb);
    MOSSQ_HIP(hipMemcpyAsync(h.data(), scr, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    out[0] = 0;
    out[1] = INFINITY;
    out[2] = 1;
    for (unsigned b = 0; b < nb; ++b) {
        out[0] = std::max(out[0], h[3 * b]);
        out[1] = std::min(out[1], h[3 * b + 1]);
        out[2] = std::min(out[2], h[3 * b + 2]);
    }
    return MOSSQ_OK;
}

}  // extern "C"
