// Device form of an observable (QubitsOperator grouped by site tuple).
#pragma once

#include "mossq_internal.h"

namespace mossq {

constexpr int kMaxObsSites = 5;  // reference LARGEST_SUPPORTED_NTERMS, backends/cpu/operators/generic.jl:1
constexpr int kObsFlips = 4;     // nonzero flip patterns listed per group
// A global-memory c128 vector: its loads are global (not flat) loads, which
// wait on the memory counter alone.
typedef const __attribute__((address_space(1))) c128 *gptr_c128;

struct ObsGroup {
    int k;                      // number of sites (0..5)
    int sites[kMaxObsSites];    // increasing 0-based sites
    int moff;                   // offset of the 2^k x 2^k column-major matrix
    unsigned xmask;             // bit x set: M[r, r ^ x] != 0 for some local row r
    // The nonzero flip patterns x != 0 of xmask with their global flip masks
    // (a Pauli term flips one pattern: an XX+YY pair or an X site has one),
    // or nf = -1 when there are more than kObsFlips.
    int nf;
    int fx[kObsFlips];
    // Sites s0, s0 + 1, ..., s0 + k - 1 (a chain's neighbour pairs): rshift =
    // s0 and the local index is one shift and mask; -1 for other site sets.
    int rshift;
    u64 fm[kObsFlips];
};

// Row-wise action of a group sum, out_i = sum_g sum_x A_g[r, r ^ x] psi[i ^ flip_g(x)]
// with r the local index of i on g's sites: only the flip patterns x in
// xmask load, and every group's x = 0 (diagonal) term shares the one load of
// psi[i].
// (Loops over the site slots with constant indices: a dynamic index into the
// record would put it in scratch memory.)
__host__ __device__ __forceinline__ u64 group_flip(const ObsGroup &g, int x) {
    u64 o = 0;
#pragma unroll
    for (int j = 0; j < kMaxObsSites; ++j)
        if (j < g.k && (x & (1 << j))) o |= u64(1) << g.sites[j];
    return o;
}

__host__ __device__ __forceinline__ int group_row(const ObsGroup &g, u64 i) {
    if (g.rshift >= 0) return int(i >> g.rshift) & ((1 << g.k) - 1);
    int r = 0;
#pragma unroll
    for (int j = 0; j < kMaxObsSites; ++j)
        if (j < g.k) r |= int((i >> g.sites[j]) & 1) << j;
    return r;
}

// Row i of the group sum: dh += the diagonal entries M_g[r, r] (their term
// shares psi_i), hp += sum over the nonzero flip patterns x of
// M_g[r, r ^ x] psi[i ^ flip_g(x)]; with TWO the same with the matrices W
// into dw, wp.  Group records are read uniformly (one per group for the
// whole wave), matrices at the row's local index from M / W.
template <bool TWO>
__device__ __forceinline__ void obs_row(u64 i, const c128 *__restrict__ psi, int G, const ObsGroup *groups,
                                        const c128 *M, const c128 *W, c128 &hp, c128 &dh, c128 &wp, c128 &dw) {
    const gptr_c128 gpsi = (gptr_c128)psi;
    for (int gi = 0; gi < G; ++gi) {
        const ObsGroup g = groups[gi];
        const int D = 1 << g.k;
        const int r = group_row(g, i);
        const c128 *m = M + g.moff + r;
        const c128 *w = W + g.moff + r;
        if (g.xmask & 1u) {
            dh = cadd(dh, m[D * r]);
            if (TWO) dw = cadd(dw, w[D * r]);
        }
        // A row whose entry is zero skips its partner load (the sum is
        // unchanged): XX + YY pairs have nonzero entries on half the rows, and
        // for sites at or above the wave's 64 rows whole waves skip.
        if (g.nf >= 0) {
#pragma unroll
            for (int q = 0; q < kObsFlips; ++q)
                if (q < g.nf) {
                    const c128 mq = m[D * (r ^ g.fx[q])];
                    const c128 wq = TWO ? w[D * (r ^ g.fx[q])] : mq;
                    if (mq.re != 0.0 || mq.im != 0.0 || (TWO && (wq.re != 0.0 || wq.im != 0.0))) {
                        const u64 pi = i ^ g.fm[q];
                        const c128 y = mk(gpsi[pi].re, gpsi[pi].im);
                        hp = cfma(mq, y, hp);
                        if (TWO) wp = cfma(wq, y, wp);
                    }
                }
        } else {
            for (int x = 1; x < D; ++x)
                if ((g.xmask >> x) & 1u) {
                    const c128 mq = m[D * (r ^ x)];
                    const c128 wq = TWO ? w[D * (r ^ x)] : mq;
                    if (mq.re != 0.0 || mq.im != 0.0 || (TWO && (wq.re != 0.0 || wq.im != 0.0))) {
                        const u64 pi = i ^ group_flip(g, x);
                        const c128 y = mk(gpsi[pi].re, gpsi[pi].im);
                        hp = cfma(mq, y, hp);
                        if (TWO) wp = cfma(wq, y, wp);
                    }
                }
        }
    }
}

// Rows ia and ib of the group sum (one sum, matrices M), interleaved so the
// two rows' partner loads are in flight together.  A flip pattern is skipped
// when no lane of the wave has a nonzero entry for it in either row (for
// sites at or above the wave's 64 rows whole waves skip); otherwise every
// lane loads its partners and adds entry x partner (a zero entry adds zero).
__device__ __forceinline__ void obs_pair_term(const gptr_c128 gpsi, u64 ia, u64 ib, u64 fm, c128 qa, c128 qb,
                                              c128 &hpa, c128 &hpb) {
    const bool nz = qa.re != 0.0 || qa.im != 0.0 || qb.re != 0.0 || qb.im != 0.0;
    if (__any(nz)) {
        const u64 pa = ia ^ fm, pb = ib ^ fm;
        const c128 ya = mk(gpsi[pa].re, gpsi[pa].im), yb = mk(gpsi[pb].re, gpsi[pb].im);
        hpa = cfma(qa, ya, hpa);
        hpb = cfma(qb, yb, hpb);
    }
}

__device__ __forceinline__ void obs_row_pair(u64 ia, u64 ib, const c128 *__restrict__ psi, int G,
                                             const ObsGroup *groups, const c128 *M, c128 &hpa, c128 &dha,
                                             c128 &hpb, c128 &dhb) {
    const gptr_c128 gpsi = (gptr_c128)psi;
    for (int gi = 0; gi < G; ++gi) {
        const ObsGroup g = groups[gi];
        const int D = 1 << g.k;
        const int ra = group_row(g, ia), rb = group_row(g, ib);
        const c128 *ma = M + g.moff + ra, *mb = M + g.moff + rb;
        if (g.xmask & 1u) {
            dha = cadd(dha, ma[D * ra]);
            dhb = cadd(dhb, mb[D * rb]);
        }
        if (g.nf >= 0) {
#pragma unroll
            for (int q = 0; q < kObsFlips; ++q)
                if (q < g.nf)
                    obs_pair_term(gpsi, ia, ib, g.fm[q], ma[D * (ra ^ g.fx[q])], mb[D * (rb ^ g.fx[q])], hpa, hpb);
        } else {
            for (int x = 1; x < D; ++x)
                if ((g.xmask >> x) & 1u)
                    obs_pair_term(gpsi, ia, ib, group_flip(g, x), ma[D * (ra ^ x)], mb[D * (rb ^ x)], hpa, hpb);
        }
    }
}

struct ObsSites {
    int s[40];
};

struct Obs {
    Ctx *ctx = nullptr;
    int n = 0;
    int G = 0;
    std::vector<ObsGroup> h_groups;
    std::vector<c128> h_mats;  // M_g
    std::vector<c128> h_seed;  // M_g + M_g^dagger (pure-state adjoint seed)
    std::vector<c128> h_adj;   // M_g^dagger (density-matrix adjoint seed)
    ObsGroup *d_groups = nullptr;
    c128 *d_mats = nullptr;
    c128 *d_mats_seed = nullptr;
    c128 *d_mats_adj = nullptr;
    // Every seed entry is exactly twice its M_g entry (a Hermitian observable):
    // the seed (M + M^dagger) psi is then 2 (M psi), computed from one sum.
    bool seed_twice = false;
};

// Partial sums the one-pass observation kernels write (at most).
constexpr int kObsParts = 2048;
// One pass over psi for every group: dev_out[0] = <phi|H|psi> = sum_i
// conj(phi_i) (H psi)_i, and when seed != nullptr also seed = (H + H^dagger)
// psi (the pure-state adjoint seed); part needs kObsParts entries.
int observe_dev(Ctx *c, const Obs *o, const c128 *phi, const c128 *psi, c128 *seed, c128 *dev_out, c128 *part);
// <phi|H|psi> copied to the host (out[0] re, out[1] im).
int expect_total_host(Ctx *c, const Obs *o, const c128 *phi, const c128 *psi, double *out);
// out = sum_g A_g psi on a 2^nbits vector with A_g = M_g, M_g + M_g^dagger or M_g^dagger.
enum class OpMats { Plain, Seed, Adjoint };
int apply_op_dev(Ctx *c, const Obs *o, const c128 *psi, c128 *out, int nbits, OpMats which);
// Per-group tr(M_g rho) to dev_out[G]; part needs G * 64 entries.
int dm_expect_groups_dev(Ctx *c, const Obs *o, const c128 *rho, c128 *dev_out, c128 *part);

}  // namespace mossq
