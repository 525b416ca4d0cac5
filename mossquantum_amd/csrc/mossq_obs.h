// Device form of an observable (QubitsOperator grouped by site tuple).
#pragma once

#include "mossq_internal.h"

namespace mossq {

constexpr int kMaxObsSites = 5;  // reference LARGEST_SUPPORTED_NTERMS, backends/cpu/operators/generic.jl:1

struct ObsGroup {
    int k;                      // number of sites (0..5)
    int sites[kMaxObsSites];    // increasing 0-based sites
    int moff;                   // offset of the 2^k x 2^k column-major matrix
    unsigned xmask;             // bit x set: M[r, r ^ x] != 0 for some local row r
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
    int r = 0;
#pragma unroll
    for (int j = 0; j < kMaxObsSites; ++j)
        if (j < g.k) r |= int((i >> g.sites[j]) & 1) << j;
    return r;
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
