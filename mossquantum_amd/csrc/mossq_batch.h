// Batched adjoint engine: many independent pure-state programs (uneven qubit
// counts, own parameters, own observables) advanced together.
//
// Every item runs the same micro-program: nfwd forward gates, one observe step
// (energy partials and the adjoint seed (H + H^dagger) psi), then nfwd reverse
// steps (U^dagger on psi and lambda, gradient partials).  Item b started at
// iteration t0_b is at micro-op pc = t - t0_b of iteration t, so one iteration
// is two launches for all active items -- k_biter (the work) and k_breduce
// (deterministic per-item reductions of the partial sums) -- with no host
// round trip; new items join at any iteration (continuous batching).
#pragma once

#include "mossq_program.h"

namespace mossq {

constexpr int kBatchPairsPerChunk = 2048;  // one workgroup of 256 threads x 8 pairs

struct BOp {
    int kind;  // 1: dense 1-bit, 2: dense 2-bit
    int b0, b1;
    int fwd, rev, grad;  // pool offsets; grad = -1 when the gate has no parameter
    int gslot;           // gradient slot relative to the item's base, or -1
    int pad;
};

struct BItemDev {
    c128 *psi;
    c128 *lam;
    const BOp *ops;
    const c128 *pool;
    const ObsGroup *groups;
    const c128 *mats;  // M_g
    const c128 *seed;  // M_g + M_g^dagger
    int n, nfwd, G, t0;
    int chunk0, nchunk;
    int gslot0, vslot;
    int value_only;  // stop after the observe step (no seed, no reverse pass)
    int preset;      // psi already holds the initial state (k_binit leaves it)
    int seed_twice;  // the observable's seed matrices are 2 M_g: seed = 2 (M psi)
};

// Host-side description of one item.
struct BItemHost {
    const Program *prog = nullptr;
    const Obs *obs = nullptr;
    std::vector<BOp> ops;
    BOp *d_ops = nullptr;
    c128 *psi = nullptr, *lam = nullptr;
    int t0 = 0;
    int nfwd = 0;
    int nchunk = 1;
    int gslot0 = 0, vslot = 0;
};

// Lower a pure-state program to engine ops (1- and 2-bit gates only).
int batch_lower(const Program *p, std::vector<BOp> &ops, int *nderiv);
int batch_chunks(int n);

// Enqueue one iteration t for items (device table `items`, `nitems` entries,
// chunk prefix `chunk0` in the table, total chunks `nchunks_total`).
int batch_iteration(Ctx *c, const BItemDev *items, int nitems, int nchunks_total, int t, c128 *part,
                    double *values, double *grads);
// The same on stream st.
int batch_iteration_on(hipStream_t st, const BItemDev *items, int nitems, int nchunks_total, int t, c128 *part,
                       double *values, double *grads);

}  // namespace mossq
