// Lowered circuits: launch plans over a device matrix pool.
#pragma once

#include <memory>

#include "mossq_obs.h"

namespace mossq {

// Classifications of a split step's column-pass matrices (kept out of Step,
// whose other instances do not need them).
struct SplitInfo {
    MatInfo fwd, rev;
};

// One unfused instruction, expressed on the bits of the stored vector
// (state vector: the qubit bits; density matrix: the 2n bits of vec(rho)).
struct Step {
    bool channel = false;
    int k = 0;        // register qubits of the instruction
    int kb = 0;       // bits of the stored vector touched (k or 2k)
    int bits[10] = {};
    int fwd = -1;     // forward matrix offset in the device pool (2^kb x 2^kb)
    MatInfo ifwd;
    int rev = -1;     // reverse matrix: U^dagger (SV), its superoperator (DM), S^dagger (channel)
    MatInfo irev;
    int grad = -1;    // gradient matrix: D (SV) or D^dagger(x)I + I(x)D^T (DM), 2^kb x 2^kb
    int grad_index = -1;
    int u_src = -1;   // density 2-qubit gates: offset of U itself (for the fused row/column split)
    // density gates wider than 2 qubits: applied as row pass then column pass
    bool split = false;
    int fwd_cols = -1, rev_cols = -1;
    std::shared_ptr<const SplitInfo> cols;  // split steps: the column passes' classifications
};

struct FusedPlan;           // LDS-tile sweep plan (tile.hip)
struct DensityAdjointPlan;  // chunk plans of the density-matrix adjoint (tile.hip)

struct Program {
    Ctx *ctx = nullptr;
    int n = 0;
    int is_density = 0;
    int nbits = 0;
    bool has_channel = false;
    int nderiv = 0;
    std::vector<Step> steps;
    std::vector<c128> h_pool;
    c128 *d_pool = nullptr;
    FusedPlan *fused = nullptr;
    FusedPlan *fused_rev = nullptr;  // pure-state adjoint reverse pass (with gradient probes)
    DensityAdjointPlan *dm_adj = nullptr;
    // Graphs captured from this program hold its plan tables' device
    // addresses: while any is alive, a plan the autotuner replaces is kept in
    // `retired` (released when the last graph is destroyed, or with the
    // program).
    mutable int live_graphs = 0;
    // of which program graphs, which may launch unfused steps with their
    // monomial entries as kernel arguments (not rewritten by an update)
    mutable int live_step_graphs = 0;
    std::vector<FusedPlan *> retired;
    // A fused pure-state program of fewer than 12 qubits also carries itself
    // as a 12-qubit program (the same steps; the extra qubits stay |0>) with
    // fused plans, so the batched engine runs it as tile sweeps with the
    // others.  Built on the first batched call (fused_pad); it shares d_pool
    // with this program.
    bool fuse_requested = false;
    mutable bool pad_tried = false;
    mutable Program *pad = nullptr;
};

struct BatchRun;  // a prepared batch of the batched engine (batch.hip)
void batch_run_free(BatchRun *r);

}  // namespace mossq

// A captured graph: one program run (prog) or one batched-engine call (batch,
// over the programs in progs), with what it keeps alive.
struct mossq_graph {
    mossq::Ctx *ctx = nullptr;
    const mossq::Program *prog = nullptr;  // its plan stays alive while the graph is (Program::retired)
    std::vector<const mossq::Program *> progs;
    mossq::BatchRun *batch = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace mossq {

// Host side of mossq_program_create: validate records, build steps and pool.
int program_build_host(Program *p, int n, int is_density, int nops, const mossq_op *ops, const double *pool,
                       uint64_t pool_len);

// Unfused forward application of steps [lo, hi).
int run_steps(Ctx *c, const Program *p, c128 *state, int lo, int hi);
// Step i on v alone with its forward or reverse matrices (program.hip).
int apply_step(Ctx *c, const Program *p, int i, c128 *v, bool reverse);

// Fused plans (tile.hip).
int fused_plan_build(Program *p);
void fused_plan_free(Program *p);
int fused_plan_sweeps(const Program *p);
int fused_run(Ctx *c, const Program *p, c128 *state);
// Time every forward-plan variant on a scratch state (reps runs each) and keep
// the fastest as p->fused; tried = distinct variants, best_ms = its run time.
int fused_plan_autotune(Program *p, int reps, int *tried, double *best_ms);
// New matrix values (q: the program rebuilt from the new records), in three
// steps.  prepare (host work only): every plan of p rebuilt with the variant
// it came from, which must keep its tables (MOSSQ_EINVAL, p unchanged,
// otherwise).  copies: the new values as host-to-device copies onto the old
// plans' device arrays.  commit: the new values into the old plans' host
// tables (the copies' sources move with them and stay valid).
struct PlanUpdate {
    std::vector<std::pair<FusedPlan *, FusedPlan *>> pairs;  // (old, rebuilt)
    DensityAdjointPlan *ndp = nullptr;                       // rebuilt density adjoint plans
    bool dm_paired = false;                                  // ndp's plans are in pairs
    PlanUpdate() = default;
    PlanUpdate(const PlanUpdate &) = delete;
    PlanUpdate &operator=(const PlanUpdate &) = delete;
    ~PlanUpdate();  // frees the rebuilt plans
};
struct DevCopy {
    void *dst;
    const void *src;
    u64 bytes;
};
int fused_plan_update_prepare(const Program *p, const Program *q, PlanUpdate *u);
void fused_plan_update_copies(const PlanUpdate &u, std::vector<DevCopy> &out);
void fused_plan_update_commit(PlanUpdate *u);
// Whether the forward run of p launches step i on its own (an unfused item of
// the fused plan, or every step of an unfused program), passing a monomial
// step's entries as kernel arguments.
bool fused_run_launches_step(const Program *p, int i);
// The 12-qubit form of a fused pure-state program below 12 qubits
// (Program::pad), built on first use; null when there is none.
const Program *fused_pad(const Program *p);
// Releases the plans kept for captured graphs (call with no graph alive).
void fused_plan_release_retired(Program *p);
// Reverse pass of the pure-state adjoint on (y, l) with the adjoint plan;
// gradient partials into rows of `row` entries (zeroed by the caller).  y and
// l are left undefined (the last sweep does not store them).
int fused_adjoint_run(Ctx *c, const Program *p, c128 *y, c128 *l, c128 *part, int row);
// Partials per gradient row the adjoint plan writes (0 without one).
int fused_adjoint_parts(const Program *p);
// One unfused reverse step of the pure-state adjoint (program.hip).
int adjoint_rev_step(Ctx *c, const Program *p, int i, c128 *y, c128 *l, c128 *part, int row);
// Gradient partials per row for p's pure-state adjoint.
int adjoint_part_row(const Program *p);
// Fused density-matrix adjoint (tile.hip): chunk plans cut at the
// parameterized steps, built by fused_plan_build for density programs with
// parameters and at least 12 bits.  layout gives the checkpoint segment
// length seg and count nseg (0 without a plan); forward runs the chunks on
// work (the input on entry, the final state on exit) keeping nseg
// checkpoints; reverse takes Delta (the seed at the end) down to the first
// parameterized step with seg - 1 recompute buffers, writing gradient
// partials into rows of `row` entries.
int density_adjoint_plan_build(Program *p);
void density_adjoint_plan_free(Program *p);
void density_adjoint_layout(const Program *p, int *seg, int *nseg);
int density_adjoint_forward(Ctx *c, const Program *p, c128 *work, c128 *const *ckpt);
// The probe table lists the states the reverse plans' fixed probes read
// (nseg x seg entries); reverse takes it in device memory.
void density_adjoint_probe_table(const Program *p, c128 *const *ckpt, c128 *const *stored,
                                 std::vector<const c128 *> &tab);
int density_adjoint_reverse(Ctx *c, const Program *p, c128 *delta, c128 *const *ckpt, c128 *const *stored,
                            const c128 *const *d_tab, c128 *part, int row);
// Gradient partials per row the fused density adjoint writes (zero them first).
int density_adjoint_parts(const Program *p);
// Batched fused sweeps (tile.hip): the plans of many programs advanced
// together, launch s running sweep s of every program that has one, each
// program on its own vectors.  A program qualifies when every item of its
// forward plan (and, for adjoint sweeps, of its adjoint plan) is a sweep.
bool fused_plan_batchable(const Program *p, bool adjoint);
constexpr int kBatchLaunchItems = 65536;  // items per batched launch (16-bit item field)
// Workgroup table of one batched launch (items with nt[i] tiles, round robin).
int deal_tiles(const std::vector<u64> &nt, std::vector<unsigned> &wg);
struct BatchSweeps {
    void *d_items = nullptr;  // device launch items (layout private to tile.hip)
    unsigned *d_wg = nullptr; // workgroup -> item of its launch << 16 | tile
    c128 *vec = nullptr, *part = nullptr;
    std::vector<int> item0, wg0, tiles, rb;  // per launch (rb: register bits of its plans)
    std::vector<char> h_items;                // host tables of an upload not waited for
    std::vector<unsigned> h_wg;
};
// Builds and uploads the launch tables into al's memory (synchronises c's
// stream; with sync false the copies are only enqueued there, and the host
// tables stay in *out until it is destroyed).  Every vector lies in the allocation vec (vec_len entries), given
// by its offset: forward, x[b] is program b's state; adjoint, x[b] and l[b]
// are y and lambda, and the gradient partials go to rows of row[b] entries
// at part + part_off[b] (part holds part_len entries).
int batch_sweeps_build(Ctx *c, Allocs &al, int m, const Program *const *progs, c128 *vec, u64 vec_len,
                       const u64 *x, const u64 *l, c128 *part, u64 part_len, const u64 *part_off, const int *row,
                       bool adjoint, BatchSweeps *out, bool sync = true);
int batch_sweeps_run(Ctx *c, const BatchSweeps &bs, bool adjoint);
// Zero `count` complex entries with a kernel (safe inside captured graphs).
int zero_fill(Ctx *c, c128 *v, u64 count);

// Pure-state adjoint enqueued on the stream, no host synchronisation.
// in == nullptr starts from |0...0>.  Buffers a, b hold 2^n amplitudes; part
// holds max(kObsParts, nderiv * adjoint_part_row) entries.  value_dev receives
// <H> (one complex); grads_dev nderiv complex.
struct AdjointScratch {
    c128 *a = nullptr, *b = nullptr, *part = nullptr, *value_dev = nullptr, *grads_dev = nullptr;
};
int adjoint_pure_enqueue(Ctx *c, const Program *p, const Obs *o, const c128 *in, c128 *out_state,
                         const AdjointScratch &s);
// Pure-state adjoint with its own scratch; synchronises c's stream.
int adjoint_pure(Ctx *c, const Program *p, const Obs *o, const c128 *in, c128 *out_state, double *value,
                 double *grads);
// Density-matrix adjoint (segment checkpoints + recompute); synchronises c's stream.
int adjoint_density(Ctx *c, const Program *p, const Obs *o, const c128 *in, int in_is_density, c128 *out_state,
                    double *value, double *grads);
// Device bytes adjoint_density allocates for p (states, partial sums).
u64 adjoint_density_bytes(const Program *p, const Obs *o);
// Upload a host-built program's pool to c's device memory (synchronises c's stream).
int program_upload(Ctx *c, Program *p);

}  // namespace mossq
