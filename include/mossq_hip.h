/*
 * mossq_hip.h -- C ABI of libmossq_hip.so, the MI355X (gfx950) dense
 * simulation backend for MOSSQuantum.
 *
 * The Julia side binds every entry point with `ccall((:name, libmossq_hip), ...)`;
 * the Python host in mossquantum_amd/ binds the same symbols through ctypes.
 *
 * Conventions (all entry points):
 *   - Return an int status: MOSSQ_OK (0) or a nonzero MOSSQ_E* code.  The
 *     message of the last failure on the calling thread is mossq_last_error().
 *   - Complex numbers are complex128, interleaved (re, im), i.e. Julia's
 *     ComplexF64 / numpy complex128 memory layout.
 *   - Qubit positions at this boundary are 0-based bit indices, qubit 1 of the
 *     reference = bit 0 = least significant index bit
 *     (reference docs/conventions.md:7-13).  The Julia side subtracts one.
 *   - Local matrices are column-major 2^k x 2^k in the reference's canonical
 *     "ordered" basis: positions sorted ascending, the first (lowest) position
 *     is the least significant local bit
 *     (reference src/quantum_circuits/gates/abstract_gates.jl:88-121,162-166).
 *   - A density matrix of n qubits is vec(rho) of length 4^n, rho[i,j] at
 *     i + 2^n * j (reference src/states/density_matrix.jl:12-22,74).
 *   - Device buffers are owned by the caller (allocated with mossq_alloc);
 *     the library keeps no host pointer after a call returns.
 *   - Every context owns one in-order HIP stream.  Calls on one context are
 *     ordered; calls on distinct contexts are thread-safe.
 *   - Nothing here falls back to the CPU: with no usable device every entry
 *     point returns MOSSQ_EDEVICE.
 */
#ifndef MOSSQ_HIP_H
#define MOSSQ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOSSQ_OK 0
#define MOSSQ_EINVAL 1   /* invalid argument (reference: ArgumentError / DimensionMismatch) */
#define MOSSQ_EOOM 2     /* device memory or memory budget exhausted */
#define MOSSQ_EDEVICE 3  /* HIP runtime / device failure */
#define MOSSQ_EUNSUPPORTED 4 /* reference: UnsupportedOperation, backend_protocols.jl:16-36 */

typedef struct mossq_ctx mossq_ctx;
typedef struct mossq_program mossq_program;
typedef struct mossq_obs mossq_obs;
typedef struct mossq_graph mossq_graph;

/* Instruction kinds of a program (a lowered QCircuit). */
#define MOSSQ_OP_GATE 0    /* unitary gate: mat = 2^k x 2^k ordered_mat */
#define MOSSQ_OP_CHANNEL 1 /* Kraus channel: mat = 4^k x 4^k ordered_supermat */

/* One instruction of a lowered circuit.  Replaces the per-instruction loop of
 * reference src/execution/gates.jl:21-36 (gates) and
 * src/execution/channels.jl:8-17 (channels). */
typedef struct mossq_op {
    int32_t kind;   /* MOSSQ_OP_GATE or MOSSQ_OP_CHANNEL */
    int32_t k;      /* number of register qubits the instruction acts on (1..5; channels 1..2) */
    int32_t pos[5]; /* 0-based qubit positions, strictly increasing */
    int32_t mat;    /* offset (in complex elements) of the matrix in the pool */
    int32_t dmat;   /* offset of the derivative gate D = (dU/dtheta)^dagger U, or -1
                       (reference src/quantum_circuits/gates/parametric_gates.jl:76-91) */
    int32_t reserved;
} mossq_op;

/* ---- library / context --------------------------------------------------- */
int mossq_version(void);
const char *mossq_last_error(void);
/* Device count visible to HIP (0 when no GPU). */
int mossq_device_count(int *count);
/* Create a context on `device` with a device-memory budget in bytes (0 = no limit).
 * Backs the HIP storage family's executor (reference: backend(state) selection,
 * src/api/backend_protocols.jl:38-70). */
int mossq_ctx_create(int device, uint64_t mem_budget_bytes, mossq_ctx **out);
int mossq_ctx_destroy(mossq_ctx *ctx);
int mossq_sync(mossq_ctx *ctx);
/* Bytes currently allocated through this context, including the blocks the
 * library keeps for reuse (its per-call scratch and table memory is released
 * to a per-context cache rather than to the device; a budget overrun or a
 * failed allocation releases the cache first, and mossq_ctx_destroy frees it). */
int mossq_mem_in_use(mossq_ctx *ctx, uint64_t *bytes);

/* ---- device buffers (back the HIPArray storage family: similar/copyto!/copy,
 *      reference src/states/array_storage.jl:15-64) ------------------------- */
int mossq_alloc(mossq_ctx *ctx, uint64_t bytes, void **dptr);
int mossq_free(mossq_ctx *ctx, void *dptr);
int mossq_memset_zero(mossq_ctx *ctx, void *dptr, uint64_t bytes);
int mossq_copy_h2d(mossq_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int mossq_copy_d2h(mossq_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int mossq_copy_d2d(mossq_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* ---- states ------------------------------------------------------------- */
/* |0...0> (reference StateVector zero state, src/states/state_vector.jl:36-40). */
int mossq_sv_init_zero(mossq_ctx *ctx, void *psi, int n);
/* |0...0><0...0| (reference DensityMatrix(n), src/states/density_matrix.jl:12-22). */
int mossq_dm_init_zero(mossq_ctx *ctx, void *rho, int n);
/* rho = psi psi^dagger (reference DensityMatrix(::StateVector),
 * src/states/density_matrix.jl:65-71; SV->DM promotion of execution/channels.jl:19). */
int mossq_dm_from_sv(mossq_ctx *ctx, void *rho, const void *psi, int n);

/* ---- gates --------------------------------------------------------------- */
/* Dense k-qubit gate on a state vector of n qubits, one read+write sweep
 * (reference _apply_gate_2!/apply_threaded!, src/backends/cpu/gates/
 * serial_short_range.jl:2-71, threaded_short_range.jl, threaded_long_range.jl,
 * generic.jl:20).  Diagonal and permutation matrices (CZ, CNOT, PHASE, X, SWAP,
 * iSWAP, FREDKIN, CCPHASE, Rz ...) are detected and touch only the amplitudes
 * they change (reference src/backends/cpu/gates/specific/ cz_gate.jl, phase_gate.jl, swap_gates.jl, fredkin_gate.jl).  k <= 5. */
int mossq_sv_apply_gate(mossq_ctx *ctx, void *psi, int n, int k, const int32_t *pos,
                        const double *mat);
/* U rho U^dagger in ONE sweep of U (x) conj(U) on bits (pos, pos+n)
 * (reference _dm_apply_threaded!, src/backends/cpu/gates/generic.jl:24-34, which
 * takes two sweeps).  k <= 5. */
int mossq_dm_apply_gate(mossq_ctx *ctx, void *rho, int n, int k, const int32_t *pos,
                        const double *mat);
/* Kraus channel via its 4^k x 4^k supermatrix on bits (pos, pos+n); dagger=1
 * applies S^dagger (reference _qmap_apply_threaded! / _qmap_apply_dagger_threaded!,
 * src/backends/cpu/gates/channels.jl:2-24).  k <= 2. */
int mossq_dm_apply_channel(mossq_ctx *ctx, void *rho, int n, int k, const int32_t *pos,
                           const double *supermat, int dagger);

/* ---- observables (QubitsOperator grouped by site tuple) ------------------ */
/* ngroups groups; group g acts on group_k[g] (<=5) sites listed consecutively in
 * sites (0-based, increasing) with a 2^k x 2^k column-major matrix stored
 * consecutively in mats (reference _get_mat, src/backends/cpu/operators/
 * utils.jl:3-20; QubitsOperator data, quantum_circuits/operators/
 * qubits_operator.jl:1-50). A group with k=0 is c * identity (1x1 matrix). */
int mossq_obs_create(mossq_ctx *ctx, int n, int ngroups, const int32_t *group_k,
                     const int32_t *sites, const double *mats, mossq_obs **out);
int mossq_obs_destroy(mossq_obs *obs);

/* <psi|H|psi> = sum_g sum_blocks v_b^dagger M_g v_b, complex result in out[2]
 * (reference _cpu_expectation(::QubitsOperator, ::StateVector),
 * src/observables/expectation.jl:12-33, one sweep per site group; here one
 * pass over psi for all groups). */
int mossq_sv_expectation(mossq_ctx *ctx, const mossq_obs *obs, const void *psi, double *out);
/* <phi|H|psi> (reference _cpu_cross_expectation, src/observables/expectation.jl:48-66). */
int mossq_sv_cross_expectation(mossq_ctx *ctx, const mossq_obs *obs, const void *phi,
                               const void *psi, double *out);
/* out = H psi (reference apply(::QubitsOperator, ::StateVector),
 * src/execution/operators.jl:3-64, src/backends/cpu/operators/generic.jl:12-28). */
int mossq_sv_apply_operator(mossq_ctx *ctx, const mossq_obs *obs, const void *psi, void *out);
/* <psi_b|H|psi_b> for the B columns of a batch of state vectors stored
 * contiguously ((2^n, B) column-major: column b at psi + b 2^n), one launch
 * over all columns; out[2b], out[2b+1] = Re, Im of column b (reference
 * batch expectation, src/observables/batch_expectation.jl:1-179,
 * backends/cpu/expectations/batch.jl:17-273).  B <= 65535. */
int mossq_sv_expectation_batch(mossq_ctx *ctx, const mossq_obs *obs, const void *psi, int B, double *out);
/* tr(H rho_b) for B density matrices stored contiguously (column b at
 * rho + b 4^n), one launch (reference observables/batch_density_expectation.jl:4-63). */
int mossq_dm_expectation_batch(mossq_ctx *ctx, const mossq_obs *obs, const void *rho, int B, double *out);
/* tr(H rho) (reference _cpu_expectation(::QubitsOperator, ::DensityMatrix),
 * src/observables/density_matrix_expectation.jl:12-25). */
int mossq_dm_expectation(mossq_ctx *ctx, const mossq_obs *obs, const void *rho, double *out);
/* out = H rho, left action on the row bits (reference operators/density_matrix.jl:1-5). */
int mossq_dm_apply_operator(mossq_ctx *ctx, const mossq_obs *obs, const void *rho, void *out);
/* sum_i conj(a_i) b_i over len complex entries (reference LinearAlgebra.dot on storage;
 * purity = Re dot(rho, rho), src/states/density_matrix.jl:106). */
int mossq_dot(mossq_ctx *ctx, const void *a, const void *b, uint64_t len, double *out);
/* tr(rho) (reference src/states/density_matrix.jl:105). */
int mossq_dm_trace(mossq_ctx *ctx, const void *rho, int n, double *out);
/* Marginal readout probabilities of `k` qubits (first listed = least significant
 * output bit), out has 2^k doubles (reference probabilities,
 * src/execution/probabilities.jl:42-87). is_density selects the diagonal of rho. */
int mossq_probabilities(mossq_ctx *ctx, const void *state, int n, int is_density, int k,
                        const int32_t *qubits, double *out);
/* Bounds of the diagonal of a density matrix of n qubits: out[0] = max |Im rho_ii|,
 * out[1] = min Re rho_ii, out[2] = 1 when every diagonal entry is finite, else 0.
 * The host checks them against the reference's tolerance before taking a
 * marginal (src/execution/probabilities.jl:55-64,91-103). */
int mossq_dm_diag_bounds(mossq_ctx *ctx, const void *rho, int n, double *out);

/* ---- programs (lowered circuits) ----------------------------------------- */
/* Lower a circuit once: matrices are uploaded and the launch plan is built.
 * fuse=1 lets the planner merge consecutive instructions into LDS-resident
 * tile sweeps (one HBM read+write per sweep).  pool holds pool_len complex
 * numbers.  Replaces the instruction loop of src/execution/gates.jl:21-36. */
int mossq_program_create(mossq_ctx *ctx, int n, int is_density, int nops, const mossq_op *ops,
                         const double *pool, uint64_t pool_len, int fuse, mossq_program **out);
int mossq_program_destroy(mossq_program *prog);
/* Number of HBM sweeps (kernel launches that stream the whole state) the plan makes. */
int mossq_program_sweeps(const mossq_program *prog, int *sweeps);
/* Apply the program in place (reference apply!(circuit, state),
 * src/api/backend_protocols.jl:246-254). */
int mossq_program_run(mossq_ctx *ctx, const mossq_program *prog, void *state);
/* New matrix values for a program: ops/pool as for mossq_program_create, with
 * the same instructions (kinds, positions, offsets) and new matrices, such as
 * the template's rotations at new angles: the device form of the reference's
 * in-place reset_parameters! (src/quantum_circuits/instructions.jl:62-80) and
 * of bind_parameters (src/api/parameter_binding.jl:29) on a lowered circuit.
 * Every fused plan is rebuilt with the variant it came from and must keep its
 * tables (a value that turns a gate into the identity or a permutation
 * changes them): the new values are then written over the program's device
 * tables in place, so graphs captured from it (mossq_program_capture,
 * mossq_batch_capture) replay the new values.  MOSSQ_EINVAL and the program
 * unchanged when the records describe another program or a plan would change;
 * also when a live program graph launches a monomial step on its own (an
 * unfused step: its entries are kernel arguments) whose entries change, and
 * after an autotune replaced a plan that a live graph still uses.  Updates of
 * different programs of one context may run on several host threads at once
 * (host replanning, then copies on the context's stream). */
int mossq_program_update(mossq_ctx *ctx, mossq_program *prog, int nops, const mossq_op *ops, const double *pool,
                         uint64_t pool_len);
/* mossq_program_update of m distinct programs of one context at once (an
 * optimizer rebinding a batch before the next mossq_batch_graph_run): program
 * i takes nops[i] records ops[i] and pool pools[i] of pool_lens[i] complex
 * entries.
 * The host replanning runs on `threads` host threads (0: one per hardware
 * thread, at most 64); the new values of all programs then go to the device
 * in one host-to-device copy and one scatter launch on the context's stream,
 * and the call returns after them.  A refused program keeps its old values,
 * the others take their new ones.  status (NULL, or m entries) receives each
 * program's status; the call returns the first refused program's (its index
 * in the message).  Calls on one context must not overlap (they share its
 * pinned staging).  No counterpart in the reference (its parameters are
 * rebound per circuit). */
int mossq_program_update_many(mossq_ctx *ctx, int m, mossq_program *const *progs, const int *nops,
                              const mossq_op *const *ops, const double *const *pools, const uint64_t *pool_lens,
                              int threads, int *status);
/* Autotune a fused program: plan every forward-plan variant (tail thresholds
 * 0/3/5/8, with and without pairing one-qubit gates into 4x4 ops, with 4 or 3
 * register bits per layout: 256- or 512-thread tiles), time each
 * distinct one `reps` times on a scratch state of the program's size, and keep
 * the fastest (the planner's cost-model choice is one candidate).  variants =
 * distinct plans timed (0 for an unfused program), best_ms = the kept plan's
 * time per run.  Graphs captured earlier keep the plan they recorded: a
 * replaced plan's device tables stay allocated until the last such graph is
 * destroyed (or the program is). */
int mossq_program_autotune(mossq_ctx *ctx, mossq_program *prog, int reps, int *variants, double *best_ms);
/* Record one run of the program on `state` as a HIP graph (stream capture of
 * the same launches mossq_program_run makes) and instantiate it; each
 * mossq_graph_launch replays the whole run with one launch call on the
 * context's stream.  The program and the state buffer must outlive the graph.
 * This is the bulk-replay form of the reference's simulate_circuit! hook
 * (src/execution/simulation.jl:166-175) for circuits run many times on one
 * buffer; the host launch cost of a run drops from one call per sweep to one. */
int mossq_program_capture(mossq_ctx *ctx, const mossq_program *prog, void *state, mossq_graph **out);
int mossq_graph_launch(mossq_ctx *ctx, mossq_graph *graph);
int mossq_graph_destroy(mossq_graph *graph);
/* Host execution of the fused plan the planner builds for these records, with
 * the tile kernel's index arithmetic; needs no device.  Used to test the
 * planner (the reference's fuse_gates, src/utilities/algorithms/
 * gate_fusion.jl:10-62, is its CPU counterpart).  state (complex, in place) has
 * 2^n (or 4^n) entries; stats[0..9] = sweeps, unfused launches, segments,
 * LDS transpositions, tile ops, of which one-bit dense, two-bit dense,
 * diagonal; transpositions without a bank-conflict-free LDS slot map; groups
 * that run as register permutations (two-bit ops such as CNOT and SWAP).
 * state == NULL returns the statistics only.
 * Registers of fewer than 12 bits are rejected. */
int mossq_plan_emulate(int n, int is_density, int nops, const mossq_op *ops, const double *pool,
                       uint64_t pool_len, double *state, int *stats);
/* Host execution of the pure-state adjoint's fused reverse plan (the steps
 * in reverse with their inverses, each parameterized step followed by its
 * probe Re<y|D|l>, as the reverse pass of the reference's
 * differentiation/circuit.jl:136-144), with the adjoint tile kernel's index
 * arithmetic; needs no device.  y and l (complex, 2^n entries, in place) are
 * the forward output and the seed; grads[j] (nderiv entries) receives the
 * gradient row j.  stats as mossq_plan_emulate for the reverse plan.  Used to
 * test the planner's adjoint plans. */
int mossq_plan_emulate_adjoint(int n, int nops, const mossq_op *ops, const double *pool, uint64_t pool_len,
                               double *y, double *l, double *grads, int *stats);

/* Host execution of the fused density-matrix adjoint (density program of n
 * qubits, 2n >= 12 bits, at least one parameterized gate): the program is cut
 * at its parameterized steps into forward and reverse chunk plans (plus one
 * forward plan per checkpoint segment), run with the device schedule
 * (checkpoints every seg parameters, recompute per segment).  rho (4^n complex) holds the input on entry and the final state on
 * exit; delta holds the seed at the program's end on entry (the reference's
 * z * H^dagger, src/differentiation/expectation.jl:69-72) and the adjoint
 * state at the first parameterized step on exit; grads[j] = derivative j.
 * stats[0..5] = parameterized steps, checkpoint segment length, sweeps of the
 * forward pass (chunk 0 and the segment plans), sweeps of the chunk plans the
 * reverse pass runs (recompute and reverse), unfused launches, plans.  With
 * rho == NULL only the plans are built and counted.  whole_segments selects
 * the reverse plans: 1 one plan per checkpoint segment (where every
 * parameterized gate acts on one qubit), 0 one per parameter, -1 as the
 * device path chooses for this size.  Used to test the chunk
 * plans and the schedule. */
int mossq_plan_emulate_density_adjoint(int n, int nops, const mossq_op *ops, const double *pool,
                                       uint64_t pool_len, int whole_segments, double *rho, double *delta,
                                       double *grads, int *stats);

/* Host only (needs no device): the workgroup tables of the batched tile
 * sweeps (mossq_batch_adjoint, mossq_batch_expectation) for one sweep index
 * over m items of ntiles[i] tiles each.  Items are split into launches of at
 * most 65536; launch_items[L] = items of launch L (max_launches entries), and
 * entries (cap entries) receives every launch's workgroup entries in order,
 * each (item within its launch) << 16 | tile, dealt round robin over the
 * launch's items.  *nlaunches and *nentries are set even when the arrays are
 * too small (then MOSSQ_EINVAL).  No counterpart in the reference; it tests
 * the packing of the launch tables. */
int mossq_batch_launch_tables(int m, const int32_t *ntiles, int max_launches, int32_t *launch_items,
                              uint64_t cap, uint32_t *entries, int *nlaunches, uint64_t *nentries);

/* ---- adjoint gradients ---------------------------------------------------- */
/* value = Re <H> at C|in>, grads[j] for the j-th derivative gate in forward
 * order (reference _circuit_adjoint + back_propagate,
 * src/differentiation/circuit.jl:77-229, driven by value_and_pullback,
 * src/differentiation/objectives.jl:183-218).  For a pure program (no
 * channels) `in` is a state vector of n qubits; with channels `in` is a
 * density matrix (in_is_density=1) or a state vector promoted at the first
 * channel.  The input buffer is not modified.  The final state is written to
 * `out_state` when it is non-NULL (size of the final representation). */
int mossq_adjoint(mossq_ctx *ctx, const mossq_program *prog, const mossq_obs *obs,
                  const void *in, int in_is_density, void *out_state, double *value,
                  double *grads);

/* mossq_adjoint for a state-vector program that allocates nothing: the same
 * kernels, with the forward state, lambda, gradient partials and sums in a
 * 64 MiB device array the library holds for the life of the process (one
 * per device, set up by mossq_ctx_create).  The caller passes the program,
 * the observable, the input (or NULL for |0...0>) and the outputs.  Programs
 * whose scratch does not fit (beyond 20 qubits, or 20 qubits with very many
 * parameters) return MOSSQ_EINVAL with the size needed; density programs
 * return MOSSQ_EUNSUPPORTED.  All calls on one device share the array:
 * calls that overlap on one device race (serialise them, e.g. one context
 * per device).  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving loops, code run under stream
 * capture, harnesses that forbid hipMalloc). */
int mossq_adjoint_noalloc(mossq_ctx *ctx, const mossq_program *prog, const mossq_obs *obs, const void *in,
                          void *out_state, double *value, double *grads);

/* Batched adjoint gradients of B independent programs (uneven qubit counts,
 * own parameters, own observable), all started from |0...0>.  grads of item b
 * start at grads[grad_offset[b]].  Programs created with fuse=1 (whose fused
 * plans are all sweeps; below 12 qubits a 12-qubit form with the extra qubits
 * idle, built on first use) run their forward and reverse passes as batched
 * LDS-tile sweeps, one launch per sweep index for the whole batch; the other
 * items (gates on at most two qubits) run gate by gate.
 * New capability: the reference's batch reverse is per column with one
 * shared circuit (src/differentiation/batch.jl:35-66). */
int mossq_batch_adjoint(mossq_ctx *ctx, int B, mossq_program *const *progs,
                        mossq_obs *const *obs, const int64_t *grad_offset, double *values,
                        double *grads);

/* mossq_batch_adjoint with item b starting from init[b] (a device vector of
 * 2^n_b amplitudes, not modified) instead of |0...0> (NULL init or init[b]:
 * |0...0>).  The gradients of a batch objective over many initial states of
 * one circuit (reference _batch_circuit_adjoint / _batch_expectation_adjoint,
 * src/differentiation/batch.jl:35-66,128-204, one column at a time) in one
 * call: the same program may appear in every item. */
int mossq_batch_adjoint_init(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                             const void *const *init, const int64_t *grad_offset, double *values,
                             double *grads);

/* Record one mossq_batch_adjoint call (items from |0...0>) as a HIP graph: the
 * launch tables, item tables and scratch vectors are built once and held by
 * the graph; mossq_batch_graph_run replays the whole batch (tiled sweeps,
 * observe, any gate-by-gate items on the second stream) with one launch call and
 * copies the values and gradients out as mossq_batch_adjoint does.  With
 * mossq_program_update between runs, the same graph evaluates new parameters
 * (an optimizer's repeated batch gradients).  The programs and observables
 * must outlive the graph; mossq_graph_destroy releases it.  No counterpart in
 * the reference (its batch reverse runs per column, src/differentiation/
 * batch.jl:35-66). */
int mossq_batch_capture(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                        const int64_t *grad_offset, mossq_graph **out);
int mossq_batch_graph_run(mossq_ctx *ctx, mossq_graph *graph, double *values, double *grads);

/* Values Re <H_b> of B independent state-vector programs in one engine run
 * (forward and observe steps only), item b starting from init[b] (a device
 * vector of 2^n_b amplitudes) or from |0...0> when init or init[b] is NULL.
 * Fused programs of 12 or more qubits run their forward pass as batched
 * tile sweeps, as in mossq_batch_adjoint.
 * Serves the sampled-gradient methods, whose shifted or perturbed circuits
 * are evaluated together (reference ParameterShift / FiniteDifference /
 * SPSAGradient, src/differentiation/sampled_gradients.jl:323-406, which run
 * one objective evaluation at a time). */
int mossq_batch_expectation(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                            const void *const *init, double *values);

/* ---- request server (continuous batching) ----------------------------------
 * New capability (the reference evaluates one objective at a time,
 * src/differentiation/objectives.jl:183-218).  A request is an energy +
 * adjoint-gradient evaluation of a lowered circuit (same records and pool as
 * mossq_program_create) started from |0...0> (or |0><0| when is_density),
 * for an observable created on the server's context.  Requests are admitted
 * first-in first-out while their device memory fits the budget; state-vector
 * requests join the batched engine at the next iteration and stay resident on
 * the device until they finish; density-matrix requests run on a worker stream
 * concurrently.  The budget bounds all device memory the requests use,
 * including the blocks the server keeps for reuse between requests (they are
 * released before an allocation would pass the budget); the server's fixed
 * tables (about 40 MiB) come on top.  step/poll/completed are called from one
 * thread; the observable must outlive its requests. */
typedef struct mossq_server mossq_server;
int mossq_server_create(mossq_ctx *ctx, uint64_t budget_bytes, mossq_server **out);
int mossq_server_destroy(mossq_server *srv);
/* Queue a request; MOSSQ_EOOM if it can never fit the budget. */
int mossq_server_submit(mossq_server *srv, int n, int is_density, int nops, const mossq_op *ops,
                        const double *pool, uint64_t pool_len, const mossq_obs *obs, int64_t *ticket);
/* One scheduling iteration: collect finished results, retire, admit, enqueue
 * one engine iteration (at most 8 iterations are queued ahead of the device).
 * active = admitted unfinished requests, queued = waiting for admission. */
int mossq_server_step(mossq_server *srv, int *active, int *queued);
/* Tickets finished since the last call (up to max), count in *count. */
int mossq_server_completed(mossq_server *srv, int64_t *tickets, int max, int *count);
/* Device bytes the server's requests hold now (in_use, cached blocks
 * included) and at most since creation (peak); both stay within the budget. */
int mossq_server_memory(mossq_server *srv, uint64_t *in_use, uint64_t *peak);
/* done=1 when finished: value and gradients (in the program's derivative
 * order) are copied out and the ticket is released. */
int mossq_server_poll(mossq_server *srv, int64_t ticket, int *done, double *value, double *grads,
                      int max_grads);

/* ---- timing ----------------------------------------------------------------
 * HIP events recorded on the context's stream (the stream every kernel of the
 * context is launched on). */
int mossq_event_record(mossq_ctx *ctx, int slot);
int mossq_event_elapsed_ms(mossq_ctx *ctx, int slot0, int slot1, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* MOSSQ_HIP_H */
