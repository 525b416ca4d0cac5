/*
 * cpu_ref.c -- CPU restatement of MOSSQuantum's dense CPU backend.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle for the HIP backend and
 * the CPU baseline timed beside it.  It is never linked into, loaded by or
 * called from the product library (libmossq_hip.so) or its Python host.
 *
 * Every function restates the reference algorithm it names (file:line relative
 * to the MOSSQuantum checkout).  Layout and conventions are the reference's:
 * little-endian qubit order, column-major local matrices in the ordered basis,
 * vec(rho) with rho[i,j] at i + 2^n j.  Work is partitioned over OpenMP threads
 * the way parallel_run partitions it (src/utils/parallel_for.jl:9-42): the
 * independent blocks of one sweep are split into contiguous chunks.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
typedef int64_t i64;

static i64 insert0(i64 x, int q) {
    i64 lo = x & ((((i64)1) << q) - 1);
    return ((x >> q) << (q + 1)) | lo;
}

int ref_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void ref_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* k-qubit dense gate on sorted bit positions: v_b <- U v_b for every block b
 * of 2^k amplitudes (reference _apply_gate_2_impl!,
 * src/backends/cpu/gates/serial_short_range.jl:2-71; the threaded variants
 * threaded_short_range.jl / threaded_long_range.jl compute the same blocks). */
void ref_apply_gate(cplx *v, int nbits, int k, const int *pos, const cplx *U) {
    const int d = 1 << k;
    const i64 nblocks = ((i64)1) << (nbits - k);
    i64 off[32];
    for (int c = 0; c < d; ++c) {
        i64 o = 0;
        for (int j = 0; j < k; ++j)
            if (c & (1 << j)) o |= ((i64)1) << pos[j];
        off[c] = o;
    }
#pragma omp parallel for schedule(static) if (nblocks >= 1024)
    for (i64 t = 0; t < nblocks; ++t) {
        i64 base = t;
        for (int j = 0; j < k; ++j) base = insert0(base, pos[j]);
        cplx in[32];
        for (int c = 0; c < d; ++c) in[c] = v[base + off[c]];
        for (int r = 0; r < d; ++r) {
            cplx acc = 0;
            for (int c = 0; c < d; ++c) acc += U[r + d * c] * in[c];
            v[base + off[r]] = acc;
        }
    }
}

/* CZ: negate the |11> quarter (reference src/backends/cpu/gates/specific/cz_gate.jl:6-26). */
void ref_apply_cz(cplx *v, int nbits, int q1, int q2) {
    int a = q1 < q2 ? q1 : q2, b = q1 < q2 ? q2 : q1;
    const i64 n4 = ((i64)1) << (nbits - 2);
    const i64 set = (((i64)1) << a) | (((i64)1) << b);
#pragma omp parallel for schedule(static) if (n4 >= 4096)
    for (i64 t = 0; t < n4; ++t) {
        i64 i = insert0(insert0(t, a), b) | set;
        v[i] = -v[i];
    }
}

/* Density-matrix gate: U on the row bits, then conj(U) on the column bits
 * (reference _dm_apply_threaded!, src/backends/cpu/gates/generic.jl:24-34). */
void ref_dm_apply_gate(cplx *rho, int n, int k, const int *pos, const cplx *U) {
    int d = 1 << k;
    int cpos[5];
    cplx *Uc = (cplx *)malloc(sizeof(cplx) * d * d);
    for (int i = 0; i < d * d; ++i) Uc[i] = conj(U[i]);
    for (int j = 0; j < k; ++j) cpos[j] = pos[j] + n;
    ref_apply_gate(rho, 2 * n, k, pos, U);
    ref_apply_gate(rho, 2 * n, k, cpos, Uc);
    free(Uc);
}

/* Supermatrix of a Kraus list in the ordered basis:
 * S[(a + d c), (b + d e)] = sum_K K[a,b] conj(K[c,e])
 * (reference _compute_supermat, src/quantum_circuits/channels/quantum_map.jl:59-67). */
void ref_supermatrix(int k, int nkraus, const cplx *kraus, cplx *S) {
    int d = 1 << k, L = d * d;
    memset(S, 0, sizeof(cplx) * L * L);
    for (int q = 0; q < nkraus; ++q) {
        const cplx *K = kraus + (i64)q * d * d;
        for (int a = 0; a < d; ++a)
            for (int c = 0; c < d; ++c)
                for (int b = 0; b < d; ++b)
                    for (int e = 0; e < d; ++e) S[(a + d * c) + L * (b + d * e)] += K[a + d * b] * conj(K[c + d * e]);
    }
}

/* Channel on a density matrix: S (or S^dagger) on bits (pos..., pos+n...)
 * (reference _qmap_apply_threaded! / _qmap_apply_dagger_threaded!,
 * src/backends/cpu/gates/channels.jl:2-24). */
void ref_dm_apply_channel(cplx *rho, int n, int k, const int *pos, const cplx *S, int dagger) {
    int L = 1 << (2 * k);
    int all[4];
    for (int j = 0; j < k; ++j) {
        all[j] = pos[j];
        all[j + k] = pos[j] + n;
    }
    if (!dagger) {
        ref_apply_gate(rho, 2 * n, 2 * k, all, S);
        return;
    }
    cplx *Sd = (cplx *)malloc(sizeof(cplx) * L * L);
    for (int r = 0; r < L; ++r)
        for (int c = 0; c < L; ++c) Sd[c + L * r] = conj(S[r + L * c]);
    ref_apply_gate(rho, 2 * n, 2 * k, all, Sd);
    free(Sd);
}

/* Observable groups: group g has gk[g] sites (listed consecutively in `sites`)
 * and a 2^k x 2^k column-major matrix (listed consecutively in `mats`), the
 * group matrix of _get_mat (src/backends/cpu/operators/utils.jl:3-20). */

/* sum_g sum_b phi_b^dagger M_g psi_b (reference expectation_value /
 * _expectation_value_util, src/backends/cpu/expectations/serial.jl:2-128;
 * summed over groups as _cpu_expectation, src/observables/expectation.jl:12-33). */
void ref_cross_expect_groups(const cplx *phi, const cplx *psi, int n, int G, const int *gk, const int *sites,
                             const cplx *mats, double *out) {
    cplx total = 0;
    int so = 0, mo = 0;
    for (int g = 0; g < G; ++g) {
        int k = gk[g], d = 1 << k;
        const int *s = sites + so;
        const cplx *M = mats + mo;
        i64 off[32];
        for (int c = 0; c < d; ++c) {
            i64 o = 0;
            for (int j = 0; j < k; ++j)
                if (c & (1 << j)) o |= ((i64)1) << s[j];
            off[c] = o;
        }
        const i64 nb = ((i64)1) << (n - k);
        double re = 0, im = 0;
#pragma omp parallel for schedule(static) reduction(+ : re, im) if (nb >= 4096)
        for (i64 t = 0; t < nb; ++t) {
            i64 base = t;
            for (int j = 0; j < k; ++j) base = insert0(base, s[j]);
            cplx acc = 0;
            for (int r = 0; r < d; ++r) {
                cplx mv = 0;
                for (int c = 0; c < d; ++c) mv += M[r + d * c] * psi[base + off[c]];
                acc += conj(phi[base + off[r]]) * mv;
            }
            re += creal(acc);
            im += cimag(acc);
        }
        total += re + I * im;
        so += k;
        mo += d * d;
    }
    out[0] = creal(total);
    out[1] = cimag(total);
}

void ref_expect_groups(const cplx *psi, int n, int G, const int *gk, const int *sites, const cplx *mats,
                       double *out) {
    ref_cross_expect_groups(psi, psi, n, G, gk, sites, mats, out);
}

/* out = sum_g M_g psi on a 2^nbits vector, groups on bits < nbits (reference
 * _apply_util! / _apply_qterm_util!, src/backends/cpu/operators/generic.jl:12-28;
 * density matrices: left action on the row bits, operators/density_matrix.jl:1-5). */
void ref_apply_op(cplx *out, const cplx *psi, int nbits, int G, const int *gk, const int *sites,
                  const cplx *mats) {
    const i64 len = ((i64)1) << nbits;
    memset(out, 0, sizeof(cplx) * len);
    int so = 0, mo = 0;
    for (int g = 0; g < G; ++g) {
        int k = gk[g], d = 1 << k;
        const int *s = sites + so;
        const cplx *M = mats + mo;
        i64 off[32];
        for (int c = 0; c < d; ++c) {
            i64 o = 0;
            for (int j = 0; j < k; ++j)
                if (c & (1 << j)) o |= ((i64)1) << s[j];
            off[c] = o;
        }
        const i64 nb = ((i64)1) << (nbits - k);
#pragma omp parallel for schedule(static) if (nb >= 4096)
        for (i64 t = 0; t < nb; ++t) {
            i64 base = t;
            for (int j = 0; j < k; ++j) base = insert0(base, s[j]);
            for (int r = 0; r < d; ++r) {
                cplx acc = 0;
                for (int c = 0; c < d; ++c) acc += M[r + d * c] * psi[base + off[c]];
                out[base + off[r]] += acc;
            }
        }
        so += k;
        mo += d * d;
    }
}

/* tr(H rho) = sum_g sum_b tr(M_g rho_b) (reference _expectation_value_util on
 * matrices, src/backends/cpu/expectations/density_matrix_serial.jl:23-79). */
void ref_dm_expect_groups(const cplx *rho, int n, int G, const int *gk, const int *sites, const cplx *mats,
                          double *out) {
    const i64 N = ((i64)1) << n;
    cplx total = 0;
    int so = 0, mo = 0;
    for (int g = 0; g < G; ++g) {
        int k = gk[g], d = 1 << k;
        const int *s = sites + so;
        const cplx *M = mats + mo;
        i64 off[32];
        for (int c = 0; c < d; ++c) {
            i64 o = 0;
            for (int j = 0; j < k; ++j)
                if (c & (1 << j)) o |= ((i64)1) << s[j];
            off[c] = o;
        }
        const i64 nb = ((i64)1) << (n - k);
        for (i64 t = 0; t < nb; ++t) {
            i64 b = t;
            for (int j = 0; j < k; ++j) b = insert0(b, s[j]);
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) total += M[r + d * c] * rho[(b + off[c]) + N * (b + off[r])];
        }
        so += k;
        mo += d * d;
    }
    out[0] = creal(total);
    out[1] = cimag(total);
}

/* ---------------------------------------------------------------- programs ----
 * op records: kind (0 gate, 1 channel), k, pos[5], mat offset, dmat offset
 * (same layout as mossq_op). */
typedef struct {
    int kind, k, pos[5], mat, dmat, reserved;
} ref_op;

static void apply_op_fwd(cplx *state, int n, int is_dm, const ref_op *op, const cplx *pool) {
    if (op->kind == 1) {
        ref_dm_apply_channel(state, n, op->k, op->pos, pool + op->mat, 0);
    } else if (is_dm) {
        ref_dm_apply_gate(state, n, op->k, op->pos, pool + op->mat);
    } else {
        ref_apply_gate(state, n, op->k, op->pos, pool + op->mat);
    }
}

/* apply!(circuit, state): the instruction loop of src/execution/gates.jl:21-36. */
void ref_run_program(cplx *state, int n, int is_dm, int nops, const ref_op *ops, const cplx *pool) {
    for (int i = 0; i < nops; ++i) apply_op_fwd(state, n, is_dm, &ops[i], pool);
}

static void conj_transpose(const cplx *m, int d, cplx *out) {
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) out[c + d * r] = conj(m[r + d * c]);
}

/* Adjoint gradient of real(<H>) (reference _circuit_adjoint, back_propagate,
 * _density_gate_gradient!, src/differentiation/circuit.jl:77-229, with the
 * seeds of src/differentiation/expectation.jl:32-35 (pure: (H + H^dag) psi) and
 * :69-72 (density: H^dag as a dense matrix)).  Pure programs reconstruct the
 * state by inverse gates; density programs keep a copy of the state before
 * every channel, exactly as the reference does.  in == NULL starts from
 * |0...0> (or |0><0|).  grads receives one entry per op with dmat >= 0, in
 * forward order. */
int ref_adjoint(int n, int is_dm, int nops, const ref_op *ops, const cplx *pool, int G, const int *gk,
                const int *sites, const cplx *mats, const cplx *in, cplx *out_state, double *value,
                double *grads) {
    const int nbits = is_dm ? 2 * n : n;
    const i64 len = ((i64)1) << nbits;
    cplx *y = (cplx *)malloc(sizeof(cplx) * len);
    cplx *delta = (cplx *)malloc(sizeof(cplx) * len);
    if (!y || !delta) return 2;
    if (in)
        memcpy(y, in, sizeof(cplx) * len);
    else {
        memset(y, 0, sizeof(cplx) * len);
        y[0] = 1;
    }
    int nder = 0;
    for (int i = 0; i < nops; ++i)
        if (ops[i].dmat >= 0) ++nder;
    double val[2];
    if (!is_dm) {
        ref_run_program(y, n, 0, nops, ops, pool);
        if (out_state) memcpy(out_state, y, sizeof(cplx) * len);
        ref_expect_groups(y, n, G, gk, sites, mats, val);
        /* seed (H + H^dag) y */
        int mo = 0;
        for (int g = 0; g < G; ++g) mo += 1 << (2 * gk[g]);
        cplx *seedm = (cplx *)malloc(sizeof(cplx) * (mo > 0 ? mo : 1));
        mo = 0;
        for (int g = 0; g < G; ++g) {
            int d = 1 << gk[g];
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) seedm[mo + r + d * c] = mats[mo + r + d * c] + conj(mats[mo + c + d * r]);
            mo += d * d;
        }
        ref_apply_op(delta, y, n, G, gk, sites, seedm);
        free(seedm);
        int gi = nder - 1;
        cplx Ud[1024], tmp[1024];
        cplx *dy = (cplx *)malloc(sizeof(cplx) * len);
        for (int i = nops - 1; i >= 0; --i) {
            const ref_op *op = &ops[i];
            int d = 1 << op->k;
            conj_transpose(pool + op->mat, d, Ud);
            ref_apply_gate(delta, n, op->k, op->pos, Ud);
            ref_apply_gate(y, n, op->k, op->pos, Ud);
            if (op->dmat >= 0) {
                /* g = Re <y|D|delta> (reference back_propagate, circuit.jl:136-144) */
                memcpy(tmp, pool + op->dmat, sizeof(cplx) * d * d);
                memcpy(dy, delta, sizeof(cplx) * len);
                ref_apply_gate(dy, n, op->k, op->pos, tmp);
                double re = 0;
#pragma omp parallel for schedule(static) reduction(+ : re) if (len >= 4096)
                for (i64 t = 0; t < len; ++t) re += creal(conj(y[t]) * dy[t]);
                grads[gi--] = re;
            }
        }
        free(dy);
    } else {
        /* forward with a checkpoint before every channel (circuit.jl:90-98) */
        cplx **ck = (cplx **)calloc(nops > 0 ? nops : 1, sizeof(cplx *));
        for (int i = 0; i < nops; ++i) {
            if (ops[i].kind == 1) {
                ck[i] = (cplx *)malloc(sizeof(cplx) * len);
                memcpy(ck[i], y, sizeof(cplx) * len);
            }
            apply_op_fwd(y, n, 1, &ops[i], pool);
        }
        if (out_state) memcpy(out_state, y, sizeof(cplx) * len);
        ref_dm_expect_groups(y, n, G, gk, sites, mats, val);
        /* seed: H^dag applied to the identity matrix */
        int mo = 0;
        for (int g = 0; g < G; ++g) mo += 1 << (2 * gk[g]);
        cplx *adjm = (cplx *)malloc(sizeof(cplx) * (mo > 0 ? mo : 1));
        mo = 0;
        for (int g = 0; g < G; ++g) {
            int d = 1 << gk[g];
            conj_transpose(mats + mo, d, adjm + mo);
            mo += d * d;
        }
        cplx *ident = (cplx *)calloc(len, sizeof(cplx));
        const i64 N = ((i64)1) << n;
        for (i64 i = 0; i < N; ++i) ident[i + N * i] = 1;
        ref_apply_op(delta, ident, nbits, G, gk, sites, adjm);
        free(adjm);
        cplx *left = ident; /* reuse as workspace */
        cplx *right = (cplx *)malloc(sizeof(cplx) * len);
        int gi = nder - 1;
        cplx Ud[1024], Dd[1024], Dt[1024];
        for (int i = nops - 1; i >= 0; --i) {
            const ref_op *op = &ops[i];
            if (op->kind == 1) {
                ref_dm_apply_channel(delta, n, op->k, op->pos, pool + op->mat, 1);
                memcpy(y, ck[i], sizeof(cplx) * len);
                free(ck[i]);
                ck[i] = NULL;
                continue;
            }
            int d = 1 << op->k;
            conj_transpose(pool + op->mat, d, Ud);
            ref_dm_apply_gate(delta, n, op->k, op->pos, Ud);
            ref_dm_apply_gate(y, n, op->k, op->pos, Ud);
            if (op->dmat >= 0) {
                /* g = Re[<Delta, D^dag rho> + <Delta, rho D>] (circuit.jl:155-172) */
                const cplx *D = pool + op->dmat;
                int bra[5];
                conj_transpose(D, d, Dd);
                for (int r = 0; r < d; ++r)
                    for (int c = 0; c < d; ++c) Dt[r + d * c] = D[c + d * r];
                for (int j = 0; j < op->k; ++j) bra[j] = op->pos[j] + n;
                memcpy(left, y, sizeof(cplx) * len);
                memcpy(right, y, sizeof(cplx) * len);
                ref_apply_gate(left, nbits, op->k, op->pos, Dd);
                ref_apply_gate(right, nbits, op->k, bra, Dt);
                double re = 0;
#pragma omp parallel for schedule(static) reduction(+ : re) if (len >= 4096)
                for (i64 t = 0; t < len; ++t) re += creal(conj(delta[t]) * (left[t] + right[t]));
                grads[gi--] = re;
            }
        }
        free(right);
        free(ident);
        for (int i = 0; i < nops; ++i) free(ck[i]);
        free(ck);
    }
    *value = val[0];
    free(y);
    free(delta);
    return 0;
}

/* Readout marginals (reference _cpu_probabilities, src/execution/probabilities.jl:42-63):
 * out[m] = sum over basis states whose bits at qubits[] spell m of |psi|^2 (or Re rho_ii).
 * Unnormalised; the host normalises as the reference does. */
void ref_probabilities(const cplx *s, int n, int is_dm, int k, const int *qubits, double *out) {
    const i64 N = ((i64)1) << n;
    memset(out, 0, sizeof(double) * (((i64)1) << k));
    for (i64 i = 0; i < N; ++i) {
        i64 m = 0;
        for (int j = 0; j < k; ++j) m |= ((i >> qubits[j]) & 1) << j;
        double p = is_dm ? creal(s[i + N * i]) : creal(s[i]) * creal(s[i]) + cimag(s[i]) * cimag(s[i]);
        out[m] += p;
    }
}

/* ------------------------------------------------------ finite differences ----
 * Energies for central finite differences, the method of the reference's
 * gradient regression tests (test/regression/gradients.jl:4-83,
 * test/support/finite_differences.jl:8-46).  Two forms:
 *   ref_expect_groups_compensated: <psi|H|psi> with every sum (the blocks of a
 *     group, the threads' partials, the groups) compensated (Neumaier), so the
 *     summation adds no error beyond the terms' own rounding;
 *   ref_energy_ld: the whole forward run and <H> of a state-vector program in
 *     x87 extended precision (64-bit significand), with the gate matrices
 *     given in extended precision, so an FD quotient at h = 1e-6 keeps about
 *     13 significant digits instead of about 7. */
typedef struct {
    double s, c;
} neum;

static void neum_add(neum *a, double x) {
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x))
        a->c += (a->s - t) + x;
    else
        a->c += (x - t) + a->s;
    a->s = t;
}

void ref_expect_groups_compensated(const cplx *psi, int n, int G, const int *gk, const int *sites, const cplx *mats,
                                   double *out) {
    neum tre = {0, 0}, tim = {0, 0};
    int so = 0, mo = 0;
    for (int g = 0; g < G; ++g) {
        int k = gk[g], d = 1 << k;
        const int *s = sites + so;
        const cplx *M = mats + mo;
        i64 off[32];
        for (int c = 0; c < d; ++c) {
            i64 o = 0;
            for (int j = 0; j < k; ++j)
                if (c & (1 << j)) o |= ((i64)1) << s[j];
            off[c] = o;
        }
        const i64 nb = ((i64)1) << (n - k);
        int nt = 1;
#ifdef _OPENMP
        nt = omp_get_max_threads();
#endif
        neum *pr = (neum *)calloc((size_t)nt, sizeof(neum)), *pi = (neum *)calloc((size_t)nt, sizeof(neum));
#pragma omp parallel num_threads(nt)
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
#pragma omp for schedule(static)
            for (i64 t = 0; t < nb; ++t) {
                i64 base = t;
                for (int j = 0; j < k; ++j) base = insert0(base, s[j]);
                for (int r = 0; r < d; ++r) {
                    cplx mv = 0;
                    for (int c = 0; c < d; ++c) mv += M[r + d * c] * psi[base + off[c]];
                    cplx term = conj(psi[base + off[r]]) * mv;
                    neum_add(&pr[tid], creal(term));
                    neum_add(&pi[tid], cimag(term));
                }
            }
        }
        for (int t = 0; t < nt; ++t) {  /* fixed thread order */
            neum_add(&tre, pr[t].s);
            neum_add(&tre, pr[t].c);
            neum_add(&tim, pi[t].s);
            neum_add(&tim, pi[t].c);
        }
        free(pr);
        free(pi);
        so += k;
        mo += d * d;
    }
    out[0] = tre.s + tre.c;
    out[1] = tim.s + tim.c;
}

typedef long double complex lcplx;

/* ref_apply_gate in extended precision. */
static void ld_apply_gate(lcplx *v, int nbits, int k, const int *pos, const lcplx *U) {
    const int d = 1 << k;
    const i64 nblocks = ((i64)1) << (nbits - k);
    i64 off[32];
    for (int c = 0; c < d; ++c) {
        i64 o = 0;
        for (int j = 0; j < k; ++j)
            if (c & (1 << j)) o |= ((i64)1) << pos[j];
        off[c] = o;
    }
#pragma omp parallel for schedule(static) if (nblocks >= 4096)
    for (i64 t = 0; t < nblocks; ++t) {
        i64 base = t;
        for (int j = 0; j < k; ++j) base = insert0(base, pos[j]);
        lcplx in[32];
        for (int c = 0; c < d; ++c) in[c] = v[base + off[c]];
        for (int r = 0; r < d; ++r) {
            lcplx acc = 0;
            for (int c = 0; c < d; ++c) acc += U[r + d * c] * in[c];
            v[base + off[r]] = acc;
        }
    }
}

/* <H> after a state-vector program started from |0...0>, in extended
 * precision: pool holds the program's matrices as long double complex (same
 * offsets as the double pool); mats are the group matrices (double, exact for
 * Pauli strings with binary coefficients).  out[0], out[1] = Re, Im. */
int ref_energy_ld(int n, int nops, const ref_op *ops, const lcplx *pool, int G, const int *gk, const int *sites,
                  const cplx *mats, long double *out) {
    const i64 len = ((i64)1) << n;
    lcplx *v = (lcplx *)calloc((size_t)len, sizeof(lcplx));
    if (!v) return 2;
    v[0] = 1;
    for (int i = 0; i < nops; ++i) {
        if (ops[i].kind != 0) {
            free(v);
            return 1;
        }
        ld_apply_gate(v, n, ops[i].k, ops[i].pos, pool + ops[i].mat);
    }
    long double re = 0, im = 0;
    int so = 0, mo = 0;
    for (int g = 0; g < G; ++g) {
        int k = gk[g], d = 1 << k;
        const int *s = sites + so;
        const cplx *M = mats + mo;
        i64 off[32];
        for (int c = 0; c < d; ++c) {
            i64 o = 0;
            for (int j = 0; j < k; ++j)
                if (c & (1 << j)) o |= ((i64)1) << s[j];
            off[c] = o;
        }
        const i64 nb = ((i64)1) << (n - k);
        long double gre = 0, gim = 0;
#pragma omp parallel for schedule(static) reduction(+ : gre, gim) if (nb >= 4096)
        for (i64 t = 0; t < nb; ++t) {
            i64 base = t;
            for (int j = 0; j < k; ++j) base = insert0(base, s[j]);
            for (int r = 0; r < d; ++r) {
                lcplx mv = 0;
                for (int c = 0; c < d; ++c) mv += (lcplx)M[r + d * c] * v[base + off[c]];
                lcplx term = conjl(v[base + off[r]]) * mv;
                gre += creall(term);
                gim += cimagl(term);
            }
        }
        re += gre;
        im += gim;
        so += k;
        mo += d * d;
    }
    out[0] = re;
    out[1] = im;
    free(v);
    return 0;
}
