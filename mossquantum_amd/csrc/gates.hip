// Per-instruction sweep kernels for gfx950: dense k-bit blocks, monomial
// (diagonal / permutation) blocks, the fused reverse-pass pair kernel, state
// initialisation and the per-call gate / channel entry points of the C ABI.
//
// Layout: a state is a flat array of complex128 (16 B, one dwordx4 access per
// amplitude).  A k-bit block kernel maps thread -> block base index by
// inserting zero bits at the k target positions, so consecutive lanes touch
// consecutive amplitudes whenever the lowest target bit is >= 2 and runs of
// 2^q amplitudes otherwise; every byte of the state is read and written once.
#include <cmath>
#include <cstring>

#include "mossq_internal.h"

namespace mossq {

typedef double double4_t __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;

// Grid size for a sweep of `work` items with `items` per thread.
static inline unsigned grid_for(u64 work, int items) {
    u64 per_block = u64(kBlock) * items;
    u64 g = (work + per_block - 1) / per_block;
    if (g == 0) g = 1;
    if (g > 0x7fffffffULL) g = 0x7fffffffULL;
    return unsigned(g);
}

// ------------------------------------------------------------ 1-bit dense ----
struct M2 {
    c128 m00, m10, m01, m11;  // column-major 2x2
};

template <int ITEMS>
__global__ __launch_bounds__(kBlock) void k_gate1(c128 *__restrict__ v, u64 npairs, int q, M2 m) {
    const u64 base = u64(blockIdx.x) * kBlock * ITEMS + threadIdx.x;
    const u64 lo_mask = (u64(1) << q) - 1;
    const u64 stride = u64(1) << q;
    c128 a[ITEMS], b[ITEMS];
    u64 idx[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        u64 p = base + u64(j) * kBlock;
        idx[j] = ((p & ~lo_mask) << 1) | (p & lo_mask);
        if (p < npairs) {
            a[j] = load_nt(v + idx[j]);
            b[j] = load_nt(v + idx[j] + stride);
        }
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
        u64 p = base + u64(j) * kBlock;
        if (p < npairs) {
            c128 x = cfma(m.m01, b[j], cmul(m.m00, a[j]));
            c128 y = cfma(m.m11, b[j], cmul(m.m10, a[j]));
            store_nt(v + idx[j], x);
            store_nt(v + idx[j] + stride, y);
        }
    }
}

// ------------------------------------------------------------ k-bit dense ----
struct Bits {
    int b[10];
};

// Offset of local index c of a k-bit block (bit j of c on global bit b[j]).
template <int K>
__device__ __forceinline__ u64 block_off(int c, const Bits &bits) {
    u64 o = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) o |= u64((c >> j) & 1) << bits.b[j];
    return o;
}

// One block of 2^K amplitudes per thread, the matrix in LDS (read as a
// wave-wide broadcast).  The block's inputs stay in registers; outputs are
// produced one row at a time (rows not unrolled), so K = 3..5 keep 4 * 2^K
// data VGPRs and no per-slot offset table (no scratch, several waves per SIMD).
template <int K>
__global__ __launch_bounds__(kBlock) void k_dense(c128 *__restrict__ v, u64 nblocks, Bits bits,
                                                   const c128 *__restrict__ mat) {
    constexpr int D = 1 << K;
    __shared__ c128 sm[D * D];
    for (int i = threadIdx.x; i < D * D; i += kBlock) sm[i] = mat[i];
    __syncthreads();
    for (u64 t = u64(blockIdx.x) * kBlock + threadIdx.x; t < nblocks; t += u64(gridDim.x) * kBlock) {
        u64 base = insert0s<K>(t, bits.b);
        c128 in[D];
#pragma unroll
        for (int c = 0; c < D; ++c) in[c] = load_nt(v + base + block_off<K>(c, bits));
#pragma unroll(K <= 2 ? D : 1)
        for (int r = 0; r < D; ++r) {
            c128 acc = mk(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) acc = cfma(sm[r + D * c], in[c], acc);
            store_nt(v + base + block_off<K>(r, bits), acc);
        }
    }
}

// K = 4, 5 dense blocks on the FP64 matrix cores (v_mfma_f64_16x16x4f64).
// The complex update out = U in is the real product
//     [out_re; out_im] = [[U_re, -U_im], [U_im, U_re]] [in_re; in_im]
// of a (2D x 2D) matrix with 2D-long columns, D = 2^K.  Each wave takes 16
// blocks at once (the 16 columns of the MFMA's B and D operands), so one
// wave's matrix fragments (2D/16 row tiles x D/2 k-steps, held in VGPRs for
// the whole sweep) serve every block it visits.  Lane l = 16 q + j loads
// amplitudes c = 4 t + q (t < D/4) of block j; k-step s consumes the real
// parts of c = 4 s + q for s < D/4 and the imaginary parts after that, which
// is exactly operand B[k = 4 s + q][j].  Result element e of row tile rb is
// real-matrix row 16 rb + q + 4 e: the real part of amplitude row
// 16 rb + q + 4 e for rb < D/16, its imaginary part D/16 tiles later.
// The matrix cores give the same FP64 peak as the vector units but take the
// 2^K-wide complex dot products off the VALU, which leaves the sweep HBM
// bound (measured at 26 qubits: 0.38-0.45 ms against 0.48-3.3 ms for the
// vector kernel over target sets from {0..4} to {21..25}).
template <int K>
__global__ __launch_bounds__(kBlock) void k_dense_mfma(c128 *__restrict__ v, u64 ngroups, Bits bits,
                                                        const c128 *__restrict__ mat) {
    constexpr int D = 1 << K, RB = D / 8, S = D / 2, T = D / 4;
    const int l = threadIdx.x & 63, q = l >> 4, j = l & 15;
    double afr[RB][S];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = 16 * rb + j, k = 4 * s + q;
            const c128 m = mat[(i & (D - 1)) + D * (k & (D - 1))];
            afr[rb][s] = i < D ? (k < D ? m.re : -m.im) : (k < D ? m.im : m.re);
        }
    const u64 stride = u64(gridDim.x) * (kBlock / 64);
    for (u64 g = u64(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6); g < ngroups; g += stride) {
        const u64 base = insert0s<K>(g * 16 + j, bits.b);
        c128 x[T];
#pragma unroll
        for (int t = 0; t < T; ++t) x[t] = load_nt(v + base + block_off<K>(4 * t + q, bits));
        double4_t acc[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[rb] = double4_t{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int rb = 0; rb < RB; ++rb)
                acc[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(afr[rb][s], s < T ? x[s].re : x[s - T].im, acc[rb],
                                                               0, 0, 0);
#pragma unroll
        for (int rb = 0; rb < RB / 2; ++rb)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                store_nt(v + base + block_off<K>(16 * rb + q + 4 * e, bits), mk(acc[rb][e], acc[rb + RB / 2][e]));
    }
}

// One-bit dense sweep, one amplitude pair per thread (the widest grid streams
// HBM best on MI355X: ~6 TB/s for an in-place read+write); the matrix comes
// through the scalar cache, no LDS staging.
__global__ __launch_bounds__(kBlock) void k_dense1(c128 *__restrict__ v, u64 npairs, int q,
                                                    const c128 *__restrict__ mat) {
    const c128 m00 = mat[0], m10 = mat[1], m01 = mat[2], m11 = mat[3];
    for (u64 p = u64(blockIdx.x) * kBlock + threadIdx.x; p < npairs; p += u64(gridDim.x) * kBlock) {
        const u64 i0 = insert0(p, q), i1 = i0 | (u64(1) << q);
        const c128 a = load_nt(v + i0), b = load_nt(v + i1);
        store_nt(v + i0, cfma(m01, b, cmul(m00, a)));
        store_nt(v + i1, cfma(m11, b, cmul(m10, a)));
    }
}

// -------------------------------------------------------- k-bit monomial ----
// Matrices with one nonzero per row and column (diagonal and permutation
// gates).  Only the changed rows are read and written: CZ touches a quarter of
// the state, CNOT a half, a general diagonal all of it but with no matrix
// product.
struct Mono {
    int k;
    int nchanged;
    int bits[10];
    int rows[32];
    int src[32];
    c128 val[32];
};

template <int NCH>
__global__ __launch_bounds__(kBlock) void k_monomial(c128 *__restrict__ v, u64 nblocks, Mono m) {
    // Offsets of the changed rows and of their sources (loop-invariant).
    u64 orow[NCH], osrc[NCH];
#pragma unroll
    for (int j = 0; j < NCH; ++j) {
        u64 a = 0, b = 0;
        for (int q = 0; q < m.k; ++q) {
            if (m.rows[j] & (1 << q)) a |= u64(1) << m.bits[q];
            if (m.src[j] & (1 << q)) b |= u64(1) << m.bits[q];
        }
        orow[j] = a;
        osrc[j] = b;
    }
    for (u64 t = u64(blockIdx.x) * kBlock + threadIdx.x; t < nblocks; t += u64(gridDim.x) * kBlock) {
        u64 base = insert0s_dyn(t, m.bits, m.k);
        c128 in[NCH];
#pragma unroll
        for (int j = 0; j < NCH; ++j)
            if (j < m.nchanged) in[j] = load_nt(v + base + osrc[j]);
#pragma unroll
        for (int j = 0; j < NCH; ++j)
            if (j < m.nchanged) store_nt(v + base + orow[j], cmul(m.val[j], in[j]));
    }
}

// Specialisations for the common single-row-changed case (CZ, CPHASE, CCPHASE):
// each thread touches exactly one amplitude, with the fixed bits set.
// Launched with one amplitude per thread (the loop runs once).
__global__ __launch_bounds__(kBlock) void k_phase1(c128 *__restrict__ v, u64 nblocks, Bits bits, int k,
                                                    u64 set_mask, c128 val) {
    for (u64 t = u64(blockIdx.x) * kBlock + threadIdx.x; t < nblocks; t += u64(gridDim.x) * kBlock) {
        const u64 i = insert0s_dyn(t, bits.b, k) | set_mask;
        store_nt(v + i, cmul(val, load_nt(v + i)));
    }
}

MatInfo classify(const c128 *m, int dim) {
    MatInfo info;
    info.monomial = false;
    if (dim > 32) return info;
    int src_of_row[32];
    bool used_col[32] = {false};
    for (int r = 0; r < dim; ++r) {
        int cnt = 0, col = -1;
        for (int c = 0; c < dim; ++c) {
            const c128 &e = m[r + dim * c];
            if (e.re != 0.0 || e.im != 0.0) {
                ++cnt;
                col = c;
            }
        }
        if (cnt != 1 || used_col[col]) return info;
        used_col[col] = true;
        src_of_row[r] = col;
    }
    info.monomial = true;
    info.nchanged = 0;
    for (int r = 0; r < dim; ++r) {
        const c128 &e = m[r + dim * src_of_row[r]];
        // A diagonal entry within a few ulps of 1 counts as 1: |u|^2 = c^2 + s^2
        // in the density superoperator of a phase gate rounds to 1 - 2^-53,
        // 1 or 1 + 2^-52 depending on the angle, and the kernels chosen (and
        // the tables an in-place rebinding must keep) would follow that rounding.
        const double tol = 4 * 2.220446049250313e-16;
        bool ident = (src_of_row[r] == r) && std::fabs(e.re - 1.0) <= tol && std::fabs(e.im) <= tol;
        if (!ident) {
            int j = info.nchanged++;
            info.rows[j] = r;
            info.src[j] = src_of_row[r];
            info.val[j] = e;
        }
    }
    return info;
}

static int launch_mono(Ctx *c, c128 *v, int nbits, int k, const int *bits, const MatInfo &info) {
    if (info.nchanged == 0) return MOSSQ_OK;  // identity
    u64 nblocks = u64(1) << (nbits - k);
    if (info.nchanged == 1 && info.rows[0] == info.src[0]) {
        Bits b{};
        for (int j = 0; j < k; ++j) b.b[j] = bits[j];
        u64 set = 0;
        for (int j = 0; j < k; ++j)
            if (info.rows[0] & (1 << j)) set |= u64(1) << bits[j];
        hipLaunchKernelGGL(k_phase1, dim3(grid_for(nblocks, 1)), dim3(kBlock), 0, c->stream, v, nblocks, b,
                           k, set, info.val[0]);
    } else {
        Mono m{};
        m.k = k;
        m.nchanged = info.nchanged;
        for (int j = 0; j < k; ++j) m.bits[j] = bits[j];
        for (int j = 0; j < info.nchanged; ++j) {
            m.rows[j] = info.rows[j];
            m.src[j] = info.src[j];
            m.val[j] = info.val[j];
        }
        dim3 g(grid_for(nblocks, 4)), bl(kBlock);
        if (m.nchanged <= 2)
            hipLaunchKernelGGL(k_monomial<2>, g, bl, 0, c->stream, v, nblocks, m);
        else if (m.nchanged <= 4)
            hipLaunchKernelGGL(k_monomial<4>, g, bl, 0, c->stream, v, nblocks, m);
        else if (m.nchanged <= 8)
            hipLaunchKernelGGL(k_monomial<8>, g, bl, 0, c->stream, v, nblocks, m);
        else if (m.nchanged <= 16)
            hipLaunchKernelGGL(k_monomial<16>, g, bl, 0, c->stream, v, nblocks, m);
        else
            hipLaunchKernelGGL(k_monomial<32>, g, bl, 0, c->stream, v, nblocks, m);
    }
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// Workgroups for k_dense_mfma over `nblocks` blocks: one 16-block group per
// wave, at most 4096 workgroups (16384 waves, so each wave loads its matrix
// fragments once and reuses them over several groups on large states).
static inline unsigned mfma_grid(u64 nblocks) {
    u64 waves = nblocks / 16, g = (waves + kBlock / 64 - 1) / (kBlock / 64);
    return unsigned(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}

int launch_dense(Ctx *c, c128 *v, int nbits, int k, const int *bits, const c128 *dmat,
                 const MatInfo &info) {
    MOSSQ_CHECK(k >= 1 && k <= 5 && k <= nbits, MOSSQ_EINVAL, "dense block size out of range");
    if (info.monomial) return launch_mono(c, v, nbits, k, bits, info);
    u64 nblocks = u64(1) << (nbits - k);
    Bits b{};
    for (int j = 0; j < k; ++j) b.b[j] = bits[j];
    unsigned grid = grid_for(nblocks, 2);
    switch (k) {
        case 1:
            hipLaunchKernelGGL(k_dense1, dim3(grid_for(nblocks, 1)), dim3(kBlock), 0, c->stream, v, nblocks, bits[0],
                               dmat);
            break;
        case 2:
            hipLaunchKernelGGL(k_dense<2>, dim3(grid), dim3(kBlock), 0, c->stream, v, nblocks, b, dmat);
            break;
        case 3:
            hipLaunchKernelGGL(k_dense<3>, dim3(grid), dim3(kBlock), 0, c->stream, v, nblocks, b, dmat);
            break;
        case 4:
            if (nblocks >= 16)
                hipLaunchKernelGGL(k_dense_mfma<4>, dim3(mfma_grid(nblocks)), dim3(kBlock), 0, c->stream, v,
                                   nblocks / 16, b, dmat);
            else
                hipLaunchKernelGGL(k_dense<4>, dim3(grid), dim3(kBlock), 0, c->stream, v, nblocks, b, dmat);
            break;
        default:
            if (nblocks >= 16)
                hipLaunchKernelGGL(k_dense_mfma<5>, dim3(mfma_grid(nblocks)), dim3(kBlock), 0, c->stream, v,
                                   nblocks / 16, b, dmat);
            else
                hipLaunchKernelGGL(k_dense<5>, dim3(grid), dim3(kBlock), 0, c->stream, v, nblocks, b, dmat);
            break;
    }
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// 1-bit dense with the matrix known on the host: kernel-argument matrix, no upload.
static int launch_gate1_host(Ctx *c, c128 *v, int nbits, int q, const c128 *m) {
    u64 npairs = u64(1) << (nbits - 1);
    M2 mm{m[0], m[1], m[2], m[3]};
    hipLaunchKernelGGL(k_gate1<4>, dim3(grid_for(npairs, 4)), dim3(kBlock), 0, c->stream, v, npairs, q, mm);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// Apply a host matrix on `k` sorted bits: monomial -> sparse kernel; 1 bit ->
// kernel-argument matrix; otherwise stage the matrix and run the dense kernel.
int apply_host_matrix(Ctx *c, c128 *v, int nbits, int k, const int *bits, const c128 *m) {
    int dim = 1 << k;
    MatInfo info = classify(m, dim);
    if (info.monomial) return launch_mono(c, v, nbits, k, bits, info);
    if (k == 1) return launch_gate1_host(c, v, nbits, bits[0], m);
    const c128 *dm = nullptr;
    MOSSQ_TRY(stage_matrix(c, m, u64(dim) * dim, &dm));
    return launch_dense(c, v, nbits, k, bits, dm, info);
}

// -------------------------------------------------------- reverse-pass pair ----
// x <- V x and (unless Y_FIXED) y <- V y on the same k bits, then (if GRAD)
// accumulate sum_blocks x_b^dagger D y_b with the updated vectors.  One
// read+write of both vectors; the gradient reduction costs no extra sweep.
// Per-workgroup partial sums go to part[blockIdx.x].  With y updated, D (V y)
// is formed as (D V) y from the product staged in LDS, so a block's outputs
// go out row by row (rows not unrolled for K >= 2: no per-slot offset table
// or output arrays, no scratch, and the LDS matrices are not hoisted into
// registers).
template <int K, bool GRAD, bool Y_FIXED>
__global__ __launch_bounds__(kBlock) void k_pair(c128 *__restrict__ x, c128 *__restrict__ y, u64 nblocks,
                                                  Bits bits, const c128 *__restrict__ V,
                                                  const c128 *__restrict__ Dm, c128 *__restrict__ part) {
    constexpr int D = 1 << K;
    __shared__ c128 sV[D * D];
    __shared__ c128 sD[GRAD ? D * D : 1];
    __shared__ c128 sG[GRAD && !Y_FIXED ? D * D : 1];
    __shared__ c128 red[kBlock / 64];
    for (int i = threadIdx.x; i < D * D; i += kBlock) {
        sV[i] = V[i];
        if (GRAD) sD[i] = Dm[i];
    }
    __syncthreads();
    if (GRAD && !Y_FIXED) {
        // G = D V (column-major), the gradient matrix on the input of y
        for (int i = threadIdx.x; i < D * D; i += kBlock) {
            const int r = i % D, c = i / D;
            c128 a = mk(0.0, 0.0);
            for (int q = 0; q < D; ++q) a = cfma(sD[r + D * q], sV[q + D * c], a);
            sG[i] = a;
        }
        __syncthreads();
    }
    const c128 *G = Y_FIXED ? sD : sG;
    c128 acc = mk(0.0, 0.0);
    for (u64 t = u64(blockIdx.x) * kBlock + threadIdx.x; t < nblocks; t += u64(gridDim.x) * kBlock) {
        const u64 base = insert0s<K>(t, bits.b);
        c128 xi[D], yi[D];
#pragma unroll
        for (int c = 0; c < D; ++c) xi[c] = load_nt(x + base + block_off<K>(c, bits));
        if (GRAD || !Y_FIXED) {
#pragma unroll
            for (int c = 0; c < D; ++c) yi[c] = load_nt(y + base + block_off<K>(c, bits));
        }
#pragma unroll(K <= 1 ? D : 1)
        for (int r = 0; r < D; ++r) {
            const u64 o = base + block_off<K>(r, bits);
            c128 ax = mk(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) ax = cfma(sV[r + D * c], xi[c], ax);
            store_nt(x + o, ax);
            if (!Y_FIXED) {
                c128 ay = mk(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < D; ++c) ay = cfma(sV[r + D * c], yi[c], ay);
                store_nt(y + o, ay);
            }
            if (GRAD) {
                c128 dy = mk(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < D; ++c) dy = cfma(G[r + D * c], yi[c], dy);
                acc = cfmaconj(ax, dy, acc);
            }
        }
    }
    if (GRAD) {
        for (int o = 32; o > 0; o >>= 1) {
            acc.re += __shfl_xor(acc.re, o, 64);
            acc.im += __shfl_xor(acc.im, o, 64);
        }
        int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        if (lane == 0) red[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            c128 s = red[0];
            for (int w = 1; w < kBlock / 64; ++w) s = cadd(s, red[w]);
            part[blockIdx.x] = s;
        }
    }
}

// Number of partial sums a gradient pair launch writes for a 2^nbits vector.
int pair_parts(int nbits) {
    u64 nblocks = nbits > 4 ? (u64(1) << (nbits - 4)) : 1;
    u64 g = (nblocks + 2 * kBlock - 1) / (2 * kBlock);
    if (g < 1) g = 1;
    if (g > 1024) g = 1024;
    return int(g);
}

int launch_dense_pair(Ctx *c, c128 *x, c128 *y, int nbits, int k, const int *bits, const c128 *dV,
                      const c128 *dD, c128 *part, bool y_fixed) {
    MOSSQ_CHECK(k >= 1 && k <= 4 && k <= nbits, MOSSQ_EINVAL, "pair block size out of range");
    u64 nblocks = u64(1) << (nbits - k);
    Bits b{};
    for (int j = 0; j < k; ++j) b.b[j] = bits[j];
    unsigned grid = dD ? unsigned(pair_parts(nbits)) : grid_for(nblocks, 2);
#define MOSSQ_PAIR(KK)                                                                                       \
    if (dD && y_fixed)                                                                                       \
        hipLaunchKernelGGL((k_pair<KK, true, true>), dim3(grid), dim3(kBlock), 0, c->stream, x, y, nblocks, \
                           b, dV, dD, part);                                                                 \
    else if (dD)                                                                                             \
        hipLaunchKernelGGL((k_pair<KK, true, false>), dim3(grid), dim3(kBlock), 0, c->stream, x, y, nblocks, \
                           b, dV, dD, part);                                                                 \
    else                                                                                                     \
        hipLaunchKernelGGL((k_pair<KK, false, false>), dim3(grid), dim3(kBlock), 0, c->stream, x, y,        \
                           nblocks, b, dV, dD, part);
    switch (k) {
        case 1:
            MOSSQ_PAIR(1);
            break;
        case 2:
            MOSSQ_PAIR(2);
            break;
        case 3:
            MOSSQ_PAIR(3);
            break;
        default:
            MOSSQ_PAIR(4);
            break;
    }
#undef MOSSQ_PAIR
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// ------------------------------------------------------------------- init ----
__global__ void k_set_basis0(c128 *__restrict__ v, u64 len) {
    for (u64 i = u64(blockIdx.x) * kBlock + threadIdx.x; i < len; i += u64(gridDim.x) * kBlock)
        v[i] = mk(i == 0 ? 1.0 : 0.0, 0.0);
}

__global__ void k_outer(c128 *__restrict__ rho, const c128 *__restrict__ psi, int n) {
    const u64 N = u64(1) << n;
    const u64 len = N * N;
    for (u64 t = u64(blockIdx.x) * kBlock + threadIdx.x; t < len; t += u64(gridDim.x) * kBlock) {
        u64 i = t & (N - 1), j = t >> n;
        c128 a = psi[i], b = psi[j];
        rho[t] = cmul(a, cconj(b));
    }
}

int fill_basis0(Ctx *c, c128 *v, u64 len) {
    hipLaunchKernelGGL(k_set_basis0, dim3(grid_for(len, 4)), dim3(kBlock), 0, c->stream, v, len);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

int outer_product(Ctx *c, c128 *rho, const c128 *psi, int n) {
    u64 len = u64(1) << (2 * n);
    hipLaunchKernelGGL(k_outer, dim3(grid_for(len, 4)), dim3(kBlock), 0, c->stream, rho, psi, n);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}

// kron(conj(U), U) in the local order (row bits low, column bits high).
void dm_superop_of_gate(const c128 *U, int k, std::vector<c128> &out) {
    int d = 1 << k, D = d * d;
    out.assign(size_t(D) * D, mk(0, 0));
    for (int a = 0; a < d; ++a)
        for (int cc = 0; cc < d; ++cc)
            for (int b = 0; b < d; ++b)
                for (int e = 0; e < d; ++e) {
                    // S[(a + d c), (b + d e)] = U[a,b] conj(U[c,e])
                    int row = a + d * cc, col = b + d * e;
                    out[size_t(row) + size_t(D) * col] = cmul(U[a + d * b], cconj(U[cc + d * e]));
                }
}

void conj_transpose(const c128 *m, int dim, std::vector<c128> &out) {
    out.resize(size_t(dim) * dim);
    for (int r = 0; r < dim; ++r)
        for (int c2 = 0; c2 < dim; ++c2) out[size_t(c2) + size_t(dim) * r] = cconj(m[r + size_t(dim) * c2]);
}

static int check_positions(int n, int k, const int32_t *pos, int kmax) {
    MOSSQ_CHECK(n >= 1 && n <= 40, MOSSQ_EINVAL, "qubit count out of range");
    MOSSQ_CHECK(k >= 1 && k <= kmax && k <= n, MOSSQ_EINVAL, "gate size out of range");
    MOSSQ_CHECK(pos != nullptr, MOSSQ_EINVAL, "null positions");
    for (int j = 0; j < k; ++j) {
        MOSSQ_CHECK(pos[j] >= 0 && pos[j] < n, MOSSQ_EINVAL, "position out of range");
        MOSSQ_CHECK(j == 0 || pos[j] > pos[j - 1], MOSSQ_EINVAL, "positions must be strictly increasing");
    }
    return MOSSQ_OK;
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_sv_init_zero(mossq_ctx *ctx, void *psi, int n) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(psi && n >= 0 && n <= 40, MOSSQ_EINVAL, "bad state arguments");
    return fill_basis0(C(ctx), (c128 *)psi, u64(1) << n);
}

int mossq_dm_init_zero(mossq_ctx *ctx, void *rho, int n) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(rho && n >= 0 && n <= 20, MOSSQ_EINVAL, "bad state arguments");
    return fill_basis0(C(ctx), (c128 *)rho, u64(1) << (2 * n));
}

int mossq_dm_from_sv(mossq_ctx *ctx, void *rho, const void *psi, int n) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(rho && psi && n >= 0 && n <= 20, MOSSQ_EINVAL, "bad state arguments");
    return outer_product(C(ctx), (c128 *)rho, (const c128 *)psi, n);
}

int mossq_sv_apply_gate(mossq_ctx *ctx, void *psi, int n, int k, const int32_t *pos, const double *mat) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(psi && mat, MOSSQ_EINVAL, "null buffer");
    MOSSQ_TRY(check_positions(n, k, pos, 5));
    int bits[5];
    for (int j = 0; j < k; ++j) bits[j] = pos[j];
    return apply_host_matrix(C(ctx), (c128 *)psi, n, k, bits, (const c128 *)mat);
}

int mossq_dm_apply_gate(mossq_ctx *ctx, void *rho, int n, int k, const int32_t *pos, const double *mat) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(rho && mat, MOSSQ_EINVAL, "null buffer");
    MOSSQ_TRY(check_positions(n, k, pos, 5));
    Ctx *c = C(ctx);
    const c128 *U = (const c128 *)mat;
    if (k <= 2) {
        std::vector<c128> S;
        dm_superop_of_gate(U, k, S);
        int bits[4];
        for (int j = 0; j < k; ++j) {
            bits[j] = pos[j];
            bits[j + k] = pos[j] + n;
        }
        return apply_host_matrix(c, (c128 *)rho, 2 * n, 2 * k, bits, S.data());
    }
    // Wider gates: U on the row bits, then conj(U) on the column bits.
    int d = 1 << k;
    std::vector<c128> Uc(size_t(d) * d);
    for (size_t i = 0; i < Uc.size(); ++i) Uc[i] = cconj(U[i]);
    int rb[5], cb[5];
    for (int j = 0; j < k; ++j) {
        rb[j] = pos[j];
        cb[j] = pos[j] + n;
    }
    MOSSQ_TRY(apply_host_matrix(c, (c128 *)rho, 2 * n, k, rb, U));
    // The staging area is rewritten by a copy queued behind the first kernel.
    return apply_host_matrix(c, (c128 *)rho, 2 * n, k, cb, Uc.data());
}

int mossq_dm_apply_channel(mossq_ctx *ctx, void *rho, int n, int k, const int32_t *pos, const double *supermat,
                           int dagger) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(rho && supermat, MOSSQ_EINVAL, "null buffer");
    MOSSQ_TRY(check_positions(n, k, pos, 2));
    int D = 1 << (2 * k);
    std::vector<c128> S((const c128 *)supermat, (const c128 *)supermat + size_t(D) * D);
    if (dagger) {
        std::vector<c128> St;
        conj_transpose(S.data(), D, St);
        S.swap(St);
    }
    int bits[4];
    for (int j = 0; j < k; ++j) {
        bits[j] = pos[j];
        bits[j + k] = pos[j] + n;
    }
    return apply_host_matrix(C(ctx), (c128 *)rho, 2 * n, 2 * k, bits, S.data());
}

}  // extern "C"
