// Fused sweeps: LDS-resident tiles of 2^12 amplitudes.
//
// A sweep streams the state once (one HBM read + one write).  Each workgroup
// owns one tile: the 2^12 amplitudes whose indices agree outside the sweep's
// 12 "tile bits".  The tile bits always contain the lowest global bits
// (coalesced runs of >= 8 amplitudes = 128 B) plus the bits of the fused gates.
// Every thread holds 16 amplitudes in registers, spanning 4 tile bits (the
// "register bits").  Gates on register bits are applied in registers; the
// register bits are changed by a transposition through LDS (padded one slot per
// 16 to spread ds_read_b128 across banks).  Diagonal gates (CZ, Rz, phases,
// their density-matrix forms) are applied from the amplitude index and may use
// any global bit, inside or outside the tile: sign-only diagonals flip signs,
// diagonals on <= 2 bits select among four uniform values, wider ones gather.
//
// The planner (host) takes the program's instructions in order, moves
// instructions past earlier ones they commute with (disjoint supports, or both
// diagonal), and packs them greedily into sweeps of at most 12 tile bits; it
// then orders each sweep's instructions into register-bit segments, multiplies
// adjacent instructions on the same bits into one matrix, and absorbs
// diagonals on register bits into the dense gates before them.  A sweep loads
// and stores in the layout of its first and last segment when that layout
// keeps the three lowest tile bits on lanes (coalescing), saving transpositions.
//
// emulate_plan() executes a plan on the host with the same index arithmetic,
// so the planner is tested on a CPU against the oracle.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "mossq_program.h"

namespace mossq {

constexpr int TM = 12;         // tile bits
constexpr int kRegMax = 4;     // register bits of a layout: 4 (16 slots, 256 threads) or 3 (8 slots, 512 threads)
constexpr int kThreadMax = 9;  // thread-index bits: TM - register bits
constexpr int kCoalesceBits = 3;

// Tile geometry for RB register bits: TN amplitudes per thread, TT threads
// per tile (TTB thread-index bits), WPT waves per tile.  With 3 register bits a
// thread holds half the amplitudes, so a tile of the same 2^12 amplitudes has
// twice the threads at about half the registers per thread: four waves per
// SIMD instead of two.
template <int RB>
struct TGeo {
    static constexpr int TN = 1 << RB;
    static constexpr int TTB = TM - RB;
    static constexpr int TT = 1 << TTB;
    static constexpr int WPT = TT / 64;
};

enum TOpType { T_DENSE1 = 0, T_DENSE2 = 1, T_SIGN = 2, T_SELECT = 3, T_GATHER = 4, T_PROBE1 = 5, T_PROBE2 = 6 };

// Diagonal index of register slot j: idx = f | rp(j).  f collects the op bits
// that are not register bits: from the thread's tile-local index (lbit[i] >= 0)
// or, for bits outside the tile, from the tile's base index (qbit[i] >= 0).
// rp(j) (4 bits per slot, precomputed by the planner for the segment's register
// layout) collects the register bits.
struct TOp {
    int type;
    int a, b;      // dense: register-bit indices (a < b)
    int k;         // diagonal: number of bits
    int lbit[4];   // diagonal: tile-local bit of op bit i if it is a thread bit, else -1
    int qbit[4];   // diagonal: global bit of op bit i if it lies outside the tile, else -1
    unsigned rp_lo, rp_hi;  // diagonal: rp(j) for j < 8 (lo) and j >= 8 (hi), 4 bits each
    int negmask;   // T_SIGN: bit idx set -> negate
    int moff;      // offset in the plan's device pool (matrix / diagonal values)
    int rmask;     // T_SIGN: register-bit mask of the op's register bits
    int probe;     // T_PROBE*: parameter (gradient row) index
};

// Device form of a segment's ops: groups, each one out-of-place step
// (in -> out) of the two-register-set op loop, so the loop never copies.
//   C_DENSE1 + a, C_DENSE2 + pair: dense 2x2 on register bit a / 4x4 on a
//     register-bit pair (kPairA/kPairB; with 3 register bits only the pairs
//     of bits 0..2);
//   C_FLIP + any of those: the dense op, then sign flips of its outputs;
//   C_COPY: copy with sign flips (flips at the start of a segment);
//   C_DIAG: a general diagonal (TDiag record x);
//   C_PERM + 6 * pair + kind: a 4x4 permutation on a register-bit pair
//     (CNOT either way, SWAP, X on either bit or both; perm_src), which only
//     moves slots; + kPermFlip: the same, then sign flips;
//   C_DIAG1 + a (adjoint plans): a diagonal 2x2 on register bit a, one
//     complex product per slot (no flips ride on it).
// Sign flips are controlled-Z-type diagonals (-1 exactly where all their bits
// are 1), up to kGroupCz per group: flip i negates the slots in the slot
// pattern pat[i] wherever its condition bits are all 1.  cond[i] packs up to
// three condition bits, 7 bits each from bit 0: 0x7F none; 0x40 | b: global bit
// b outside the tile; else tile-local bit b.  A segment with an odd number of
// groups ends with a plain C_COPY, so groups always come in pairs.
enum : int { C_DENSE1 = 0, C_DENSE2 = 4, C_FLIP = 16, C_COPY = 30, C_DIAG = 31, C_PERM = 32, C_DIAG1 = 112 };
constexpr int kPermFlip = 36;
enum : int { D_SIGN = 0, D_SELECT = 1, D_GATHER = 2 };
constexpr int kPairA[6] = {0, 0, 0, 1, 1, 2};
constexpr int kPairB[6] = {1, 2, 3, 2, 3, 3};
constexpr int kGroupCz = 3;
constexpr int kCzNone = 0x7F;

// After a group's op and flips come its probes, np of them at probes[p0 ..]
// (up to kGroupProbes in adjoint plans, one fixed probe in density plans):
// the contribution of this tile to Re<y|D|l> (or Re<v|D|fixed>) for
// parameter `probe`, D a 2x2 / 4x4 on register bits (pcode as a dense code) at
// pool offset pmoff.  The probes of one group are summed over the wave
// together.
constexpr int kGroupProbes = 4;
struct TGroup {
    int code;
    int moff;  // dense: matrix offset in the pool
    int ncz;
    int x;     // C_DIAG: TDiag index
    unsigned pat[kGroupCz];
    int cond[kGroupCz];
    int np;  // probes after the op
    int p0;  // first probe (TProbe index)
};

// Probe kinds (pcode bits 8 and up; the low bits are the dense code): the
// general form walks the nonzero parts of D; a rotation's generator is one
// of two two-part forms with its own code path.
enum : int { PK_GENERAL = 0, PK_IMDIAG = 1, PK_REANTI = 2 };
struct TProbe {
    int probe;  // parameter (gradient row)
    int pcode;  // dense code | kind << 8
    int pmoff;
    unsigned pmask;  // nonzero parts of D: bit 2e real, bit 2e + 1 imaginary part of entry e
    // the factors of the first two nonzero parts in mask order (Re D for a
    // real part, -Im D for an imaginary one), so a rotation's probe (two
    // parts) reads no matrix
    double coef[2];
};

struct TDiag {
    int code;  // D_SIGN, D_SELECT, D_GATHER
    int lbit[4];
    int qbit[4];
    unsigned rp_lo, rp_hi;
    int negmask;
    int moff;  // D_GATHER: offset of the 2^k values in the pool
    int pad;
    c128 val[4];  // D_SELECT: the four values
};

// A register layout and its instructions.  The layout's index maps are
// precomputed on the host: thread-index bit k sits at tile-local bit tpos[k]
// (global bit gpos[k]); register bit i at tile-local bit reg[i] (global gbit[i]).
// A plan's layouts all hold the same number of register bits (3 or 4); the
// arrays are sized for the larger thread and register counts.
//
// Entering a segment whose layout differs from the previous one transposes
// the tile through LDS.  The LDS slot of tile-local index l is the XOR of one
// column per set bit of l; the columns are chosen per transposition (unit
// lower-triangular, hence a bijection) so that the first 8 and the first 16
// lanes of a wave address distinct 16-byte bank groups on both the write (old
// layout) and the read (new layout) side.  wt/wr and rt/rr are those columns
// at the old and new layouts' thread and register bits.
struct TSeg {
    int reg[kRegMax];  // tile-local bits held in registers (increasing; -1 past the layout's count)
    int op0, op1;      // planner ops (TOp) of the segment
    int g0, g1;        // device groups of the segment
    int tpos[kThreadMax];
    int gpos[kThreadMax];
    int gbit[kRegMax];
    int transpose;
    int wt[kThreadMax], wr[kRegMax];
    int rt[kThreadMax], rr[kRegMax];
};

struct TSweepArgs {
    int tbits[TM];  // global bits of the tile, increasing
    int cbits[40];  // remaining global bits, increasing (tile id bits)
    int ncbits;
    int seg0, seg1;
};


// Sign bit of slot j under the per-lane flip pattern F: bit j of F shifted to
// bit 31, then hi ^ (s & 0x80000000) as one three-input bit operation per half
// (truth table 0x78 = A ^ (B & C)); three VALU ops per slot instead of four.
__device__ __forceinline__ c128 flip_slot(c128 x, unsigned F, int j) {
    const unsigned s = F << (31 - j);
    const unsigned rh = __builtin_amdgcn_bitop3_b32(unsigned(__double2hiint(x.re)), s, 0x80000000u, 0x78);
    const unsigned ih = __builtin_amdgcn_bitop3_b32(unsigned(__double2hiint(x.im)), s, 0x80000000u, 0x78);
    return c128{__hiloint2double(int(rh), __double2loint(x.re)), __hiloint2double(int(ih), __double2loint(x.im))};
}

// Dense ops are out of place (in -> out); with FLIP the outputs are then
// sign-flipped per the lane's pattern F.
template <int RB, int A, bool FLIP>
__device__ __forceinline__ void reg_apply1(const c128 *in, c128 *out, const c128 *__restrict__ M, unsigned F) {
    const c128 m00 = M[0], m10 = M[1], m01 = M[2], m11 = M[3];
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) {
        if (j & (1 << A)) continue;
        const c128 x = in[j], y = in[j | (1 << A)];
        const c128 y0 = cfma(m01, y, cmul(m00, x)), y1 = cfma(m11, y, cmul(m10, x));
        out[j] = FLIP ? flip_slot(y0, F, j) : y0;
        out[j | (1 << A)] = FLIP ? flip_slot(y1, F, j | (1 << A)) : y1;
    }
}

template <int RB, int A, int B, bool FLIP>
__device__ __forceinline__ void reg_apply2(const c128 *in, c128 *out, const c128 *__restrict__ M, unsigned F) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const c128 m0 = M[r], m1 = M[r + 4], m2 = M[r + 8], m3 = M[r + 12];
        const int orow = ((r & 1) << A) | ((r >> 1) << B);
#pragma unroll
        for (int j = 0; j < TGeo<RB>::TN; ++j) {
            if (j & ((1 << A) | (1 << B))) continue;
            c128 acc = cmul(m0, in[j]);
            acc = cfma(m1, in[j | (1 << A)], acc);
            acc = cfma(m2, in[j | (1 << B)], acc);
            acc = cfma(m3, in[j | (1 << A) | (1 << B)], acc);
            out[j | orow] = FLIP ? flip_slot(acc, F, j | orow) : acc;
        }
    }
}

// Permutation ops on register bits (A, B): slot j takes slot perm_src(j).
// With r = bit A + 2 bit B the local index, kind 0 is CNOT controlled by A
// (r: 1 <-> 3), kind 1 CNOT controlled by B (2 <-> 3), kind 2 SWAP (1 <-> 2),
// kinds 3, 4, 5 flip bit A, bit B, both.
__host__ __device__ constexpr int perm_src(int j, int A, int B, int kind) {
    const int a = (j >> A) & 1, b = (j >> B) & 1;
    return kind == 0   ? (a ? j ^ (1 << B) : j)
           : kind == 1 ? (b ? j ^ (1 << A) : j)
           : kind == 2 ? (a != b ? j ^ ((1 << A) | (1 << B)) : j)
           : kind == 3 ? j ^ (1 << A)
           : kind == 4 ? j ^ (1 << B)
                       : j ^ ((1 << A) | (1 << B));
}

// In place: each instantiation reads through its own local copy (with one
// source array for all 36 cases, the compiler merges their moves into moves
// from a selected address, which takes the register arrays to scratch).
template <int RB, int A, int B, int K, bool FLIP>
__device__ __forceinline__ void reg_perm(const c128 *in, c128 *out, unsigned F) {
    constexpr int TN = TGeo<RB>::TN;
    c128 t[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) t[j] = in[j];
#pragma unroll
    for (int j = 0; j < TN; ++j) out[j] = FLIP ? flip_slot(t[perm_src(j, A, B, K)], F, j) : t[perm_src(j, A, B, K)];
}

// Permutation groups (code >= C_PERM): pk = code - C_PERM below kPermFlip.
// Pairs that name register bit 3 exist only in 4-bit layouts.
#define MOSSQ_PERM6(P, STMT) \
    STMT(P, 0)               \
    STMT(P, 1)               \
    STMT(P, 2)               \
    STMT(P, 3)               \
    STMT(P, 4)               \
    STMT(P, 5)
#define MOSSQ_PERM36(STMT) \
    MOSSQ_PERM6(0, STMT)   \
    MOSSQ_PERM6(1, STMT)   \
    MOSSQ_PERM6(2, STMT)   \
    MOSSQ_PERM6(3, STMT)   \
    MOSSQ_PERM6(4, STMT)   \
    MOSSQ_PERM6(5, STMT)

template <int RB, bool FLIP>
__device__ __forceinline__ void perm_ip(int pk, c128 *v, unsigned F) {
#define MOSSQ_PERM_IP(P, K)                                                                      \
    case 6 * P + K:                                                                              \
        if constexpr (kPairB[P] < RB) reg_perm<RB, kPairA[P], kPairB[P], K, FLIP>(v, v, F); \
        break;
    switch (pk) { MOSSQ_PERM36(MOSSQ_PERM_IP) }
#undef MOSSQ_PERM_IP
}

// Tile-local index of thread t's register slot 0 in a layout.
template <int RB>
__host__ __device__ __forceinline__ int seg_tdep(int t, const TSeg &sg) {
    int d = 0;
#pragma unroll
    for (int k = 0; k < TGeo<RB>::TTB; ++k) d |= ((t >> k) & 1) << sg.tpos[k];
    return d;
}

// Global offset of thread t's register slot 0 within its tile.
template <int RB>
__host__ __device__ __forceinline__ u64 seg_gthread(int t, const TSeg &sg) {
    u64 g = 0;
#pragma unroll
    for (int k = 0; k < TGeo<RB>::TTB; ++k) g |= u64((t >> k) & 1) << sg.gpos[k];
    return g;
}

// Global offset of register slot j (its register bits) in a layout.
template <int RB>
__host__ __device__ __forceinline__ u64 slot_goff(int j, const TSeg &sg) {
    u64 o = 0;
#pragma unroll
    for (int i = 0; i < RB; ++i)
        if (j & (1 << i)) o |= u64(1) << sg.gbit[i];
    return o;
}

// LDS byte offsets of thread t's register slots: the XOR of the columns of
// the set thread bits, then one column per set bit of j (TN - 1 XORs in all).
template <int RB>
__host__ __device__ __forceinline__ void lds_slots(int t, const int *tcol, const int *rcol, int *a) {
    int x = 0;
#pragma unroll
    for (int k = 0; k < TGeo<RB>::TTB; ++k)
        if ((t >> k) & 1) x ^= tcol[k];
    a[0] = x * int(sizeof(c128));
#pragma unroll
    for (int j = 1; j < TGeo<RB>::TN; ++j) a[j] = a[j & (j - 1)] ^ (rcol[__builtin_ctz(j)] * int(sizeof(c128)));
}

// Non-register part of a diagonal index.
template <class D>
__host__ __device__ __forceinline__ int diag_f(const D &op, int tdep, u64 base) {
    int f = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (op.lbit[i] >= 0) f |= ((tdep >> op.lbit[i]) & 1) << i;
        if (op.qbit[i] >= 0) f |= int((base >> op.qbit[i]) & 1) << i;
    }
    return f;
}

template <class D>
__host__ __device__ __forceinline__ int diag_rp(const D &op, int j) {
    return int(((j < 8 ? op.rp_lo : op.rp_hi) >> (4 * (j & 7))) & 15u);
}

__device__ __forceinline__ double flip_sign(double x, unsigned flip) {
    return __longlong_as_double(__double_as_longlong(x) ^ (static_cast<long long>(flip) << 63));
}

// General diagonal, out of place.
template <int RB>
__device__ __forceinline__ void diag_op(const TDiag &d, const c128 *__restrict__ pool, const c128 *in, c128 *out,
                                        int tdep, u64 base) {
    constexpr int TN = TGeo<RB>::TN;
    const int f = diag_f(d, tdep, base);
    if (d.code == D_SIGN) {
        const unsigned neg = unsigned(d.negmask);
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const unsigned flip = (neg >> (f | diag_rp(d, j))) & 1u;
            out[j].re = flip_sign(in[j].re, flip);
            out[j].im = flip_sign(in[j].im, flip);
        }
    } else if (d.code == D_SELECT) {
        const c128 d0 = d.val[0], d1 = d.val[1], d2 = d.val[2], d3 = d.val[3];
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int idx = f | diag_rp(d, j);
            const c128 lo = (idx & 1) ? d1 : d0, hi = (idx & 1) ? d3 : d2;
            out[j] = cmul((idx & 2) ? hi : lo, in[j]);
        }
    } else {
        const c128 *D = pool + d.moff;
#pragma unroll
        for (int j = 0; j < TN; ++j) out[j] = cmul(D[f | diag_rp(d, j)], in[j]);
    }
}

// Condition of a packed flip for this thread (0 or 1).
__host__ __device__ __forceinline__ unsigned flip_cond(int cw, int tdep, u64 base) {
    unsigned c = 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int f = (cw >> (7 * i)) & 0x7F;
        if (f == kCzNone) continue;
        c &= (f & 0x40) ? unsigned(base >> (f & 63)) & 1u : unsigned(tdep >> f) & 1u;
    }
    return c;
}

// The lane's flip pattern of a group: bit j set = slot j changes sign.
__host__ __device__ __forceinline__ unsigned group_flips(const TGroup &g, int tdep, u64 base) {
    unsigned F = 0;
#pragma unroll
    for (int i = 0; i < kGroupCz; ++i)
        if (i < g.ncz && flip_cond(g.cond[i], tdep, base)) F ^= g.pat[i];
    return F;
}

// Dense cases of a group code (code < C_COPY) for RB register bits; pair
// cases that name bit 3 exist only with 4 register bits.
#define MOSSQ_DENSE_CASES(RB, CALL1, CALL2)                                             \
    case C_DENSE1 + 0: CALL1(RB, 0, false); break;                                      \
    case C_DENSE1 + 1: CALL1(RB, 1, false); break;                                      \
    case C_DENSE1 + 2: CALL1(RB, 2, false); break;                                      \
    case C_DENSE1 + 3: if constexpr (RB > 3) CALL1(RB, 3, false); break;                \
    case C_DENSE2 + 0: CALL2(RB, 0, 1, false); break;                                   \
    case C_DENSE2 + 1: CALL2(RB, 0, 2, false); break;                                   \
    case C_DENSE2 + 2: if constexpr (RB > 3) CALL2(RB, 0, 3, false); break;             \
    case C_DENSE2 + 3: CALL2(RB, 1, 2, false); break;                                   \
    case C_DENSE2 + 4: if constexpr (RB > 3) CALL2(RB, 1, 3, false); break;             \
    case C_DENSE2 + 5: if constexpr (RB > 3) CALL2(RB, 2, 3, false); break;             \
    case C_FLIP + C_DENSE1 + 0: CALL1(RB, 0, true); break;                              \
    case C_FLIP + C_DENSE1 + 1: CALL1(RB, 1, true); break;                              \
    case C_FLIP + C_DENSE1 + 2: CALL1(RB, 2, true); break;                              \
    case C_FLIP + C_DENSE1 + 3: if constexpr (RB > 3) CALL1(RB, 3, true); break;        \
    case C_FLIP + C_DENSE2 + 0: CALL2(RB, 0, 1, true); break;                           \
    case C_FLIP + C_DENSE2 + 1: CALL2(RB, 0, 2, true); break;                           \
    case C_FLIP + C_DENSE2 + 2: if constexpr (RB > 3) CALL2(RB, 0, 3, true); break;     \
    case C_FLIP + C_DENSE2 + 3: CALL2(RB, 1, 2, true); break;                           \
    case C_FLIP + C_DENSE2 + 4: if constexpr (RB > 3) CALL2(RB, 1, 3, true); break;     \
    case C_FLIP + C_DENSE2 + 5: if constexpr (RB > 3) CALL2(RB, 2, 3, true); break;

// One group, out of place.
// `in` is scratch afterwards (permutations run in place on it).
template <int RB>
__device__ __forceinline__ void group_op(const TGroup &g, const TDiag *__restrict__ diag,
                                         const c128 *__restrict__ pool, c128 *in, c128 *out, int tdep,
                                         u64 base) {
    constexpr int TN = TGeo<RB>::TN;
    const c128 *M = pool + g.moff;
    const unsigned F = group_flips(g, tdep, base);
    if (g.code >= C_PERM) {  // C_DIAG1 groups occur in adjoint plans only (dense_ip)
        // permuted in place, then moved: out-of-place permutation switches
        // spill (see reg_perm)
        if (g.code >= C_PERM + kPermFlip)
            perm_ip<RB, true>(g.code - C_PERM - kPermFlip, in, F);
        else
            perm_ip<RB, false>(g.code - C_PERM, in, F);
#pragma unroll
        for (int j = 0; j < TN; ++j) out[j] = in[j];
        return;
    }
#define MOSSQ_OOP1(R, A, FL) reg_apply1<R, A, FL>(in, out, M, F)
#define MOSSQ_OOP2(R, A, B, FL) reg_apply2<R, A, B, FL>(in, out, M, F)
    switch (g.code) {
        MOSSQ_DENSE_CASES(RB, MOSSQ_OOP1, MOSSQ_OOP2)
        case C_COPY:
#pragma unroll
            for (int j = 0; j < TN; ++j) out[j] = flip_slot(in[j], F, j);
            break;
        default: diag_op<RB>(diag[g.x], pool, in, out, tdep, base); break;
    }
#undef MOSSQ_OOP1
#undef MOSSQ_OOP2
}

template <int RB>
__device__ __forceinline__ double probe_op_re(int code, unsigned mask, const c128 *y, const c128 *l,
                                              const c128 *__restrict__ D, double c0, double c1);
__device__ __forceinline__ double wave_sum_re(double a) { return wave_sum_f64(a); }

// Probe of a group against a fixed state: this wave's share of
// sum conj(cur) (D fixed) over the tile, fixed read in the current layout
// into f (the register set the group's output left free).
template <int RB>
__device__ __forceinline__ void fixed_probe(const TProbe &pr, const c128 *cur, c128 *f, const c128 *fixed,
                                            const TSeg &sg, int t, u64 base, const c128 *__restrict__ pool,
                                            c128 *part, int row, int slot) {
    const u64 gt = base | seg_gthread<RB>(t, sg);
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) f[j] = load_nt(fixed + (gt | slot_goff<RB>(j, sg)));
    const double acc = wave_sum_re(probe_op_re<RB>(pr.pcode, pr.pmask, cur, f, pool + pr.pmoff, pr.coef[0],
                                                   pr.coef[1]));
    if ((t & 63) == 0) part[size_t(pr.probe) * row + slot] = c128{acc, 0.0};
}

// Four per-lane values summed over the wave in one pass: the halves exchange
// (a, b) against (c, d), the row pairs one of the two remaining values, then
// each row of 16 lanes sums its own value.  Afterwards lane 16 k holds the
// sum of the k-th argument (k < 4); one reduction chain for four probes.
__device__ __forceinline__ double swap_add_f64(double p, double q, bool halves) {
    const long long pv = __double_as_longlong(p), qv = __double_as_longlong(q);
    const auto lo = halves ? __builtin_amdgcn_permlane32_swap(int(pv), int(qv), false, false)
                           : __builtin_amdgcn_permlane16_swap(int(pv), int(qv), false, false);
    const auto hi = halves ? __builtin_amdgcn_permlane32_swap(int(pv >> 32), int(qv >> 32), false, false)
                           : __builtin_amdgcn_permlane16_swap(int(pv >> 32), int(qv >> 32), false, false);
    return __longlong_as_double((long long)(unsigned)lo[0] | ((long long)hi[0] << 32)) +
           __longlong_as_double((long long)(unsigned)lo[1] | ((long long)hi[1] << 32));
}

__device__ __forceinline__ double wave_sum4(double a, double b, double c, double d) {
    // lanes 0..31: a (+ lane + 32), lanes 32..63: c (+ lane - 32); likewise b / d
    const double u = swap_add_f64(a, c, true), w = swap_add_f64(b, d, true);
    // rows 0 and 2 keep u summed over rows (0, 1) / (2, 3); rows 1 and 3 keep w
    double x = swap_add_f64(u, w, false);
    x += dpp_f64<0x128>(x);  // row_ror:8
    x += dpp_f64<0x124>(x);  // row_ror:4
    x += dpp_f64<0x122>(x);  // row_ror:2
    x += dpp_f64<0x121>(x);  // row_ror:1
    return x;
}

// Global index bits of a tile: tile-id bit i sits at global bit cbits[i].
__device__ __forceinline__ u64 tile_base(const TSweepArgs &sw, u64 tile) {
    u64 base = 0;
    for (int i = 0; i < sw.ncbits; ++i)
        if ((tile >> i) & 1) base |= u64(1) << sw.cbits[i];
    return base;
}

// LDS transposition of one register set into the layout of segment sg.
template <int RB>
__device__ __forceinline__ void transpose_lds(c128 *v, char *ldsb, int t, const TSeg &sg) {
    constexpr int TN = TGeo<RB>::TN;
    int a[TN];
    lds_slots<RB>(t, sg.wt, sg.wr, a);
#pragma unroll
    for (int j = 0; j < TN; ++j) *reinterpret_cast<c128 *>(ldsb + a[j]) = v[j];
    __syncthreads();
    lds_slots<RB>(t, sg.rt, sg.rr, a);
#pragma unroll
    for (int j = 0; j < TN; ++j) v[j] = *reinterpret_cast<const c128 *>(ldsb + a[j]);
    __syncthreads();
}

// Reads the tile from src and writes it to dst (src == dst for an in-place
// sweep; a workgroup stores only after it has loaded its whole tile).  With
// FIXP, a group's probe adds Re<cur|D|S> partials (one per wave per tile)
// for a state S = ftab[probe - fbase] that the sweep does not change.
template <int RB, bool FIXP>
__device__ __forceinline__ void tile_sweep(const c128 *src, c128 *dst, const TSweepArgs &sw,
                                           const TSeg *__restrict__ segs, const TGroup *__restrict__ groups,
                                           const TDiag *__restrict__ diag, const c128 *__restrict__ pool,
                                           const TProbe *__restrict__ probes, const c128 *const *ftab, int fbase,
                                           c128 *part, int row, u64 tile) {
    constexpr int TN = TGeo<RB>::TN;
    __shared__ c128 lds[1 << TM];
    char *ldsb = reinterpret_cast<char *>(lds);
    const int t = threadIdx.x;
    const u64 base = tile_base(sw, tile);
    const int seg0 = sw.seg0, seg1 = sw.seg1;
    const int slot = int(tile) * TGeo<RB>::WPT + (t >> 6);

    TSeg sg = segs[seg0];
    int tdep = seg_tdep<RB>(t, sg);
    c128 v[TN], w[TN];
    {
        const u64 g = base | seg_gthread<RB>(t, sg);
#pragma unroll
        for (int j = 0; j < TN; ++j) v[j] = load_nt(src + (g | slot_goff<RB>(j, sg)));
    }
    for (int s = seg0; s < seg1; ++s) {
        sg = segs[s];
        if (sg.transpose) {
            transpose_lds<RB>(v, ldsb, t, sg);
            tdep = seg_tdep<RB>(t, sg);
        }
        // Groups in pairs, v -> w -> v (a segment holds an even number): the
        // loop carries v only.
        for (int g = sg.g0; g < sg.g1; g += 2) {
            const TGroup ga = groups[g], gb = groups[g + 1];
            group_op<RB>(ga, diag, pool, v, w, tdep, base);
            // (a density plan's group holds at most one probe)
            if (FIXP && ga.np > 0) {
                const TProbe pr = probes[ga.p0];
                fixed_probe<RB>(pr, w, v, ftab[pr.probe - fbase], sg, t, base, pool, part, row, slot);
            }
            group_op<RB>(gb, diag, pool, w, v, tdep, base);
            if (FIXP && gb.np > 0) {
                const TProbe pr = probes[gb.p0];
                fixed_probe<RB>(pr, v, w, ftab[pr.probe - fbase], sg, t, base, pool, part, row, slot);
            }
        }
    }
    {
        // sg is the last segment: its layout is the current one.
        const u64 g = base | seg_gthread<RB>(t, sg);
#pragma unroll
        for (int j = 0; j < TN; ++j) store_nt(dst + (g | slot_goff<RB>(j, sg)), v[j]);
    }
}

// --------------------------------------------------------- adjoint sweep ----
// The reverse pass of the pure-state adjoint over LDS tiles: each thread holds
// TN amplitudes of y and the same TN of l; every op applies to both (in
// place), and a probe adds this wave's share of Re<y|D|l> to the partials of
// its parameter (one entry per wave per tile, reduced afterwards in a fixed
// order, so gradients are reproducible).  Adjoint plans hold dense ops,
// controlled-Z-type sign flips and probes only (no general diagonals).

template <int RB, int A, int B, bool FLIP>
__device__ __forceinline__ void ip_apply2(c128 *v, const c128 *__restrict__ M, unsigned F) {
    c128 m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = M[i];
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) {
        if (j & ((1 << A) | (1 << B))) continue;
        const int idx[4] = {j, j | (1 << A), j | (1 << B), j | (1 << A) | (1 << B)};
        const c128 x0 = v[idx[0]], x1 = v[idx[1]], x2 = v[idx[2]], x3 = v[idx[3]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c128 acc = cmul(m[r], x0);
            acc = cfma(m[r + 4], x1, acc);
            acc = cfma(m[r + 8], x2, acc);
            acc = cfma(m[r + 12], x3, acc);
            v[idx[r]] = FLIP ? flip_slot(acc, F, idx[r]) : acc;
        }
    }
}

template <int RB, int A>
__device__ __forceinline__ void diag1_ip(c128 *v, const c128 *__restrict__ M) {
    const c128 d0 = M[0], d1 = M[3];
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) v[j] = cmul((j >> A) & 1 ? d1 : d0, v[j]);
}

template <int RB>
__device__ __forceinline__ void dense_ip(int code, c128 *v, const c128 *__restrict__ M, unsigned F) {
    if (code >= C_DIAG1) {
        switch (code - C_DIAG1) {
            case 0: diag1_ip<RB, 0>(v, M); break;
            case 1: diag1_ip<RB, 1>(v, M); break;
            case 2: diag1_ip<RB, 2>(v, M); break;
            default:
                if constexpr (RB > 3) diag1_ip<RB, 3>(v, M);
                break;
        }
        return;
    }
    if (code >= C_PERM) {
        if (code >= C_PERM + kPermFlip)
            perm_ip<RB, true>(code - C_PERM - kPermFlip, v, F);
        else
            perm_ip<RB, false>(code - C_PERM, v, F);
        return;
    }
#define MOSSQ_IP1(R, A, FL) reg_apply1<R, A, FL>(v, v, M, F)
#define MOSSQ_IP2(R, A, B, FL) ip_apply2<R, A, B, FL>(v, M, F)
    switch (code) {
        MOSSQ_DENSE_CASES(RB, MOSSQ_IP1, MOSSQ_IP2)
        default:  // C_COPY: flips only
#pragma unroll
            for (int j = 0; j < TGeo<RB>::TN; ++j) v[j] = flip_slot(v[j], F, j);
            break;
    }
#undef MOSSQ_IP1
#undef MOSSQ_IP2
}

// Re sum over the thread's slots of conj(y) (D l), accumulating only the
// parts of D that are nonzero (mask bit 2e: real, 2e + 1: imaginary part of
// entry e = r + dim * c, which pairs conj(y) at row r with l at column c):
// Re(D conj(a) b) = Re D Re(conj(a) b) - Im D Im(conj(a) b).  Gradient probes
// of rotations touch two of the eight parts of a 2x2 D.
template <int RB>
__device__ __forceinline__ double probe_part(const c128 *y, const c128 *l, int msk, int orow, int ocol, bool imag) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) {
        if (j & msk) continue;
        const c128 a = y[j | orow], b = l[j | ocol];
        acc = imag ? fma(a.re, b.im, fma(-a.im, b.re, acc)) : fma(a.re, b.re, fma(a.im, b.im, acc));
    }
    return acc;
}

// One-bit probe forms on register bit A, factors c0, c1 of the two parts:
// imaginary diagonal (R_z generators), c0 Im sum_{a=0} conj(y) l + c1 Im
// sum_{a=1} conj(y) l; real antidiagonal (R_y generators), c0 Re sum
// conj(y_{j|a}) l_j + c1 Re sum conj(y_j) l_{j|a}.
template <int RB, int A>
__device__ __forceinline__ double probe_imdiag(const c128 *y, const c128 *l, double c0, double c1) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) {
        const c128 a = y[j], b = l[j];
        if (j & (1 << A))
            s1 = fma(a.re, b.im, fma(-a.im, b.re, s1));
        else
            s0 = fma(a.re, b.im, fma(-a.im, b.re, s0));
    }
    return fma(c0, s0, c1 * s1);
}

template <int RB, int A>
__device__ __forceinline__ double probe_reanti(const c128 *y, const c128 *l, double c0, double c1) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int j = 0; j < TGeo<RB>::TN; ++j) {
        if (j & (1 << A)) continue;
        const c128 y0 = y[j], y1 = y[j | (1 << A)], l0 = l[j], l1 = l[j | (1 << A)];
        s0 = fma(y1.re, l0.re, fma(y1.im, l0.im, s0));
        s1 = fma(y0.re, l1.re, fma(y0.im, l1.im, s1));
    }
    return fma(c0, s0, c1 * s1);
}

// The first two nonzero parts take their factors from c0, c1; further ones
// read D.
template <int RB, int A, int B>
__device__ __forceinline__ double probe_re(const c128 *y, const c128 *l, const c128 *__restrict__ D, unsigned mask,
                                           double c0, double c1) {
    constexpr int dim = B < 0 ? 2 : 4;
    constexpr int msk = B < 0 ? (1 << A) : ((1 << A) | (1 << B));
    double acc = 0.0;
    int used = 0;
#pragma unroll
    for (int e = 0; e < dim * dim; ++e) {
        const int r = e % dim, cc = e / dim;
        const int orow = B < 0 ? (r << A) : (((r & 1) << A) | ((r >> 1) << (B < 0 ? 0 : B)));
        const int ocol = B < 0 ? (cc << A) : (((cc & 1) << A) | ((cc >> 1) << (B < 0 ? 0 : B)));
#pragma unroll
        for (int im = 0; im < 2; ++im) {
            if (!((mask >> (2 * e + im)) & 1u)) continue;
            const double c = used == 0 ? c0 : used == 1 ? c1 : (im ? -D[e].im : D[e].re);
            ++used;
            acc = fma(c, probe_part<RB>(y, l, msk, orow, ocol, im != 0), acc);
        }
    }
    return acc;
}

template <int RB>
__device__ __forceinline__ double probe_op_re(int code, unsigned mask, const c128 *y, const c128 *l,
                                              const c128 *__restrict__ D, double c0, double c1) {
    switch (code) {
        case C_DENSE1 + 0 + (PK_IMDIAG << 8): return probe_imdiag<RB, 0>(y, l, c0, c1);
        case C_DENSE1 + 1 + (PK_IMDIAG << 8): return probe_imdiag<RB, 1>(y, l, c0, c1);
        case C_DENSE1 + 2 + (PK_IMDIAG << 8): return probe_imdiag<RB, 2>(y, l, c0, c1);
        case C_DENSE1 + 0 + (PK_REANTI << 8): return probe_reanti<RB, 0>(y, l, c0, c1);
        case C_DENSE1 + 1 + (PK_REANTI << 8): return probe_reanti<RB, 1>(y, l, c0, c1);
        case C_DENSE1 + 2 + (PK_REANTI << 8): return probe_reanti<RB, 2>(y, l, c0, c1);
        case C_DENSE1 + 0: return probe_re<RB, 0, -1>(y, l, D, mask, c0, c1);
        case C_DENSE1 + 1: return probe_re<RB, 1, -1>(y, l, D, mask, c0, c1);
        case C_DENSE1 + 2: return probe_re<RB, 2, -1>(y, l, D, mask, c0, c1);
        case C_DENSE2 + 0: return probe_re<RB, 0, 1>(y, l, D, mask, c0, c1);
        case C_DENSE2 + 1: return probe_re<RB, 0, 2>(y, l, D, mask, c0, c1);
        case C_DENSE2 + 3: return probe_re<RB, 1, 2>(y, l, D, mask, c0, c1);
        default:
            if constexpr (RB > 3) {
                switch (code) {
                    case C_DENSE1 + 3 + (PK_IMDIAG << 8): return probe_imdiag<RB, 3>(y, l, c0, c1);
                    case C_DENSE1 + 3 + (PK_REANTI << 8): return probe_reanti<RB, 3>(y, l, c0, c1);
                    case C_DENSE1 + 3: return probe_re<RB, 3, -1>(y, l, D, mask, c0, c1);
                    case C_DENSE2 + 2: return probe_re<RB, 0, 3>(y, l, D, mask, c0, c1);
                    case C_DENSE2 + 4: return probe_re<RB, 1, 3>(y, l, D, mask, c0, c1);
                    default: return probe_re<RB, 2, 3>(y, l, D, mask, c0, c1);
                }
            }
            return 0.0;
    }
}

template <int RB>
__device__ __forceinline__ void tile_adj_sweep(c128 *__restrict__ ys, c128 *__restrict__ ls, const TSweepArgs &sw,
                                               const TSeg *__restrict__ segs, const TGroup *__restrict__ groups,
                                               const c128 *__restrict__ pool, const TProbe *__restrict__ probes,
                                               c128 *__restrict__ part, int row, u64 tile, bool last = false) {
    constexpr int TN = TGeo<RB>::TN;
    __shared__ c128 lds[1 << TM];
    char *ldsb = reinterpret_cast<char *>(lds);
    const int t = threadIdx.x;
    const u64 base = tile_base(sw, tile);
    const int seg0 = sw.seg0, seg1 = sw.seg1;

    TSeg sg = segs[seg0];
    int tdep = seg_tdep<RB>(t, sg);
    c128 y[TN], l[TN];
    {
        const u64 g = base | seg_gthread<RB>(t, sg);
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const u64 o = g | slot_goff<RB>(j, sg);
            y[j] = load_nt(ys + o);
            l[j] = load_nt(ls + o);
        }
    }
    const int slot = int(tile) * TGeo<RB>::WPT + (t >> 6);
    for (int s = seg0; s < seg1; ++s) {
        sg = segs[s];
        if (sg.transpose) {
            transpose_lds<RB>(y, ldsb, t, sg);
            transpose_lds<RB>(l, ldsb, t, sg);
            tdep = seg_tdep<RB>(t, sg);
        }
        for (int gi = sg.g0; gi < sg.g1; ++gi) {
            const TGroup g = groups[gi];
            if (g.code != C_COPY || g.ncz > 0) {
                const unsigned F = group_flips(g, tdep, base);
                dense_ip<RB>(g.code, y, pool + g.moff, F);
                dense_ip<RB>(g.code, l, pool + g.moff, F);
            }
            if (g.np > 0) {
                // all of the group's probe records in one round trip (the
                // table holds kGroupProbes entries of padding), this thread's
                // share of each probe, then one reduction for all
                TProbe pr[kGroupProbes];
#pragma unroll
                for (int k = 0; k < kGroupProbes; ++k) pr[k] = probes[g.p0 + k];
                double v[kGroupProbes];
#pragma unroll
                for (int k = 0; k < kGroupProbes; ++k)
                    v[k] = k < g.np ? probe_op_re<RB>(pr[k].pcode, pr[k].pmask, y, l, pool + pr[k].pmoff, pr[k].coef[0],
                                                      pr[k].coef[1])
                                    : 0.0;
                const int id0 = pr[0].probe, id1 = pr[1].probe, id2 = pr[2].probe, id3 = pr[3].probe;
                const double acc = wave_sum4(v[0], v[1], v[2], v[3]);
                const int k = (t >> 4) & 3;
                if ((t & 15) == 0 && k < g.np) {
                    const int id = k == 0 ? id0 : k == 1 ? id1 : k == 2 ? id2 : id3;
                    part[size_t(id) * row + slot] = c128{acc, 0.0};
                }
            }
        }
    }
    if (!last) {  // after an item's last reverse sweep only its gradient partials are read
        const u64 g = base | seg_gthread<RB>(t, sg);
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const u64 o = g | slot_goff<RB>(j, sg);
            store_nt(ys + o, y[j]);
            store_nt(ls + o, l[j]);
        }
    }
}

// Kernels per register-bit count: 4 bits, 256 threads (two waves per SIMD);
// 3 bits, 512 threads (four waves per SIMD).
#define MOSSQ_TILE_KERNELS(RB, TTHREADS, WAVES)                                                                      \
    __global__ __launch_bounds__(TTHREADS, WAVES) void k_tile##RB(                                                   \
        const c128 *src, c128 *dst, TSweepArgs sw, const TSeg *__restrict__ segs, const TGroup *__restrict__ groups, \
        const TDiag *__restrict__ diag, const c128 *__restrict__ pool) {                                             \
        tile_sweep<RB, false>(src, dst, sw, segs, groups, diag, pool, nullptr, nullptr, 0, nullptr, 0, blockIdx.x); \
    }                                                                                                                \
    /* the density adjoint's reverse chunks: Delta in place, each probe against */                                   \
    /* the state before its parameterized gate (ftab[probe - fbase]) */                                              \
    __global__ __launch_bounds__(TTHREADS, WAVES) void k_tile_fixp##RB(                                              \
        c128 *delta, TSweepArgs sw, const TSeg *__restrict__ segs, const TGroup *__restrict__ groups,                \
        const TDiag *__restrict__ diag, const c128 *__restrict__ pool, const TProbe *__restrict__ probes,            \
        const c128 *const *__restrict__ ftab, int fbase, c128 *__restrict__ part, int row) {                         \
        tile_sweep<RB, true>(delta, delta, sw, segs, groups, diag, pool, probes, ftab, fbase, part, row,            \
                             blockIdx.x);                                                                            \
    }                                                                                                                \
    __global__ __launch_bounds__(TTHREADS, WAVES) void k_tile_adj##RB(                                               \
        c128 *__restrict__ ys, c128 *__restrict__ ls, TSweepArgs sw, const TSeg *__restrict__ segs,                  \
        const TGroup *__restrict__ groups, const c128 *__restrict__ pool, const TProbe *__restrict__ probes,         \
        c128 *__restrict__ part, int row, int last) {                                                                \
        tile_adj_sweep<RB>(ys, ls, sw, segs, groups, pool, probes, part, row, blockIdx.x, last != 0);               \
    }
MOSSQ_TILE_KERNELS(4, 256, 2)
MOSSQ_TILE_KERNELS(3, 512, 4)

// ------------------------------------------------------ batched sweeps ----
// Sweeps of many programs in one launch: item i is one sweep of one program's
// plan on that program's own vectors, one workgroup per tile; wg_item gives
// each workgroup its item and tile (unsigned item << 16 | tile, both below
// 2^16; a launch holds at most 65536 items, all with plans of one register-bit
// count; tile0 is the item's first tile in a count of the launch's tiles, used
// on the host).  The vectors are offsets into one allocation `vec` (partials
// into `part`), which the host checks against that allocation's size.
// The plan tables' addresses become constant-memory pointers in the kernel
// (the tables are read-only while a sweep runs).  Generic pointers read from
// memory would load through the flat path, into vector registers; constant
// ones load uniform records as scalars, as kernel-argument tables do.
template <class T>
using cptr = __attribute__((address_space(4))) const T *;

template <class T>
__device__ __forceinline__ const T *table(u64 addr) {
    return (const T *)(cptr<T>)addr;
}

struct TBItem {
    u64 x;  // the state (forward sweeps, in place) or y (adjoint sweeps)
    u64 l;  // lambda (adjoint sweeps)
    u64 part;  // gradient partial rows (adjoint sweeps)
    u64 segs, groups, diag, pool, probes;  // plan tables (device addresses)
    int row;
    int tile0;
    int last;  // adjoint: the item's last reverse sweep (y and lambda are not stored)
    TSweepArgs args;
};

#define MOSSQ_BTILE_KERNELS(RB, TTHREADS, WAVES)                                                                    \
    __global__ __launch_bounds__(TTHREADS, WAVES) void k_btile##RB(                                                 \
        const TBItem *__restrict__ items, const unsigned *__restrict__ wg_item, c128 *__restrict__ vec) {            \
        const unsigned e = __builtin_amdgcn_readfirstlane(wg_item[blockIdx.x]);                                      \
        const TBItem &it = items[e >> 16];                                                                          \
        tile_sweep<RB, false>(vec + it.x, vec + it.x, it.args, table<TSeg>(it.segs), table<TGroup>(it.groups),      \
                              table<TDiag>(it.diag), table<c128>(it.pool), nullptr, nullptr, 0, nullptr, 0,        \
                              u64(e & 0xFFFF));                                                                     \
    }                                                                                                               \
    __global__ __launch_bounds__(TTHREADS, WAVES) void k_btile_adj##RB(const TBItem *__restrict__ items,             \
                                                                       const unsigned *__restrict__ wg_item,         \
                                                                       c128 *__restrict__ vec,                      \
                                                                       c128 *__restrict__ part) {                   \
        const unsigned e = __builtin_amdgcn_readfirstlane(wg_item[blockIdx.x]);                                      \
        const TBItem &it = items[e >> 16];                                                                          \
        tile_adj_sweep<RB>(vec + it.x, vec + it.l, it.args, table<TSeg>(it.segs), table<TGroup>(it.groups),         \
                           table<c128>(it.pool), table<TProbe>(it.probes), part + it.part, it.row,                  \
                           u64(e & 0xFFFF), it.last != 0);                                                          \
    }
MOSSQ_BTILE_KERNELS(4, 256, 2)
MOSSQ_BTILE_KERNELS(3, 512, 4)

// ------------------------------------------------------------------ planner ----
// kind 3 (adjoint plans): probe Re<y|D|l> of parameter `probe` on the two
// states at this point; it changes neither state.  A `fixed` probe reads a
// state the plan does not change (Re<v|D|fixed>), so it depends on every op
// before it and commutes with none.
struct POp {
    int kind = 0;  // 0 dense, 1 diagonal, 2 unfusable (run as its own launch), 3 probe
    int nb = 0;
    int bits[4] = {};
    std::vector<c128> m;  // dense / probe: 2^nb x 2^nb column-major; diagonal: 2^nb values
    int step = -1;
    int probe = -1;
    bool fixed = false;
    u64 mask() const {
        u64 r = 0;
        for (int i = 0; i < nb; ++i) r |= u64(1) << bits[i];
        return r;
    }
};

struct PlanItem {
    bool sweep = false;
    int step = -1;  // unfused step index
    TSweepArgs args{};
};

struct FusedPlan {
    std::vector<PlanItem> items;
    std::vector<TSeg> h_segs;
    std::vector<TOp> h_ops;
    std::vector<c128> h_pool;
    std::vector<TGroup> h_groups;  // device encoding of h_ops
    std::vector<TDiag> h_diag;
    std::vector<TProbe> h_probes;
    TSeg *d_segs = nullptr;
    TGroup *d_groups = nullptr;
    TDiag *d_diag = nullptr;
    c128 *d_pool = nullptr;
    TProbe *d_probes = nullptr;
    int rb = 4;            // register bits of every layout (4: 256 threads per tile, 3: 512)
    int min_tail = 0;      // planner variant the plan came from (rebuilt alike by
    bool kron = true;      // fused_plan_update)
    std::vector<int> reg_trace;  // every register set the scheduler picked, in order (rb each)
    bool adjoint = false;  // reverse plan of the pure-state adjoint (probes, no general diagonals)
    bool reverse_steps = false;  // unfused items apply the steps' reverse matrices
    bool shared_device = false;  // device arrays live in a block owned elsewhere
    int nsweeps = 0;
    int ntransitions = 0;
    int nconflicted = 0;  // transitions without a conflict-free LDS slot map
};

static bool is_diag(const MatInfo &info) {
    if (!info.monomial) return false;
    for (int j = 0; j < info.nchanged; ++j)
        if (info.rows[j] != info.src[j]) return false;
    return true;
}

static std::vector<c128> diag_values(const MatInfo &info, int dim) {
    std::vector<c128> d(dim, mk(1, 0));
    for (int j = 0; j < info.nchanged; ++j) d[info.rows[j]] = info.val[j];
    return d;
}

static std::vector<c128> mat_of(const Program *p, int off, int dim) {
    return std::vector<c128>(p->h_pool.begin() + off, p->h_pool.begin() + off + size_t(dim) * dim);
}

// Lower steps [lo, hi) to plan operations: in program order with the forward
// matrices, or (reverse) from hi-1 down to lo with the reverse matrices
// (U^dagger, S^dagger), as the density adjoint's chunk plans run them.
static void lower_range(const Program *p, int lo, int hi, bool reverse, std::vector<POp> &out) {
    for (int q = 0; q < hi - lo; ++q) {
        const int si = reverse ? hi - 1 - q : lo + q;
        const Step &st = p->steps[si];
        const MatInfo &info = reverse ? st.irev : st.ifwd;
        const int dim = 1 << st.kb;
        if (!st.split && is_diag(info) && st.kb <= 4) {
            if (info.nchanged == 0) continue;  // identity
            POp d;
            d.kind = 1;
            d.nb = st.kb;
            for (int i = 0; i < st.kb; ++i) d.bits[i] = st.bits[i];
            d.m = diag_values(info, dim);
            d.step = si;
            out.push_back(d);
            continue;
        }
        if (!st.split && st.kb <= 2) {
            POp d;
            d.kind = 0;
            d.nb = st.kb;
            for (int i = 0; i < st.kb; ++i) d.bits[i] = st.bits[i];
            d.m = mat_of(p, reverse ? st.rev : st.fwd, dim);
            d.step = si;
            out.push_back(d);
            continue;
        }
        if (p->is_density && !st.split && !st.channel && st.k == 2 && st.u_src >= 0) {
            // U(x)conj(U) = (U on the row bits) then (conj(U) on the column bits);
            // the reverse superoperator takes U^dagger in place of U.
            std::vector<c128> U = mat_of(p, st.u_src, 4), Ur(16), Uc(16);
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) Ur[r + 4 * c] = reverse ? cconj(U[c + 4 * r]) : U[r + 4 * c];
            for (int q2 = 0; q2 < 16; ++q2) Uc[q2] = cconj(Ur[q2]);
            POp r, cl;
            r.kind = cl.kind = 0;
            r.nb = cl.nb = 2;
            r.bits[0] = st.bits[0];
            r.bits[1] = st.bits[1];
            cl.bits[0] = st.bits[2];
            cl.bits[1] = st.bits[3];
            r.m = Ur;
            cl.m = Uc;
            r.step = cl.step = si;
            out.push_back(r);
            out.push_back(cl);
            continue;
        }
        POp u;
        u.kind = 2;
        u.step = si;
        out.push_back(u);
    }
}

static void lower_steps(const Program *p, std::vector<POp> &out) { lower_range(p, 0, int(p->steps.size()), false, out); }

static bool sign_only(const std::vector<c128> &d, int *negmask);

// A B == B A for column-major d x d matrices, to rounding.
static bool commute_mats(const std::vector<c128> &A, const std::vector<c128> &B, int d) {
    double scale = 0.0, diff = 0.0;
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) {
            c128 ab = mk(0, 0), ba = mk(0, 0);
            for (int k = 0; k < d; ++k) {
                ab = cfma(A[r + d * k], B[k + d * c], ab);
                ba = cfma(B[r + d * k], A[k + d * c], ba);
            }
            diff = std::max(diff, std::max(std::fabs(ab.re - ba.re), std::fabs(ab.im - ba.im)));
            scale = std::max(scale, std::max(std::hypot(A[r + d * c].re, A[r + d * c].im),
                                             std::hypot(B[r + d * c].re, B[r + d * c].im)));
        }
    return diff <= 1e-14 * std::max(scale * scale, 1e-300);
}

// Adjoint plan of a pure-state program: the steps in reverse order with
// their inverse matrices, each parameterized step followed by the probe of its
// derivative matrix D (differentiation/circuit.jl:136-144: y <- U^dag y,
// l <- U^dag l, g = Re<y|D|l>).  Wider or unfusable steps keep kind 2 and run
// as their own paired launch.
static void lower_steps_rev(const Program *p, std::vector<POp> &out) {
    for (int si = int(p->steps.size()) - 1; si >= 0; --si) {
        const Step &st = p->steps[si];
        const int dim = 1 << st.kb;
        // Diagonals the adjoint kernel takes as sign flips: -1 exactly at the
        // all-ones index (CZ-type) on at most 3 bits.  Other diagonals on <= 2
        // bits become dense matrices (one-bit ones then run as diagonal groups,
        // C_DIAG1); everything wider runs on its own.
        int neg = 0;
        const bool diag = !st.split && is_diag(st.irev);
        const bool cz_type = diag && sign_only(diag_values(st.irev, dim), &neg) && neg == (1 << (dim - 1)) &&
                             st.kb <= 3 && st.grad < 0;
        if (st.split || (st.kb > 2 && !cz_type)) {
            POp u;
            u.kind = 2;
            u.step = si;
            out.push_back(u);
            continue;
        }
        POp d;
        d.nb = st.kb;
        for (int i = 0; i < st.kb; ++i) d.bits[i] = st.bits[i];
        d.step = si;
        if (cz_type) {
            d.kind = 1;
            d.m = diag_values(st.irev, dim);
            out.push_back(d);
        } else if (!(diag && st.irev.nchanged == 0)) {
            d.kind = 0;
            d.m = mat_of(p, st.rev, dim);
            out.push_back(d);
        }
        if (st.grad >= 0) {
            POp pr;
            pr.kind = 3;
            pr.nb = st.kb;
            for (int i = 0; i < st.kb; ++i) pr.bits[i] = st.bits[i];
            pr.m = mat_of(p, st.grad, dim);
            pr.step = si;
            pr.probe = st.grad_index;
            out.push_back(pr);
        }
    }
    // A probe whose D commutes with its own step's reverse matrix U^dag takes
    // the same value before U^dag as after it: Re<U^dag y|D|U^dag l> =
    // Re<y|U D U^dag|l> = Re<y|D|l>.  Where the next op is a dense op on the
    // same bits, the probe moves ahead of its own op, so the two dense ops
    // meet and the segment peephole multiplies them into one (a rotation
    // layer's R_z^dag and R_y^dag on one qubit: one 2x2 instead of two).
    for (size_t i = 0; i + 2 < out.size(); ++i) {
        POp &A = out[i], &P = out[i + 1];
        const POp &B = out[i + 2];
        if (A.kind != 0 || P.kind != 3 || B.kind != 0 || P.step != A.step || A.mask() != P.mask() ||
            B.mask() != A.mask() || A.nb != P.nb || B.nb != A.nb)
            continue;
        if (commute_mats(A.m, P.m, 1 << A.nb)) std::swap(A, P);
    }
}

// out = A * B (column-major d x d)
static void product(const std::vector<c128> &A, const std::vector<c128> &B, int d, std::vector<c128> &out) {
    out.assign(size_t(d) * d, mk(0, 0));
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) {
            c128 acc = mk(0, 0);
            for (int k = 0; k < d; ++k) acc = cfma(A[r + d * k], B[k + d * c], acc);
            out[r + d * c] = acc;
        }
}

// kron(Bm, Am) with Am on the low bit: 4x4 on (a, b), a < b.
static std::vector<c128> kron2(const std::vector<c128> &Am, const std::vector<c128> &Bm) {
    std::vector<c128> out(16);
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[r + 4 * c] = cmul(Am[(r & 1) + 2 * (c & 1)], Bm[(r >> 1) + 2 * (c >> 1)]);
    return out;
}

// Kind (perm_src) of a 4x4 column-major matrix that is one of the slot
// permutations the tile kernels move without arithmetic, else -1.
static int perm_kind(const c128 *M) {
    int src[4];
    for (int r = 0; r < 4; ++r) {
        src[r] = -1;
        for (int c = 0; c < 4; ++c) {
            const c128 v = M[r + 4 * c];
            if (v.re == 1.0 && v.im == 0.0 && src[r] < 0)
                src[r] = c;
            else if (v.re != 0.0 || v.im != 0.0)
                return -1;
        }
        if (src[r] < 0) return -1;
    }
    for (int kind = 0; kind < 6; ++kind) {
        bool ok = true;
        for (int r = 0; r < 4; ++r) ok &= perm_src(r, 0, 1, kind) == src[r];
        if (ok) return kind;
    }
    return -1;
}

static bool is_perm(const POp &o) { return o.kind == 0 && o.nb == 2 && perm_kind(o.m.data()) >= 0; }

static int popc(u64 x) { return __builtin_popcountll(x); }

// diag(D) * M for a diagonal D on bits Dbits and dense M on Mbits (Dbits subset of Mbits).
static void absorb_diag(POp &M, const POp &D) {
    const int d = 1 << M.nb;
    for (int r = 0; r < d; ++r) {
        int di = 0;
        for (int i = 0; i < D.nb; ++i)
            for (int q = 0; q < M.nb; ++q)
                if (D.bits[i] == M.bits[q] && ((r >> q) & 1)) di |= 1 << i;
        for (int c = 0; c < d; ++c) M.m[r + d * c] = cmul(D.m[di], M.m[r + d * c]);
    }
}

// Peephole on one segment's op list (in execution order): merge dense ops on
// identical bits, absorb register-bit diagonals into the dense ops before them
// (pairing two one-bit gates into a two-bit gate when the diagonal spans both).
static void peephole(std::vector<POp> &ops, const int *reg_global, int nreg, bool kron = true, bool adjoint = false) {
    auto in_reg = [&](int gb) {
        for (int i = 0; i < nreg; ++i)
            if (reg_global[i] == gb) return true;
        return false;
    };
    bool changed = true;
    while (changed) {
        changed = false;
        for (size_t i = 0; i < ops.size() && !changed; ++i) {
            POp &D = ops[i];
            if (D.kind != 1 || D.nb > 2) continue;
            bool all_reg = true;
            for (int q = 0; q < D.nb; ++q) all_reg &= in_reg(D.bits[q]);
            if (!all_reg) continue;
            // last op before i touching each diagonal bit
            int last[2] = {-1, -1};
            for (int q = 0; q < D.nb; ++q)
                for (int j = int(i) - 1; j >= 0; --j)
                    if (ops[j].mask() & (u64(1) << D.bits[q])) {
                        last[q] = j;
                        break;
                    }
            if (D.nb == 1 || last[0] == last[1]) {
                int j = last[0];
                // a permutation stays arithmetic-free: the diagonal keeps its own op
                if (j >= 0 && ops[j].kind == 0 && !is_perm(ops[j]) && (ops[j].mask() & D.mask()) == D.mask()) {
                    absorb_diag(ops[j], D);
                    ops.erase(ops.begin() + i);
                    changed = true;
                }
                continue;
            }
            int ja = last[0], jb = last[1];
            if (ja < 0 || jb < 0) continue;
            POp &A = ops[ja], &B = ops[jb];
            if (A.kind != 0 || B.kind != 0 || A.nb != 1 || B.nb != 1) continue;
            // both one-bit gates are the last to touch their bit before D: fuse at max(ja, jb)
            POp F;
            F.kind = 0;
            F.nb = 2;
            F.bits[0] = D.bits[0];
            F.bits[1] = D.bits[1];
            F.m = kron2(A.bits[0] == D.bits[0] ? A.m : B.m, A.bits[0] == D.bits[0] ? B.m : A.m);
            absorb_diag(F, D);
            int keep = std::max(ja, jb), drop = std::min(ja, jb);
            ops[keep] = F;
            ops.erase(ops.begin() + i);
            ops.erase(ops.begin() + drop);
            changed = true;
        }
        // merge adjacent dense ops on identical bits, adjacent diagonals on identical bits
        for (size_t i = 1; i < ops.size(); ++i) {
            POp &a = ops[i - 1], &b = ops[i];
            if (a.kind == b.kind && a.kind != 2 && a.kind != 3 && a.mask() == b.mask() && a.nb == b.nb) {
                if (a.kind == 0) {
                    std::vector<c128> pr;
                    product(b.m, a.m, 1 << a.nb, pr);
                    a.m.swap(pr);
                } else {
                    for (size_t q = 0; q < a.m.size(); ++q) a.m[q] = cmul(a.m[q], b.m[q]);
                }
                ops.erase(ops.begin() + i);
                changed = true;
                break;
            }
        }
    }
    // Adjoint plans: a probe changes neither state, and an op on other bits
    // (applied to both states) leaves its value alone, so probes move ahead
    // of such ops and of other probes.  Gathered probes share a group (and
    // one wave reduction), and the dense ops they no longer separate can be
    // paired below.
    if (adjoint)
        for (size_t i = 1; i < ops.size(); ++i) {
            if (ops[i].kind != 3 || ops[i].fixed) continue;
            for (size_t j = i; j > 0; --j) {
                const POp &prev = ops[j - 1];
                const bool moves = prev.kind == 3 ? !prev.fixed : (prev.mask() & ops[j].mask()) == 0;
                if (!moves) break;
                std::swap(ops[j - 1], ops[j]);
            }
        }
    // Fold one-bit dense ops into the two-bit dense op they meet first on
    // their bit (same flops, one op fewer; not into permutations, which cost
    // no arithmetic), then pair the remaining one-bit dense ops on different
    // bits into Kronecker products: a two-bit op costs the same flops as two
    // one-bit ops but half the dispatch work.
    auto embed1 = [](const POp &one, const POp &two) {
        std::vector<c128> I = {mk(1, 0), mk(0, 0), mk(0, 0), mk(1, 0)};
        return one.bits[0] == two.bits[0] ? kron2(one.m, I) : kron2(I, one.m);
    };
    changed = true;
    while (changed) {
        changed = false;
        for (size_t i = 0; i < ops.size() && !changed; ++i) {
            if (ops[i].kind != 0) continue;
            size_t j = i + 1;
            while (j < ops.size() && (ops[j].mask() & ops[i].mask()) == 0) ++j;
            if (j >= ops.size() || ops[j].kind != 0) continue;
            POp &A = ops[i], &B = ops[j];
            if (is_perm(A) || is_perm(B)) continue;
            if (A.nb == 1 && B.nb == 2) {  // B * (A on one of B's bits)
                std::vector<c128> pr;
                product(B.m, embed1(A, B), 4, pr);
                B.m.swap(pr);
                ops.erase(ops.begin() + i);
                changed = true;
            } else if (A.nb == 2 && B.nb == 1) {  // (B on one of A's bits) * A, B moved left
                std::vector<c128> pr;
                product(embed1(B, A), A.m, 4, pr);
                A.m.swap(pr);
                ops.erase(ops.begin() + j);
                changed = true;
            }
        }
    }
    // (A diagonal one-bit op stays alone: adjoint plans run it as one complex
    // product per amplitude, cheaper than its share of a 4x4.)
    auto diag1 = [](const POp &o) { return o.m[1].re == 0 && o.m[1].im == 0 && o.m[2].re == 0 && o.m[2].im == 0; };
    for (size_t i = 0; kron && i < ops.size(); ++i) {
        if (ops[i].kind != 0 || ops[i].nb != 1 || diag1(ops[i])) continue;
        const int a = ops[i].bits[0];
        if (!in_reg(a)) continue;
        u64 touched = 0;
        for (size_t k = i + 1; k < ops.size(); ++k) {
            const POp &B = ops[k];
            const bool pairable = B.kind == 0 && B.nb == 1 && !diag1(B) && B.bits[0] != a && in_reg(B.bits[0]) &&
                                  ((u64(1) << a) & B.mask()) == 0;
            // the pair runs where B runs (A moves later past ops off its bit)
            // or, in adjoint plans, where A runs (B moves earlier past ops
            // off its bit, such as the probes that follow A)
            const bool at_b = pairable && (touched & ((u64(1) << a) | B.mask())) == 0;
            const bool at_a = pairable && adjoint && (touched & B.mask()) == 0;
            if (at_b || at_a) {
                POp F;
                F.kind = 0;
                F.nb = 2;
                const POp &lo = a < B.bits[0] ? ops[i] : B, &hi = a < B.bits[0] ? B : ops[i];
                F.bits[0] = lo.bits[0];
                F.bits[1] = hi.bits[0];
                F.m = kron2(lo.m, hi.m);
                F.step = ops[i].step;
                if (at_b) {
                    ops[k] = F;
                    ops.erase(ops.begin() + i);
                    --i;
                } else {
                    ops[i] = F;
                    ops.erase(ops.begin() + k);
                }
                break;
            }
            touched |= B.mask();
            if (!adjoint && (touched & (u64(1) << a))) break;
        }
    }
}

static bool sign_only(const std::vector<c128> &d, int *negmask) {
    int m = 0;
    for (size_t i = 0; i < d.size(); ++i) {
        if (d[i].im != 0.0 || (d[i].re != 1.0 && d[i].re != -1.0)) return false;
        if (d[i].re == -1.0) m |= 1 << i;
    }
    *negmask = m;
    return true;
}

static TSeg make_seg(const int *reg, int rb, int op0, int op1, const int *tbits) {
    TSeg sg{};
    for (int i = 0; i < kRegMax; ++i) sg.reg[i] = sg.gbit[i] = -1;
    for (int i = 0; i < rb; ++i) {
        sg.reg[i] = reg[i];
        sg.gbit[i] = tbits[reg[i]];
    }
    int k = 0;
    for (int b = 0; b < TM; ++b) {
        if (std::find(reg, reg + rb, b) != reg + rb) continue;
        sg.tpos[k] = b;
        sg.gpos[k] = tbits[b];
        ++k;
    }
    sg.op0 = op0;
    sg.op1 = op1;
    return sg;
}

// Layout of a sweep's loads and stores when no segment layout keeps the
// lowest tile bits on lanes: the top rb tile bits in registers.
static const int *store_layout(int rb) {
    static const int l4[kRegMax] = {8, 9, 10, 11}, l3[kRegMax] = {9, 10, 11, -1};
    return rb == 3 ? l3 : l4;
}

// Rank of up to 4 vectors over GF(2).
static int gf2_rank(int *v, int n) {
    int r = 0;
    for (int bit = 3; bit >= 0 && r < n; --bit) {
        int piv = -1;
        for (int i = r; i < n; ++i)
            if ((v[i] >> bit) & 1) piv = i;
        if (piv < 0) continue;
        std::swap(v[r], v[piv]);
        for (int i = 0; i < n; ++i)
            if (i != r && ((v[i] >> bit) & 1)) v[i] ^= v[r];
        ++r;
    }
    return r;
}

// Bank-group constraints of one side of a transposition: the slots of lanes
// 0..7 must differ in their low 3 bits (8 lanes x 16 B cover 32 banks), those
// of lanes 0..15 in their low 4 bits.  Columns not yet chosen (col < 0) are
// skipped, so the check also prunes partial assignments.
static bool lanes_ok(const TSeg &L, const int *col) {
    int v4[4], n4 = 0, v3[3], n3 = 0;
    for (int k = 0; k < 4; ++k) {
        const int c = col[L.tpos[k]];
        if (c < 0) continue;
        v4[n4++] = c & 15;
        if (k < 3) v3[n3++] = c & 7;
    }
    return gf2_rank(v4, n4) == n4 && gf2_rank(v3, n3) == n3;
}

// Columns col[b] = (1 << b) ^ x_b with x_b < 2^min(b, 4): a unit
// lower-triangular map of the tile-local index, so slots stay a bijection.
static bool swizzle_search(const TSeg &A, const TSeg &B, const std::vector<int> &vars, size_t i, int *col) {
    if (i == vars.size()) return true;
    const int b = vars[i];
    const int nx = 1 << std::min(b, 4);
    for (int x = 0; x < nx; ++x) {
        col[b] = (1 << b) ^ x;
        if (lanes_ok(A, col) && lanes_ok(B, col) && swizzle_search(A, B, vars, i + 1, col)) return true;
    }
    col[b] = -1;
    return false;
}

// Returns false when no conflict-free columns exist (the identity map is used).
static bool set_transpose(const TSeg &prev, TSeg &next, int rb) {
    int col[TM];
    std::vector<int> vars;
    for (int b = 0; b < TM; ++b) col[b] = 1 << b;
    for (const TSeg *L : {&prev, static_cast<const TSeg *>(&next)})
        for (int k = 0; k < 4; ++k) {
            const int b = L->tpos[k];
            if (b >= 3 && std::find(vars.begin(), vars.end(), b) == vars.end()) vars.push_back(b);
        }
    std::sort(vars.begin(), vars.end());
    for (int b : vars) col[b] = -1;
    const bool found = swizzle_search(prev, next, vars, 0, col);
    if (!found)
        for (int b = 0; b < TM; ++b) col[b] = 1 << b;  // plain layout (bank conflicts, still correct)
    next.transpose = 1;
    for (int k = 0; k < TM - rb; ++k) {
        next.wt[k] = col[prev.tpos[k]];
        next.rt[k] = col[next.tpos[k]];
    }
    for (int i = 0; i < rb; ++i) {
        next.wr[i] = col[prev.reg[i]];
        next.rr[i] = col[next.reg[i]];
    }
    return found;
}

static bool safe_layout(const int *reg, int rb) {
    for (int i = 0; i < rb; ++i)
        if (reg[i] < kCoalesceBits) return false;
    return true;
}

static int encode_ops(FusedPlan *fp, std::string *err);

// The ops one sweep over tile bits T executes, in program order, from the
// not-yet-done ops: an op runs when it fits (dense ops and probes need their
// bits in T; diagonals are applied from the index and fit any T) and no
// earlier op left undone fails to commute with it.  Undone ops are tracked per
// bit: bd[q] = an undone non-diagonal op on bit q (blocks every later op on
// q), bg[q] = an undone diagonal on q (blocks later non-diagonal ops on q);
// unfusable steps and fixed probes commute with nothing, so an undone one
// ends the scan.  Returns the work executed (2 per dense op or probe, 1 per
// diagonal); `out` (if given) receives the ops.
static int sweep_ops(const std::vector<POp> &ops, const std::vector<char> &done, size_t first, int nbits, u64 T,
                     std::vector<int> *out) {
    u64 bd = 0, bg = 0;
    const u64 all = nbits >= 64 ? ~u64(0) : (u64(1) << nbits) - 1;
    int work = 0;
    for (size_t i = first; i < ops.size(); ++i) {
        if (done[i]) continue;
        const POp &o = ops[i];
        const u64 m = o.mask();
        if (o.kind == 2 || o.fixed) {
            // commutes with nothing: runs only if nothing before it is undone
            if (o.kind != 2 && !bd && !bg && (m & ~T) == 0) {
                work += 2;
                if (out) out->push_back(int(i));
                continue;
            }
            break;
        }
        if (o.kind == 1) {
            if (m & bd) {
                bg |= m;
                continue;
            }
            work += 1;
            if (out) out->push_back(int(i));
            continue;
        }
        if ((m & (bd | bg)) || (m & ~T)) {
            bd |= m;
            if ((bd & all) == all) break;  // every later op is blocked
            continue;
        }
        work += 2;
        if (out) out->push_back(int(i));
    }
    return work;
}

// Tile bits of the next sweep: the lowest kCoalesceBits bits plus up to
// TM - kCoalesceBits more.  A first choice takes the bits of the undone dense
// ops in program order while they fit; then single-bit exchanges (one chosen
// bit out, one other in) are applied while one increases the work the sweep
// executes (a few rounds).  The exchanges let a sweep follow the diagonal
// front of layered circuits (CNOT ladders) instead of the bits that happen to
// come first.
static u64 choose_tile_bits(const std::vector<POp> &ops, const std::vector<char> &done, size_t first, int nbits) {
    const u64 coal = (u64(1) << kCoalesceBits) - 1;
    u64 T = coal, bd = 0, bg = 0;
    for (size_t i = first; i < ops.size() && popc(T) < TM; ++i) {
        if (done[i]) continue;
        const POp &o = ops[i];
        const u64 m = o.mask();
        if (o.kind == 2 || o.fixed) {
            if (o.kind != 2 && !bd && !bg && popc(T | m) <= TM) T |= m;
            break;
        }
        if (o.kind == 1) {
            if (m & bd) bg |= m;
            continue;
        }
        if (!(m & (bd | bg)) && popc(T | m) <= TM) {
            T |= m;
            continue;
        }
        bd |= m;
    }
    for (int b = 0; b < nbits && popc(T) < TM; ++b) T |= u64(1) << b;
    if (nbits <= TM) return T;
    int best = sweep_ops(ops, done, first, nbits, T, nullptr);
    for (int round = 0; round < 6; ++round) {
        u64 bestT = T;
        for (int b = kCoalesceBits; b < nbits; ++b) {
            if (!((T >> b) & 1)) continue;
            for (int c = kCoalesceBits; c < nbits; ++c) {
                if ((T >> c) & 1) continue;
                const u64 T2 = (T & ~(u64(1) << b)) | (u64(1) << c);
                const int w = sweep_ops(ops, done, first, nbits, T2, nullptr);
                if (w > best) {
                    best = w;
                    bestT = T2;
                }
            }
        }
        if (bestT == T) break;
        T = bestT;
    }
    return T;
}

// With `hint` (a plan of the same program structure), the tile bits of each
// sweep and every register set come from the hint instead of their searches,
// which are most of the planning time; the values (and everything built from
// them) are planned afresh, so the tables match the hint's whenever the
// structure does.
static FusedPlan *plan_with_tail(const Program *p, const std::vector<POp> &ops, int min_tail, bool adjoint, bool kron,
                                 int rb, std::string *err, const FusedPlan *hint = nullptr) {
    FusedPlan *fp = new FusedPlan();
    std::vector<const PlanItem *> hint_sweeps;
    if (hint)
        for (const PlanItem &it : hint->items)
            if (it.sweep) hint_sweeps.push_back(&it);
    size_t hint_reg = 0;
    fp->adjoint = adjoint;
    fp->rb = rb;
    fp->min_tail = min_tail;
    fp->kron = kron;
    const int tn = 1 << rb;
    std::vector<char> done(ops.size(), 0);
    size_t first = 0;
    auto failp = [&](const char *m) {
        *err = m;
        delete fp;
        return (FusedPlan *)nullptr;
    };
    while (true) {
        while (first < ops.size() && done[first]) ++first;
        if (first >= ops.size()) break;
        if (ops[first].kind == 2) {
            PlanItem it;
            it.sweep = false;
            it.step = ops[first].step;
            fp->items.push_back(it);
            done[first] = 1;
            continue;
        }
        u64 T = 0;
        if (hint) {
            if (size_t(fp->nsweeps) >= hint_sweeps.size()) return failp("the plan has more sweeps than its hint");
            for (int q = 0; q < TM; ++q) T |= u64(1) << hint_sweeps[fp->nsweeps]->args.tbits[q];
        } else {
            T = choose_tile_bits(ops, done, first, p->nbits);
        }
        std::vector<int> chosen;
        sweep_ops(ops, done, first, p->nbits, T, &chosen);
        for (int i : chosen) done[i] = 1;
        for (int b = 0; b < p->nbits && popc(T) < TM; ++b) T |= u64(1) << b;
        PlanItem it;
        it.sweep = true;
        int nt = 0, nc = 0;
        int local_of[64];
        for (int b = 0; b < p->nbits; ++b) {
            if (T & (u64(1) << b)) {
                local_of[b] = nt;
                it.args.tbits[nt++] = b;
            } else {
                local_of[b] = -1;
                it.args.cbits[nc++] = b;
            }
        }
        it.args.ncbits = nc;
        // Schedule the chosen ops into register segments.
        // Op ci waits for every earlier chosen op it does not commute with
        // (two ops commute when their bits are disjoint or both are diagonal;
        // unfusable steps and fixed ops never move past another op).  The
        // waits are kept as a transitive reduction with the same closure: per
        // bit, an op waits for the last non-diagonal op on that bit before it
        // and, if it is not diagonal itself, for the diagonals since then
        // (everything earlier on the bit precedes that last non-diagonal op).
        // Unfusable and fixed ops act as non-diagonal ops on one extra bit
        // that every other op holds as a diagonal.  pending[ci] counts the
        // waits not yet scheduled, so an op is ready exactly when every op
        // it does not commute with has been scheduled.
        std::vector<char> sched(chosen.size(), 0);
        std::vector<std::vector<int>> succ(chosen.size());
        std::vector<int> pending(chosen.size(), 0);
        std::vector<unsigned char> cdiag(chosen.size());
        std::vector<unsigned> lmask(chosen.size());  // tile-local bits; bit 31: a bit outside the tile
        {
            const int kAll = 64;  // the extra bit
            int last_nd[kAll + 1];
            std::vector<int> diags_since[kAll + 1];
            std::vector<int> stamp(chosen.size(), -1);
            for (int b = 0; b <= kAll; ++b) last_nd[b] = -1;
            for (size_t ci = 0; ci < chosen.size(); ++ci) {
                const POp &o = ops[chosen[ci]];
                const bool any = o.kind == 2 || o.fixed;
                cdiag[ci] = o.kind == 1;
                unsigned lm = 0;
                for (int q = 0; q < o.nb; ++q) lm |= local_of[o.bits[q]] >= 0 ? 1u << local_of[o.bits[q]] : 1u << 31;
                lmask[ci] = lm;
                auto wait_for = [&](int pj) {
                    if (stamp[pj] == int(ci)) return;
                    stamp[pj] = int(ci);
                    succ[pj].push_back(int(ci));
                    pending[ci]++;
                };
                auto on_bit = [&](int b, bool nd) {
                    if (last_nd[b] >= 0) wait_for(last_nd[b]);
                    if (nd) {
                        for (int pj : diags_since[b]) wait_for(pj);
                        diags_since[b].clear();
                        last_nd[b] = int(ci);
                    } else {
                        diags_since[b].push_back(int(ci));
                    }
                };
                for (int q = 0; q < o.nb; ++q) on_bit(o.bits[q], !cdiag[ci] || any);
                on_bit(kAll, any);
            }
        }
        auto mark = [&](size_t ci, std::vector<char> &s, std::vector<int> &pend) {
            s[ci] = 1;
            for (int x : succ[ci]) pend[x]--;
        };
        size_t nsched = 0;
        auto ready = [&](size_t ci) { return pending[ci] == 0; };
        auto reg_bits = [&](const int *reg) {
            unsigned r = 0;
            for (int i = 0; i < rb; ++i) r |= reg[i] >= 0 ? 1u << reg[i] : 1u << 31;
            return r;
        };
        // every bit of a non-diagonal op is one of the register bits
        auto fits = [&](size_t ci, unsigned rmask) { return cdiag[ci] || (lmask[ci] & ~rmask) == 0; };
        // Ops a segment with register set `reg` would schedule from the current state.
        // the chosen ops not yet scheduled, in order (compacted after each segment)
        std::vector<int> todo(chosen.size());
        for (size_t ci = 0; ci < chosen.size(); ++ci) todo[ci] = int(ci);
        auto simulate = [&](const int *reg) {
            std::vector<char> s2(sched);
            std::vector<int> p2(pending);
            const unsigned rm = reg_bits(reg);
            int count = 0;
            bool progress = true;
            while (progress) {
                progress = false;
                for (int ci : todo) {
                    if (s2[ci] || p2[ci] || !fits(ci, rm)) continue;
                    mark(ci, s2, p2);
                    count += ops[chosen[ci]].kind == 1 ? 1 : 2;
                    progress = true;
                }
            }
            return count;
        };
        // Next register set: candidates seeded by each ready dense op and grown
        // greedily with the bits of further ready ops; the candidate that
        // schedules the most work wins (ties: fewer bits changed).
        auto next_reg = [&](int *reg) {
            if (hint) {
                if (hint_reg + rb > hint->reg_trace.size()) return false;
                std::copy(hint->reg_trace.begin() + hint_reg, hint->reg_trace.begin() + hint_reg + rb, reg);
                hint_reg += rb;
                fp->reg_trace.insert(fp->reg_trace.end(), reg, reg + rb);
                return true;
            }
            std::vector<int> best;
            int best_score = -1, best_keep = -1;
            int seeds = 0;
            for (int c0 : todo) {
                if (seeds >= 16) break;
                if (sched[c0] || (ops[chosen[c0]].kind != 0 && ops[chosen[c0]].kind != 3) || !ready(c0)) continue;
                ++seeds;
                std::vector<int> want;
                auto try_add = [&](const POp &o) {
                    std::vector<int> add;
                    for (int q = 0; q < o.nb; ++q) {
                        int l = local_of[o.bits[q]];
                        if (std::find(want.begin(), want.end(), l) == want.end()) add.push_back(l);
                    }
                    if (want.size() + add.size() <= size_t(rb)) want.insert(want.end(), add.begin(), add.end());
                };
                try_add(ops[chosen[c0]]);
                for (int ci : todo) {
                    if (want.size() >= size_t(rb)) break;
                    if (!sched[ci] && ci != c0 && ops[chosen[ci]].kind != 1) try_add(ops[chosen[ci]]);
                }
                for (int i = 0; i < rb && want.size() < size_t(rb); ++i)  // keep current bits
                    if (std::find(want.begin(), want.end(), reg[i]) == want.end()) want.push_back(reg[i]);
                std::sort(want.begin(), want.end());
                int score = simulate(want.data());
                int keep = 0;
                for (int i = 0; i < rb; ++i) keep += std::find(want.begin(), want.end(), reg[i]) != want.end();
                if (score > best_score || (score == best_score && keep > best_keep)) {
                    best = want;
                    best_score = score;
                    best_keep = keep;
                }
            }
            if (best.empty()) {  // no dense op is ready: any layout works
                best.assign(reg, reg + rb);
                std::sort(best.begin(), best.end());
            }
            if (best.size() != size_t(rb) || best[0] < 0 || best[rb - 1] >= TM ||
                std::adjacent_find(best.begin(), best.end()) != best.end())
                return false;
            for (int i = 0; i < rb; ++i) reg[i] = best[i];
            fp->reg_trace.insert(fp->reg_trace.end(), reg, reg + rb);
            return true;
        };
        int reg0[kRegMax];
        std::copy(store_layout(rb), store_layout(rb) + kRegMax, reg0);
        if (!next_reg(reg0)) return failp("fusion planner produced an invalid register set");
        // Schedule: segments as (register set, chosen indices in execution order).
        struct SegPlan {
            int reg[kRegMax];
            std::vector<int> idx;
            int work;
        };
        std::vector<SegPlan> plan;
        while (nsched < chosen.size()) {
            SegPlan sp;
            memcpy(sp.reg, reg0, sizeof(reg0));
            sp.work = 0;
            const unsigned rm0 = reg_bits(reg0);
            bool progress = true;
            while (progress) {
                progress = false;
                for (int ci : todo) {
                    if (sched[ci] || !fits(ci, rm0) || !ready(ci)) continue;
                    mark(ci, sched, pending);
                    ++nsched;
                    sp.idx.push_back(ci);
                    sp.work += ops[chosen[ci]].kind == 1 ? 1 : 2;
                    progress = true;
                }
            }
            if (sp.idx.empty()) return failp("fusion planner made no progress");
            todo.erase(std::remove_if(todo.begin(), todo.end(), [&](int ci) { return sched[ci] != 0; }), todo.end());
            plan.push_back(sp);
            if (nsched >= chosen.size()) break;
            if (!next_reg(reg0)) return failp("fusion planner produced an invalid register set");
        }
        // A final segment with little work costs a transposition (and maybe the
        // one back to the store layout): its ops wait for the next sweep.
        while (plan.size() > 1 && plan.back().work < min_tail) {
            for (int ci : plan.back().idx) done[chosen[ci]] = 0;
            plan.pop_back();
        }
        it.args.seg0 = int(fp->h_segs.size());
        if (!safe_layout(plan.front().reg, rb)) {
            const int n0 = int(fp->h_ops.size());
            fp->h_segs.push_back(make_seg(store_layout(rb), rb, n0, n0, it.args.tbits));
        }
        for (const SegPlan &sp : plan) {
            const int *reg = sp.reg;
            const int op0 = int(fp->h_ops.size());
            std::vector<POp> seg_ops;
            for (int ci : sp.idx) seg_ops.push_back(ops[chosen[ci]]);
            int reg_global[kRegMax];
            for (int i = 0; i < rb; ++i) reg_global[i] = it.args.tbits[reg[i]];
            peephole(seg_ops, reg_global, rb, kron, adjoint);
            for (const POp &o : seg_ops) {
                TOp t{};
                t.moff = int(fp->h_pool.size());
                if (o.kind == 1) {
                    t.k = o.nb;
                    int ridx[4] = {-1, -1, -1, -1};
                    for (int q = 0; q < 4; ++q) t.qbit[q] = t.lbit[q] = -1;
                    for (int q = 0; q < o.nb; ++q) {
                        int l = local_of[o.bits[q]];
                        for (int i = 0; i < rb; ++i)
                            if (reg[i] == l && l >= 0) ridx[q] = i;
                        if (ridx[q] < 0) {
                            if (l >= 0)
                                t.lbit[q] = l;
                            else
                                t.qbit[q] = o.bits[q];
                        }
                    }
                    t.rmask = 0;
                    for (int q = 0; q < o.nb; ++q)
                        if (ridx[q] >= 0) t.rmask |= 1 << ridx[q];
                    t.rp_lo = t.rp_hi = 0;
                    for (int j = 0; j < tn; ++j) {
                        unsigned rp = 0;
                        for (int q = 0; q < o.nb; ++q)
                            if (ridx[q] >= 0 && ((j >> ridx[q]) & 1)) rp |= 1u << q;
                        if (j < 8)
                            t.rp_lo |= rp << (4 * j);
                        else
                            t.rp_hi |= rp << (4 * (j - 8));
                    }
                    int neg = 0;
                    if (sign_only(o.m, &neg)) {
                        t.type = T_SIGN;
                        t.negmask = neg;
                        t.moff = 0;
                    } else if (o.nb <= 2) {
                        t.type = T_SELECT;
                        for (int q = 0; q < 4; ++q) fp->h_pool.push_back(q < int(o.m.size()) ? o.m[q] : mk(1, 0));
                    } else {
                        t.type = T_GATHER;
                        fp->h_pool.insert(fp->h_pool.end(), o.m.begin(), o.m.end());
                    }
                } else {
                    fp->h_pool.insert(fp->h_pool.end(), o.m.begin(), o.m.end());
                    auto ridx = [&](int gb) {
                        int l = local_of[gb];
                        for (int i = 0; i < rb; ++i)
                            if (reg[i] == l) return i;
                        return -1;
                    };
                    t.probe = o.probe;
                    if (o.nb == 1) {
                        t.type = o.kind == 3 ? T_PROBE1 : T_DENSE1;
                        t.a = ridx(o.bits[0]);
                        if (t.a < 0) return failp("dense op outside the register bits");
                    } else {
                        t.type = o.kind == 3 ? T_PROBE2 : T_DENSE2;
                        t.a = ridx(o.bits[0]);
                        t.b = ridx(o.bits[1]);
                        if (t.a < 0 || t.b <= t.a) return failp("dense op outside the register bits");
                    }
                }
                fp->h_ops.push_back(t);
            }
            fp->h_segs.push_back(make_seg(reg, rb, op0, int(fp->h_ops.size()), it.args.tbits));
        }
        if (!safe_layout(fp->h_segs.back().reg, rb)) {
            const int n1 = int(fp->h_ops.size());
            fp->h_segs.push_back(make_seg(store_layout(rb), rb, n1, n1, it.args.tbits));
        }
        it.args.seg1 = int(fp->h_segs.size());
        if (getenv("MOSSQ_PLAN_DEBUG")) {
            fprintf(stderr, "sweep %d tbits:", fp->nsweeps);
            for (int q = 0; q < TM; ++q) fprintf(stderr, " %d", it.args.tbits[q]);
            fprintf(stderr, "\n");
            for (int q = it.args.seg0; q < it.args.seg1; ++q) {
                const TSeg &g = fp->h_segs[q];
                int changed = 0;
                if (q > it.args.seg0)
                    for (int i = 0; i < rb; ++i) {
                        bool found = false;
                        for (int k = 0; k < rb; ++k) found |= fp->h_segs[q - 1].reg[k] == g.reg[i];
                        changed += !found;
                    }
                int nd1 = 0, nd2 = 0, ndg = 0;
                for (int o = g.op0; o < g.op1; ++o) {
                    const int ty = fp->h_ops[o].type;
                    (ty == T_DENSE1 ? nd1 : ty == T_DENSE2 ? nd2 : ndg)++;
                }
                fprintf(stderr, "  seg reg {%d,%d,%d,%d} ops %d changed %d dense1 %d dense2 %d diag %d\n", g.reg[0],
                        g.reg[1], g.reg[2], g.reg[3], g.op1 - g.op0, changed, nd1, nd2, ndg);
                for (int o = g.op0; o < g.op1; ++o) {
                    const TOp &op = fp->h_ops[o];
                    if (op.type == T_DENSE1 || op.type == T_DENSE2) continue;
                    int nr = 0, nl = 0, nq = 0;
                    for (int q = 0; q < op.k; ++q) {
                        if (op.lbit[q] >= 0)
                            ++nl;
                        else if (op.qbit[q] >= 0)
                            ++nq;
                        else
                            ++nr;
                    }
                    fprintf(stderr, "    diag type %d k %d reg %d lane %d out %d neg %x\n", op.type, op.k, nr, nl, nq,
                            op.negmask);
                }
            }
        }
        for (int q = it.args.seg0 + 1; q < it.args.seg1; ++q)
            if (memcmp(fp->h_segs[q].reg, fp->h_segs[q - 1].reg, sizeof(int) * rb) != 0) {
                if (!set_transpose(fp->h_segs[q - 1], fp->h_segs[q], rb)) fp->nconflicted++;
                fp->ntransitions++;
            }
        fp->items.push_back(it);
        fp->nsweeps++;
    }
    if (fp->h_pool.empty()) fp->h_pool.push_back(mk(0, 0));
    if (encode_ops(fp, err) != MOSSQ_OK) {
        delete fp;
        return nullptr;
    }
    return fp;
}

// Packed form of a controlled-Z-type sign flip (a sign-only diagonal whose
// only -1 is at the all-ones index): its slot pattern and condition bits.
// False for other diagonals or more than three bits outside the registers.
static bool pack_flip(const TOp &o, int tn, unsigned *pat, int *cond) {
    if (o.type != T_SIGN || o.negmask != (1 << ((1 << o.k) - 1))) return false;
    int w = 0, n = 0;
    for (int q = 0; q < o.k; ++q) {
        int f;
        if (o.lbit[q] >= 0)
            f = o.lbit[q];
        else if (o.qbit[q] >= 0)
            f = 0x40 | o.qbit[q];
        else
            continue;
        if (n == 3) return false;
        w |= f << (7 * n++);
    }
    for (; n < 3; ++n) w |= kCzNone << (7 * n);
    unsigned pm = 0;
    for (int j = 0; j < tn; ++j)
        if ((j & o.rmask) == o.rmask) pm |= 1u << j;
    *pat = pm;
    *cond = w;
    return true;
}

// Device encoding of the planned ops: per segment, groups (one out-of-place
// step each).  Diagonals commute with each other, so the flips after a dense op
// are packed into its group in any order.
static int encode_ops(FusedPlan *fp, std::string *err) {
    fp->h_groups.clear();
    fp->h_diag.clear();
    fp->h_probes.clear();
    for (TSeg &sg : fp->h_segs) {
        sg.g0 = int(fp->h_groups.size());
        auto add = [&](int code, int moff) {
            TGroup g{};
            g.code = code;
            g.moff = moff;
            fp->h_groups.push_back(g);
        };
        auto pair_code = [&](const TOp &o, int *code) {
            for (int q = 0; q < 6; ++q)
                if (kPairA[q] == o.a && kPairB[q] == o.b) {
                    *code = C_DENSE2 + q;
                    return true;
                }
            *err = "dense op on an invalid register pair";
            return false;
        };
        for (int oi = sg.op0; oi < sg.op1; ++oi) {
            const TOp &o = fp->h_ops[oi];
            if (o.type == T_DENSE1) {
                const c128 *M = fp->h_pool.data() + o.moff;
                const bool diag2 = M[1].re == 0.0 && M[1].im == 0.0 && M[2].re == 0.0 && M[2].im == 0.0;
                add((fp->adjoint && diag2 ? C_DIAG1 : C_DENSE1) + o.a, o.moff);
                continue;
            }
            if (o.type == T_DENSE2) {
                int code;
                if (!pair_code(o, &code)) return MOSSQ_EINVAL;
                const int kind = perm_kind(fp->h_pool.data() + o.moff);
                if (kind >= 0) code = C_PERM + 6 * (code - C_DENSE2) + kind;
                add(code, o.moff);
                continue;
            }
            const bool here = int(fp->h_groups.size()) > sg.g0;
            if (o.type == T_PROBE1 || o.type == T_PROBE2) {
                // probes join the group before them (adjoint plans: up to
                // kGroupProbes, summed together; fixed probes: one)
                TGroup *g = here ? &fp->h_groups.back() : nullptr;
                if (!g || g->code == C_DIAG || g->np >= (fp->adjoint ? kGroupProbes : 1)) {
                    add(C_COPY, 0);
                    g = &fp->h_groups.back();
                }
                TProbe pr{};
                if (o.type == T_PROBE1) {
                    pr.pcode = C_DENSE1 + o.a;
                } else if (!pair_code(o, &pr.pcode)) {
                    return MOSSQ_EINVAL;
                }
                pr.probe = o.probe;
                pr.pmoff = o.moff;
                // D = (dU)^dagger U formed as a product leaves rounding
                // residues (~1e-17) where the exact generator is zero (R_y:
                // i Y / 2, R_z: i Z / 2); parts below 64 eps of D's largest
                // entry count as zero, so a rotation's probe takes its two-part
                // form.  The dropped terms change a gradient by ~1e-17.
                const int ne = o.type == T_PROBE1 ? 4 : 16;
                double dmax = 0.0;
                for (int e = 0; e < ne; ++e)
                    dmax = std::max(dmax, std::max(std::fabs(fp->h_pool[o.moff + e].re),
                                                   std::fabs(fp->h_pool[o.moff + e].im)));
                const double tiny = 64.0 * 2.220446049250313e-16 * dmax;
                int nc = 0;
                for (int e = 0; e < ne; ++e) {
                    const c128 d = fp->h_pool[o.moff + e];
                    if (std::fabs(d.re) > tiny) {
                        pr.pmask |= 1u << (2 * e);
                        if (nc < 2) pr.coef[nc++] = d.re;
                    }
                    if (std::fabs(d.im) > tiny) {
                        pr.pmask |= 1u << (2 * e + 1);
                        if (nc < 2) pr.coef[nc++] = -d.im;
                    }
                }
                if (o.type == T_PROBE1 && pr.pmask == 0x82u)
                    pr.pcode |= PK_IMDIAG << 8;  // parts (0, 0) and (1, 1), imaginary
                else if (o.type == T_PROBE1 && pr.pmask == 0x14u)
                    pr.pcode |= PK_REANTI << 8;  // parts (1, 0) and (0, 1), real
                if (g->np == 0) g->p0 = int(fp->h_probes.size());
                fp->h_probes.push_back(pr);
                g->np++;
                continue;
            }
            if (o.type == T_SIGN && o.negmask == 0) continue;  // identity (e.g. two equal CZs merged)
            unsigned pat;
            int cond;
            if (pack_flip(o, 1 << fp->rb, &pat, &cond)) {
                TGroup *g = here ? &fp->h_groups.back() : nullptr;
                if (!g || g->code == C_DIAG || g->code >= C_DIAG1 || g->ncz == kGroupCz || g->np > 0) {
                    add(C_COPY, 0);
                    g = &fp->h_groups.back();
                }
                if (g->code < C_FLIP)
                    g->code += C_FLIP;
                else if (g->code >= C_PERM && g->code < C_PERM + kPermFlip)
                    g->code += kPermFlip;
                g->pat[g->ncz] = pat;
                g->cond[g->ncz] = cond;
                g->ncz++;
                continue;
            }
            TDiag d{};
            d.code = o.type == T_SIGN ? D_SIGN : o.type == T_SELECT ? D_SELECT : D_GATHER;
            for (int q = 0; q < 4; ++q) {
                d.lbit[q] = o.lbit[q];
                d.qbit[q] = o.qbit[q];
            }
            d.rp_lo = o.rp_lo;
            d.rp_hi = o.rp_hi;
            d.negmask = o.negmask;
            d.moff = o.moff;
            if (o.type == T_SELECT)
                for (int q = 0; q < 4; ++q) d.val[q] = fp->h_pool[o.moff + q];
            if (fp->adjoint) {
                *err = "adjoint plan holds a general diagonal";
                return MOSSQ_EINVAL;
            }
            add(C_DIAG, 0);  // a group of its own: diagonals never follow a probe inside a group
            fp->h_groups.back().x = int(fp->h_diag.size());
            fp->h_diag.push_back(d);
        }
        if ((fp->h_groups.size() - sg.g0) & 1) add(C_COPY, 0);  // groups run in pairs
        sg.g1 = int(fp->h_groups.size());
    }
    if (fp->h_groups.empty()) fp->h_groups.push_back(TGroup{});
    if (fp->h_diag.empty()) fp->h_diag.push_back(TDiag{});
    // padding: a group reads kGroupProbes records from its first probe on
    fp->h_probes.resize(fp->h_probes.size() + kGroupProbes, TProbe{});
    return MOSSQ_OK;
}

// Modelled cost of a plan, in microseconds per sweep-equivalent at 26 qubits
// (k_tile timings on MI355X): a sweep's HBM round trip, an LDS transposition,
// and each tile op by type.
static double plan_cost(const FusedPlan *fp) {
    double t = 300.0 * fp->nsweeps + 87.0 * fp->ntransitions;
    for (const TOp &o : fp->h_ops)
        t += o.type == T_DENSE2 ? (perm_kind(fp->h_pool.data() + o.moff) >= 0 ? 6.0 : 27.0)
             : o.type == T_DENSE1 ? 20.0
                                  : 10.0;
    for (const PlanItem &it : fp->items)
        if (!it.sweep) t += 300.0;
    return t;
}

// The tail threshold trades a transposition against the ops it carries into
// the next sweep; which side wins depends on the circuit, so each candidate is
// planned and the cheapest plan under the cost model is kept.
static FusedPlan *best_plan(const Program *p, const std::vector<POp> &ops, bool adjoint, int rb, std::string *err) {
    if (const char *mt = getenv("MOSSQ_PLAN_MIN_TAIL")) return plan_with_tail(p, ops, atoi(mt), adjoint, true, rb, err);
    // One tile holds the whole state: every op fits the first sweep, and a
    // tail threshold could only push ops into an extra pass.
    if (p->nbits <= TM) return plan_with_tail(p, ops, 0, adjoint, true, rb, err);
    FusedPlan *best = nullptr;
    double best_cost = 0;
    for (int min_tail : {0, 3, 5, 8}) {
        FusedPlan *fp = plan_with_tail(p, ops, min_tail, adjoint, true, rb, err);
        if (!fp) {
            delete best;
            return nullptr;
        }
        // a one-sweep plan (short chunks of the density adjoint, small
        // circuits) cannot be beaten by moving ops into further sweeps
        if (min_tail == 0 && fp->nsweeps <= 1) return fp;
        const double c = plan_cost(fp);
        if (!best || c < best_cost) {
            delete best;
            best = fp;
            best_cost = c;
        } else {
            delete fp;
        }
    }
    return best;
}

// Register bits of the plans a program gets.
static int plan_rb(bool adjoint) {
    const char *e = getenv(adjoint ? "MOSSQ_TILE_RB_ADJ" : "MOSSQ_TILE_RB_FWD");
    return e && atoi(e) == 3 ? 3 : 4;
}

FusedPlan *fused_plan_build_host(const Program *p, std::string *err, bool adjoint, int rb) {
    std::vector<POp> ops;
    if (adjoint)
        lower_steps_rev(p, ops);
    else
        lower_steps(p, ops);
    return best_plan(p, ops, adjoint, rb > 0 ? rb : plan_rb(adjoint), err);
}

// Host execution of a plan with the kernel's index arithmetic (tiles, layouts,
// LDS transpositions, bit sources); unfused items use the oracle-free generic
// block update.
static void host_block(c128 *v, int nbits, int k, const int *bits, const c128 *m) {
    const int d = 1 << k;
    const u64 nb = u64(1) << (nbits - k);
    std::vector<c128> in(d);
    for (u64 t = 0; t < nb; ++t) {
        u64 base = insert0s_dyn(t, bits, k);
        for (int c = 0; c < d; ++c) {
            u64 o = 0;
            for (int j = 0; j < k; ++j)
                if (c & (1 << j)) o |= u64(1) << bits[j];
            in[c] = v[base + o];
        }
        for (int r = 0; r < d; ++r) {
            c128 acc = mk(0, 0);
            for (int c = 0; c < d; ++c) acc = cfma(m[r + d * c], in[c], acc);
            u64 o = 0;
            for (int j = 0; j < k; ++j)
                if (r & (1 << j)) o |= u64(1) << bits[j];
            v[base + o] = acc;
        }
    }
}

// One tile of a sweep for one state (forward plans) or the two states y, l of
// an adjoint plan (probes add Re<y|D|l> contributions to grads).
template <int RB>
static void emulate_tile_rb(const std::vector<c128 *> &states, const TSweepArgs &sw, const FusedPlan *fp, u64 tile,
                            c128 *grads, const c128 *const *ftab, int fbase) {
    constexpr int TN = TGeo<RB>::TN, TT = TGeo<RB>::TT;
    const size_t ns = states.size();
    std::vector<c128> lds(1 << TM);
    std::vector<std::vector<c128>> V(ns, std::vector<c128>(size_t(TT) * TN));
    u64 base = 0;
    for (int i = 0; i < sw.ncbits; ++i)
        if ((tile >> i) & 1) base |= u64(1) << sw.cbits[i];
    // Slot j's register-bit offsets, as the kernel forms them.
    auto slot_g = [](int j, const TSeg &sg) { return slot_goff<RB>(j, sg); };
    const TSeg *lay = &fp->h_segs[sw.seg0];
    std::vector<int> tdep(TT);
    for (int t = 0; t < TT; ++t) {
        tdep[t] = seg_tdep<RB>(t, *lay);
        const u64 gth = base | seg_gthread<RB>(t, *lay);
        for (size_t q = 0; q < ns; ++q)
            for (int j = 0; j < TN; ++j) V[q][t * TN + j] = states[q][gth | slot_g(j, *lay)];
    }
    auto dense_host = [](int dense, const c128 *M, c128 *v) {
        if (dense < C_DENSE2) {
            const int a = dense - C_DENSE1;
            for (int j = 0; j < TN; ++j) {
                if (j & (1 << a)) continue;
                c128 x = v[j], y = v[j | (1 << a)];
                v[j] = cfma(M[2], y, cmul(M[0], x));
                v[j | (1 << a)] = cfma(M[3], y, cmul(M[1], x));
            }
        } else if (dense < C_DENSE2 + 6) {
            const int a = kPairA[dense - C_DENSE2], b = kPairB[dense - C_DENSE2];
            for (int j = 0; j < TN; ++j) {
                if (j & ((1 << a) | (1 << b))) continue;
                int ii[4] = {j, j | (1 << a), j | (1 << b), j | (1 << a) | (1 << b)};
                c128 in[4] = {v[ii[0]], v[ii[1]], v[ii[2]], v[ii[3]]};
                for (int r = 0; r < 4; ++r) {
                    c128 acc = cmul(M[r], in[0]);
                    for (int c = 1; c < 4; ++c) acc = cfma(M[r + 4 * c], in[c], acc);
                    v[ii[r]] = acc;
                }
            }
        }
    };
    for (int s = sw.seg0; s < sw.seg1; ++s) {
        const TSeg &sg = fp->h_segs[s];
        if (sg.transpose) {
            int a[TN];
            for (size_t q = 0; q < ns; ++q) {
                std::vector<char> written(lds.size(), 0);
                for (int t = 0; t < TT; ++t) {
                    lds_slots<RB>(t, sg.wt, sg.wr, a);
                    for (int j = 0; j < TN; ++j) {
                        lds[a[j] / sizeof(c128)] = V[q][t * TN + j];
                        written[a[j] / sizeof(c128)] = 1;
                    }
                }
                if (std::count(written.begin(), written.end(), 1) != int(lds.size()))
                    fprintf(stderr, "mossq: LDS slot map is not a bijection\n");
                for (int t = 0; t < TT; ++t) {
                    lds_slots<RB>(t, sg.rt, sg.rr, a);
                    for (int j = 0; j < TN; ++j) V[q][t * TN + j] = lds[a[j] / sizeof(c128)];
                }
            }
            for (int t = 0; t < TT; ++t) tdep[t] = seg_tdep<RB>(t, sg);
        }
        // Executes the device encoding (groups, packed flips, diagonal records,
        // probes), as the kernels do.
        for (int gi = sg.g0; gi < sg.g1; ++gi) {
            const TGroup &g = fp->h_groups[gi];
            const c128 *M = fp->h_pool.data() + g.moff;
            const bool perm = g.code >= C_PERM && g.code < C_DIAG1;
            const int pk = perm ? (g.code - C_PERM) % kPermFlip : 0;
            const int dense = perm                  ? -1
                              : g.code >= C_DIAG1   ? C_DENSE1 + (g.code - C_DIAG1)
                              : g.code >= C_FLIP && g.code < C_COPY ? g.code - C_FLIP
                                                                     : g.code;
            const bool flips = perm ? g.code - C_PERM >= kPermFlip : g.code >= C_FLIP && g.code <= C_COPY;
            for (int t = 0; t < TT; ++t) {
                for (size_t q = 0; q < ns; ++q) {
                    c128 *v = &V[q][t * TN];
                    if (perm) {
                        c128 in[TN];
                        std::copy(v, v + TN, in);
                        for (int j = 0; j < TN; ++j) v[j] = in[perm_src(j, kPairA[pk / 6], kPairB[pk / 6], pk % 6)];
                    } else {
                        dense_host(dense, M, v);
                    }
                    if (flips) {
                        const unsigned F = group_flips(g, tdep[t], base);
                        for (int j = 0; j < TN; ++j)
                            if ((F >> j) & 1) v[j] = mk(-v[j].re, -v[j].im);
                    }
                    if (g.code == C_DIAG) {
                        const TDiag &d = fp->h_diag[g.x];
                        for (int j = 0; j < TN; ++j) {
                            int idx = diag_f(d, tdep[t], base) | diag_rp(d, j);
                            if (d.code == D_SIGN) {
                                if ((d.negmask >> idx) & 1) v[j] = mk(-v[j].re, -v[j].im);
                            } else if (d.code == D_SELECT) {
                                v[j] = cmul(d.val[idx], v[j]);
                            } else {
                                v[j] = cmul(fp->h_pool[d.moff + idx], v[j]);
                            }
                        }
                    }
                }
                for (int k = 0; k < g.np && grads && (ns == 2 || ftab); ++k) {
                    // adjoint plans: Re<y|D|l>; fixed probes: Re<v|D|fixed>, fixed
                    // read in the current layout
                    const TProbe &pr = fp->h_probes[g.p0 + k];
                    const c128 *y = &V[0][t * TN], *D = fp->h_pool.data() + pr.pmoff;
                    std::vector<c128> dl(TN);
                    if (ns == 2) {
                        dl.assign(V[1].begin() + t * TN, V[1].begin() + (t + 1) * TN);
                    } else {
                        const u64 gth = base | seg_gthread<RB>(t, sg);
                        const c128 *fixed = ftab[pr.probe - fbase];
                        for (int j = 0; j < TN; ++j) dl[j] = fixed[gth | slot_g(j, sg)];
                    }
                    dense_host(pr.pcode & 0xFF, D, dl.data());
                    c128 acc = mk(0, 0);
                    for (int j = 0; j < TN; ++j) acc = cfmaconj(y[j], dl[j], acc);
                    grads[pr.probe] = cadd(grads[pr.probe], acc);
                }
            }
        }
    }
    const TSeg &last = fp->h_segs[sw.seg1 - 1];
    for (int t = 0; t < TT; ++t) {
        const u64 gth = base | seg_gthread<RB>(t, last);
        for (size_t q = 0; q < ns; ++q)
            for (int j = 0; j < TN; ++j) states[q][gth | slot_g(j, last)] = V[q][t * TN + j];
    }
}

static void emulate_tile(const std::vector<c128 *> &states, const TSweepArgs &sw, const FusedPlan *fp, u64 tile,
                         c128 *grads, const c128 *const *ftab = nullptr, int fbase = 0) {
    if (fp->rb == 3)
        emulate_tile_rb<3>(states, sw, fp, tile, grads, ftab, fbase);
    else
        emulate_tile_rb<4>(states, sw, fp, tile, grads, ftab, fbase);
}

void emulate_plan(const Program *p, const FusedPlan *fp, c128 *state, const c128 *const *ftab = nullptr,
                  int fbase = 0, c128 *grads = nullptr) {
    for (const PlanItem &it : fp->items) {
        if (!it.sweep) {
            const Step &s = p->steps[it.step];
            const c128 *P = p->h_pool.data();
            const bool r = fp->reverse_steps;
            if (s.split) {
                host_block(state, p->nbits, s.k, s.bits, P + (r ? s.rev : s.fwd));
                host_block(state, p->nbits, s.k, s.bits + s.k, P + (r ? s.rev_cols : s.fwd_cols));
            } else {
                host_block(state, p->nbits, s.kb, s.bits, P + (r ? s.rev : s.fwd));
            }
            continue;
        }
        const u64 ntiles = u64(1) << (p->nbits - TM);
        for (u64 t = 0; t < ntiles; ++t) emulate_tile({state}, it.args, fp, t, grads, ftab, fbase);
    }
}

// The adjoint plan's reverse pass on host vectors y, l; grads[j] receives
// the sum of Re<y|D|l> over parameter j's probes (complex sums).
void emulate_adjoint_plan(const Program *p, const FusedPlan *fp, c128 *y, c128 *l, c128 *grads) {
    const c128 *P = p->h_pool.data();
    const u64 len = u64(1) << p->nbits;
    for (const PlanItem &it : fp->items) {
        if (!it.sweep) {
            const Step &s = p->steps[it.step];
            host_block(y, p->nbits, s.kb, s.bits, P + s.rev);
            host_block(l, p->nbits, s.kb, s.bits, P + s.rev);
            if (s.grad >= 0) {
                std::vector<c128> dl(l, l + len);
                host_block(dl.data(), p->nbits, s.kb, s.bits, P + s.grad);
                c128 acc = mk(0, 0);
                for (u64 i = 0; i < len; ++i) acc = cfmaconj(y[i], dl[i], acc);
                grads[s.grad_index] = cadd(grads[s.grad_index], acc);
            }
            continue;
        }
        const u64 ntiles = u64(1) << (p->nbits - TM);
        for (u64 t = 0; t < ntiles; ++t) emulate_tile({y, l}, it.args, fp, t, grads);
    }
}

int fused_plan_upload(Ctx *c, FusedPlan *fp) {
    int s = cached_alloc(c, std::max<size_t>(fp->h_segs.size(), 1) * sizeof(TSeg), (void **)&fp->d_segs);
    if (s == MOSSQ_OK) s = cached_alloc(c, fp->h_groups.size() * sizeof(TGroup), (void **)&fp->d_groups);
    if (s == MOSSQ_OK) s = cached_alloc(c, std::max<size_t>(fp->h_diag.size(), 1) * sizeof(TDiag), (void **)&fp->d_diag);
    if (s == MOSSQ_OK) s = cached_alloc(c, fp->h_pool.size() * sizeof(c128), (void **)&fp->d_pool);
    if (s == MOSSQ_OK) s = cached_alloc(c, fp->h_probes.size() * sizeof(TProbe), (void **)&fp->d_probes);
    if (s != MOSSQ_OK) return s;
    if (!fp->h_segs.empty())
        MOSSQ_HIP(hipMemcpyAsync(fp->d_segs, fp->h_segs.data(), fp->h_segs.size() * sizeof(TSeg),
                                 hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipMemcpyAsync(fp->d_groups, fp->h_groups.data(), fp->h_groups.size() * sizeof(TGroup), hipMemcpyHostToDevice,
                             c->stream));
    MOSSQ_HIP(hipMemcpyAsync(fp->d_diag, fp->h_diag.data(), fp->h_diag.size() * sizeof(TDiag),
                             hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipMemcpyAsync(fp->d_pool, fp->h_pool.data(), fp->h_pool.size() * sizeof(c128), hipMemcpyHostToDevice,
                             c->stream));
    MOSSQ_HIP(hipMemcpyAsync(fp->d_probes, fp->h_probes.data(), fp->h_probes.size() * sizeof(TProbe),
                             hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    return MOSSQ_OK;
}

// Forward plan, and for pure states the adjoint (reverse) plan.
// The 12-qubit form of a small pure-state program (Program::pad); none when
// its plans are not all sweeps (the engine then runs it gate by gate).
static void pad_build(Program *p) {
    auto *q = new Program();
    q->ctx = p->ctx;
    q->n = q->nbits = TM;
    q->nderiv = p->nderiv;
    q->steps = p->steps;
    q->h_pool = p->h_pool;
    q->d_pool = p->d_pool;
    if (fused_plan_build(q) == MOSSQ_OK && fused_plan_batchable(q, q->nderiv > 0)) {
        p->pad = q;
        return;
    }
    fused_plan_free(q);
    delete q;
}

static std::mutex g_pad_mu;  // first-use builds of Program::pad

const Program *fused_pad(const Program *p) {
    if (!p->fuse_requested || p->is_density || p->nbits >= TM) return nullptr;
    std::lock_guard<std::mutex> lk(g_pad_mu);
    if (!p->pad_tried) {
        pad_build(const_cast<Program *>(p));
        p->pad_tried = true;
    }
    return p->pad;
}

int fused_plan_build(Program *p) {
    p->fuse_requested = true;
    if (p->nbits < TM) return MOSSQ_OK;  // too small to tile: run unfused (padded in batches, fused_pad)
    if (p->is_density && p->nderiv > 0) {
        int s = density_adjoint_plan_build(p);
        if (s != MOSSQ_OK) return s;
    }
    for (int adj = 0; adj < (p->is_density ? 1 : 2); ++adj) {
        std::string err;
        FusedPlan *fp = fused_plan_build_host(p, &err, adj != 0, 0);
        if (!fp && adj) break;  // no adjoint plan: the reverse pass runs step by step
        if (!fp) {
            fused_plan_free(p);
            return fail(MOSSQ_EDEVICE, err);
        }
        (adj ? p->fused_rev : p->fused) = fp;
        int s = fused_plan_upload(p->ctx, fp);
        if (s != MOSSQ_OK) {
            fused_plan_free(p);
            return s;
        }
    }
    return MOSSQ_OK;
}

static void free_plan(Ctx *c, FusedPlan *&fp) {
    if (!fp) return;
    if (c && !fp->shared_device) {
        if (fp->d_segs) cached_free(c, fp->d_segs);
        if (fp->d_groups) cached_free(c, fp->d_groups);
        if (fp->d_diag) cached_free(c, fp->d_diag);
        if (fp->d_pool) cached_free(c, fp->d_pool);
        if (fp->d_probes) cached_free(c, fp->d_probes);
    }
    delete fp;
    fp = nullptr;
}

void fused_plan_release_retired(Program *p) {
    for (FusedPlan *&fp : p->retired) free_plan(p->ctx, fp);
    p->retired.clear();
}

void fused_plan_free(Program *p) {
    free_plan(p->ctx, p->fused);
    free_plan(p->ctx, p->fused_rev);
    fused_plan_release_retired(p);
    density_adjoint_plan_free(p);
    if (p->pad) {
        fused_plan_free(p->pad);
        delete p->pad;  // its d_pool is this program's
        p->pad = nullptr;
    }
}

int fused_plan_sweeps(const Program *p) {
    if (!p->fused) return int(p->steps.size());
    int unfused = 0;
    for (auto &it : p->fused->items)
        if (!it.sweep) ++unfused;
    return p->fused->nsweeps + unfused;
}

// Runs plan fp on dst; the first pass reads src (src may equal dst).  A null
// plan stands for no steps.
static int fused_run_plan(Ctx *c, const Program *p, const FusedPlan *fp, const c128 *src, c128 *dst) {
    const u64 bytes = (u64(1) << p->nbits) * sizeof(c128);
    const unsigned ntiles = unsigned(u64(1) << (p->nbits - TM));
    const c128 *from = src;
    if (fp)
        for (const PlanItem &it : fp->items) {
            if (!it.sweep) {
                if (from != dst) MOSSQ_HIP(hipMemcpyAsync(dst, from, bytes, hipMemcpyDeviceToDevice, c->stream));
                from = dst;
                MOSSQ_TRY(apply_step(c, p, it.step, dst, fp->reverse_steps));
                continue;
            }
            if (fp->rb == 3)
                hipLaunchKernelGGL(k_tile3, dim3(ntiles), dim3(TGeo<3>::TT), 0, c->stream, from, dst, it.args,
                                   fp->d_segs, fp->d_groups, fp->d_diag, fp->d_pool);
            else
                hipLaunchKernelGGL(k_tile4, dim3(ntiles), dim3(TGeo<4>::TT), 0, c->stream, from, dst, it.args,
                                   fp->d_segs, fp->d_groups, fp->d_diag, fp->d_pool);
            MOSSQ_HIP(hipGetLastError());
            from = dst;
        }
    if (from != dst) MOSSQ_HIP(hipMemcpyAsync(dst, from, bytes, hipMemcpyDeviceToDevice, c->stream));
    return MOSSQ_OK;
}

// Runs a plan with fixed probes on delta in place; the probe of parameter j
// reads ftab[j - fbase] (a device table) and writes partials into row j of
// part (rows of `row` entries).
static int fused_run_fixp(Ctx *c, const Program *p, const FusedPlan *fp, c128 *delta, const c128 *const *ftab,
                          int fbase, c128 *part, int row) {
    const unsigned ntiles = unsigned(u64(1) << (p->nbits - TM));
    for (const PlanItem &it : fp->items) {
        if (!it.sweep) {
            MOSSQ_TRY(apply_step(c, p, it.step, delta, fp->reverse_steps));
            continue;
        }
        if (fp->rb == 3)
            hipLaunchKernelGGL(k_tile_fixp3, dim3(ntiles), dim3(TGeo<3>::TT), 0, c->stream, delta, it.args,
                               fp->d_segs, fp->d_groups, fp->d_diag, fp->d_pool, fp->d_probes, ftab, fbase, part,
                               row);
        else
            hipLaunchKernelGGL(k_tile_fixp4, dim3(ntiles), dim3(TGeo<4>::TT), 0, c->stream, delta, it.args,
                               fp->d_segs, fp->d_groups, fp->d_diag, fp->d_pool, fp->d_probes, ftab, fbase, part,
                               row);
        MOSSQ_HIP(hipGetLastError());
    }
    return MOSSQ_OK;
}

int fused_run(Ctx *c, const Program *p, c128 *state) {
    if (!p->fused) return run_steps(c, p, state, 0, int(p->steps.size()));
    return fused_run_plan(c, p, p->fused, state, state);
}

// Waves per tile of a plan's sweeps (one gradient partial each).
static int plan_wpt(const FusedPlan *fp) { return fp->rb == 3 ? TGeo<3>::WPT : TGeo<4>::WPT; }

int fused_adjoint_parts(const Program *p) {
    if (!p->fused_rev) return 0;
    return int(u64(plan_wpt(p->fused_rev)) << (p->nbits - TM));  // one partial per wave per tile
}

int fused_adjoint_run(Ctx *c, const Program *p, c128 *y, c128 *l, c128 *part, int row) {
    const FusedPlan *fp = p->fused_rev;
    const unsigned ntiles = unsigned(u64(1) << (p->nbits - TM));
    for (size_t i = 0; i < fp->items.size(); ++i) {
        const PlanItem &it = fp->items[i];
        if (!it.sweep) {
            MOSSQ_TRY(adjoint_rev_step(c, p, it.step, y, l, part, row));
            continue;
        }
        // after the plan's last item only the gradient partials are read
        const int last = i + 1 == fp->items.size() ? 1 : 0;
        if (fp->rb == 3)
            hipLaunchKernelGGL(k_tile_adj3, dim3(ntiles), dim3(TGeo<3>::TT), 0, c->stream, y, l, it.args,
                               fp->d_segs, fp->d_groups, fp->d_pool, fp->d_probes, part, row, last);
        else
            hipLaunchKernelGGL(k_tile_adj4, dim3(ntiles), dim3(TGeo<4>::TT), 0, c->stream, y, l, it.args,
                               fp->d_segs, fp->d_groups, fp->d_pool, fp->d_probes, part, row, last);
        MOSSQ_HIP(hipGetLastError());
    }
    return MOSSQ_OK;
}

// Workgroup table of one batched launch whose items have nt[i] tiles: the
// tiles are dealt round robin over the items (entry = item << 16 | tile), so
// every stretch of the grid mixes the heavy tiles of small items with the
// light ones of large items.  Items and tiles must each stay below 2^16.
int deal_tiles(const std::vector<u64> &nt, std::vector<unsigned> &wg) {
    MOSSQ_CHECK(nt.size() <= size_t(kBatchLaunchItems), MOSSQ_EINVAL, "too many items in one batched launch");
    u64 maxnt = 0;
    for (u64 t : nt) {
        MOSSQ_CHECK(t <= 65536, MOSSQ_EINVAL, "batched sweeps take at most 28 qubits");
        maxnt = std::max(maxnt, t);
    }
    for (u64 r = 0; r < maxnt; ++r)
        for (size_t i = 0; i < nt.size(); ++i)
            if (r < nt[i]) wg.push_back((unsigned(i) << 16) | unsigned(r));
    return MOSSQ_OK;
}

bool fused_plan_batchable(const Program *p, bool adjoint) {
    // (the launch tables give a tile 16 bits: up to 28 qubits)
    if (!p->fused || (adjoint && !p->fused_rev) || p->nbits - TM > 16) return false;
    for (const FusedPlan *fp : {p->fused, adjoint ? p->fused_rev : nullptr}) {
        if (!fp) continue;
        for (const PlanItem &it : fp->items)
            if (!it.sweep) return false;
    }
    return true;
}

int batch_sweeps_build(Ctx *c, Allocs &al, int m, const Program *const *progs, c128 *vec, u64 vec_len,
                       const u64 *x, const u64 *l, c128 *part, u64 part_len, const u64 *part_off, const int *row,
                       bool adjoint, BatchSweeps *out, bool sync) {
    std::vector<TBItem> items;
    out->vec = vec;
    out->part = part;
    std::vector<unsigned> wg;
    out->item0.clear();
    out->wg0.clear();
    out->tiles.clear();
    out->rb.clear();
    size_t nlaunch = 0;
    for (int b = 0; b < m; ++b) {
        MOSSQ_CHECK(fused_plan_batchable(progs[b], adjoint), MOSSQ_EINVAL, "program has no all-sweep fused plan");
        const FusedPlan *fp = adjoint ? progs[b]->fused_rev : progs[b]->fused;
        nlaunch = std::max(nlaunch, fp->items.size());
    }
    for (size_t s = 0; s < nlaunch; ++s)
     for (int rb : {4, 3}) {
      for (int b0 = 0; b0 < m;) {
        // one launch: the next kBatchLaunchItems programs whose plans have
        // sweep s and rb register bits
        const size_t i0 = items.size(), w0 = wg.size();
        u64 tiles = 0;
        int b = b0;
        for (; b < m && items.size() - i0 < size_t(kBatchLaunchItems); ++b) {
            const Program *p = progs[b];
            const FusedPlan *fp = adjoint ? p->fused_rev : p->fused;
            if (s >= fp->items.size() || fp->rb != rb) continue;
            TBItem it{};
            const u64 len = u64(1) << p->nbits, nt = len >> TM;
            it.x = x[b];
            it.l = adjoint ? l[b] : 0;
            it.part = adjoint ? part_off[b] : 0;
            it.segs = reinterpret_cast<u64>(fp->d_segs);
            it.groups = reinterpret_cast<u64>(fp->d_groups);
            it.diag = reinterpret_cast<u64>(fp->d_diag);
            it.pool = reinterpret_cast<u64>(fp->d_pool);
            it.probes = reinterpret_cast<u64>(fp->d_probes);
            it.row = adjoint ? row[b] : 0;
            it.tile0 = int(tiles);
            it.last = (adjoint && s + 1 == fp->items.size()) ? 1 : 0;
            it.args = fp->items[s].args;
            MOSSQ_CHECK(tiles + nt <= u64(INT32_MAX), MOSSQ_EINVAL, "batched sweep launch too large");
            MOSSQ_CHECK(x[b] + len <= vec_len && (!adjoint || l[b] + len <= vec_len), MOSSQ_EINVAL,
                        "batched sweep vector outside its allocation");
            MOSSQ_CHECK(!adjoint || (row[b] >= int(u64(plan_wpt(fp)) * nt) &&
                                     part_off[b] + u64(std::max(p->nderiv, 1)) * u64(row[b]) <= part_len),
                        MOSSQ_EINVAL, "gradient partial rows outside their allocation");
            MOSSQ_CHECK(nt <= 65536, MOSSQ_EINVAL, "batched sweeps take at most 28 qubits");
            tiles += nt;
            items.push_back(it);
        }
        b0 = b;
        if (items.size() == i0) continue;
        std::vector<u64> nts;
        for (size_t i = i0; i < items.size(); ++i)
            nts.push_back(i + 1 < items.size() ? u64(items[i + 1].tile0 - items[i].tile0) : tiles - items[i].tile0);
        MOSSQ_TRY(deal_tiles(nts, wg));
        out->item0.push_back(int(i0));
        out->wg0.push_back(int(w0));
        out->tiles.push_back(int(tiles));
        out->rb.push_back(rb);
      }
     }
    if (items.empty()) return MOSSQ_OK;
    MOSSQ_TRY(al.get(items.size() * sizeof(TBItem), &out->d_items));
    MOSSQ_TRY(al.get(wg.size() * sizeof(unsigned), (void **)&out->d_wg));
    if (!sync) {  // the copies read the host tables after this returns: keep them
        out->h_items.assign(reinterpret_cast<const char *>(items.data()),
                            reinterpret_cast<const char *>(items.data() + items.size()));
        out->h_wg.swap(wg);
        MOSSQ_HIP(hipMemcpyAsync(out->d_items, out->h_items.data(), out->h_items.size(), hipMemcpyHostToDevice,
                                 c->stream));
        MOSSQ_HIP(hipMemcpyAsync(out->d_wg, out->h_wg.data(), out->h_wg.size() * sizeof(unsigned),
                                 hipMemcpyHostToDevice, c->stream));
        return MOSSQ_OK;
    }
    MOSSQ_HIP(hipMemcpyAsync(out->d_items, items.data(), items.size() * sizeof(TBItem), hipMemcpyHostToDevice,
                             c->stream));
    MOSSQ_HIP(hipMemcpyAsync(out->d_wg, wg.data(), wg.size() * sizeof(unsigned), hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));  // the host tables go out of scope
    return MOSSQ_OK;
}

int batch_sweeps_run(Ctx *c, const BatchSweeps &bs, bool adjoint) {
    const TBItem *items = static_cast<const TBItem *>(bs.d_items);
    for (size_t s = 0; s < bs.tiles.size(); ++s) {
        if (bs.tiles[s] == 0) continue;
        const dim3 grid(unsigned(bs.tiles[s]));
        const TBItem *it = items + bs.item0[s];
        const unsigned *wg = bs.d_wg + bs.wg0[s];
        if (adjoint && bs.rb[s] == 3)
            hipLaunchKernelGGL(k_btile_adj3, grid, dim3(TGeo<3>::TT), 0, c->stream, it, wg, bs.vec, bs.part);
        else if (adjoint)
            hipLaunchKernelGGL(k_btile_adj4, grid, dim3(TGeo<4>::TT), 0, c->stream, it, wg, bs.vec, bs.part);
        else if (bs.rb[s] == 3)
            hipLaunchKernelGGL(k_btile3, grid, dim3(TGeo<3>::TT), 0, c->stream, it, wg, bs.vec);
        else
            hipLaunchKernelGGL(k_btile4, grid, dim3(TGeo<4>::TT), 0, c->stream, it, wg, bs.vec);
        MOSSQ_HIP(hipGetLastError());
    }
    return MOSSQ_OK;
}

// ------------------------------------------------------------- autotuner ----
// The planner's cost model picks one forward plan; the autotuner plans every
// variant (tail thresholds 0/3/5/8, with and without pairing one-qubit gates
// into 4x4 ops), times each distinct one on the device, and keeps the fastest.
static bool same_plan(const FusedPlan *a, const FusedPlan *b) {
    if (a->rb != b->rb || a->nsweeps != b->nsweeps || a->ntransitions != b->ntransitions || a->items.size() != b->items.size() ||
        a->h_groups.size() != b->h_groups.size() || a->h_segs.size() != b->h_segs.size())
        return false;
    for (size_t i = 0; i < a->h_groups.size(); ++i)
        if (a->h_groups[i].code != b->h_groups[i].code || a->h_groups[i].ncz != b->h_groups[i].ncz) return false;
    return true;
}

int fused_plan_autotune(Program *p, int reps, int *tried, double *best_ms) {
    *tried = 0;
    *best_ms = 0;
    if (!p->fused) return MOSSQ_OK;  // unfused program: nothing to choose
    Ctx *c = p->ctx;
    std::vector<POp> ops;
    lower_steps(p, ops);
    std::vector<FusedPlan *> cand;
    std::string err;
    for (int rb : {4, 3})
      for (bool kron : {true, false})
        for (int tail : {0, 3, 5, 8}) {
            FusedPlan *fp = plan_with_tail(p, ops, tail, false, kron, rb, &err);
            if (!fp) continue;
            bool dup = same_plan(fp, p->fused);
            for (const FusedPlan *q : cand) dup = dup || same_plan(q, fp);
            if (dup)
                delete fp;
            else
                cand.push_back(fp);
        }
    cand.insert(cand.begin(), p->fused);  // the model's choice is candidate 0
    const u64 len = u64(1) << p->nbits;
    c128 *scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int s = cached_alloc(c, len * sizeof(c128), (void **)&scratch);
    std::vector<double> ms(cand.size(), 1e300);
    if (s == MOSSQ_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        s = fail(MOSSQ_EDEVICE, "hipEventCreate failed");
    if (s == MOSSQ_OK) s = fill_basis0(c, scratch, len);
    for (size_t i = 0; i < cand.size() && s == MOSSQ_OK; ++i) {
        if (i > 0) s = fused_plan_upload(c, cand[i]);
        if (s == MOSSQ_OK) s = fused_run_plan(c, p, cand[i], scratch, scratch);  // warm-up
        if (s != MOSSQ_OK) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) s = fail(MOSSQ_EDEVICE, "hipEventRecord failed");
        for (int r = 0; r < reps && s == MOSSQ_OK; ++r) s = fused_run_plan(c, p, cand[i], scratch, scratch);
        float t = 0;
        if (s == MOSSQ_OK && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&t, e0, e1) != hipSuccess))
            s = fail(MOSSQ_EDEVICE, "timing a plan variant failed");
        ms[i] = double(t) / std::max(reps, 1);
    }
    size_t best = 0;
    if (s == MOSSQ_OK)
        for (size_t i = 1; i < cand.size(); ++i)
            if (ms[i] < ms[best]) best = i;
    for (size_t i = 1; i < cand.size(); ++i)
        if (i != best) free_plan(c, cand[i]);
    if (s == MOSSQ_OK && best != 0) {
        if (p->live_graphs > 0)
            p->retired.push_back(p->fused);  // a captured graph still reads its tables
        else
            free_plan(c, p->fused);
        p->fused = cand[best];
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (scratch) cached_free(c, scratch);
    if (s != MOSSQ_OK) return s;
    *tried = int(cand.size());
    *best_ms = ms[best];
    return MOSSQ_OK;
}

__global__ void k_zero(c128 *__restrict__ v, u64 count) {
    for (u64 i = u64(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += u64(gridDim.x) * blockDim.x)
        v[i] = c128{0.0, 0.0};
}

int zero_fill(Ctx *c, c128 *v, u64 count) {
    if (count == 0) return MOSSQ_OK;
    const u64 blocks = std::min<u64>((count + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero, dim3(unsigned(blocks)), dim3(256), 0, c->stream, v, count);
    MOSSQ_HIP(hipGetLastError());
    return MOSSQ_OK;
}


// ----------------------------------------------- fused density adjoint ----
// Channels have no inverse, so the density-matrix reverse pass needs the state
// before every parameterized step.  The program is cut at its parameterized
// steps p_1 < ... < p_m:
//   forward chunk j (j = 0..m): steps [b_j, b_{j+1}) with b_0 = 0, b_j = p_j,
//     b_{m+1} = N; it maps S_j (the state before p_j, S_0 the input) to S_{j+1};
//   reverse chunk j (j = 1..m): steps (p_j, b_{j+1}) from last to first with the
//     reverse matrices S^dagger and U^dagger (.) U.
// Every chunk is an LDS-tile plan of its own.  The forward pass keeps S_j at
// j = 1, 1 + seg, 1 + 2 seg, ... (seg = ceil(sqrt(m))), running each segment
// of seg chunks as one plan; the reverse pass recomputes one segment of S_j
// at a time from its checkpoint, then per parameter takes Delta through the
// reverse chunk and the update Delta <- V Delta with Re<V Delta | G | S_j>
// (differentiation/circuit.jl:90-172).  For a one-qubit parameterized gate
// the reverse chunk plan ends with V and a probe of G against S_j, so the
// whole step is one pass over Delta plus one read of S_j; wider ones run the
// reverse chunk, then a paired launch.
struct DensityAdjointPlan {
    std::vector<int> psteps;       // p_1..p_m
    std::vector<FusedPlan *> fwd;  // m + 1 chunks (null: no steps)
    std::vector<FusedPlan *> rev;   // 1..m (rev[0] unused): reverse chunk without p_j
    std::vector<FusedPlan *> revp;  // 1..m: reverse chunk, V_j and the probe of G_j (null: use rev[j])
    std::vector<FusedPlan *> segfwd;  // segment k: chunks 1 + k seg .. min(m, (k + 1) seg) as one plan
    // segment k's whole reverse pass (every reverse chunk, V_j and fixed probe,
    // from the last parameter down) as one plan; null: per parameter
    std::vector<FusedPlan *> segrev;
    int seg = 1, nseg = 1;
    void *d_block = nullptr;  // device arrays of every chunk plan
};

static FusedPlan *plan_range(const Program *p, int lo, int hi, bool reverse, std::string *err, bool *ok) {
    *ok = true;
    std::vector<POp> ops;
    lower_range(p, lo, hi, reverse, ops);
    if (ops.empty()) return nullptr;
    FusedPlan *fp = best_plan(p, ops, false, 4, err);
    if (!fp) {
        *ok = false;
        return nullptr;
    }
    fp->reverse_steps = reverse;
    return fp;
}

static void dm_plan_delete(DensityAdjointPlan *dp) {
    for (auto *fp : dp->fwd) delete fp;
    for (auto *fp : dp->rev) delete fp;
    for (auto *fp : dp->revp) delete fp;
    for (auto *fp : dp->segfwd) delete fp;
    for (auto *fp : dp->segrev) delete fp;
    delete dp;
}

// Reverse chunk of parameterized step si through its own reverse matrix, then
// the fixed probe of its gradient matrix; null when the step is wider than a
// tile probe takes (two register bits).
static bool probe_fits(const Step &st) { return !st.split && st.kb <= 2; }

// Steps [si, hi) from last to first with their reverse matrices, then the
// fixed probe of step si's gradient matrix.
static void lower_reverse_probe(const Program *p, int si, int hi, std::vector<POp> &ops) {
    const Step &st = p->steps[si];
    lower_range(p, si, hi, true, ops);
    POp pr;
    pr.kind = 3;
    pr.nb = st.kb;
    for (int i = 0; i < st.kb; ++i) pr.bits[i] = st.bits[i];
    pr.m = mat_of(p, st.grad, 1 << st.kb);
    pr.step = si;
    pr.probe = st.grad_index;
    pr.fixed = true;
    ops.push_back(pr);
}

static FusedPlan *plan_reverse_probe(const Program *p, int si, int hi, std::string *err, bool *ok) {
    *ok = true;
    if (!probe_fits(p->steps[si])) return nullptr;
    std::vector<POp> ops;
    lower_reverse_probe(p, si, hi, ops);
    FusedPlan *fp = best_plan(p, ops, false, 4, err);
    if (!fp) {
        *ok = false;
        return nullptr;
    }
    fp->reverse_steps = true;
    return fp;
}

// Every plan of a density adjoint plan (null entries skipped).
static std::vector<FusedPlan *> dm_plans(const DensityAdjointPlan *dp) {
    std::vector<FusedPlan *> out;
    for (const auto *v : {&dp->fwd, &dp->rev, &dp->revp, &dp->segfwd, &dp->segrev})
        for (FusedPlan *fp : *v)
            if (fp) out.push_back(fp);
    return out;
}

// whole_segments: 1 plans each segment's reverse pass as one plan where its
// parameterized gates allow, 0 plans it per parameter, -1 picks by size.
static DensityAdjointPlan *dm_plan_build_host(const Program *p, std::string *err, int whole_segments = -1) {
    auto *dp = new DensityAdjointPlan();
    const int N = int(p->steps.size());
    for (int i = 0; i < N; ++i)
        if (p->steps[i].grad >= 0) dp->psteps.push_back(i);
    const int m = int(dp->psteps.size());
    auto b = [&](int j) { return j == 0 ? 0 : j <= m ? dp->psteps[j - 1] : N; };
    for (int j = 2; j <= m; ++j)
        if (p->steps[dp->psteps[j - 1]].grad_index != p->steps[dp->psteps[j - 2]].grad_index + 1) {
            *err = "gradient rows of a density program are not consecutive";
            delete dp;
            return nullptr;
        }
    dp->fwd.assign(m + 1, nullptr);
    dp->rev.assign(m + 1, nullptr);
    dp->revp.assign(m + 1, nullptr);
    dp->seg = std::max(1, int(std::ceil(std::sqrt(double(std::max(m, 1))))));
    dp->nseg = std::max(1, (m + dp->seg - 1) / dp->seg);
    dp->segfwd.assign(dp->nseg, nullptr);
    dp->segrev.assign(dp->nseg, nullptr);
    bool ok = true;
    for (int j = 0; j <= m && ok; ++j) dp->fwd[j] = plan_range(p, b(j), b(j + 1), false, err, &ok);
    for (int k = 0; k < dp->nseg && ok; ++k) {
        const int a = 1 + k * dp->seg, e = std::min(m, (k + 1) * dp->seg);
        dp->segfwd[k] = plan_range(p, b(a), b(e + 1), false, err, &ok);
        // One plan for the segment's reverse pass reads Delta once per sweep
        // instead of once per parameter, but its probes wait on their loads
        // mid-sweep; below 2^22 entries (2^10 tiles, under 4 per CU) those
        // waits are not hidden and the per-parameter plans measure faster.
        bool fits = whole_segments < 0 ? p->nbits >= 22 : whole_segments > 0;
        for (int j = a; j <= e; ++j) fits &= probe_fits(p->steps[dp->psteps[j - 1]]);
        if (ok && fits) {
            std::vector<POp> ops;
            for (int j = e; j >= a; --j) lower_reverse_probe(p, dp->psteps[j - 1], b(j + 1), ops);
            dp->segrev[k] = best_plan(p, ops, false, 4, err);
            if (!dp->segrev[k]) {
                ok = false;
                break;
            }
            dp->segrev[k]->reverse_steps = true;
            continue;
        }
        for (int j = a; j <= e && ok; ++j) {
            dp->revp[j] = plan_reverse_probe(p, dp->psteps[j - 1], b(j + 1), err, &ok);
            if (ok && !dp->revp[j]) dp->rev[j] = plan_range(p, dp->psteps[j - 1] + 1, b(j + 1), true, err, &ok);
        }
    }
    if (!ok) {
        dm_plan_delete(dp);
        return nullptr;
    }
    return dp;
}

// Forward pass: chunk 0 takes S_0 (in work) to S_1 in ckpt[0]; segment k
// takes S_{1 + k seg} in ckpt[k] to the next checkpoint, the last one to
// S_{m+1} in work.
template <class Run>
static int dm_forward_sched(const DensityAdjointPlan *dp, c128 *work, c128 *const *ckpt, Run run) {
    MOSSQ_TRY(run(dp->fwd[0], work, ckpt[0]));
    for (int k = 0; k < dp->nseg; ++k)
        MOSSQ_TRY(run(dp->segfwd[k], ckpt[k], k + 1 < dp->nseg ? ckpt[k + 1] : work));
    return MOSSQ_OK;
}

// Reverse pass from Delta at the program's end down to p_1; stored holds
// seg - 1 states.
// run_probe(fp, delta, k, j) runs a plan with fixed probes whose lowest
// parameter is j in segment k (probe of parameter i reads S_i: ckpt[k] for
// i = 1 + k seg, else stored[i - 2 - k seg]).
template <class Run, class Pair, class RunProbe>
static int dm_reverse_sched(const DensityAdjointPlan *dp, c128 *delta, c128 *const *ckpt, c128 *const *stored,
                            Run run, Pair pair, RunProbe run_probe) {
    const int m = int(dp->psteps.size()), seg = dp->seg;
    for (int k = dp->nseg - 1; k >= 0; --k) {
        const int a = 1 + k * seg, e = std::min(m, (k + 1) * seg);
        auto S = [&](int j) { return j == a ? ckpt[k] : stored[j - a - 1]; };
        for (int j = a + 1; j <= e; ++j) MOSSQ_TRY(run(dp->fwd[j - 1], S(j - 1), S(j)));
        if (dp->segrev[k]) {
            MOSSQ_TRY(run_probe(dp->segrev[k], delta, k, a));
            continue;
        }
        for (int j = e; j >= a; --j) {
            if (dp->revp[j]) {
                MOSSQ_TRY(run_probe(dp->revp[j], delta, k, j));
            } else {
                MOSSQ_TRY(run(dp->rev[j], delta, delta));
                MOSSQ_TRY(pair(dp->psteps[j - 1], delta, S(j)));
            }
        }
    }
    return MOSSQ_OK;
}

// One device block for the arrays of every chunk plan (one allocation and one
// copy per program).
static int upload_plan_block(Ctx *c, const std::vector<FusedPlan *> &plans, void **block) {
    auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
    size_t total = 0;
    for (const FusedPlan *fp : plans)
        total += al(fp->h_segs.size() * sizeof(TSeg)) + al(fp->h_groups.size() * sizeof(TGroup)) +
                 al(fp->h_diag.size() * sizeof(TDiag)) + al(fp->h_pool.size() * sizeof(c128)) +
                 al(fp->h_probes.size() * sizeof(TProbe));
    if (total == 0) return MOSSQ_OK;
    std::vector<char> h(total);
    MOSSQ_TRY(cached_alloc(c, total, block));
    char *d = static_cast<char *>(*block);
    size_t off = 0;
    auto put = [&](const void *src, size_t bytes) {
        std::memcpy(h.data() + off, src, bytes);
        char *at = d + off;
        off += al(bytes);
        return at;
    };
    for (FusedPlan *fp : plans) {
        fp->d_segs = reinterpret_cast<TSeg *>(put(fp->h_segs.data(), fp->h_segs.size() * sizeof(TSeg)));
        fp->d_groups = reinterpret_cast<TGroup *>(put(fp->h_groups.data(), fp->h_groups.size() * sizeof(TGroup)));
        fp->d_diag = reinterpret_cast<TDiag *>(put(fp->h_diag.data(), fp->h_diag.size() * sizeof(TDiag)));
        fp->d_pool = reinterpret_cast<c128 *>(put(fp->h_pool.data(), fp->h_pool.size() * sizeof(c128)));
        fp->d_probes = reinterpret_cast<TProbe *>(put(fp->h_probes.data(), fp->h_probes.size() * sizeof(TProbe)));
        fp->shared_device = true;
    }
    MOSSQ_HIP(hipMemcpyAsync(d, h.data(), total, hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    return MOSSQ_OK;
}

int density_adjoint_plan_build(Program *p) {
    if (!p->is_density || p->nderiv == 0 || p->nbits < TM || p->dm_adj) return MOSSQ_OK;
    std::string err;
    DensityAdjointPlan *dp = dm_plan_build_host(p, &err);
    if (!dp) return MOSSQ_OK;  // no chunk plans: the reverse pass runs step by step
    int s = upload_plan_block(p->ctx, dm_plans(dp), &dp->d_block);
    if (s != MOSSQ_OK) {
        if (dp->d_block) cached_free(p->ctx, dp->d_block);
        dm_plan_delete(dp);
        return s;
    }
    p->dm_adj = dp;
    return MOSSQ_OK;
}

void density_adjoint_plan_free(Program *p) {
    if (!p->dm_adj) return;
    if (p->ctx && p->dm_adj->d_block) cached_free(p->ctx, p->dm_adj->d_block);
    dm_plan_delete(p->dm_adj);
    p->dm_adj = nullptr;
}

void density_adjoint_layout(const Program *p, int *seg, int *nseg) {
    *seg = p->dm_adj ? p->dm_adj->seg : 0;
    *nseg = p->dm_adj ? p->dm_adj->nseg : 0;
}

int density_adjoint_forward(Ctx *c, const Program *p, c128 *work, c128 *const *ckpt) {
    return dm_forward_sched(p->dm_adj, work, ckpt, [&](const FusedPlan *fp, c128 *src, c128 *dst) {
        return fused_run_plan(c, p, fp, src, dst);
    });
}

// Table of the states the fixed probes read: entry k seg + i holds S_{1 + k seg + i}.
void density_adjoint_probe_table(const Program *p, c128 *const *ckpt, c128 *const *stored,
                                 std::vector<const c128 *> &tab) {
    const DensityAdjointPlan *dp = p->dm_adj;
    tab.assign(size_t(dp->nseg) * dp->seg, nullptr);
    for (int k = 0; k < dp->nseg; ++k)
        for (int i = 0; i < dp->seg; ++i) tab[size_t(k) * dp->seg + i] = i == 0 ? ckpt[k] : stored[i - 1];
}

int density_adjoint_reverse(Ctx *c, const Program *p, c128 *delta, c128 *const *ckpt, c128 *const *stored,
                            const c128 *const *d_tab, c128 *part, int row) {
    const DensityAdjointPlan *dp = p->dm_adj;
    return dm_reverse_sched(
        p->dm_adj, delta, ckpt, stored,
        [&](const FusedPlan *fp, c128 *src, c128 *dst) { return fused_run_plan(c, p, fp, src, dst); },
        [&](int si, c128 *d, c128 *S) {
            const Step &st = p->steps[si];
            return launch_dense_pair(c, d, S, p->nbits, st.kb, st.bits, p->d_pool + st.rev, p->d_pool + st.grad,
                                     part + size_t(st.grad_index) * row, true);
        },
        [&](const FusedPlan *fp, c128 *d, int k, int j) {
            const int a = 1 + k * dp->seg;
            return fused_run_fixp(c, p, fp, d, d_tab + size_t(k) * dp->seg + (j - a),
                                  p->steps[dp->psteps[j - 1]].grad_index, part, row);
        });
}

int density_adjoint_parts(const Program *p) {
    if (!p->dm_adj) return 0;
    int wpt = 1;
    for (const FusedPlan *fp : dm_plans(p->dm_adj)) wpt = std::max(wpt, plan_wpt(fp));
    return std::max(pair_parts(p->nbits), int(u64(wpt) << (p->nbits - TM)));
}

// Host execution of the fused density adjoint with the same schedule:
// rho holds S_0 on entry and the final state on exit; delta holds the seed at
// the program's end on entry; grads[j] receives parameter j's complex sum.
static void emulate_density_adjoint(const Program *p, const DensityAdjointPlan *dp, c128 *rho, c128 *delta,
                                    c128 *grads) {
    const u64 len = u64(1) << p->nbits;
    std::vector<std::vector<c128>> ck(dp->nseg, std::vector<c128>(len)), sto(dp->seg - 1, std::vector<c128>(len));
    std::vector<c128 *> ckpt, stored;
    for (auto &v : ck) ckpt.push_back(v.data());
    for (auto &v : sto) stored.push_back(v.data());
    auto run = [&](const FusedPlan *fp, c128 *src, c128 *dst) {
        if (src != dst) std::memcpy(dst, src, len * sizeof(c128));
        if (fp) emulate_plan(p, fp, dst);
        return MOSSQ_OK;
    };
    dm_forward_sched(dp, rho, ckpt.data(), run);
    const c128 *P = p->h_pool.data();
    dm_reverse_sched(dp, delta, ckpt.data(), stored.data(), run, [&](int si, c128 *d, c128 *S) {
        const Step &st = p->steps[si];
        host_block(d, p->nbits, st.kb, st.bits, P + st.rev);
        std::vector<c128> dy(S, S + len);
        host_block(dy.data(), p->nbits, st.kb, st.bits, P + st.grad);
        c128 acc = mk(0, 0);
        for (u64 i = 0; i < len; ++i) acc = cfmaconj(d[i], dy[i], acc);
        grads[st.grad_index] = cadd(grads[st.grad_index], acc);
        return MOSSQ_OK;
    }, [&](const FusedPlan *fp, c128 *d, int k, int j) {
        std::vector<const c128 *> tab(size_t(dp->nseg) * dp->seg);
        for (int kk = 0; kk < dp->nseg; ++kk)
            for (int i = 0; i < dp->seg; ++i) tab[size_t(kk) * dp->seg + i] = i == 0 ? ckpt[kk] : stored[i - 1];
        const int a = 1 + k * dp->seg;
        emulate_plan(p, fp, d, tab.data() + size_t(k) * dp->seg + (j - a), p->steps[dp->psteps[j - 1]].grad_index,
                     grads);
        return MOSSQ_OK;
    });
}
// ------------------------------------------------------ parameter update ----
// New matrix values for a program whose structure stays: every plan is
// rebuilt from q (the program rebuilt from the new records) with the variant
// it came from, and must come out with the same tables but for matrix values,
// diagonal values and probe factors, which are then written over the old
// plan's device arrays in place (captured graphs keep reading them).  The
// rebuild is host work only, so updates of different programs can be
// prepared on several threads and their copies issued together.
static bool same_tables(const FusedPlan *a, const FusedPlan *b) {
    if (a->rb != b->rb || a->adjoint != b->adjoint || a->items.size() != b->items.size() ||
        a->h_segs.size() != b->h_segs.size() || a->h_groups.size() != b->h_groups.size() ||
        a->h_diag.size() != b->h_diag.size() || a->h_probes.size() != b->h_probes.size() ||
        a->h_pool.size() != b->h_pool.size() || a->reverse_steps != b->reverse_steps)
        return false;
    for (size_t i = 0; i < a->items.size(); ++i)
        if (a->items[i].sweep != b->items[i].sweep || a->items[i].step != b->items[i].step ||
            memcmp(&a->items[i].args, &b->items[i].args, sizeof(TSweepArgs)) != 0)
            return false;
    if (memcmp(a->h_segs.data(), b->h_segs.data(), a->h_segs.size() * sizeof(TSeg)) != 0 ||
        memcmp(a->h_groups.data(), b->h_groups.data(), a->h_groups.size() * sizeof(TGroup)) != 0)
        return false;
    for (size_t i = 0; i < a->h_diag.size(); ++i) {
        const TDiag &x = a->h_diag[i], &y = b->h_diag[i];
        if (x.code != y.code || memcmp(x.lbit, y.lbit, sizeof(x.lbit)) != 0 || memcmp(x.qbit, y.qbit, sizeof(x.qbit)) ||
            x.rp_lo != y.rp_lo || x.rp_hi != y.rp_hi || x.negmask != y.negmask || x.moff != y.moff)
            return false;
    }
    for (size_t i = 0; i < a->h_probes.size(); ++i) {
        const TProbe &x = a->h_probes[i], &y = b->h_probes[i];
        if (x.probe != y.probe || x.pcode != y.pcode || x.pmoff != y.pmoff || x.pmask != y.pmask) return false;
    }
    return true;
}

bool fused_run_launches_step(const Program *p, int i) {
    if (!p->fused) return true;
    for (const PlanItem &it : p->fused->items)
        if (!it.sweep && it.step == i) return true;
    return false;
}

PlanUpdate::~PlanUpdate() {
    for (auto &pr : pairs) delete pr.second;
    if (ndp) {
        // paired density plans are deleted above, not with ndp
        if (dm_paired)
            for (auto *v : {&ndp->fwd, &ndp->rev, &ndp->revp, &ndp->segfwd, &ndp->segrev}) v->clear();
        dm_plan_delete(ndp);
    }
}

int fused_plan_update_prepare(const Program *p, const Program *q, PlanUpdate *u) {
    MOSSQ_CHECK(p->retired.empty(), MOSSQ_EINVAL, "a captured graph still uses a plan the autotuner replaced");
    std::string err;
    for (FusedPlan *old : {p->fused, p->fused_rev}) {
        if (!old) continue;
        std::vector<POp> ops;
        if (old->adjoint)
            lower_steps_rev(q, ops);
        else
            lower_steps(q, ops);
        FusedPlan *nw = plan_with_tail(q, ops, old->min_tail, old->adjoint, old->kron, old->rb, &err, old);
        if (!nw) return fail(MOSSQ_EINVAL, "rebuilding a plan failed: " + err);
        u->pairs.push_back({old, nw});
        if (!same_tables(old, nw)) return fail(MOSSQ_EINVAL, "the new values change the program's fused plan");
    }
    if (p->dm_adj) {
        u->ndp = dm_plan_build_host(q, &err);
        if (!u->ndp) return fail(MOSSQ_EINVAL, "rebuilding the density adjoint plans failed: " + err);
        const std::vector<FusedPlan *> a = dm_plans(p->dm_adj), b = dm_plans(u->ndp);
        if (a.size() != b.size()) return fail(MOSSQ_EINVAL, "the new values change the density adjoint plans");
        for (size_t i = 0; i < a.size(); ++i)
            if (!same_tables(a[i], b[i])) return fail(MOSSQ_EINVAL, "the new values change the density adjoint plans");
        for (size_t i = 0; i < a.size(); ++i) u->pairs.push_back({a[i], b[i]});
        u->dm_paired = true;
    }
    return MOSSQ_OK;
}

void fused_plan_update_copies(const PlanUpdate &u, std::vector<DevCopy> &out) {
    for (const auto &pr : u.pairs) {
        const FusedPlan *old = pr.first, *nw = pr.second;
        out.push_back({old->d_pool, nw->h_pool.data(), nw->h_pool.size() * sizeof(c128)});
        out.push_back({old->d_diag, nw->h_diag.data(), nw->h_diag.size() * sizeof(TDiag)});
        out.push_back({old->d_probes, nw->h_probes.data(), nw->h_probes.size() * sizeof(TProbe)});
    }
}

void fused_plan_update_commit(PlanUpdate *u) {
    for (auto &pr : u->pairs) {
        pr.first->h_pool.swap(pr.second->h_pool);
        pr.first->h_diag.swap(pr.second->h_diag);
        pr.first->h_probes.swap(pr.second->h_probes);
    }
}

}  // namespace mossq

using namespace mossq;

static void plan_stats(const FusedPlan *fp, int *stats) {
    int unfused = 0;
    for (auto &it : fp->items)
        if (!it.sweep) ++unfused;
    stats[0] = fp->nsweeps;
    stats[1] = unfused;
    stats[2] = int(fp->h_segs.size());
    stats[3] = fp->ntransitions;
    stats[4] = int(fp->h_ops.size());
    stats[5] = stats[6] = stats[7] = 0;
    for (const TOp &o : fp->h_ops) {
        if (o.type == T_DENSE1)
            stats[5]++;
        else if (o.type == T_DENSE2)
            stats[6]++;
        else if (o.type != T_PROBE1 && o.type != T_PROBE2)
            stats[7]++;
    }
    stats[8] = fp->nconflicted;
    stats[9] = 0;
    for (const TGroup &g : fp->h_groups) stats[9] += g.code >= C_PERM && g.code < C_DIAG1;
}

extern "C" {

int mossq_batch_launch_tables(int m, const int32_t *ntiles, int max_launches, int32_t *launch_items, uint64_t cap,
                              uint32_t *entries, int *nlaunches, uint64_t *nentries) {
    MOSSQ_CHECK(m >= 0 && (m == 0 || ntiles) && nlaunches && nentries && (max_launches == 0 || launch_items),
                MOSSQ_EINVAL, "bad arguments");
    std::vector<unsigned> wg;
    int L = 0;
    for (int b0 = 0; b0 < m; b0 += kBatchLaunchItems) {
        const int b1 = std::min(m, b0 + kBatchLaunchItems);
        std::vector<u64> nt;
        for (int b = b0; b < b1; ++b) {
            MOSSQ_CHECK(ntiles[b] >= 1, MOSSQ_EINVAL, "an item has at least one tile");
            nt.push_back(u64(ntiles[b]));
        }
        MOSSQ_TRY(deal_tiles(nt, wg));
        if (L < max_launches) launch_items[L] = b1 - b0;
        ++L;
    }
    *nlaunches = L;
    *nentries = wg.size();
    MOSSQ_CHECK(L <= max_launches && wg.size() <= cap && (wg.empty() || entries), MOSSQ_EINVAL,
                "output arrays too small");
    std::copy(wg.begin(), wg.end(), entries);
    return MOSSQ_OK;
}

int mossq_plan_emulate(int n, int is_density, int nops, const mossq_op *ops, const double *pool, uint64_t pool_len,
                       double *state, int *stats) {
    MOSSQ_CHECK(stats, MOSSQ_EINVAL, "null argument");
    Program p;
    MOSSQ_TRY(program_build_host(&p, n, is_density, nops, ops, pool, pool_len));
    MOSSQ_CHECK(p.nbits >= TM, MOSSQ_EINVAL, "plans tile registers of at least 12 bits");
    std::string err;
    FusedPlan *fp = fused_plan_build_host(&p, &err, false, 0);
    MOSSQ_CHECK(fp, MOSSQ_EINVAL, err);
    if (state) emulate_plan(&p, fp, (c128 *)state);
    plan_stats(fp, stats);
    delete fp;
    return MOSSQ_OK;
}

int mossq_plan_emulate_adjoint(int n, int nops, const mossq_op *ops, const double *pool, uint64_t pool_len,
                               double *y, double *l, double *grads, int *stats) {
    MOSSQ_CHECK(stats && y && l && grads, MOSSQ_EINVAL, "null argument");
    Program p;
    MOSSQ_TRY(program_build_host(&p, n, 0, nops, ops, pool, pool_len));
    MOSSQ_CHECK(p.nbits >= TM, MOSSQ_EINVAL, "plans tile registers of at least 12 bits");
    std::string err;
    FusedPlan *fp = fused_plan_build_host(&p, &err, true, 0);
    MOSSQ_CHECK(fp, MOSSQ_EINVAL, err);
    std::vector<c128> g(std::max(p.nderiv, 1), mk(0, 0));
    emulate_adjoint_plan(&p, fp, (c128 *)y, (c128 *)l, g.data());
    for (int j = 0; j < p.nderiv; ++j) grads[j] = g[j].re;
    plan_stats(fp, stats);
    delete fp;
    return MOSSQ_OK;
}

int mossq_plan_emulate_density_adjoint(int n, int nops, const mossq_op *ops, const double *pool, uint64_t pool_len,
                                       int whole_segments, double *rho, double *delta, double *grads, int *stats) {
    MOSSQ_CHECK(stats && (!rho || (delta && grads)), MOSSQ_EINVAL, "null argument");
    Program p;
    MOSSQ_TRY(program_build_host(&p, n, 1, nops, ops, pool, pool_len));
    MOSSQ_CHECK(p.nbits >= TM, MOSSQ_EINVAL, "plans tile registers of at least 12 bits");
    MOSSQ_CHECK(p.nderiv > 0, MOSSQ_EINVAL, "the program has no parameterized steps");
    std::string err;
    DensityAdjointPlan *dp = dm_plan_build_host(&p, &err, whole_segments < 0 ? -1 : whole_segments > 0);
    MOSSQ_CHECK(dp, MOSSQ_EINVAL, err);
    if (rho) {
        std::vector<c128> g(p.nderiv, mk(0, 0));
        emulate_density_adjoint(&p, dp, (c128 *)rho, (c128 *)delta, g.data());
        for (int j = 0; j < p.nderiv; ++j) grads[j] = g[j].re;
    }
    for (int i = 0; i < 6; ++i) stats[i] = 0;
    stats[0] = int(dp->psteps.size());
    stats[1] = dp->seg;
    // forward: chunk 0 and the segment plans; reverse: the chunk plans the
    // recompute and the reverse pass run
    auto count = [&](const FusedPlan *fp, int slot) {
        if (!fp) return;
        stats[slot] += fp->nsweeps;
        for (auto &it : fp->items) stats[4] += !it.sweep;
    };
    count(dp->fwd[0], 2);
    for (const FusedPlan *fp : dp->segfwd) count(fp, 2);
    for (size_t j = 1; j < dp->fwd.size(); ++j) count(dp->fwd[j], 3);
    for (const auto *v : {&dp->rev, &dp->revp, &dp->segrev})
        for (const FusedPlan *fp : *v) count(fp, 3);
    stats[5] = int(dm_plans(dp).size());
    dm_plan_delete(dp);
    return MOSSQ_OK;
}

}  // extern "C"
