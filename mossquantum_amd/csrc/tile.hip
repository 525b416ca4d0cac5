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
// any global bit, inside or outside the tile.
//
// The planner (host) takes the program's instructions in order, moves
// instructions past earlier ones they commute with (disjoint supports, or both
// diagonal), and packs them greedily into sweeps of at most 12 tile bits, then
// orders each sweep's instructions into register-bit segments to minimise LDS
// transpositions, and multiplies adjacent instructions on the same bits into
// one matrix.
#include <algorithm>
#include <set>

#include "mossq_program.h"

namespace mossq {

constexpr int TM = 12;   // tile bits
constexpr int TRB = 4;   // register bits
constexpr int TT = 256;  // threads per tile
constexpr int TN = 16;   // amplitudes per thread
constexpr int kCoalesceBits = 3;

struct TOp {
    int type;  // 0 dense 1-bit, 1 dense 2-bit, 2 diagonal
    int a, b;  // dense: register-bit indices (a < b)
    int k;     // diagonal: number of bits
    int dsrc[4];  // diagonal bit sources: 0..11 tile-local bit, >= 100: global bit (src - 100)
    int moff;     // offset in the plan's device pool
};

struct TSeg {
    int reg[TRB];  // tile-local bits held in registers (increasing)
    int op0, op1;
};

struct TSweepArgs {
    int tbits[TM];  // global bits of the tile, increasing
    int cbits[40];  // remaining global bits, increasing (tile id bits)
    int ncbits;
    int seg0, seg1;
};

__device__ __forceinline__ int lds_pad(int l) { return l + (l >> 4); }

template <int A>
__device__ __forceinline__ void reg_apply1(c128 *v, c128 m00, c128 m10, c128 m01, c128 m11) {
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        if (j & (1 << A)) continue;
        c128 x = v[j], y = v[j | (1 << A)];
        v[j] = cfma(m01, y, cmul(m00, x));
        v[j | (1 << A)] = cfma(m11, y, cmul(m10, x));
    }
}

template <int A, int B>
__device__ __forceinline__ void reg_apply2(c128 *v, const c128 *__restrict__ M) {
    c128 m[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = M[i];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        if (j & ((1 << A) | (1 << B))) continue;
        const int i0 = j, i1 = j | (1 << A), i2 = j | (1 << B), i3 = j | (1 << A) | (1 << B);
        c128 in0 = v[i0], in1 = v[i1], in2 = v[i2], in3 = v[i3];
        c128 o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            c128 acc = cmul(m[r], in0);
            acc = cfma(m[r + 4], in1, acc);
            acc = cfma(m[r + 8], in2, acc);
            acc = cfma(m[r + 12], in3, acc);
            o[r] = acc;
        }
        v[i0] = o[0];
        v[i1] = o[1];
        v[i2] = o[2];
        v[i3] = o[3];
    }
}

__device__ __forceinline__ int deposit4(int j, const int *reg) {
    return (((j >> 0) & 1) << reg[0]) | (((j >> 1) & 1) << reg[1]) | (((j >> 2) & 1) << reg[2]) |
           (((j >> 3) & 1) << reg[3]);
}

// Local index contribution of thread t when `reg` are the register bits: the
// 8 thread bits are the tile bits not in reg, in increasing order.
__device__ __forceinline__ int thread_deposit(int t, const int *reg) {
    int out = 0, k = 0;
#pragma unroll
    for (int b = 0; b < TM; ++b) {
        bool isreg = (b == reg[0]) | (b == reg[1]) | (b == reg[2]) | (b == reg[3]);
        if (!isreg) {
            out |= ((t >> k) & 1) << b;
            ++k;
        }
    }
    return out;
}

__global__ __launch_bounds__(TT, 2) void k_tile(c128 *__restrict__ state, TSweepArgs sw,
                                                  const TSeg *__restrict__ segs, const TOp *__restrict__ ops,
                                                  const c128 *__restrict__ pool) {
    __shared__ c128 lds[(1 << TM) + (1 << TM) / 16];
    const int t = threadIdx.x;
    const u64 tile = blockIdx.x;
    u64 base = 0;
    for (int i = 0; i < sw.ncbits; ++i)
        if ((tile >> i) & 1) base |= u64(1) << sw.cbits[i];
    u64 gt = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
        if ((t >> i) & 1) gt |= u64(1) << sw.tbits[i];
    u64 gj[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        u64 g = 0;
#pragma unroll
        for (int i = 0; i < TRB; ++i)
            if ((j >> i) & 1) g |= u64(1) << sw.tbits[8 + i];
        gj[j] = base | gt | g;
    }
    c128 v[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) v[j] = state[gj[j]];

    int cur[TRB] = {8, 9, 10, 11};
    int tdep = t;  // thread bits are tile bits 0..7 in the load layout
    for (int s = sw.seg0; s < sw.seg1; ++s) {
        const TSeg sg = segs[s];
        if (sg.reg[0] != cur[0] || sg.reg[1] != cur[1] || sg.reg[2] != cur[2] || sg.reg[3] != cur[3]) {
#pragma unroll
            for (int j = 0; j < TN; ++j) lds[lds_pad(deposit4(j, cur) | tdep)] = v[j];
            __syncthreads();
#pragma unroll
            for (int i = 0; i < TRB; ++i) cur[i] = sg.reg[i];
            tdep = thread_deposit(t, cur);
#pragma unroll
            for (int j = 0; j < TN; ++j) v[j] = lds[lds_pad(deposit4(j, cur) | tdep)];
            __syncthreads();
        }
        for (int o = sg.op0; o < sg.op1; ++o) {
            const TOp op = ops[o];
            const c128 *M = pool + op.moff;
            if (op.type == 0) {
                c128 m00 = M[0], m10 = M[1], m01 = M[2], m11 = M[3];
                switch (op.a) {
                    case 0: reg_apply1<0>(v, m00, m10, m01, m11); break;
                    case 1: reg_apply1<1>(v, m00, m10, m01, m11); break;
                    case 2: reg_apply1<2>(v, m00, m10, m01, m11); break;
                    default: reg_apply1<3>(v, m00, m10, m01, m11); break;
                }
            } else if (op.type == 1) {
                switch (op.a * 4 + op.b) {
                    case 1: reg_apply2<0, 1>(v, M); break;
                    case 2: reg_apply2<0, 2>(v, M); break;
                    case 3: reg_apply2<0, 3>(v, M); break;
                    case 6: reg_apply2<1, 2>(v, M); break;
                    case 7: reg_apply2<1, 3>(v, M); break;
                    default: reg_apply2<2, 3>(v, M); break;
                }
            } else {
#pragma unroll
                for (int j = 0; j < TN; ++j) {
                    const int l = deposit4(j, cur) | tdep;
                    int idx = 0;
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        if (i < op.k) {
                            const int sbit = op.dsrc[i];
                            const int bit = sbit < 100 ? ((l >> sbit) & 1) : int((base >> (sbit - 100)) & 1);
                            idx |= bit << i;
                        }
                    }
                    v[j] = cmul(M[idx], v[j]);
                }
            }
        }
    }
    if (cur[0] != 8 || cur[1] != 9 || cur[2] != 10 || cur[3] != 11) {
#pragma unroll
        for (int j = 0; j < TN; ++j) lds[lds_pad(deposit4(j, cur) | tdep)] = v[j];
        __syncthreads();
        // load layout: registers on tile bits 8..11, thread bits 0..7
#pragma unroll
        for (int j = 0; j < TN; ++j) v[j] = lds[lds_pad((j << 8) | t)];
    }
#pragma unroll
    for (int j = 0; j < TN; ++j) state[gj[j]] = v[j];
}

// ------------------------------------------------------------------ planner ----
struct POp {
    int kind = 0;  // 0 dense, 1 diagonal, 2 unfusable (run as its own launch)
    int nb = 0;
    int bits[4] = {};
    std::vector<c128> m;  // dense: 2^nb x 2^nb column-major; diagonal: 2^nb values
    int step = -1;
    u64 mask() const {
        u64 r = 0;
        for (int i = 0; i < nb; ++i) r |= u64(1) << bits[i];
        return r;
    }
};

struct PlanItem {
    bool sweep = false;
    int step = -1;   // unfused step index (split-half flag in `half`)
    TSweepArgs args{};
};

struct FusedPlan {
    std::vector<PlanItem> items;
    std::vector<TSeg> h_segs;
    std::vector<TOp> h_ops;
    std::vector<c128> h_pool;
    TSeg *d_segs = nullptr;
    TOp *d_ops = nullptr;
    c128 *d_pool = nullptr;
    int nsweeps = 0;
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

// Lower the program's forward steps to plan operations.
static void lower_steps(const Program *p, std::vector<POp> &out) {
    for (size_t si = 0; si < p->steps.size(); ++si) {
        const Step &st = p->steps[si];
        const int dim = 1 << st.kb;
        if (st.split) {
            POp u;
            u.kind = 2;
            u.step = int(si);
            u.nb = 0;
            out.push_back(u);
            continue;
        }
        if (is_diag(st.ifwd) && st.kb <= 4) {
            if (st.ifwd.nchanged == 0) continue;  // identity
            POp d;
            d.kind = 1;
            d.nb = st.kb;
            for (int i = 0; i < st.kb; ++i) d.bits[i] = st.bits[i];
            d.m = diag_values(st.ifwd, dim);
            d.step = int(si);
            out.push_back(d);
            continue;
        }
        if (st.kb <= 2) {
            POp d;
            d.kind = 0;
            d.nb = st.kb;
            for (int i = 0; i < st.kb; ++i) d.bits[i] = st.bits[i];
            d.m = mat_of(p, st.fwd, dim);
            d.step = int(si);
            out.push_back(d);
            continue;
        }
        if (p->is_density && !st.channel && st.k == 2 && st.u_src >= 0) {
            // U(x)conj(U) = (U on the row bits) then (conj(U) on the column bits)
            std::vector<c128> U = mat_of(p, st.u_src, 4), Uc(16);
            for (int q = 0; q < 16; ++q) Uc[q] = cconj(U[q]);
            POp r, cl;
            r.kind = cl.kind = 0;
            r.nb = cl.nb = 2;
            r.bits[0] = st.bits[0];
            r.bits[1] = st.bits[1];
            cl.bits[0] = st.bits[2];
            cl.bits[1] = st.bits[3];
            r.m = U;
            cl.m = Uc;
            r.step = cl.step = int(si);
            out.push_back(r);
            out.push_back(cl);
            continue;
        }
        POp u;
        u.kind = 2;
        u.step = int(si);
        out.push_back(u);
    }
}

static void product(const std::vector<c128> &A, const std::vector<c128> &B, int d, std::vector<c128> &out) {
    // out = A * B (column-major d x d)
    out.assign(size_t(d) * d, mk(0, 0));
    for (int r = 0; r < d; ++r)
        for (int c = 0; c < d; ++c) {
            c128 acc = mk(0, 0);
            for (int k = 0; k < d; ++k) acc = cfma(A[r + d * k], B[k + d * c], acc);
            out[r + d * c] = acc;
        }
}

static bool commutes(const POp &a, const POp &b) {
    if (a.kind == 2 || b.kind == 2) {
        // unfusable steps: conservative (they commute only with disjoint supports;
        // their support is unknown here, so never)
        return false;
    }
    if ((a.mask() & b.mask()) == 0) return true;
    return a.kind == 1 && b.kind == 1;
}

static int popc(u64 x) { return __builtin_popcountll(x); }

int fused_plan_build(Program *p) {
    if (p->nbits < TM) return MOSSQ_OK;  // too small to tile: run unfused
    std::vector<POp> ops;
    lower_steps(p, ops);
    FusedPlan *fp = new FusedPlan();
    const u64 coal = (u64(1) << kCoalesceBits) - 1;
    std::vector<char> done(ops.size(), 0);
    size_t first = 0;
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
        // Greedy sweep with commutation lookahead.
        u64 T = coal;
        std::vector<int> chosen;
        std::vector<int> deferred;
        for (size_t i = first; i < ops.size(); ++i) {
            if (done[i]) continue;
            const POp &o = ops[i];
            bool blocked = false;
            for (int d : deferred)
                if (!commutes(ops[d], o)) {
                    blocked = true;
                    break;
                }
            if (!blocked && o.kind != 2) {
                u64 need = o.kind == 0 ? o.mask() : 0;
                if (popc(T | need) <= TM) {
                    T |= need;
                    chosen.push_back(int(i));
                    continue;
                }
            }
            deferred.push_back(int(i));
            if (deferred.size() > 64) break;  // bounded lookahead
        }
        for (int i : chosen) done[i] = 1;
        // fill the tile with the lowest unused bits
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
        std::vector<char> sched(chosen.size(), 0);
        int reg[TRB] = {8, 9, 10, 11};
        it.args.seg0 = int(fp->h_segs.size());
        size_t nsched = 0;
        auto ready = [&](size_t ci) {
            for (size_t pj = 0; pj < ci; ++pj)
                if (!sched[pj] && !commutes(ops[chosen[pj]], ops[chosen[ci]])) return false;
            return true;
        };
        auto fits = [&](size_t ci) {
            const POp &o = ops[chosen[ci]];
            if (o.kind == 1) return true;
            for (int q = 0; q < o.nb; ++q) {
                int l = local_of[o.bits[q]];
                if (l != reg[0] && l != reg[1] && l != reg[2] && l != reg[3]) return false;
            }
            return true;
        };
        bool first_seg = true;
        while (nsched < chosen.size()) {
            TSeg sg;
            for (int i = 0; i < TRB; ++i) sg.reg[i] = reg[i];
            sg.op0 = int(fp->h_ops.size());
            // emit every ready op that fits, repeatedly
            std::vector<int> seg_ops;
            bool progress = true;
            while (progress) {
                progress = false;
                for (size_t ci = 0; ci < chosen.size(); ++ci) {
                    if (sched[ci] || !fits(ci) || !ready(ci)) continue;
                    sched[ci] = 1;
                    ++nsched;
                    seg_ops.push_back(chosen[ci]);
                    progress = true;
                }
            }
            // merge adjacent dense ops on identical bits, lower to TOps
            std::vector<POp> merged;
            for (int oi : seg_ops) {
                const POp &o = ops[oi];
                if (!merged.empty() && o.kind == 0 && merged.back().kind == 0 && merged.back().nb == o.nb &&
                    merged.back().mask() == o.mask()) {
                    std::vector<c128> pr;
                    product(o.m, merged.back().m, 1 << o.nb, pr);
                    merged.back().m.swap(pr);
                    continue;
                }
                if (!merged.empty() && o.kind == 1 && merged.back().kind == 1 && merged.back().nb == o.nb &&
                    merged.back().mask() == o.mask()) {
                    for (size_t q = 0; q < o.m.size(); ++q) merged.back().m[q] = cmul(merged.back().m[q], o.m[q]);
                    continue;
                }
                merged.push_back(o);
            }
            for (const POp &o : merged) {
                TOp t{};
                t.moff = int(fp->h_pool.size());
                fp->h_pool.insert(fp->h_pool.end(), o.m.begin(), o.m.end());
                if (o.kind == 1) {
                    t.type = 2;
                    t.k = o.nb;
                    for (int q = 0; q < o.nb; ++q) {
                        int l = local_of[o.bits[q]];
                        t.dsrc[q] = l >= 0 ? l : 100 + o.bits[q];
                    }
                } else {
                    auto ridx = [&](int gb) {
                        int l = local_of[gb];
                        for (int i = 0; i < TRB; ++i)
                            if (reg[i] == l) return i;
                        return -1;
                    };
                    if (o.nb == 1) {
                        t.type = 0;
                        t.a = ridx(o.bits[0]);
                    } else {
                        t.type = 1;
                        t.a = ridx(o.bits[0]);
                        t.b = ridx(o.bits[1]);
                    }
                }
                fp->h_ops.push_back(t);
            }
            sg.op1 = int(fp->h_ops.size());
            if (sg.op1 > sg.op0 || first_seg) fp->h_segs.push_back(sg);
            first_seg = false;
            if (nsched >= chosen.size()) break;
            // next register set: bits of the first ready unscheduled dense op, then
            // bits of further ready dense ops, then fill
            std::vector<int> want;
            for (size_t ci = 0; ci < chosen.size() && want.size() < TRB; ++ci) {
                if (sched[ci] || !ready(ci)) continue;
                const POp &o = ops[chosen[ci]];
                if (o.kind != 0) continue;
                std::vector<int> add;
                for (int q = 0; q < o.nb; ++q) {
                    int l = local_of[o.bits[q]];
                    if (std::find(want.begin(), want.end(), l) == want.end()) add.push_back(l);
                }
                if (want.size() + add.size() <= TRB) want.insert(want.end(), add.begin(), add.end());
            }
            if (want.empty()) {
                // only blocked ops remain: cannot happen (a ready op always exists)
                delete fp;
                return fail(MOSSQ_EDEVICE, "fusion planner made no progress");
            }
            for (int l = TM - 1; l >= 0 && want.size() < TRB; --l)
                if (std::find(want.begin(), want.end(), l) == want.end()) want.push_back(l);
            std::sort(want.begin(), want.end());
            if (want.size() != TRB || want[0] < 0 || want[TRB - 1] >= TM ||
                std::adjacent_find(want.begin(), want.end()) != want.end()) {
                delete fp;
                return fail(MOSSQ_EDEVICE, "fusion planner produced an invalid register set");
            }
            for (int i = 0; i < TRB; ++i) reg[i] = want[i];
        }
        it.args.seg1 = int(fp->h_segs.size());
        fp->items.push_back(it);
        fp->nsweeps++;
    }
    // upload
    Ctx *c = p->ctx;
    int s = MOSSQ_OK;
    s = dev_alloc(c, std::max<size_t>(fp->h_segs.size(), 1) * sizeof(TSeg), (void **)&fp->d_segs);
    if (s == MOSSQ_OK) s = dev_alloc(c, std::max<size_t>(fp->h_ops.size(), 1) * sizeof(TOp), (void **)&fp->d_ops);
    if (s == MOSSQ_OK) s = dev_alloc(c, std::max<size_t>(fp->h_pool.size(), 1) * sizeof(c128), (void **)&fp->d_pool);
    if (s != MOSSQ_OK) {
        if (fp->d_segs) dev_free(c, fp->d_segs);
        if (fp->d_ops) dev_free(c, fp->d_ops);
        delete fp;
        return s;
    }
    if (!fp->h_segs.empty())
        MOSSQ_HIP(hipMemcpyAsync(fp->d_segs, fp->h_segs.data(), fp->h_segs.size() * sizeof(TSeg),
                                 hipMemcpyHostToDevice, c->stream));
    if (!fp->h_ops.empty())
        MOSSQ_HIP(hipMemcpyAsync(fp->d_ops, fp->h_ops.data(), fp->h_ops.size() * sizeof(TOp), hipMemcpyHostToDevice,
                                 c->stream));
    if (!fp->h_pool.empty())
        MOSSQ_HIP(hipMemcpyAsync(fp->d_pool, fp->h_pool.data(), fp->h_pool.size() * sizeof(c128),
                                 hipMemcpyHostToDevice, c->stream));
    MOSSQ_HIP(hipStreamSynchronize(c->stream));
    p->fused = fp;
    return MOSSQ_OK;
}

void fused_plan_free(Program *p) {
    if (!p->fused) return;
    FusedPlan *fp = p->fused;
    dev_free(p->ctx, fp->d_segs);
    dev_free(p->ctx, fp->d_ops);
    dev_free(p->ctx, fp->d_pool);
    delete fp;
    p->fused = nullptr;
}

int fused_plan_sweeps(const Program *p) {
    if (!p->fused) return int(p->steps.size());
    int unfused = 0;
    for (auto &it : p->fused->items)
        if (!it.sweep) ++unfused;
    return p->fused->nsweeps + unfused;
}

int fused_run(Ctx *c, const Program *p, c128 *state) {
    const FusedPlan *fp = p->fused;
    if (!fp) return run_steps(c, p, state, 0, int(p->steps.size()));
    const unsigned ntiles = unsigned(u64(1) << (p->nbits - TM));
    for (const PlanItem &it : fp->items) {
        if (!it.sweep) {
            MOSSQ_TRY(run_steps(c, p, state, it.step, it.step + 1));
            continue;
        }
        hipLaunchKernelGGL(k_tile, dim3(ntiles), dim3(TT), 0, c->stream, state, it.args, fp->d_segs, fp->d_ops,
                           fp->d_pool);
        MOSSQ_HIP(hipGetLastError());
    }
    return MOSSQ_OK;
}

}  // namespace mossq
