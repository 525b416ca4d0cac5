// Batched adjoint gradients over independent programs with uneven qubit counts.
#include <algorithm>

#include "mossq_program.h"

namespace mossq {

// Plain per-circuit path: every item's adjoint is enqueued back to back on the
// context stream into device result slots; one host synchronisation at the end.
static int batch_adjoint_sequential(Ctx *c, int B, const Program *const *progs, const Obs *const *obs,
                                    const int64_t *grad_offset, double *values, double *grads) {
    int nmax = 0, Gmax = 1, Dmax = 1;
    i64 Gtot = 0, Dtot = 0;
    for (int b = 0; b < B; ++b) {
        nmax = std::max(nmax, progs[b]->n);
        Gmax = std::max(Gmax, obs[b]->G);
        Dmax = std::max(Dmax, progs[b]->nderiv);
        Gtot += obs[b]->G;
        Dtot += progs[b]->nderiv;
    }
    const u64 len = u64(1) << nmax;
    const int P = pair_parts(nmax);
    u64 nparts = std::max<u64>(u64(Gmax) * 512, u64(Dmax) * P);
    c128 *a = nullptr, *bb = nullptr, *part = nullptr, *vals = nullptr, *grd = nullptr;
    int st = dev_alloc(c, len * sizeof(c128), (void **)&a);
    if (st == MOSSQ_OK) st = dev_alloc(c, len * sizeof(c128), (void **)&bb);
    if (st == MOSSQ_OK) st = dev_alloc(c, nparts * sizeof(c128), (void **)&part);
    if (st == MOSSQ_OK) st = dev_alloc(c, std::max<i64>(Gtot, 1) * sizeof(c128), (void **)&vals);
    if (st == MOSSQ_OK) st = dev_alloc(c, std::max<i64>(Dtot, 1) * sizeof(c128), (void **)&grd);
    i64 go = 0, dofs = 0;
    for (int b = 0; st == MOSSQ_OK && b < B; ++b) {
        AdjointScratch s;
        s.a = a;
        s.b = bb;
        s.part = part;
        s.value_dev = vals + go;
        s.grads_dev = grd + dofs;
        st = adjoint_pure_enqueue(c, progs[b], obs[b], nullptr, nullptr, s);
        go += obs[b]->G;
        dofs += progs[b]->nderiv;
    }
    if (st == MOSSQ_OK) {
        std::vector<c128> hv(std::max<i64>(Gtot, 1)), hg(std::max<i64>(Dtot, 1));
        hipError_t e = hipMemcpyAsync(hv.data(), vals, Gtot * sizeof(c128), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && Dtot)
            e = hipMemcpyAsync(hg.data(), grd, Dtot * sizeof(c128), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) {
            st = fail(MOSSQ_EDEVICE, std::string("batch adjoint copy: ") + hipGetErrorString(e));
        } else {
            go = dofs = 0;
            for (int b = 0; b < B; ++b) {
                double v = 0;
                for (int g = 0; g < obs[b]->G; ++g) v += hv[go + g].re;
                values[b] = v;
                for (int j = 0; j < progs[b]->nderiv; ++j) grads[grad_offset[b] + j] = hg[dofs + j].re;
                go += obs[b]->G;
                dofs += progs[b]->nderiv;
            }
        }
    }
    for (c128 *p : {a, bb, part, vals, grd})
        if (p) dev_free(c, p);
    return st;
}

}  // namespace mossq

using namespace mossq;

extern "C" {

int mossq_batch_adjoint(mossq_ctx *ctx, int B, mossq_program *const *progs, mossq_obs *const *obs,
                        const int64_t *grad_offset, double *values, double *grads) {
    MOSSQ_TRY(ctx_check(ctx));
    MOSSQ_CHECK(B >= 0 && (B == 0 || (progs && obs && grad_offset && values && grads)), MOSSQ_EINVAL,
                "bad batch arguments");
    Ctx *c = C(ctx);
    std::vector<const Program *> P(B);
    std::vector<const Obs *> O(B);
    for (int b = 0; b < B; ++b) {
        P[b] = reinterpret_cast<const Program *>(progs[b]);
        O[b] = reinterpret_cast<const Obs *>(obs[b]);
        MOSSQ_CHECK(P[b] && O[b], MOSSQ_EINVAL, "null program or observable in batch");
        MOSSQ_CHECK(P[b]->ctx == c && O[b]->ctx == c, MOSSQ_EINVAL, "batch item belongs to another context");
        MOSSQ_CHECK(!P[b]->is_density, MOSSQ_EUNSUPPORTED, "batched adjoint takes state-vector programs");
        MOSSQ_CHECK(O[b]->n == P[b]->n, MOSSQ_EINVAL, "observable and program qubit counts differ");
    }
    if (B == 0) return MOSSQ_OK;
    return batch_adjoint_sequential(c, B, P.data(), O.data(), grad_offset, values, grads);
}

}  // extern "C"
