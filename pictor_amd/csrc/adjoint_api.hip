// adjoint_api.hip -- the adjoint engine of a context and its extern "C" entry points.
//
// Segment checkpointing: the forward pass keeps the particle state and E, B
// at every `segment`-th step.  The backward pass walks the segments from the
// last to the first; for each it restores the checkpoint, recomputes the
// segment's forward steps storing every step's particle state and node
// fields, then runs the reverse steps of that segment.  Device memory:
// (segment + nsteps/segment + 2) particle states plus one set of node fields
// per step of a segment.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>

#include "adjoint.h"
#include "ctx.h"

using namespace pictor;

namespace {

constexpr double kTwo36 = 68719476736.0;

template <typename R>
hipError_t alloc_part(EPart<R> &p, long long n, long long &bytes)
{
    hipError_t e;
    p = EPart<R>{};
    if ((e = dalloc(&p.cs, n)) != hipSuccess) return e;
    R **arr[6] = {&p.ox, &p.oy, &p.oz, &p.u, &p.v, &p.w};
    for (int i = 0; i < 6; i++)
        if ((e = dalloc(arr[i], n)) != hipSuccess) return e;
    bytes += n * (4 + 6 * (long long)sizeof(R));
    return hipSuccess;
}

template <typename R>
void free_part(EPart<R> &p)
{
    dfree(p.cs); dfree(p.ox); dfree(p.oy); dfree(p.oz); dfree(p.u); dfree(p.v); dfree(p.w);
    p = EPart<R>{};
}

template <typename R>
hipError_t copy_part(EPart<R> dst, EPart<R> src, long long n, hipStream_t s)
{
    hipError_t e;
    if ((e = hipMemcpyAsync(dst.cs, src.cs, n * 4, hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    R *d[6] = {dst.ox, dst.oy, dst.oz, dst.u, dst.v, dst.w};
    R *a[6] = {src.ox, src.oy, src.oz, src.u, src.v, src.w};
    for (int i = 0; i < 6; i++)
        if ((e = hipMemcpyAsync(d[i], a[i], n * sizeof(R), hipMemcpyDeviceToDevice, s)) != hipSuccess) return e;
    return hipSuccess;
}

template <typename R>
struct AdjEngine : AdjEngineBase {
    pictor_ctx *c = nullptr;
    AdjP ap{};
    double ctx_jscale = 1.0;   // scale of J in the E update (split form: 2^-36 qi)
    long long n = 0;
    int nsteps = 0, seg = 1, nseg = 0;
    long long ncell = 0, bytes = 0;
    bool have_forward = false;

    R *E[3] = {}, *B[3] = {};
    double *J[3] = {};
    std::vector<R *> ckE, ckB;          // per checkpoint: 3 + 3 arrays
    std::vector<EPart<R>> ck;           // particle checkpoints
    std::vector<EPart<R>> st;           // states of one segment (seg + 1)
    std::vector<void *> nodE, nodB;     // node fields of one segment (seg)
    EPart<R> adj{};
    double *Eb[3] = {}, *Bb[3] = {}, *gE[3] = {}, *gB[3] = {};
    double *partial = nullptr;
    int *d_err = nullptr;
    double objective = 0;
    double t_fwd = 0, t_rec = 0, t_adj = 0;
    hipEvent_t ev[4] = {};

    ~AdjEngine() override { release(); }

    void release()
    {
        if (c && c->stream) (void)hipStreamSynchronize(c->stream);
        for (int i = 0; i < 3; i++) {
            dfree(E[i]); dfree(B[i]); dfree(J[i]); dfree(Eb[i]); dfree(Bb[i]); dfree(gE[i]); dfree(gB[i]);
            E[i] = B[i] = nullptr; J[i] = Eb[i] = Bb[i] = gE[i] = gB[i] = nullptr;
        }
        for (R *p : ckE) dfree(p);
        for (R *p : ckB) dfree(p);
        ckE.clear(); ckB.clear();
        for (auto &p : ck) free_part(p);
        for (auto &p : st) free_part(p);
        ck.clear(); st.clear();
        for (void *p : nodE) dfree(p);
        for (void *p : nodB) dfree(p);
        nodE.clear(); nodB.clear();
        free_part(adj);
        dfree(partial); partial = nullptr;
        dfree(d_err); d_err = nullptr;
        for (auto &e : ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
        bytes = 0;
        have_forward = false;
    }

    int setup(pictor_ctx *ctx, long long np, int ns, int sg)
    {
        release();
        c = ctx;
        n = np; nsteps = ns; seg = sg;
        nseg = (nsteps + seg - 1) / seg;
        const GridP &g = c->g;
        const PhysP &ph = c->ph;
        ncell = g.ncell;
        ap.nx = g.nx; ap.ny = g.ny; ap.mx = g.mx; ap.my = g.my;
        ap.c = ph.c; ap.sqc = ph.sqc; ap.cinv = ph.cinv;
        const bool f32 = std::is_same<R, float>::value;
        for (int s = 0; s < 2; s++) {
            ap.qm[s] = ph.flvrqm[s] * 0.5;
            ap.qs[s] = f32 ? ph.flvq[s] * kTwo36 : ph.flvq[s] * ph.qi;
        }
        // the field update scales J by ctx_jscale; the adjoint current is -ctx_jscale * Eb
        ctx_jscale = f32 ? ph.qi / kTwo36 : 1.0;
        ap.jscale = ctx_jscale;
        for (int i = 0; i < 3; i++) {
            HIPTRY(dalloc(&E[i], ncell)); HIPTRY(dalloc(&B[i], ncell)); HIPTRY(dalloc(&J[i], ncell));
            HIPTRY(dalloc(&Eb[i], ncell)); HIPTRY(dalloc(&Bb[i], ncell));
            HIPTRY(dalloc(&gE[i], ncell)); HIPTRY(dalloc(&gB[i], ncell));
        }
        bytes += ncell * (6 * (long long)sizeof(R) + 15 * 8);
        ckE.assign((size_t)nseg * 3, nullptr);
        ckB.assign((size_t)nseg * 3, nullptr);
        for (size_t i = 0; i < ckE.size(); i++) {
            HIPTRY(dalloc(&ckE[i], ncell)); HIPTRY(dalloc(&ckB[i], ncell));
        }
        bytes += (long long)nseg * 6 * ncell * sizeof(R);
        ck.assign(nseg, EPart<R>{});
        for (auto &p : ck) HIPTRY(alloc_part(p, n, bytes));
        st.assign(seg + 1, EPart<R>{});
        for (int k = 1; k <= seg; k++) HIPTRY(alloc_part(st[k], n, bytes));
        nodE.assign(seg, nullptr);
        nodB.assign(seg, nullptr);
        for (int k = 0; k < seg; k++) {
            HIPTRY(hipMalloc(&nodE[k], ncell * 4 * sizeof(R)));
            HIPTRY(hipMalloc(&nodB[k], ncell * 4 * sizeof(R)));
            HIPTRY(hipMemsetAsync(nodE[k], 0, ncell * 4 * sizeof(R), c->stream));
            HIPTRY(hipMemsetAsync(nodB[k], 0, ncell * 4 * sizeof(R), c->stream));
        }
        bytes += (long long)seg * 8 * ncell * sizeof(R);
        HIPTRY(alloc_part(adj, n, bytes));
        HIPTRY(dalloc(&partial, 2048));
        HIPTRY(dalloc(&d_err, 1));
        for (auto &e : ev) HIPTRY(hipEventCreate(&e));
        return PICTOR_OK;
    }

    int load(const double *q, const double *x, const double *y, const double *z, const double *u, const double *v,
             const double *w, const int32_t *flv)
    {
        for (long long i = 0; i < n; i++) {
            if (flv[i] != 1 && flv[i] != 2) return fail(PICTOR_ERR_ARG, "adjoint: flavour must be 1 or 2");
            if (q[i] != c->ph.flvq[flv[i] - 1])
                return fail(PICTOR_ERR_ARG, "adjoint: q must be the flavour charge (+1 / -1)");
        }
        double *tmp[6] = {};
        int *tf = nullptr;
        const double *src[6] = {x, y, z, u, v, w};
        for (int i = 0; i < 6; i++) {
            HIPTRY(dalloc(&tmp[i], n));
            HIPTRY(hipMemcpyAsync(tmp[i], src[i], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        }
        HIPTRY(dalloc(&tf, n));
        HIPTRY(hipMemcpyAsync(tf, flv, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPTRY(hipMemsetAsync(d_err, 0, sizeof(int), c->stream));
        const double *const raw[6] = {tmp[0], tmp[1], tmp[2], tmp[3], tmp[4], tmp[5]};
        launch_eng_load<R>(ap, raw, tf, ck[0], n, d_err, c->stream);
        HIPTRY(hipGetLastError());
        int err = 0;
        HIPTRY(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));
        for (int i = 0; i < 6; i++) dfree(tmp[i]);
        dfree(tf);
        if (err) return fail(PICTOR_ERR_ARG, "adjoint: particle outside the physical cells or 2-D z outside [1,2)");
        return PICTOR_OK;
    }

    // one forward step: in -> out, node fields into (nE, nB)
    int fwd_step(EPart<R> in, EPart<R> out, void *nE, void *nB)
    {
        hipStream_t s = c->stream;
        launch_update_b_half<R>(c->g, c->ph, E, B, s);
        launch_ghost_copy<R>(c->g, B, s);
        launch_eng_nodes<R>(ap, E, B, nE, nB, s);
        for (int i = 0; i < 3; i++) launch_fill_zero<double>(J[i], ncell, s);
        launch_eng_push<R>(ap, nE, nB, in, out, J, n, s);
        launch_update_b_half<R>(c->g, c->ph, E, B, s);
        launch_ghost_copy<R>(c->g, B, s);
        launch_current_fold<double>(c->g, J, s);
        launch_update_e_add_current<R, double>(c->g, c->ph, E, B, J, ctx_jscale, s);
        launch_ghost_copy<R>(c->g, E, s);
        HIPTRY(hipGetLastError());
        return PICTOR_OK;
    }

    // one reverse step with the forward state of that step
    int rev_step(EPart<R> stn, void *nE, void *nB)
    {
        hipStream_t s = c->stream;
        const GridP &g = c->g;
        PhysP pa = c->ph, pb = c->ph;
        pa.fld_halfc = -c->ph.fldc;      // B-half kernel as the transpose of the E update's curl
        pb.fldc = -c->ph.fld_halfc;      // E kernel (curl only) as the transpose of a B half step
        launch_ghost_copy<double>(g, Eb, s);
        launch_update_b_half<double>(g, pa, Eb, Bb, s);
        for (int i = 0; i < 3; i++) {
            launch_fill_zero<double>(gE[i], ncell, s);
            launch_fill_zero<double>(gB[i], ncell, s);
        }
        launch_eng_adjoint<R>(ap, nE, nB, stn, Eb, adj, gE, gB, n, s);
        launch_ghost_copy<double>(g, Bb, s);
        launch_update_e<double>(g, pb, Eb, Bb, s);
        launch_current_fold<double>(g, gE, s);
        launch_current_fold<double>(g, gB, s);
        launch_ghost_copy<double>(g, gE, s);
        launch_ghost_copy<double>(g, gB, s);
        launch_node_to_stag(ap, gE, gB, Eb, Bb, s);
        launch_ghost_copy<double>(g, Bb, s);
        launch_update_e<double>(g, pb, Eb, Bb, s);
        HIPTRY(hipGetLastError());
        return PICTOR_OK;
    }

    int save_ck(int k)
    {
        for (int i = 0; i < 3; i++) {
            HIPTRY(hipMemcpyAsync(ckE[3 * k + i], E[i], ncell * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
            HIPTRY(hipMemcpyAsync(ckB[3 * k + i], B[i], ncell * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
        }
        return PICTOR_OK;
    }

    int restore_ck(int k)
    {
        for (int i = 0; i < 3; i++) {
            HIPTRY(hipMemcpyAsync(E[i], ckE[3 * k + i], ncell * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
            HIPTRY(hipMemcpyAsync(B[i], ckB[3 * k + i], ncell * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
        }
        return PICTOR_OK;
    }

    int forward()
    {
        hipStream_t s = c->stream;
        // initial fields: the context's E and B
        for (int i = 0; i < 3; i++) {
            HIPTRY(hipMemcpyAsync(E[i], c->E[i], ncell * sizeof(R), hipMemcpyDeviceToDevice, s));
            HIPTRY(hipMemcpyAsync(B[i], c->B[i], ncell * sizeof(R), hipMemcpyDeviceToDevice, s));
        }
        launch_ghost_copy<R>(c->g, E, s);
        launch_ghost_copy<R>(c->g, B, s);
        HIPTRY(hipEventRecord(ev[0], s));
        int rc;
        // ping-pong between st[1] and st[2] (st[1] alone when a segment is one step)
        EPart<R> a = ck[0];
        for (int t = 0; t < nsteps; t++) {
            if (t % seg == 0) {
                const int k = t / seg;
                if (k > 0 && (rc = copy_part_rc(ck[k], a))) return rc;
                if ((rc = save_ck(k))) return rc;
                if (k > 0) a = ck[k];
            }
            EPart<R> b = (seg >= 2) ? st[1 + (t % 2)] : st[1];
            if ((rc = fwd_step(a, b, nodE[0], nodB[0]))) return rc;
            a = b;
        }
        HIPTRY(hipEventRecord(ev[1], s));
        // objective: 1/2 sum B^2 over the physical cells
        launch_field_energy<R>(c->g, E, B, partial, 1024, s);
        std::vector<double> part(2048);
        HIPTRY(hipMemcpyAsync(part.data(), partial, 2048 * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPTRY(hipStreamSynchronize(s));
        double be = 0;
        for (int b = 0; b < 1024; b++) be += part[2 * b];
        objective = 0.5 * be;
        float ms = 0;
        HIPTRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        t_fwd = ms;
        have_forward = true;
        return PICTOR_OK;
    }

    int copy_part_rc(EPart<R> dst, EPart<R> src)
    {
        HIPTRY(copy_part(dst, src, n, c->stream));
        return PICTOR_OK;
    }

    int backward(double scale, double *du, double *dv, double *dw)
    {
        if (!have_forward) return fail(PICTOR_ERR_ARG, "pictor_adjoint_backward without pictor_adjoint_forward");
        hipStream_t s = c->stream;
        int rc;
        // seeds: dL/dB = B at the final step, dL/dE = 0, particles 0
        for (int i = 0; i < 3; i++) {
            launch_fill_zero<double>(Eb[i], ncell, s);
            launch_seed_b<R>(B[i], Bb[i], ncell, s);
        }
        launch_eng_zero<R>(adj, n, s);
        t_rec = t_adj = 0;
        for (int k = nseg - 1; k >= 0; k--) {
            const int t0 = k * seg, t1 = std::min(nsteps, t0 + seg), L = t1 - t0;
            HIPTRY(hipEventRecord(ev[0], s));
            if ((rc = restore_ck(k))) return rc;
            // state j of the segment: the checkpoint for j = 0, st[j] after that
            auto S = [&](int j) { return j == 0 ? ck[k] : st[j]; };
            for (int j = 0; j < L; j++)
                if ((rc = fwd_step(S(j), st[j + 1], nodE[j], nodB[j]))) return rc;
            HIPTRY(hipEventRecord(ev[1], s));
            for (int j = L - 1; j >= 0; j--)
                if ((rc = rev_step(S(j), nodE[j], nodB[j]))) return rc;
            HIPTRY(hipEventRecord(ev[2], s));
            HIPTRY(hipEventSynchronize(ev[2]));
            float m1 = 0, m2 = 0;
            HIPTRY(hipEventElapsedTime(&m1, ev[0], ev[1]));
            HIPTRY(hipEventElapsedTime(&m2, ev[1], ev[2]));
            t_rec += m1;
            t_adj += m2;
        }
        if (du || dv || dw) {
            double *d[3] = {};
            for (int i = 0; i < 3; i++) HIPTRY(dalloc(&d[i], n));
            launch_eng_store_grad<R>(adj, scale, d[0], d[1], d[2], n, s);
            double *h[3] = {du, dv, dw};
            for (int i = 0; i < 3; i++)
                if (h[i]) HIPTRY(hipMemcpyAsync(h[i], d[i], n * sizeof(double), hipMemcpyDeviceToHost, s));
            HIPTRY(hipStreamSynchronize(s));
            for (int i = 0; i < 3; i++) dfree(d[i]);
        }
        HIPTRY(hipStreamSynchronize(s));
        return PICTOR_OK;
    }
};

template <typename R>
AdjEngine<R> *engine_of(pictor_ctx *c)
{
    return dynamic_cast<AdjEngine<R> *>(c->adj);
}

int check_ctx2(pictor_ctx *c)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    if (hipSetDevice(c->dev) != hipSuccess) return fail(PICTOR_ERR_HIP, "hipSetDevice failed");
    return PICTOR_OK;
}

template <typename R>
int adj_forward(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y, const double *z,
                const double *u, const double *v, const double *w, const int32_t *flv, int32_t nsteps,
                int32_t segment, double *objective)
{
    AdjEngine<R> *e = engine_of<R>(c);
    if (!e) {
        delete c->adj;
        c->adj = nullptr;
        e = new AdjEngine<R>();
        c->adj = e;
    }
    int rc = e->setup(c, n, nsteps, segment);
    if (!rc) rc = e->load(q, x, y, z, u, v, w, flv);
    if (!rc) rc = e->forward();
    if (rc) {
        e->release();
        return rc;
    }
    if (objective) *objective = e->objective;
    return PICTOR_OK;
}

}  // namespace

extern "C" {

int pictor_adjoint_forward(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                           const double *z, const double *u, const double *v, const double *w, const int32_t *flv,
                           int32_t nsteps, int32_t segment, double *objective)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (c->g.ndim != 2) return fail(PICTOR_ERR_ARG, "the adjoint is available for ndim = 2");
    if (c->ph.fdtd != PICTOR_FDTD_YEE) return fail(PICTOR_ERR_ARG, "the adjoint is available for the Yee solver");
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "adjoint inside a split step");
    if (n < 1 || nsteps < 1 || segment < 1) return fail(PICTOR_ERR_ARG, "n, nsteps and segment must be >= 1");
    if (!q || !x || !y || !z || !u || !v || !w || !flv) return fail(PICTOR_ERR_ARG, "null particle array");
    if (segment > nsteps) segment = nsteps;
    if (c->form == PICTOR_FORM_FP64) return adj_forward<double>(c, n, q, x, y, z, u, v, w, flv, nsteps, segment, objective);
    return adj_forward<float>(c, n, q, x, y, z, u, v, w, flv, nsteps, segment, objective);
}

int pictor_adjoint_backward(pictor_ctx *c, double scale, double *du, double *dv, double *dw)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (auto *e = engine_of<double>(c)) return e->backward(scale, du, dv, dw);
    if (auto *e = engine_of<float>(c)) return e->backward(scale, du, dv, dw);
    return fail(PICTOR_ERR_ARG, "pictor_adjoint_backward without pictor_adjoint_forward");
}

int pictor_adjoint_stats(pictor_ctx *c, double out[8])
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (!out) return fail(PICTOR_ERR_ARG, "null argument");
    for (int i = 0; i < 8; i++) out[i] = 0;
    auto fill = [&](auto *e) {
        out[0] = e->t_fwd;
        out[1] = e->t_rec;
        out[2] = e->t_adj;
        out[3] = (double)e->bytes;
        out[4] = e->nsteps;
        out[5] = e->seg;
        out[6] = e->nseg;
        out[7] = e->objective;
    };
    if (auto *e = engine_of<double>(c)) fill(e);
    else if (auto *e = engine_of<float>(c)) fill(e);
    else return fail(PICTOR_ERR_ARG, "no adjoint run on this context");
    return PICTOR_OK;
}

int pictor_adjoint_release(pictor_ctx *c)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    delete c->adj;
    c->adj = nullptr;
    return PICTOR_OK;
}

}  // extern "C"
