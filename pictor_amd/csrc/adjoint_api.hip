// adjoint_api.hip -- the adjoint engine of a context and its extern "C" entry points.
//
// Segment checkpointing: the forward pass keeps the particle state and E, B
// at every `segment`-th step.  The backward pass walks the segments from the
// last to the first; for each it restores the checkpoint, recomputes the
// segment's forward steps storing every step's particle state and fields,
// then runs the reverse steps of that segment.
//
// Two engines:
//  * tile-sorted (FP32 split form only): the particles live in the
//    production tile bins and every forward step is the production push
//    kernel in its index-carrying variant, which also records for every
//    input slot the slot its particle went to.  The reverse step walks the
//    same bins (adjoint_fast.hip).  Recomputed states hold the same particles
//    as the forward pass's, possibly in another order inside a bin (slot
//    reservation order varies between runs); the particle ids map the
//    adjoint from one to the other at segment boundaries.
//  * unsorted (both forms): particles in the caller's order, one thread per
//    particle, node fields of every step (adjoint.hip).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>

#include "adjoint.h"
#include "ctx.h"

using namespace pictor;

namespace {

constexpr double kFixQ = kFixScale;   // the split form's current counts per charge cell
constexpr int kRetryLayout = -100;   // internal: tile bins overflowed, redo with larger bins

template <typename R>
hipError_t alloc_part(EPart<R> &p, long long n, long long &bytes, bool with_cs = true)
{
    hipError_t e;
    p = EPart<R>{};
    if (with_cs && (e = dalloc(&p.cs, n)) != hipSuccess) return e;
    R **arr[6] = {&p.ox, &p.oy, &p.oz, &p.u, &p.v, &p.w};
    for (int i = 0; i < 6; i++)
        if ((e = dalloc(arr[i], n)) != hipSuccess) return e;
    bytes += n * ((with_cs ? 4 : 0) + 6 * (long long)sizeof(R));
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

// split-form particle storage of `cap` slots: with_id for the engine's states, with_sp for the spill buffer
hipError_t alloc_split_state(PartSplit &p, long long cap, bool with_id, bool with_sp, long long &bytes)
{
    hipError_t e;
    p = PartSplit{};
    if ((e = dalloc(&p.cell, cap)) != hipSuccess) return e;
    // the 2-D split record: no z offset (kZMid)
    float **arr[5] = {&p.ox, &p.oy, &p.u, &p.v, &p.w};
    for (int i = 0; i < 5; i++)
        if ((e = dalloc(arr[i], cap)) != hipSuccess) return e;
    if (with_id && (e = dalloc(&p.id, cap)) != hipSuccess) return e;
    if (with_sp && (e = dalloc(&p.sp, cap)) != hipSuccess) return e;
    bytes += cap * (24LL + (with_id ? 4 : 0) + (with_sp ? 4 : 0));
    return hipSuccess;
}

void free_split_state(PartSplit &p)
{
    dfree(p.cell); dfree(p.ox); dfree(p.oy); dfree(p.oz); dfree(p.u); dfree(p.v); dfree(p.w);
    dfree(p.id); dfree(p.sp);
    p = PartSplit{};
}

// Bin layout of the tile-sorted engine from the particle count of every
// (tile, species) group: the production rule (a centre bin for the group,
// side bins for the flux through a face or corner in one step) with the
// centre scaled by cf and the sides doubled, since the engine keeps its
// layout for the whole run instead of re-laying out on overflow.
void engine_layout(const TileP &tp, double c, const std::vector<long long> &counts, double cf,
                   std::vector<long long> &base, std::vector<int> &cap, long long &total)
{
    const double frac = std::min(1.0, c / tp.T);
    base.assign((size_t)tp.ntiles * tp.NB, 0);
    cap.assign((size_t)tp.ntiles * tp.NB, 0);
    long long off = 0;
    for (int t = 0; t < tp.ntiles; t++)
        for (int s = 0; s < tp.nsp; s++) {
            const long long n = counts[(size_t)t * tp.nsp + s];
            for (int d = 0; d < tp.ND; d++) {
                const int nz = (d % 3 != 1) + (d / 3 != 1);
                double cp;
                if (nz == 0) cp = std::ceil(n * cf) + 64;
                else cp = 2.0 * (std::ceil(n * std::pow(frac, nz) * 1.6 * std::pow(2.0, nz - 1)) + (nz == 1 ? 32 : 16));
                const long long ci = ((long long)cp + 15) / 16 * 16;
                const size_t b = (size_t)t * tp.NB + s * tp.ND + d;
                base[b] = off;
                cap[b] = (int)std::min<long long>(ci, 0x7fffffff);
                off += ci;
            }
        }
    total = off;
}

// flavour and charge of every particle of caller arrays in device memory: bit 0 a
// flavour other than 1 or 2, bit 1 a charge other than the flavour charge
__global__ void k_adj_check(const double *__restrict__ q, const int32_t *__restrict__ flv, long long n, double q1,
                            double q2, int *__restrict__ err)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int f = flv[i];
    if (f != 1 && f != 2) atomicOr(err, 1);
    else if (q[i] != (f == 1 ? q1 : q2)) atomicOr(err, 2);
}

template <typename R>
struct AdjEngine : AdjEngineBase {
    pictor_ctx *c = nullptr;
    // where the caller's particle arrays (forward) and gradients (backward) live:
    // host memory, or device memory of this context's GPU
    hipMemcpyKind kin = hipMemcpyHostToDevice, kout = hipMemcpyDeviceToHost;
    AdjP ap{};
    double ctx_jscale = 1.0;   // scale of J in the E update (split form: 2^-30 qi)
    long long n = 0;
    int nsteps = 0, seg = 1, nseg = 0;
    long long ncell = 0, bytes = 0;
    bool have_forward = false;

    R *E[3] = {}, *B[3] = {};
    double *J[3] = {};
    std::vector<R *> ckE, ckB;          // per checkpoint: 3 + 3 arrays
    // FP32 split form: the low part of E (E = hi + lo), as the context keeps it
    float *Elo[3] = {};
    std::vector<float *> ckElo;
    double *Eb[3] = {}, *Bb[3] = {}, *gE[3] = {}, *gB[3] = {};
    // current filter (curr_filter > 0): scratch, and the filtered adjoint current
    // the particle reverse step reads (Jadj = Ebf, else Eb)
    double *Jt[3] = {}, *Ebf[3] = {};
    double *Jadj[3] = {};
    double *partial = nullptr;
    int *d_err = nullptr;
    double objective = 0;
    double t_fwd = 0, t_rec = 0, t_adj = 0;
    hipEvent_t ev[4] = {};

    // unsorted engine: particles in the caller's order
    std::vector<EPart<R>> ck;           // particle checkpoints
    std::vector<EPart<R>> st;           // states of one segment (seg + 1)
    std::vector<void *> nodE, nodB;     // node fields of one segment (seg)
    EPart<R> adj{};

    // tile-sorted engine (R = float): the production tile bins
    bool fast = false;
    double cap_factor = 1.5;            // centre-bin capacity / particles of the group at load time
    int relayouts = 0;                  // forward passes repeated with larger bins
    std::vector<long long> h_base;
    std::vector<int> h_cap;
    long long lay_total = 0;
    long long *d_base = nullptr;
    int *d_cap = nullptr;
    std::vector<PartSplit> ck2;         // particle checkpoints (nseg)
    std::vector<int *> ck_cnt;          // their bin counts
    std::vector<PartSplit> st2;         // states 1.. of a segment
    std::vector<int *> cntv;            // their bin counts
    std::vector<int *> dstv;            // per step of a segment: output slot of every input slot
    std::vector<float *> sfE, sfB;      // per step of a segment: E and B as the push saw them
    EPart<float> fa[2] = {};            // particle adjoints, indexed by slot
    int *pos = nullptr;                 // particle id -> slot
    PartSplit spill{};
    int spill_cap = 0;
    int *d_spill = nullptr;

    ~AdjEngine() override { release(); }

    void release()
    {
        if (c && c->stream) (void)hipStreamSynchronize(c->stream);
        for (int i = 0; i < 3; i++) {
            dfree(E[i]); dfree(B[i]); dfree(J[i]); dfree(Eb[i]); dfree(Bb[i]); dfree(gE[i]); dfree(gB[i]);
            dfree(Jt[i]); dfree(Ebf[i]);
            E[i] = B[i] = nullptr; J[i] = Eb[i] = Bb[i] = gE[i] = gB[i] = nullptr;
            Jt[i] = Ebf[i] = Jadj[i] = nullptr;
        }
        for (R *p : ckE) dfree(p);
        for (R *p : ckB) dfree(p);
        ckE.clear(); ckB.clear();
        for (int i = 0; i < 3; i++) { dfree(Elo[i]); Elo[i] = nullptr; }
        for (float *p : ckElo) dfree(p);
        ckElo.clear();
        for (auto &p : ck) free_part(p);
        for (auto &p : st) free_part(p);
        ck.clear(); st.clear();
        for (void *p : nodE) dfree(p);
        for (void *p : nodB) dfree(p);
        nodE.clear(); nodB.clear();
        free_part(adj);
        for (auto &p : ck2) free_split_state(p);
        for (auto &p : st2) free_split_state(p);
        ck2.clear(); st2.clear();
        for (int *p : ck_cnt) dfree(p);
        for (int *p : cntv) dfree(p);
        for (int *p : dstv) dfree(p);
        ck_cnt.clear(); cntv.clear(); dstv.clear();
        for (float *p : sfE) dfree(p);
        for (float *p : sfB) dfree(p);
        sfE.clear(); sfB.clear();
        free_part(fa[0]); free_part(fa[1]);
        free_split_state(spill);
        dfree(d_base); dfree(d_cap); dfree(pos); dfree(d_spill);
        d_base = nullptr; d_cap = nullptr; pos = nullptr; d_spill = nullptr;
        h_base.clear(); h_cap.clear();
        lay_total = 0;
        dfree(partial); partial = nullptr;
        dfree(d_err); d_err = nullptr;
        for (auto &e : ev) if (e) { (void)hipEventDestroy(e); e = nullptr; }
        bytes = 0;
        have_forward = false;
    }

    int setup(pictor_ctx *ctx, long long np, int ns, int sg, bool use_fast, double cf)
    {
        release();
        c = ctx;
        fast = use_fast && std::is_same<R, float>::value;
        cap_factor = cf;
        n = np; nsteps = ns; seg = sg;
        nseg = (nsteps + seg - 1) / seg;
        const GridP &g = c->g;
        const PhysP &ph = c->ph;
        ncell = g.ncell;
        ap.nx = g.nx; ap.ny = g.ny; ap.mx = g.mx; ap.my = g.my;
        ap.c = ph.c; ap.sqc = ph.sqc; ap.cinv = ph.cinv; ap.inv_d = ph.inv_dx;
        const bool f32 = std::is_same<R, float>::value;
        for (int s = 0; s < 2; s++) {
            ap.qm[s] = ph.flvrqm[s] * 0.5;
            ap.qs[s] = f32 ? ph.flvq[s] * kFixQ : ph.flvq[s] * ph.qi;
        }
        // the field update scales J by ctx_jscale; the adjoint current is -ctx_jscale * Eb
        ctx_jscale = f32 ? ph.qi / kFixQ : 1.0;
        ap.jscale = ctx_jscale;
        for (int i = 0; i < 3; i++) {
            HIPTRY(dalloc(&E[i], ncell)); HIPTRY(dalloc(&B[i], ncell)); HIPTRY(dalloc(&J[i], ncell));
            HIPTRY(dalloc(&Eb[i], ncell)); HIPTRY(dalloc(&Bb[i], ncell));
            HIPTRY(dalloc(&gE[i], ncell)); HIPTRY(dalloc(&gB[i], ncell));
            if (c->curr_filter > 0) { HIPTRY(dalloc(&Jt[i], ncell)); HIPTRY(dalloc(&Ebf[i], ncell)); }
            Jadj[i] = c->curr_filter > 0 ? Ebf[i] : Eb[i];
        }
        bytes += ncell * (6 * (long long)sizeof(R) + 15 * 8 + (c->curr_filter > 0 ? 48 : 0));
        ckE.assign((size_t)nseg * 3, nullptr);
        ckB.assign((size_t)nseg * 3, nullptr);
        for (size_t i = 0; i < ckE.size(); i++) {
            HIPTRY(dalloc(&ckE[i], ncell)); HIPTRY(dalloc(&ckB[i], ncell));
        }
        bytes += (long long)nseg * 6 * ncell * sizeof(R);
        if (f32) {
            for (int i = 0; i < 3; i++) HIPTRY(dalloc(&Elo[i], ncell));
            ckElo.assign((size_t)nseg * 3, nullptr);
            for (size_t i = 0; i < ckElo.size(); i++) HIPTRY(dalloc(&ckElo[i], ncell));
            bytes += (long long)(nseg + 1) * 3 * ncell * 4;
        }
        if (!fast) {
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
        }
        HIPTRY(dalloc(&partial, 2048));
        HIPTRY(dalloc(&d_err, 1));
        for (auto &e : ev) HIPTRY(hipEventCreate(&e));
        return PICTOR_OK;
    }

    // tile-sorted engine: storage sized by the layout of the loaded particles
    int setup_bins()
    {
        const TileP &tp = c->tp;
        const size_t nb = (size_t)tp.ntiles * tp.NB;
        HIPTRY(dalloc(&d_base, nb));
        HIPTRY(dalloc(&d_cap, nb));
        HIPTRY(hipMemcpyAsync(d_base, h_base.data(), nb * sizeof(long long), hipMemcpyHostToDevice, c->stream));
        HIPTRY(hipMemcpyAsync(d_cap, h_cap.data(), nb * sizeof(int), hipMemcpyHostToDevice, c->stream));
        ck2.assign(nseg, PartSplit{});
        ck_cnt.assign(nseg, nullptr);
        for (int k = 0; k < nseg; k++) {
            HIPTRY(alloc_split_state(ck2[k], lay_total, true, false, bytes));
            HIPTRY(dalloc(&ck_cnt[k], nb));
        }
        // states 1..seg of a segment; the first pass ping-pongs between 1 and 2
        const int ns = std::max(seg, 2) + 1;
        st2.assign(ns, PartSplit{});
        cntv.assign(ns, nullptr);
        for (int j = 1; j < ns; j++) {
            HIPTRY(alloc_split_state(st2[j], lay_total, true, false, bytes));
            HIPTRY(dalloc(&cntv[j], nb));
        }
        dstv.assign(seg, nullptr);
        for (int j = 0; j < seg; j++) HIPTRY(dalloc(&dstv[j], lay_total));
        sfE.assign((size_t)seg * 3, nullptr);
        sfB.assign((size_t)seg * 3, nullptr);
        for (size_t i = 0; i < sfE.size(); i++) {
            HIPTRY(dalloc(&sfE[i], ncell));
            HIPTRY(dalloc(&sfB[i], ncell));
        }
        HIPTRY(alloc_part(fa[0], lay_total, bytes, false));
        HIPTRY(alloc_part(fa[1], lay_total, bytes, false));
        HIPTRY(dalloc(&pos, n));
        spill_cap = 4096;
        HIPTRY(alloc_split_state(spill, spill_cap, false, true, bytes));
        HIPTRY(dalloc(&d_spill, 1));
        bytes += (long long)(nseg + ns) * 4 * nb + 12LL * nb + (long long)seg * 4 * lay_total +
                 (long long)seg * 24 * ncell + 4 * n;
        return PICTOR_OK;
    }

    int load(const double *q, const double *x, const double *y, const double *z, const double *u, const double *v,
             const double *w, const int32_t *flv)
    {
        if (kin == hipMemcpyHostToDevice) {
            for (long long i = 0; i < n; i++) {
                if (flv[i] != 1 && flv[i] != 2) return fail(PICTOR_ERR_ARG, "adjoint: flavour must be 1 or 2");
                if (q[i] != c->ph.flvq[flv[i] - 1])
                    return fail(PICTOR_ERR_ARG, "adjoint: q must be the flavour charge (+1 / -1)");
            }
        } else {
            int err = 0;
            HIPTRY(hipMemsetAsync(d_err, 0, sizeof(int), c->stream));
            hipLaunchKernelGGL(k_adj_check, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, q, flv, n,
                               c->ph.flvq[0], c->ph.flvq[1], d_err);
            HIPTRY(hipGetLastError());
            HIPTRY(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPTRY(hipStreamSynchronize(c->stream));
            if (err & 1) return fail(PICTOR_ERR_ARG, "adjoint: flavour must be 1 or 2");
            if (err & 2) return fail(PICTOR_ERR_ARG, "adjoint: q must be the flavour charge (+1 / -1)");
        }
        if (fast) return load_bins(q, x, y, z, u, v, w, flv);
        double *tmp[6] = {};
        int *tf = nullptr;
        const double *src[6] = {x, y, z, u, v, w};
        for (int i = 0; i < 6; i++) {
            HIPTRY(dalloc(&tmp[i], n));
            HIPTRY(hipMemcpyAsync(tmp[i], src[i], n * sizeof(double), kin, c->stream));
        }
        HIPTRY(dalloc(&tf, n));
        HIPTRY(hipMemcpyAsync(tf, flv, n * sizeof(int), kin, c->stream));
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

    // the production upload path (tile of every particle, layout, permuted
    // scatter into the centre bins) into checkpoint 0, the caller's index as id
    int load_bins(const double *q, const double *x, const double *y, const double *z, const double *u,
                  const double *v, const double *w, const int32_t *flv)
    {
        hipStream_t s = c->stream;
        const TileP &tp = c->tp;
        const size_t ng = (size_t)tp.ntiles * tp.nsp;
        PartF64 raw{};
        double **ds[7] = {&raw.q, &raw.x, &raw.y, &raw.z, &raw.u, &raw.v, &raw.w};
        const double *hs[7] = {q, x, y, z, u, v, w};
        int *d_tile = nullptr, *d_count = nullptr, *d_fill = nullptr;
        long long *d_off = nullptr;
        auto cleanup = [&]() {
            (void)hipStreamSynchronize(s);
            for (auto p : ds) dfree(*p);
            dfree(raw.flv); dfree(d_tile); dfree(d_count); dfree(d_fill); dfree(d_off);
        };
        int rc = PICTOR_OK;
        do {
            bool ok = true;
            for (int i = 0; i < 7 && ok; i++) ok = dalloc(ds[i], n) == hipSuccess &&
                hipMemcpyAsync(*ds[i], hs[i], n * sizeof(double), kin, s) == hipSuccess;
            ok = ok && dalloc(&raw.flv, n) == hipSuccess &&
                 hipMemcpyAsync(raw.flv, flv, n * sizeof(int), kin, s) == hipSuccess;
            ok = ok && dalloc(&d_tile, n) == hipSuccess && dalloc(&d_count, ng) == hipSuccess &&
                 dalloc(&d_fill, ng) == hipSuccess && dalloc(&d_off, ng + 1) == hipSuccess;
            ok = ok && hipMemsetAsync(d_count, 0, ng * sizeof(int), s) == hipSuccess &&
                 hipMemsetAsync(d_err, 0, sizeof(int), s) == hipSuccess;
            if (!ok) { rc = fail(PICTOR_ERR_HIP, "adjoint: device allocation or copy failed"); break; }
            launch_raw_tiles_f64(c->g, c->ph, tp, raw, n, 1, d_tile, d_count, d_err, s);
            std::vector<int> cnt(ng);
            int err = 0;
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(cnt.data(), d_count, ng * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = fail(PICTOR_ERR_HIP, "adjoint: tile count failed");
                break;
            }
            if (err) { rc = fail(PICTOR_ERR_ARG, "adjoint: particle outside the physical cells"); break; }
            std::vector<long long> counts(ng);
            for (size_t i = 0; i < ng; i++) counts[i] = cnt[i];
            engine_layout(tp, c->ph.c, counts, cap_factor, h_base, h_cap, lay_total);
            if (lay_total >= 0x7fffffffLL) { rc = fail(PICTOR_ERR_CAPACITY, "adjoint: tile bins beyond 2^31 slots"); break; }
            if ((rc = setup_bins())) break;
            // centre-bin bases of every group, fill counters, centre counts of checkpoint 0
            const int centre = (tp.ND - 1) / 2;
            std::vector<long long> off(ng + 1);
            std::vector<int> c0((size_t)tp.ntiles * tp.NB, 0);
            for (size_t gi = 0; gi < ng; gi++) {
                const size_t b = (gi / tp.nsp) * tp.NB + (gi % tp.nsp) * tp.ND + centre;
                off[gi] = h_base[b];
                c0[b] = cnt[gi];
            }
            off[ng] = lay_total;
            if (hipMemcpyAsync(d_off, off.data(), (ng + 1) * sizeof(long long), hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemsetAsync(d_fill, 0, ng * sizeof(int), s) != hipSuccess ||
                hipMemcpyAsync(ck_cnt[0], c0.data(), c0.size() * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess) {
                rc = fail(PICTOR_ERR_HIP, "adjoint: layout upload failed");
                break;
            }
            launch_raw_scatter_split(c->g, c->ph, tp, raw, n, d_tile, d_off, d_fill, d_count, ck2[0], s);
            if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
                rc = fail(PICTOR_ERR_HIP, "adjoint: particle scatter failed");
        } while (false);
        cleanup();
        return rc;
    }

    // one step of the tile-sorted engine: state (in, cin) -> (out, cout); j >= 0
    // records step j of a segment (E, B as the push reads them, slot map)
    int fwd_bins(PartSplit in, const int *cin, PartSplit out, int *cout, int j)
    {
        hipStream_t s = c->stream;
        float *const *fE = reinterpret_cast<float *const *>(E);
        float *const *fB = reinterpret_cast<float *const *>(B);
        launch_update_b_half<float>(c->g, c->ph, fE, fB, s);
        launch_ghost_copy<float>(c->g, fB, s);
        if (j >= 0) {
            float *dst[6] = {sfE[3 * j], sfE[3 * j + 1], sfE[3 * j + 2], sfB[3 * j], sfB[3 * j + 1], sfB[3 * j + 2]};
            float *src[6] = {fE[0], fE[1], fE[2], fB[0], fB[1], fB[2]};
            launch_copy_n<float>(dst, src, 6, ncell, s);
        }
        launch_fill_zero_n<double>(J, 3, ncell, s);
        BinsP bins{d_base, d_cap, cin, cout};
        SpillP sp{d_spill, spill_cap, d_err};
        launch_push2d_split_track(c->tp.T, c->g, c->ph, c->tp, fE, fB, J, in, out, bins, spill, sp,
                                  j >= 0 ? dstv[j] : nullptr, s);
        launch_update_b_half<float>(c->g, c->ph, fE, fB, s);
        launch_ghost_copy<float>(c->g, fB, s);
        launch_current_fold<double>(c->g, J, s);
        if (c->curr_filter > 0) launch_smoothen_current(c->g, c->curr_filter, J, Jt, s);
        launch_update_e_pair(c->g, c->ph, fE, Elo, fB, J, ctx_jscale, true, true, s);
        launch_ghost_copy<float>(c->g, fE, s, Elo);
        HIPTRY(hipGetLastError());
        return PICTOR_OK;
    }

    // spill / error flags of the steps since the last reset (synchronises)
    int check_bins(bool forward_pass)
    {
        int h[2] = {0, 0};
        HIPTRY(hipMemcpyAsync(&h[0], d_spill, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(hipMemcpyAsync(&h[1], d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));
        if (h[1] & kErrBounds) return fail(PICTOR_ERR_BOUNDS, "adjoint: a particle moved more than one tile");
        if (h[1] & kErrPrecision)
            return fail(PICTOR_ERR_CAPACITY, "adjoint: current at a node exceeded the exact fixed-point range");
        if (h[0] > 0 || (h[1] & kErrCapacity)) {
            if (forward_pass) return kRetryLayout;
            return fail(PICTOR_ERR_CAPACITY, "adjoint: tile bins overflowed while recomputing a segment");
        }
        return PICTOR_OK;
    }

    int reset_flags()
    {
        HIPTRY(hipMemsetAsync(d_err, 0, sizeof(int), c->stream));
        if (d_spill) HIPTRY(hipMemsetAsync(d_spill, 0, sizeof(int), c->stream));
        return PICTOR_OK;
    }

    // one forward step of the unsorted engine: in -> out; j >= 0 records step j's node fields
    int fwd_step(EPart<R> in, EPart<R> out, int j)
    {
        hipStream_t s = c->stream;
        launch_update_b_half<R>(c->g, c->ph, E, B, s);
        launch_ghost_copy<R>(c->g, B, s);
        launch_fill_zero_n<double>(J, 3, ncell, s);
        void *nE = nodE[j < 0 ? 0 : j], *nB = nodB[j < 0 ? 0 : j];
        launch_eng_nodes<R>(ap, E, B, nE, nB, s);
        launch_eng_push<R>(ap, nE, nB, in, out, J, n, s);
        launch_update_b_half<R>(c->g, c->ph, E, B, s);
        launch_ghost_copy<R>(c->g, B, s);
        launch_current_fold<double>(c->g, J, s);
        if (c->curr_filter > 0) launch_smoothen_current(c->g, c->curr_filter, J, Jt, s);
        if constexpr (std::is_same<R, float>::value) {
            launch_update_e_pair(c->g, c->ph, E, Elo, B, J, ctx_jscale, true, true, s);
            launch_ghost_copy<float>(c->g, Elo, s);
        } else {
            launch_update_e_add_current<R, double>(c->g, c->ph, E, B, J, ctx_jscale, s);
        }
        launch_ghost_copy<R>(c->g, E, s);
        HIPTRY(hipGetLastError());
        return PICTOR_OK;
    }

    // reverse step j of a segment.  Field adjoints in the order of the forward
    // step reversed: the E update, the second B half step, the particle step
    // (from -jscale Eb as the current's adjoint), the first B half step.
    template <typename PartFn>
    int rev_step(PartFn particle_adjoint)
    {
        hipStream_t s = c->stream;
        const GridP &g = c->g;
        PhysP pa = c->ph, pb = c->ph;
        pa.fld_halfc = -c->ph.fldc;      // B-half kernel as the transpose of the E update's curl
        pb.fldc = -c->ph.fld_halfc;      // E kernel (curl only) as the transpose of a B half step
        launch_ghost_copy<double>(g, Eb, s);
        launch_update_b_half<double>(g, pa, Eb, Bb, s);
        double *gEB[6] = {gE[0], gE[1], gE[2], gB[0], gB[1], gB[2]};
        launch_fill_zero_n<double>(gEB, 6, ncell, s);
        if (c->curr_filter > 0) {
            // the adjoint current is F^T(-jscale Eb) = F(-jscale Eb): the periodic 1-2-1
            // filter is symmetric; its ghosts are the images (the fold's transpose)
            launch_copy_n<double>(Ebf, Eb, 3, ncell, s);
            launch_smoothen_current(g, c->curr_filter, Ebf, Jt, s);
            launch_ghost_copy<double>(g, Ebf, s);
        }
        particle_adjoint();
        launch_ghost_copy<double>(g, Bb, s);
        launch_update_e<double>(g, pb, Eb, Bb, s);
        launch_current_fold<double>(g, gE, s, gB);
        launch_ghost_copy<double>(g, gE, s, gB);
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
            if (Elo[i]) HIPTRY(hipMemcpyAsync(ckElo[3 * k + i], Elo[i], ncell * 4, hipMemcpyDeviceToDevice, c->stream));
        }
        return PICTOR_OK;
    }

    int restore_ck(int k)
    {
        for (int i = 0; i < 3; i++) {
            HIPTRY(hipMemcpyAsync(E[i], ckE[3 * k + i], ncell * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
            HIPTRY(hipMemcpyAsync(B[i], ckB[3 * k + i], ncell * sizeof(R), hipMemcpyDeviceToDevice, c->stream));
            if (Elo[i]) HIPTRY(hipMemcpyAsync(Elo[i], ckElo[3 * k + i], ncell * 4, hipMemcpyDeviceToDevice, c->stream));
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
            if (Elo[i]) HIPTRY(hipMemcpyAsync(Elo[i], c->Elo[i], ncell * 4, hipMemcpyDeviceToDevice, s));
        }
        launch_ghost_copy<R>(c->g, E, s);
        if (Elo[0]) launch_ghost_copy<float>(c->g, Elo, s);
        launch_ghost_copy<R>(c->g, B, s);
        int rc;
        if ((rc = reset_flags())) return rc;
        HIPTRY(hipEventRecord(ev[0], s));
        if (fast) {
            // the step that ends a segment writes the next checkpoint directly
            PartSplit a = ck2[0];
            int *ac = ck_cnt[0];
            for (int t = 0; t < nsteps; t++) {
                if (t % seg == 0 && (rc = save_ck(t / seg))) return rc;
                const int k1 = (t + 1) / seg;
                const bool to_ck = (t + 1) % seg == 0 && k1 < nseg;
                PartSplit b = to_ck ? ck2[k1] : st2[1 + (t % 2)];
                int *bc = to_ck ? ck_cnt[k1] : cntv[1 + (t % 2)];
                if ((rc = fwd_bins(a, ac, b, bc, -1))) return rc;
                a = b;
                ac = bc;
            }
        } else {
            // ping-pong between st[1] and st[2] (st[1] alone when a segment is one step)
            EPart<R> a = ck[0];
            for (int t = 0; t < nsteps; t++) {
                if (t % seg == 0) {
                    const int k = t / seg;
                    if (k > 0) HIPTRY(copy_part(ck[k], a, n, s));
                    if ((rc = save_ck(k))) return rc;
                    if (k > 0) a = ck[k];
                }
                EPart<R> b = (seg >= 2) ? st[1 + (t % 2)] : st[1];
                if ((rc = fwd_step(a, b, -1))) return rc;
                a = b;
            }
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
        if (fast) {
            if ((rc = check_bins(true))) return rc;
        } else {
            int err = 0;
            HIPTRY(hipMemcpy(&err, d_err, sizeof(int), hipMemcpyDeviceToHost));
            if (err & 1) return fail(PICTOR_ERR_BOUNDS, "adjoint engine: a particle outside its tile");
        }
        float ms = 0;
        HIPTRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        t_fwd = ms;
        have_forward = true;
        return PICTOR_OK;
    }

    int backward(double scale, double *du, double *dv, double *dw)
    {
        if (!have_forward) return fail(PICTOR_ERR_ARG, "pictor_adjoint_backward without pictor_adjoint_forward");
        hipStream_t s = c->stream;
        const TileP &tp = c->tp;
        int rc;
        // seeds: dL/dB = B at the final step, dL/dE = 0, particles 0
        for (int i = 0; i < 3; i++) {
            launch_fill_zero<double>(Eb[i], ncell, s);
            launch_seed_b<R>(B[i], Bb[i], ncell, s);
        }
        if (!fast) launch_eng_zero<R>(adj, n, s);
        if ((rc = reset_flags())) return rc;
        int cur = 0;   // fast engine: fa[cur] holds the adjoint of the latest state reached
        t_rec = t_adj = 0;
        for (int k = nseg - 1; k >= 0; k--) {
            const int t0 = k * seg, t1 = std::min(nsteps, t0 + seg), L = t1 - t0;
            HIPTRY(hipEventRecord(ev[0], s));
            if ((rc = restore_ck(k))) return rc;
            if (fast) {
                auto S = [&](int j) { return j == 0 ? ck2[k] : st2[j]; };
                auto C = [&](int j) { return j == 0 ? ck_cnt[k] : cntv[j]; };
                for (int j = 0; j < L; j++)
                    if ((rc = fwd_bins(S(j), C(j), st2[j + 1], cntv[j + 1], j))) return rc;
                HIPTRY(hipEventRecord(ev[1], s));
                if ((rc = check_bins(false))) return rc;
                // adjoint of state L of this segment: zero after the last step, else the
                // adjoint of the next checkpoint moved to the recomputed slot order
                if (k == nseg - 1) {
                    for (float *p : {fa[cur].ox, fa[cur].oy, fa[cur].oz, fa[cur].u, fa[cur].v, fa[cur].w})
                        launch_fill_zero<float>(p, lay_total, s);
                } else {
                    launch_fa_slots(tp, 0, d_base, ck_cnt[k + 1], ck2[k + 1].id, pos, fa[cur], fa[cur], 1.0, nullptr,
                                    nullptr, nullptr, s);
                    launch_fa_slots(tp, 1, d_base, C(L), S(L).id, pos, fa[cur], fa[1 - cur], 1.0, nullptr, nullptr,
                                    nullptr, s);
                    cur = 1 - cur;
                }
                for (int j = L - 1; j >= 0; j--) {
                    const int src = cur;
                    float *fE[3] = {sfE[3 * j], sfE[3 * j + 1], sfE[3 * j + 2]};
                    float *fB[3] = {sfB[3 * j], sfB[3 * j + 1], sfB[3 * j + 2]};
                    rc = rev_step([&]() {
                        launch_fa_adjoint(c->g, tp, ap, fE, fB, Jadj, d_base, C(j), S(j), dstv[j], fa[src], fa[1 - src],
                                          gE, gB, s);
                    });
                    if (rc) return rc;
                    cur = 1 - cur;
                }
            } else {
                // state j of the segment: the checkpoint for j = 0, st[j] after that
                auto S = [&](int j) { return j == 0 ? ck[k] : st[j]; };
                for (int j = 0; j < L; j++)
                    if ((rc = fwd_step(S(j), st[j + 1], j))) return rc;
                HIPTRY(hipEventRecord(ev[1], s));
                for (int j = L - 1; j >= 0; j--) {
                    rc = rev_step([&]() { launch_eng_adjoint<R>(ap, nodE[j], nodB[j], S(j), Jadj, adj, gE, gB, n, s); });
                    if (rc) return rc;
                }
            }
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
            if (fast) {
                for (int i = 0; i < 3; i++) launch_fill_zero<double>(d[i], n, s);
                launch_fa_slots(tp, 2, d_base, ck_cnt[0], ck2[0].id, pos, fa[cur], fa[cur], scale, d[0], d[1], d[2], s);
            } else {
                launch_eng_store_grad<R>(adj, scale, d[0], d[1], d[2], n, s);
            }
            HIPTRY(hipGetLastError());
            double *h[3] = {du, dv, dw};
            for (int i = 0; i < 3; i++)
                if (h[i]) HIPTRY(hipMemcpyAsync(h[i], d[i], n * sizeof(double), kout, s));
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
                int32_t segment, int engine, bool device, double *objective)
{
    AdjEngine<R> *e = engine_of<R>(c);
    if (!e) {
        delete c->adj;
        c->adj = nullptr;
        e = new AdjEngine<R>();
        c->adj = e;
    }
    e->kin = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    // tile bins that overflow during the forward pass: the pass is repeated
    // with the centre bins twice as large (at most three times)
    double cf = 1.5;
    int rc = PICTOR_OK, repeats = 0;
    for (;;) {
        rc = e->setup(c, n, nsteps, segment, engine == 0, cf);
        if (!rc) rc = e->load(q, x, y, z, u, v, w, flv);
        if (!rc) rc = e->forward();
        if (rc != kRetryLayout) break;
        if (++repeats > 3) {
            rc = fail(PICTOR_ERR_CAPACITY, "adjoint: tile bins overflowed at 12x the initial tile population");
            break;
        }
        cf *= 2.0;
    }
    if (rc) {
        e->release();
        return rc;
    }
    e->relayouts = repeats;
    if (objective) *objective = e->objective;
    return PICTOR_OK;
}

// the argument checks and engine dispatch of the two forward entry points
int adjoint_forward_entry(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                          const double *z, const double *u, const double *v, const double *w, const int32_t *flv,
                          int32_t nsteps, int32_t segment, bool device, double *objective)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (c->g.ndim != 2 || c->cyl) return fail(PICTOR_ERR_ARG, "the adjoint is available for the 2-D Cartesian geometry");
    if (c->ph.fdtd != PICTOR_FDTD_YEE) return fail(PICTOR_ERR_ARG, "the adjoint is available for the Yee solver");
    if (c->bc_any) return fail(PICTOR_ERR_ARG, "the adjoint is available for the periodic box (no boundary faces)");
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "adjoint inside a split step");
    if (n < 1 || nsteps < 1 || segment < 1) return fail(PICTOR_ERR_ARG, "n, nsteps and segment must be >= 1");
    if (n >= 0x7fffffffLL) return fail(PICTOR_ERR_ARG, "adjoint: more than 2^31-1 particles");
    if (!q || !x || !y || !z || !u || !v || !w || !flv) return fail(PICTOR_ERR_ARG, "null particle array");
    if (segment > nsteps) segment = nsteps;
    const int engine = c->adj_engine;
    if (c->form == PICTOR_FORM_FP64)
        return adj_forward<double>(c, n, q, x, y, z, u, v, w, flv, nsteps, segment, engine, device, objective);
    if (engine == 0 && c->tp.T != 8 && c->tp.T != 16 && c->tp.T != 32)
        return fail(PICTOR_ERR_ARG, "adjoint: the tile-sorted engine needs a tile edge of 8, 16 or 32");
    return adj_forward<float>(c, n, q, x, y, z, u, v, w, flv, nsteps, segment, engine, device, objective);
}

template <typename R>
int adjoint_backward_entry(AdjEngine<R> *e, double scale, double *du, double *dv, double *dw, bool device)
{
    e->kout = device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    return e->backward(scale, du, dv, dw);
}

}  // namespace

extern "C" {

int pictor_adjoint_forward(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                           const double *z, const double *u, const double *v, const double *w, const int32_t *flv,
                           int32_t nsteps, int32_t segment, double *objective)
{
    return adjoint_forward_entry(c, n, q, x, y, z, u, v, w, flv, nsteps, segment, false, objective);
}

int pictor_adjoint_forward_device(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                                  const double *z, const double *u, const double *v, const double *w,
                                  const int32_t *flv, int32_t nsteps, int32_t segment, double *objective)
{
    return adjoint_forward_entry(c, n, q, x, y, z, u, v, w, flv, nsteps, segment, true, objective);
}

int pictor_adjoint_backward(pictor_ctx *c, double scale, double *du, double *dv, double *dw)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (auto *e = engine_of<double>(c)) return adjoint_backward_entry(e, scale, du, dv, dw, false);
    if (auto *e = engine_of<float>(c)) return adjoint_backward_entry(e, scale, du, dv, dw, false);
    return fail(PICTOR_ERR_ARG, "pictor_adjoint_backward without pictor_adjoint_forward");
}

int pictor_adjoint_backward_device(pictor_ctx *c, double scale, double *du, double *dv, double *dw)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (auto *e = engine_of<double>(c)) return adjoint_backward_entry(e, scale, du, dv, dw, true);
    if (auto *e = engine_of<float>(c)) return adjoint_backward_entry(e, scale, du, dv, dw, true);
    return fail(PICTOR_ERR_ARG, "pictor_adjoint_backward_device without pictor_adjoint_forward");
}

int pictor_adjoint_stats(pictor_ctx *c, double out[10])
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (!out) return fail(PICTOR_ERR_ARG, "null argument");
    for (int i = 0; i < 10; i++) out[i] = 0;
    auto fill = [&](auto *e) {
        out[0] = e->t_fwd;
        out[1] = e->t_rec;
        out[2] = e->t_adj;
        out[3] = (double)e->bytes;
        out[4] = e->nsteps;
        out[5] = e->seg;
        out[6] = e->nseg;
        out[7] = e->objective;
        out[8] = e->fast ? 0.0 : 1.0;
        out[9] = e->relayouts;
    };
    if (auto *e = engine_of<double>(c)) fill(e);
    else if (auto *e = engine_of<float>(c)) fill(e);
    else return fail(PICTOR_ERR_ARG, "no adjoint run on this context");
    return PICTOR_OK;
}

int pictor_adjoint_engine(pictor_ctx *c, int32_t engine)
{
    int rc = check_ctx2(c);
    if (rc) return rc;
    if (engine != 0 && engine != 1) return fail(PICTOR_ERR_ARG, "engine must be 0 (tile-sorted) or 1 (unsorted)");
    c->adj_engine = engine;
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
