// api.hip -- context, storage layout and the extern "C" entry points.
//
// Host-side control of the particle step.  The reference's MAIN loop calls
// one subroutine per phase (src/MAIN.F90:42-284); each pictor_* phase call
// below enqueues the kernels of that phase on the context's stream.  The
// fused pictor_step runs a whole time step with the same kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pictor_hip.h"
#include "ctx.h"
#include "pictor_common.h"

using namespace pictor;

namespace pictor {
thread_local std::string g_err;
}

namespace {

template <typename R>
void field_ptrs(pictor_ctx *c, R *E[3], R *B[3])
{
    for (int i = 0; i < 3; i++) { E[i] = (R *)c->E[i]; B[i] = (R *)c->B[i]; }
}

void free_f64(PartF64 &p)
{
    dfree(p.q); dfree(p.x); dfree(p.y); dfree(p.z);
    dfree(p.u); dfree(p.v); dfree(p.w); dfree(p.flv); dfree(p.tag);
    p = PartF64{};
}

void free_split(PartSplit &p)
{
    dfree(p.cell); dfree(p.ox); dfree(p.oy); dfree(p.oz);
    dfree(p.u); dfree(p.v); dfree(p.w); dfree(p.sp); dfree(p.id); dfree(p.wq);
    p = PartSplit{};
}

hipError_t alloc_f64(PartF64 &p, long long n, bool tagged = false)
{
    hipError_t e;
    if (tagged && (e = dalloc(&p.tag, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.q, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.x, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.y, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.z, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.u, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.v, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.w, n)) != hipSuccess) return e;
    return dalloc(&p.flv, n);
}

// oz only in 3-D: the 2-D split record has no z (kZMid)
hipError_t alloc_split(PartSplit &p, long long n, bool with_sp, int ndim, bool tagged = false, bool weighted = false)
{
    hipError_t e;
    if (tagged && (e = dalloc(&p.id, n)) != hipSuccess) return e;
    if (weighted && (e = dalloc(&p.wq, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.cell, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.ox, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.oy, n)) != hipSuccess) return e;
    if (ndim == 3 && (e = dalloc(&p.oz, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.u, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.v, n)) != hipSuccess) return e;
    if ((e = dalloc(&p.w, n)) != hipSuccess) return e;
    if (with_sp) return dalloc(&p.sp, n);
    return hipSuccess;
}

int ensure_buffer(pictor_ctx *c, int which, long long need)
{
    if (c->buf_cap[which] >= need) return PICTOR_OK;
    // the split-form kernels address a buffer with 32-bit byte offsets of 4-byte elements
    if (c->form != PICTOR_FORM_FP64 && need >= (1LL << 30))
        return fail(PICTOR_ERR_CAPACITY, "FP32 split form: more than 2^30 particle slots per buffer");
    long long n = (long long)(need * 1.1) + 1024;
    if (c->form != PICTOR_FORM_FP64) n = std::min(n, (1LL << 30) - 1);
    if (c->form == PICTOR_FORM_FP64) {
        free_f64(c->pf[which]);
        HIPTRY(alloc_f64(c->pf[which], n, c->tagged));
    } else {
        free_split(c->ps[which]);
        HIPTRY(alloc_split(c->ps[which], n, false, c->g.ndim, c->tagged, c->weighted));
    }
    c->buf_cap[which] = n;
    return PICTOR_OK;
}

int ensure_spill(pictor_ctx *c, int need)
{
    if (c->spill_cap >= need) return PICTOR_OK;
    if (c->form == PICTOR_FORM_FP64) {
        free_f64(c->spill_f);
        c->spill_cap = 0;
        HIPTRY(alloc_f64(c->spill_f, need, c->tagged));
    } else {
        free_split(c->spill_s);
        c->spill_cap = 0;
        HIPTRY(alloc_split(c->spill_s, need, true, c->g.ndim, c->tagged, c->weighted));
    }
    dfree(c->rl_grp);
    c->rl_grp = nullptr;
    c->rl_grp_cap = 0;
    HIPTRY(dalloc(&c->rl_grp, need));
    c->rl_grp_cap = need;
    c->spill_cap = need;
    return PICTOR_OK;
}

// Bin capacities from the particle count of every (tile, species) group.
// A tile can lose or gain at most the particles within c cells of a face per
// step; the side bins are sized for that flux with a margin, the centre bin
// for the whole group times capacity_factor.  A re-layout also passes the bin
// counts of the last push (`seen`, with the capacities they had): a bin that
// was full gets twice the room (the centre bin: twice its group's count, a
// side bin: twice its old capacity), and a side bin at least 1.5 times the
// arrivals it saw.  In the nonlinear Weibel phase the filaments keep
// gathering particles into the same tiles, whose centre bins would otherwise
// overflow again a few dozen steps after every re-layout.
void compute_layout_into(const pictor_ctx *c, const std::vector<long long> &counts, std::vector<long long> &h_base,
                         std::vector<int> &h_cap, long long &total, const std::vector<int> *seen = nullptr,
                         const std::vector<int> *seen_cap = nullptr)
{
    const TileP &tp = c->tp;
    // 1.75: the filaments of the nonlinear Weibel phase grow some tiles by half their
    // population within a few hundred steps (C2: 63 re-layouts over 2000 steps at 1.25,
    // 16 at 1.75, for 8 GB more of the 288 GB)
    const double cf = c->cfg.capacity_factor > 0 ? c->cfg.capacity_factor : 1.75;
    const double frac = std::min(1.0, c->ph.c / tp.T);
    h_base.assign((size_t)tp.ntiles * tp.NB, 0);
    h_cap.assign((size_t)tp.ntiles * tp.NB, 0);
    long long off = 0;
    for (int t = 0; t < tp.ntiles; t++)
        for (int s = 0; s < tp.nsp; s++) {
            const long long n = counts[(size_t)t * tp.nsp + s];
            for (int d = 0; d < tp.ND; d++) {
                int nz = 0, r = d;
                for (int a = 0; a < c->g.ndim; a++) { nz += (r % 3 != 1); r /= 3; }
                double cap;
                const size_t b = (size_t)t * tp.NB + s * tp.ND + d;
                if (nz == 0) cap = std::ceil(n * cf) + 64;
                else cap = std::ceil(n * std::pow(frac, nz) * 1.6 * std::pow(2.0, nz - 1)) + (nz == 1 ? 32 : 16);
                if (seen && seen_cap && b < seen->size() && b < seen_cap->size()) {
                    const int was = (*seen)[b], had = (*seen_cap)[b];
                    if (nz == 0) {
                        if (was >= had) cap = std::max(cap, 2.0 * n + 64);
                    } else {
                        cap = std::max(cap, std::ceil(1.5 * was) + 16);
                        if (was >= had) cap = std::max(cap, 2.0 * had);
                    }
                }
                long long ci = ((long long)cap + 15) / 16 * 16;
                h_base[b] = off;
                h_cap[b] = (int)std::min<long long>(ci, 0x7fffffff);
                off += ci;
            }
        }
    total = off;
}

void compute_layout(pictor_ctx *c, const std::vector<long long> &counts)
{
    compute_layout_into(c, counts, c->h_base, c->h_cap, c->layout_cap);
}

int upload_layout(pictor_ctx *c)
{
    const size_t nb = (size_t)c->tp.ntiles * c->tp.NB;
    HIPTRY(hipMemcpyAsync(c->d_base, c->h_base.data(), nb * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemcpyAsync(c->d_cap, c->h_cap.data(), nb * sizeof(int), hipMemcpyHostToDevice, c->stream));
    return PICTOR_OK;
}

// centre-bin counts from per-group totals
int upload_center_counts(pictor_ctx *c, int which, const std::vector<long long> &counts)
{
    const TileP &tp = c->tp;
    std::vector<int> cnt((size_t)tp.ntiles * tp.NB, 0);
    const int centre = (tp.ND - 1) / 2;
    for (int t = 0; t < tp.ntiles; t++)
        for (int s = 0; s < tp.nsp; s++)
            cnt[(size_t)t * tp.NB + s * tp.ND + centre] = (int)counts[(size_t)t * tp.nsp + s];
    HIPTRY(hipMemcpyAsync(c->d_cnt[which], cnt.data(), cnt.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));   // cnt is a local host buffer
    return PICTOR_OK;
}

std::vector<long long> group_offsets(const pictor_ctx *c, bool dense, const std::vector<long long> &counts,
                                     const std::vector<long long> *base = nullptr, long long total = -1)
{
    // dense: consecutive runs (download); otherwise the centre-bin bases of the layout
    // (the context's, or `base` with `total` slots)
    const TileP &tp = c->tp;
    const int centre = (tp.ND - 1) / 2;
    const size_t ng = (size_t)tp.ntiles * tp.nsp;
    const std::vector<long long> &hb = base ? *base : c->h_base;
    std::vector<long long> off(ng + 1, 0);
    long long acc = 0;
    for (size_t gi = 0; gi < ng; gi++) {
        if (dense) { off[gi] = acc; acc += counts[gi]; }
        else off[gi] = hb[(gi / tp.nsp) * tp.NB + (gi % tp.nsp) * tp.ND + centre];
    }
    off[ng] = dense ? acc : (base ? total : c->layout_cap);
    return off;
}

// per-group particle counts of buffer `which` (host copy of the bin counts)
int group_counts(pictor_ctx *c, int which, std::vector<long long> &counts, std::vector<int> *bins = nullptr)
{
    const TileP &tp = c->tp;
    std::vector<int> local;
    std::vector<int> &cnt = bins ? *bins : local;
    cnt.resize((size_t)tp.ntiles * tp.NB);
    HIPTRY(hipMemcpyAsync(cnt.data(), c->d_cnt[which], cnt.size() * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    counts.assign((size_t)tp.ntiles * tp.nsp, 0);
    for (int t = 0; t < tp.ntiles; t++)
        for (int s = 0; s < tp.nsp; s++) {
            long long a = 0;
            for (int d = 0; d < tp.ND; d++) a += cnt[(size_t)t * tp.NB + s * tp.ND + d];
            counts[(size_t)t * tp.nsp + s] = a;
        }
    return PICTOR_OK;
}

// Re-layout: buffer `cur` bins + `nspill` spilled particles -> new layout in the other
// buffer.  The scratch lives in the context (rl_*); the host copy of the new layout
// replaces the old one only once every allocation has succeeded, so a failure leaves
// the context on its old layout.
int relayout(pictor_ctx *c, int nspill)
{
    const TileP &tp = c->tp;
    const size_t ng = (size_t)tp.ntiles * tp.nsp;
    if (nspill > c->rl_grp_cap) return fail(PICTOR_ERR_CAPACITY, "re-layout: more spilled particles than the spill buffer");
    std::vector<long long> counts;
    std::vector<int> seen;
    int rc = group_counts(c, c->cur, counts, &seen);
    if (rc) return rc;
    HIPTRY(hipMemsetAsync(c->rl_gcount, 0, ng * sizeof(int), c->stream));
    if (nspill > 0) {
        if (c->form == PICTOR_FORM_FP64)
            launch_spill_tiles_f64(c->g, tp, c->spill_f, nspill, c->rl_grp, c->rl_gcount, c->stream);
        else
            launch_spill_tiles_split(c->g, tp, c->spill_s, nspill, c->rl_grp, c->rl_gcount, c->stream);
    }
    std::vector<int> sc(ng);
    HIPTRY(hipMemcpyAsync(sc.data(), c->rl_gcount, ng * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    std::vector<long long> total(ng);
    long long np = 0;
    for (size_t gi = 0; gi < ng; gi++) { total[gi] = counts[gi] + sc[gi]; np += total[gi]; }
    std::vector<long long> nbase;
    std::vector<int> ncap;
    long long ntotal = 0;
    compute_layout_into(c, total, nbase, ncap, ntotal, &seen, &c->h_cap);
    const int dst = 1 - c->cur;
    if ((rc = ensure_buffer(c, dst, ntotal))) return rc;
    std::vector<long long> off = group_offsets(c, false, total, &nbase, ntotal);
    // fill starts after the compacted bins of each group
    std::vector<int> fill(ng), gtot(ng);
    for (size_t gi = 0; gi < ng; gi++) { fill[gi] = (int)counts[gi]; gtot[gi] = (int)total[gi]; }
    HIPTRY(hipMemcpyAsync(c->rl_off, off.data(), (ng + 1) * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemcpyAsync(c->rl_fill, fill.data(), ng * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemcpyAsync(c->rl_gtotal, gtot.data(), ng * sizeof(int), hipMemcpyHostToDevice, c->stream));
    // the compaction reads the old layout on the device (d_base, d_cnt).  It copies each
    // group's bins in order (coalesced): unlike a fresh upload in the loader's cell order,
    // the particles of a bin are already mixed by the pushes that filled it, so the slot
    // permutation of the upload is not needed here (and its scattered writes made a
    // re-layout of C2 take 60 ms)
    if (c->form == PICTOR_FORM_FP64) {
        launch_compact_f64(tp, c->pf[c->cur], c->d_base, c->d_cnt[c->cur], c->rl_off, nullptr, c->pf[dst], c->stream);
        launch_spill_scatter_f64(tp, c->spill_f, nspill, c->rl_grp, c->rl_off, c->rl_fill, nullptr, c->pf[dst],
                                 c->stream);
    } else {
        launch_compact_split(tp, c->ps[c->cur], c->d_base, c->d_cnt[c->cur], c->rl_off, nullptr, c->ps[dst],
                             c->stream);
        launch_spill_scatter_split(tp, c->spill_s, nspill, c->rl_grp, c->rl_off, c->rl_fill, nullptr, c->ps[dst],
                                   c->stream);
    }
    HIPTRY(hipGetLastError());
    // the host vectors of the copies must live until they complete
    HIPTRY(hipStreamSynchronize(c->stream));
    // commit the new layout
    c->h_base.swap(nbase);
    c->h_cap.swap(ncap);
    c->layout_cap = ntotal;
    if ((rc = upload_layout(c))) return rc;
    if ((rc = upload_center_counts(c, dst, total))) return rc;
    HIPTRY(hipMemsetAsync(c->d_spill_count, 0, sizeof(int), c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    c->cur = dst;
    c->np = np;
    // the other buffer must hold the new layout too
    if ((rc = ensure_buffer(c, 1 - dst, c->layout_cap))) return rc;
    return ensure_spill(c, (int)std::min<long long>(std::max<long long>(65536, np / 64), 1 << 30));
}

// The spill count and error bits of the last push, copied to the pinned host flags
// behind the step's kernels; `ev` marks their arrival.
int enqueue_flags(pictor_ctx *c, hipEvent_t ev)
{
    HIPTRY(hipMemcpyAsync(c->h_flags, c->d_spill_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(c->h_flags + 1, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipEventRecord(ev, c->stream));
    return PICTOR_OK;
}

// After the flags of a push have arrived: swap the buffers, report errors, and re-lay
// out storage if particles spilled.
int finish_push(pictor_ctx *c, hipEvent_t ev)
{
    HIPTRY(hipEventSynchronize(ev));
    c->cur = 1 - c->cur;
    c->pending_swap = false;
    const int nspill = c->h_flags[0], err = c->h_flags[1];
    if (err & kErrBounds) return fail(PICTOR_ERR_BOUNDS, "a particle was outside its tile window or moved more than one tile");
    if (err & kErrPrecision)
        return fail(PICTOR_ERR_CAPACITY, "current at a node exceeded the exact fixed-point range (|J| >= 4096 q cell)");
    if ((err & kErrCapacity) || nspill > c->spill_cap)
        return fail(PICTOR_ERR_CAPACITY, "particle spill buffer exhausted");
    if (nspill > 0) {
        c->n_relayout++;
        c->n_spilled += nspill;
        return relayout(c, nspill);
    }
    return PICTOR_OK;
}

// spill / error check after a push (pictor_exchange_particles): wait for the flags
int commit_push(pictor_ctx *c)
{
    int rc = enqueue_flags(c, c->ev[22]);
    return rc ? rc : finish_push(c, c->ev[22]);
}

int check_ctx(pictor_ctx *c)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    HIPTRY(hipSetDevice(c->dev));
    return PICTOR_OK;
}

// halo = false leaves the ghost copy to the caller (the PML runs in between)
int do_b_half(pictor_ctx *c, bool halo = true)
{
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3];
        field_ptrs(c, E, B);
        if (c->ph.fdtd == PICTOR_FDTD_LEHE) launch_update_b_half_lehe<double>(c->g, c->ph, c->ph.lehe, E, B, c->stream);
        else launch_update_b_half<double>(c->g, c->ph, E, B, c->stream);
        if (halo) launch_ghost_copy<double>(c->g, B, c->stream);
    } else {
        float *E[3], *B[3];
        field_ptrs(c, E, B);
        if (c->ph.fdtd == PICTOR_FDTD_LEHE) launch_update_b_half_lehe<float>(c->g, c->ph, c->ph.lehe, E, B, c->stream);
        else launch_update_b_half<float>(c->g, c->ph, E, B, c->stream);
        if (halo) launch_ghost_copy<float>(c->g, B, c->stream);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

// timed: the push kernel is bracketed by events e0, e1 (ev[0], ev[1] unless given)
int do_push(pictor_ctx *c, bool timed, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    if (!e0) e0 = c->ev[0];
    if (!e1) e1 = c->ev[1];
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "pictor_move_deposit called twice without pictor_exchange_particles");
    HIPTRY(hipMemsetAsync(c->d_spill_count, 0, sizeof(int), c->stream));
    const int a = c->cur, b = 1 - c->cur;
    BinsP bins{c->d_base, c->d_cap, c->d_cnt[a], c->d_cnt[b]};
    SpillP sp{c->d_spill_count, c->spill_cap, c->d_err};
    if (timed) HIPTRY(hipEventRecord(e0, c->stream));
    const bool three = (c->g.ndim == 3);
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3], *J[3];
        field_ptrs(c, E, B);
        for (int i = 0; i < 3; i++) J[i] = (double *)c->J[i];
        if (three)
            launch_push3d_f64(c->tp.T, c->g, c->ph, c->tp, E, B, J, c->pf[a], c->pf[b], bins, c->spill_f, sp, c->stream);
        else
            launch_push2d_f64(c->tp.T, c->g, c->ph, c->tp, E, B, J, c->pf[a], c->pf[b], bins, c->spill_f, sp, c->stream);
    } else {
        float *E[3], *B[3];
        double *J[3];
        field_ptrs(c, E, B);
        for (int i = 0; i < 3; i++) J[i] = (double *)c->J[i];
        if (three)
            launch_push3d_split(c->tp.T, c->g, c->ph, c->tp, E, B, J, c->ps[a], c->ps[b], bins, c->spill_s, sp, c->stream);
        else
            launch_push2d_split(c->tp.T, c->g, c->ph, c->tp, E, B, J, c->ps[a], c->ps[b], bins, c->spill_s, sp, c->stream);
    }
    HIPTRY(hipGetLastError());
    if (timed) HIPTRY(hipEventRecord(e1, c->stream));
    c->pending_swap = true;
    return PICTOR_OK;
}

int do_reset_current(pictor_ctx *c)
{
    double *J[3] = {(double *)c->J[0], (double *)c->J[1], (double *)c->J[2]};
    launch_fill_zero_n<double>(J, 3, c->g.ncell, c->stream);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int do_fold(pictor_ctx *c)
{
    double *J[3];
    for (int i = 0; i < 3; i++) J[i] = (double *)c->J[i];
    launch_current_fold<double>(c->g, J, c->stream);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int do_smoothen(pictor_ctx *c)
{
    if (c->curr_filter <= 0) return PICTOR_OK;
    double *J[3] = {(double *)c->J[0], (double *)c->J[1], (double *)c->J[2]};
    launch_smoothen_current(c->g, c->curr_filter, J, c->Jt, c->stream);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

double jscale(const pictor_ctx *c) { return c->ph.qi * kFixInv; }

// SendRecvFlds on E (which = 0) or B (which = 1), Cartesian geometry
int do_halo(pictor_ctx *c, int which)
{
    void **F = which == 0 ? c->E : c->B;
    if (c->form == PICTOR_FORM_FP64) {
        double *P[3] = {(double *)F[0], (double *)F[1], (double *)F[2]};
        launch_ghost_copy<double>(c->g, P, c->stream);
    } else {
        float *P[3] = {(float *)F[0], (float *)F[1], (float *)F[2]};
        launch_ghost_copy<float>(c->g, P, c->stream, which == 0 ? c->Elo : nullptr);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int do_e_phase(pictor_ctx *c, bool curl, bool add, bool halo = true)
{
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3], *J[3];
        field_ptrs(c, E, B);
        for (int i = 0; i < 3; i++) J[i] = (double *)c->J[i];
        if (curl && add) launch_update_e_add_current<double, double>(c->g, c->ph, E, B, J, 1.0, c->stream);
        else if (curl) launch_update_e<double>(c->g, c->ph, E, B, c->stream);
        else launch_add_current<double, double>(c->g, c->ph, E, J, 1.0, c->stream);
        if (halo) launch_ghost_copy<double>(c->g, E, c->stream);
    } else {
        float *E[3], *B[3];
        double *J[3];
        field_ptrs(c, E, B);
        for (int i = 0; i < 3; i++) J[i] = (double *)c->J[i];
        launch_update_e_pair(c->g, c->ph, E, c->Elo, B, J, jscale(c), curl, add, c->stream);
        if (halo) launch_ghost_copy<float>(c->g, E, c->stream, c->Elo);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int choose_tile(const pictor_config *cfg)
{
    // 2-D: tile 16 (512-lane workgroups, 35 KB of LDS: four per CU) is the fastest push
    // in both forms once the grid has 1024 tiles of it; smaller grids take tile 8 for
    // more workgroups.  3-D windows are (T+3)^3 / (T+5)^3 nodes: 8 is the largest tile
    // whose LDS fits
    const int cands2[2] = {16, 8}, cands3[2] = {8, 4};
    const int *cands = cfg->ndim == 3 ? cands3 : cands2;
    const int ncand = 2;
    auto fits = [&](int T) {
        if (cfg->nx % T || cfg->ny % T) return false;
        if (cfg->ndim == 3 && cfg->nz % T) return false;
        return true;
    };
    int best = 0;
    for (int q = 0; q < ncand; q++) {
        const int T = cands[q];
        if (!fits(T)) continue;
        long long nt = (long long)(cfg->nx / T) * (cfg->ny / T) * (cfg->ndim == 3 ? cfg->nz / T : 1);
        if (nt >= 1024) return T;
        best = T;   // smallest fitting tile so far
    }
    return best;
}

}  // namespace

extern "C" {

const char *pictor_last_error(void) { return g_err.c_str(); }

int pictor_init(const pictor_config *cfg, pictor_ctx **out)
{
    if (!cfg || !out) return fail(PICTOR_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->ndim != 2 && cfg->ndim != 3) return fail(PICTOR_ERR_ARG, "ndim must be 2 or 3");
    const bool ny_modes = cfg->geometry == PICTOR_GEOM_CYL_MODAL;   // ny counts modes there, not cells
    if (cfg->nx < 8 || (!ny_modes && cfg->ny < 8) || (cfg->ndim == 3 && cfg->nz < 8))
        return fail(PICTOR_ERR_ARG, "nx, ny (, nz) must be >= 8");
    // the packed cell word of the split form bounds the Cartesian grids; the cylindrical
    // path keeps FP64 coordinates
    const bool cyl_geom = cfg->geometry == PICTOR_GEOM_CYL || cfg->geometry == PICTOR_GEOM_CYL_MODAL;
    if (!cyl_geom && cfg->ndim == 2 && (cfg->nx > 32768 || cfg->ny > 32768))
        return fail(PICTOR_ERR_ARG, "2-D nx, ny must be <= 32768");
    if (!cyl_geom && cfg->ndim == 3 && (cfg->nx > 2048 || cfg->ny > 2048 || cfg->nz > 512))
        return fail(PICTOR_ERR_ARG, "3-D nx, ny must be <= 2048 and nz <= 512");
    if (cyl_geom && ((long long)cfg->nx + 5) * (cfg->ny + 5) * (cfg->ndim == 3 ? cfg->nz + 5 : 1) > (1LL << 31))
        return fail(PICTOR_ERR_ARG, "cylindrical grid larger than 2^31 cells");
    if (cfg->form != PICTOR_FORM_FP64 && cfg->form != PICTOR_FORM_FP32_SPLIT) return fail(PICTOR_ERR_ARG, "unknown form");
    if (cfg->fdtd != PICTOR_FDTD_YEE && cfg->fdtd != PICTOR_FDTD_LEHE) return fail(PICTOR_ERR_ARG, "unknown fdtd");
    if (cfg->geometry != PICTOR_GEOM_CARTESIAN && cfg->geometry != PICTOR_GEOM_CYL &&
        cfg->geometry != PICTOR_GEOM_CYL_MODAL)
        return fail(PICTOR_ERR_ARG, "unknown geometry");
    const bool modal = cfg->geometry == PICTOR_GEOM_CYL_MODAL;
    if (modal) {
        if (cfg->ndim != 3 || cfg->ny != 2) return fail(PICTOR_ERR_ARG, "the modal cylinder needs ndim 3 and ny = 2 (modes m = 0, 1)");
        if (cfg->curr_filter) return fail(PICTOR_ERR_ARG, "the modal cylinder runs without the current filter");
        // the split record packs i in 12 bits and k in 19 bits of its cell word
        if (cfg->form == PICTOR_FORM_FP32_SPLIT && (cfg->nx > 4090 || cfg->nz > 500000))
            return fail(PICTOR_ERR_ARG, "modal cylinder, split form: nx <= 4090 and nz <= 500000");
        if (cfg->dr < 0.0) return fail(PICTOR_ERR_ARG, "dr must be >= 0");
        // the m = 1 terms near the axis act like a theta cell of dr / 2 at r = dr / 2
        const double dr = cfg->dr > 0.0 ? cfg->dr : 1.0;
        const double lim = 1.0 / std::sqrt(1.0 + 1.0 / (dr * dr) + 4.0 / (dr * dr));
        if (cfg->c >= lim)
            return fail(PICTOR_ERR_ARG, "modal cylinder: c must be below 1/sqrt(1 + 5/dr^2) = " + std::to_string(lim));
        // the Lehe update raises the z Nyquist difference by (1 - 4 d_z) = sin^2(pi c/2) / c^2;
        // with the axis terms it stays bounded for sin^2(pi c/2) + 1.25 c^2/dr^2 < 1
        const double sn = std::sin(0.5 * 3.14159265358979323846 * cfg->c);
        if (cfg->fdtd == PICTOR_FDTD_LEHE && sn * sn + 1.25 * cfg->c * cfg->c / (dr * dr) >= 1.0)
            return fail(PICTOR_ERR_ARG, "modal cylinder with Lehe: c must satisfy sin^2(pi c/2) + 1.25 c^2/dr^2 < 1");
    }
    const bool cyl = cfg->geometry == PICTOR_GEOM_CYL || modal;
    if (cfg->curr_filter < 0 || cfg->curr_filter > 64) return fail(PICTOR_ERR_ARG, "curr_filter must be 0..64");
    if (cyl && cfg->curr_filter) return fail(PICTOR_ERR_ARG, "the current filter is for the Cartesian geometry");
    if (cyl && !modal && cfg->form != PICTOR_FORM_FP64)
        return fail(PICTOR_ERR_ARG, "the cylindrical geometry runs in the FP64 form");
    if (cyl && !modal && cfg->ny % 2) return fail(PICTOR_ERR_ARG, "the cylindrical grid needs an even number of theta cells (ny)");
    if (cyl && !modal) {
        // the narrowest cell of the r-theta grid is the arc 0.5 * dtheta at r = 1/2 (grid index 4):
        // the FDTD step is stable for c < 1 / sqrt(1 + 1 / (0.5 dtheta)^2 (+ 1 in 3-D))
        const double arc = 0.5 * (2.0 * 3.14159265358979323846264 / cfg->ny);
        const double lim = 1.0 / std::sqrt(1.0 + 1.0 / (arc * arc) + (cfg->ndim == 3 ? 1.0 : 0.0));
        if (cfg->c >= lim)
            return fail(PICTOR_ERR_ARG, "cylindrical grid: c must be below the Courant limit of the cells next to the "
                                        "axis, 1/sqrt(1 + (ny/pi)^2" + std::string(cfg->ndim == 3 ? " + 1" : "") +
                                        ") = " + std::to_string(lim));
    }
    if (!(cfg->c > 0.0 && cfg->c < 1.0)) return fail(PICTOR_ERR_ARG, "Courant number c must lie in (0,1)");
    if (cfg->epc < 1 || cfg->c_ompe <= 0 || cfg->mi_me <= 0) return fail(PICTOR_ERR_ARG, "invalid epc, c_ompe or mi_me");
    // the cylindrical path keeps its particles unsorted: one tile over the grid
    int T = cyl ? 1 : (cfg->tile ? cfg->tile : choose_tile(cfg));
    if (!cyl) {
        if (cfg->ndim == 2 && T != 8 && T != 16 && T != 32) return fail(PICTOR_ERR_ARG, "2-D tile must be 8, 16 or 32 and divide nx, ny");
        if (cfg->ndim == 3 && T != 4 && T != 8) return fail(PICTOR_ERR_ARG, "3-D tile must be 4 or 8 and divide nx, ny, nz");
        if (cfg->nx % T || cfg->ny % T || (cfg->ndim == 3 && cfg->nz % T)) return fail(PICTOR_ERR_ARG, "tile must divide the grid");
    }

    pictor_ctx *c = new pictor_ctx();
    c->cfg = *cfg;
    c->cfg.tile = T;
    c->form = cfg->form;
    if (hipGetDevice(&c->dev) != hipSuccess) { delete c; return fail(PICTOR_ERR_HIP, "no HIP device"); }

    GridP &g = c->g;
    g.ndim = cfg->ndim;
    g.nx = cfg->nx; g.ny = cfg->ny; g.nz = (cfg->ndim == 2) ? 1 : cfg->nz;
    g.mx = g.nx + 5; g.my = g.ny + 5; g.mz = (cfg->ndim == 2) ? 1 : g.nz + 5;
    if (modal) { g.ny = 3; g.my = 3; }   // the middle index holds the three mode slots
    g.ncell = (long long)g.mx * g.my * g.mz;
    g.per[0] = g.per[1] = g.per[2] = 1;   // periodic until a boundary face is set

    // InitScale, initialise.F90:44-53; vars.F90:220-223; subdomains.F90:52-67
    PhysP &ph = c->ph;
    const double me = 1.0, mi = cfg->mi_me;
    const double ompe = cfg->c / cfg->c_ompe;
    ph.c = cfg->c;
    ph.sqc = cfg->c * cfg->c;
    ph.cinv = 1.0 / cfg->c;
    ph.qi = (cfg->g0 * (ompe * ompe)) / (((double)cfg->epc / 1.0) * (1.0 + me / mi));
    ph.fldc = cfg->c * 1.0;
    ph.fld_halfc = cfg->c * 0.5 * 1.0;
    ph.inv_dx = ph.inv_dy = ph.inv_dz = 1.0;
    ph.xmin = 3.0; ph.xmax = g.mx - 2;
    ph.ymin = 3.0; ph.ymax = g.my - 2;
    if (g.ndim == 2) { ph.zmin = 1.0; ph.zmax = 2.0; }
    else { ph.zmin = 3.0; ph.zmax = g.mz - 2; }
    ph.xlen = ph.xmax - ph.xmin; ph.ylen = ph.ymax - ph.ymin; ph.zlen = ph.zmax - ph.zmin;
    ph.bext[0] = ph.bext[1] = ph.bext[2] = 0.0;
    ph.nflvr = 2;
    ph.flvrqm[0] = 1.0 / mi;   // qmi
    ph.flvrqm[1] = -1.0;       // qme
    ph.flvq[0] = 1.0;          // FlvrCharge, initialise.F90:86-87 (unit weight)
    ph.flvq[1] = -1.0;
    ph.mass[0] = ph.qi * mi;                 // massi
    ph.mass[1] = std::fabs(-ph.qi) * me;     // masse
    // em_update_lehe.F90:8-25 (unit cells): beta_xy = beta_xz = beta_yx = beta_zx = 1/8,
    // beta_yz = beta_zy = 0, delta_x = (1 - sin^2(pi c / 2) / c^2) / 4, delta_y = delta_z = 0
    ph.fdtd = cfg->fdtd;
    {
        const double pi = 3.14159265358979323846264;   // vars.F90:224
        const double bxy = 0.125 * 1.0, bxz = 0.125 * 1.0, byx = 0.125, byz = 0.0, bzx = 0.125, bzy = 0.0;
        const double sn = std::sin(0.5 * pi * cfg->c);
        const double dx = 0.25 * (1.0 - (ph.cinv * ph.cinv) * (sn * sn)), dy = 0.0, dz = 0.0;
        double *L = ph.lehe;
        if (cfg->ndim == 3) {
            L[0] = 1.0 - 2.0 * bxy - 2.0 * bxz - 3.0 * dx;
            L[1] = 1.0 - 2.0 * byx - 2.0 * byz - 3.0 * dy;
            L[2] = 1.0 - 2.0 * bzx - 2.0 * bzy - 3.0 * dz;
        } else {
            L[0] = 1.0 - 2.0 * bxy - 3.0 * dx;
            L[1] = 1.0 - 2.0 * byx - 3.0 * dy;
            L[2] = 1.0;
        }
        L[3] = bxy; L[4] = bxz; L[5] = byx; L[6] = byz; L[7] = bzx; L[8] = bzy;
        L[9] = dx; L[10] = dy; L[11] = dz;
    }

    c->cyl = cyl;
    c->cylm = modal;
    if (modal) cylm_params(*cfg, ph, c->cq);
    TileP &tp = c->tp;
    tp.T = cyl ? g.nx : T;
    tp.ntx = cyl ? 1 : g.nx / T; tp.nty = cyl ? 1 : g.ny / T; tp.ntz = (g.ndim == 3 && !cyl) ? g.nz / T : 1;
    tp.ntiles = tp.ntx * tp.nty * tp.ntz;
    tp.ND = (g.ndim == 2) ? 9 : 27;
    tp.nsp = (c->form == PICTOR_FORM_FP64) ? 1 : ph.nflvr;
    tp.NB = tp.nsp * tp.ND;
    tp.xcd_remap = (tp.ntiles % 8 == 0) ? 1 : 0;

    auto cleanup_fail = [&](int rc) { pictor_finalize(c); return rc; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup_fail(fail(PICTOR_ERR_HIP, "hipStreamCreate failed"));
    // FP64 fields: the FP64 form and the modal cylinder (whose split form keeps FP64 fields)
    const bool f64 = c->form == PICTOR_FORM_FP64 || c->cylm;
    const size_t rs = f64 ? sizeof(double) : sizeof(float);
    for (int i = 0; i < 3; i++) {
        if (hipMalloc(&c->E[i], g.ncell * rs) != hipSuccess || hipMalloc(&c->B[i], g.ncell * rs) != hipSuccess ||
            hipMalloc(&c->J[i], g.ncell * 8) != hipSuccess)
            return cleanup_fail(fail(PICTOR_ERR_HIP, "field allocation failed"));
        if (hipMemsetAsync(c->E[i], 0, g.ncell * rs, c->stream) != hipSuccess ||
            hipMemsetAsync(c->B[i], 0, g.ncell * rs, c->stream) != hipSuccess ||
            hipMemsetAsync(c->J[i], 0, g.ncell * 8, c->stream) != hipSuccess)
            return cleanup_fail(fail(PICTOR_ERR_HIP, "field clearing failed"));
        if (!f64) {
            if (hipMalloc(&c->Elo[i], g.ncell * sizeof(float)) != hipSuccess)
                return cleanup_fail(fail(PICTOR_ERR_HIP, "field allocation failed"));
            if (hipMemsetAsync(c->Elo[i], 0, g.ncell * sizeof(float), c->stream) != hipSuccess)
                return cleanup_fail(fail(PICTOR_ERR_HIP, "field clearing failed"));
        }
        if (cfg->curr_filter > 0 && hipMalloc(&c->Jt[i], g.ncell * 8) != hipSuccess)
            return cleanup_fail(fail(PICTOR_ERR_HIP, "field allocation failed"));
    }
    c->curr_filter = cfg->curr_filter;
    const size_t nb = (size_t)tp.ntiles * tp.NB;
    const size_t ngr = (size_t)tp.ntiles * tp.nsp;
    if (dalloc(&c->d_base, nb) || dalloc(&c->d_cap, nb) || dalloc(&c->d_cnt[0], nb) || dalloc(&c->d_cnt[1], nb) ||
        dalloc(&c->d_spill_count, 1) || dalloc(&c->d_err, 1) || dalloc(&c->rl_gcount, ngr) ||
        dalloc(&c->rl_gtotal, ngr) || dalloc(&c->rl_fill, ngr) || dalloc(&c->rl_off, ngr + 1))
        return cleanup_fail(fail(PICTOR_ERR_HIP, "bin allocation failed"));
    if (hipMemsetAsync(c->d_cnt[0], 0, nb * sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_cnt[1], 0, nb * sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_spill_count, 0, sizeof(int), c->stream) != hipSuccess ||
        hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream) != hipSuccess)
        return cleanup_fail(fail(PICTOR_ERR_HIP, "bin clearing failed"));
    if (hipHostMalloc((void **)&c->h_flags, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess)
        return cleanup_fail(fail(PICTOR_ERR_HIP, "pinned allocation failed"));
    for (int i = 0; i < 23; i++)
        if (hipEventCreate(&c->ev[i]) != hipSuccess) return cleanup_fail(fail(PICTOR_ERR_HIP, "event creation failed"));
    // empty layout so that a step without particles is valid
    int rc = PICTOR_OK;
    if (modal) {
        rc = cylm_setup(c);
    } else if (cyl) {
        rc = cyl_setup(c);
    } else {
        std::vector<long long> zero((size_t)tp.ntiles * tp.nsp, 0);
        compute_layout(c, zero);
        rc = upload_layout(c);
        if (!rc) rc = ensure_buffer(c, 0, c->layout_cap);
        if (!rc) rc = ensure_buffer(c, 1, c->layout_cap);
        if (!rc) rc = ensure_spill(c, 65536);
    }
    c->partial_len = std::max(1024, tp.ntiles) * 5;
    if (!rc && (dalloc(&c->d_partial, c->partial_len) != hipSuccess || dalloc(&c->d_rho, g.ncell) != hipSuccess))
        rc = fail(PICTOR_ERR_HIP, "diagnostic allocation failed");
    if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(PICTOR_ERR_HIP, "init synchronize failed");
    if (rc) return cleanup_fail(rc);
    *out = c;
    return PICTOR_OK;
}

int pictor_finalize(pictor_ctx *c)
{
    if (!c) return PICTOR_OK;
    hipSetDevice(c->dev);
    if (c->stream) hipStreamSynchronize(c->stream);
    delete c->adj;
    c->adj = nullptr;
    for (int i = 0; i < 3; i++) { dfree(c->E[i]); dfree(c->B[i]); dfree(c->J[i]); dfree(c->Elo[i]); dfree(c->Jt[i]); }
    dfree(c->d_base); dfree(c->d_cap); dfree(c->d_cnt[0]); dfree(c->d_cnt[1]);
    dfree(c->d_spill_count); dfree(c->d_err);
    dfree(c->rl_gcount); dfree(c->rl_gtotal); dfree(c->rl_fill); dfree(c->rl_off); dfree(c->rl_grp);
    for (int i = 0; i < 2; i++) { free_f64(c->pf[i]); free_split(c->ps[i]); }
    free_f64(c->spill_f); free_split(c->spill_s);
    if (c->cylm) cylm_release(c);
    cyl_release(c);
    dfree(c->d_partial); dfree(c->d_rho);
    bc_release(c);
    if (c->h_flags) hipHostFree(c->h_flags);
    for (int i = 0; i < 23; i++) if (c->ev[i]) hipEventDestroy(c->ev[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return PICTOR_OK;
}

int pictor_get_scalars(pictor_ctx *c, double out[16])
{
    if (!c || !out) return fail(PICTOR_ERR_ARG, "null argument");
    out[0] = c->ph.qi; out[1] = -c->ph.qi; out[2] = c->ph.mass[1]; out[3] = c->ph.mass[0];
    out[4] = c->cfg.c / c->cfg.c_ompe; out[5] = c->ph.fldc;
    out[6] = c->ph.xlen; out[7] = c->ph.ylen; out[8] = c->ph.zlen;
    out[9] = c->g.mx; out[10] = c->g.my; out[11] = c->g.mz;
    out[12] = c->tp.T; out[13] = c->tp.ntiles;
    out[14] = (double)c->n_relayout; out[15] = (double)c->n_spilled;
    return PICTOR_OK;
}

int pictor_upload_fields(pictor_ctx *c, const double *ex, const double *ey, const double *ez, const double *bx,
                         const double *by, const double *bz)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    const double *src[6] = {ex, ey, ez, bx, by, bz};
    void *dst[6] = {c->E[0], c->E[1], c->E[2], c->B[0], c->B[1], c->B[2]};
    const size_t n = (size_t)c->g.ncell;
    const bool f64 = c->form == PICTOR_FORM_FP64 || c->cylm;
    double *tmp = nullptr;
    if (!f64) HIPTRY(dalloc(&tmp, n));
    for (int i = 0; i < 6; i++) {
        if (!src[i]) return fail(PICTOR_ERR_ARG, "null field pointer");
        if (f64) {
            HIPTRY(hipMemcpyAsync(dst[i], src[i], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        } else {
            HIPTRY(hipMemcpyAsync(tmp, src[i], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
            if (i < 3) launch_pair_from_f64(tmp, (float *)dst[i], c->Elo[i], (long long)n, c->stream);
            else launch_convert_fields_f32(tmp, (float *)dst[i], (long long)n, c->stream);
        }
    }
    HIPTRY(hipStreamSynchronize(c->stream));
    if (tmp) dfree(tmp);
    return PICTOR_OK;
}

int pictor_download_fields(pictor_ctx *c, double *ex, double *ey, double *ez, double *bx, double *by, double *bz,
                           double *jx, double *jy, double *jz)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    double *dst[9] = {ex, ey, ez, bx, by, bz, jx, jy, jz};
    void *src[9] = {c->E[0], c->E[1], c->E[2], c->B[0], c->B[1], c->B[2], c->J[0], c->J[1], c->J[2]};
    const size_t n = (size_t)c->g.ncell;
    const bool f64 = c->form == PICTOR_FORM_FP64 || c->cylm;
    double *tmp = nullptr;
    if (!f64) HIPTRY(dalloc(&tmp, n));
    for (int i = 0; i < 9; i++) {
        if (!dst[i]) continue;
        if (f64) {
            HIPTRY(hipMemcpyAsync(dst[i], src[i], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        } else {
            if (i < 3) launch_pair_to_f64((const float *)src[i], c->Elo[i], tmp, (long long)n, c->stream);
            else if (i < 6) launch_convert_fields_f64((const float *)src[i], tmp, (long long)n, c->stream);
            else launch_convert_jfix((const double *)src[i], tmp, (long long)n, jscale(c), c->stream);
            HIPTRY(hipMemcpyAsync(dst[i], tmp, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            HIPTRY(hipStreamSynchronize(c->stream));
        }
    }
    HIPTRY(hipStreamSynchronize(c->stream));
    if (tmp) dfree(tmp);
    return PICTOR_OK;
}

// SetFilteredEfldGPU (fields_gpu.F90:70-90): E (the FP64 value, hi + lo in the split
// form) through `passes` rounds of the 1-2-1 filter of smoothen_current (the
// reference's nMoverEMfilter passes of MovingAverageFilterGPU over x, y (, z)), then
// the periodic halo copy (the edge exchanges); the result goes to the host arrays
int pictor_filtered_efield(pictor_ctx *c, int32_t passes, double *fx, double *fy, double *fz)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->cyl) return fail(PICTOR_ERR_ARG, "filtered E: Cartesian geometry only");
    if (c->bc_any) return fail(PICTOR_ERR_ARG, "filtered E: the filter reads periodic images (no boundary faces)");
    if (passes < 0) return fail(PICTOR_ERR_ARG, "filtered E: negative number of passes");
    if (!fx || !fy || !fz) return fail(PICTOR_ERR_ARG, "null output array");
    const size_t n = (size_t)c->g.ncell;
    double *F[3] = {}, *T[3] = {};
    auto release = [&]() { for (int i = 0; i < 3; i++) { dfree(F[i]); dfree(T[i]); } };
    for (int i = 0; i < 3; i++)
        if (dalloc(&F[i], n) != hipSuccess || dalloc(&T[i], n) != hipSuccess) {
            release();
            return fail(PICTOR_ERR_HIP, "filtered E: device allocation failed");
        }
    for (int i = 0; i < 3; i++) {
        if (c->form == PICTOR_FORM_FP64)
            HIPTRY(hipMemcpyAsync(F[i], c->E[i], n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        else
            launch_pair_to_f64((const float *)c->E[i], c->Elo[i], F[i], (long long)n, c->stream);
    }
    launch_smoothen_current(c->g, passes, F, T, c->stream);
    launch_ghost_copy<double>(c->g, F, c->stream);
    double *dst[3] = {fx, fy, fz};
    for (int i = 0; i < 3; i++)
        HIPTRY(hipMemcpyAsync(dst[i], F[i], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    const hipError_t e = hipStreamSynchronize(c->stream);
    release();
    if (e != hipSuccess) return fail(PICTOR_ERR_HIP, std::string("filtered E: ") + hipGetErrorString(e));
    return PICTOR_OK;
}

// the upload of both entry points; tag (may be null): the caller's per-particle tags
static int upload_particles_impl(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                                 const double *z, const double *u, const double *v, const double *w,
                                 const int32_t *flv, const int32_t *tag)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n < 0) return fail(PICTOR_ERR_ARG, "negative particle count");
    if (n > 0 && (!q || !x || !y || !z || !u || !v || !w || !flv)) return fail(PICTOR_ERR_ARG, "null particle pointer");
    if (n >= 0x7fffffffLL) return fail(PICTOR_ERR_ARG, "more than 2^31-1 particles per device");
    c->pending_swap = false;
    if (c->cyl) {
        if (tag) return fail(PICTOR_ERR_ARG, "tagged particles: Cartesian geometry only");
        return c->cylm ? cylm_upload_particles(c, n, q, x, y, z, u, v, w, flv) : cyl_upload_particles(c, n, q, x, y, z, u, v, w, flv);
    }
    const bool tagged = tag != nullptr;
    // split form: a charge weight per particle when some particle's charge is not its
    // flavour's (the FP64 record keeps q itself)
    bool weighted = false;
    if (c->form == PICTOR_FORM_FP32_SPLIT)
        for (int64_t i = 0; i < n && !weighted; i++)
            weighted = flv[i] >= 1 && flv[i] <= c->ph.nflvr && q[i] != c->ph.flvq[flv[i] - 1];
    if (tagged != c->tagged || weighted != c->weighted) {
        // the particle buffers change their record: the next allocation brings (or drops)
        // the tags and weights
        for (int i = 0; i < 2; i++) { free_f64(c->pf[i]); free_split(c->ps[i]); c->buf_cap[i] = 0; }
        free_f64(c->spill_f); free_split(c->spill_s);
        c->spill_cap = 0;
        c->tagged = tagged;
        c->weighted = weighted;
    }
    const TileP &tp = c->tp;
    const size_t ng = (size_t)tp.ntiles * tp.nsp;
    PartF64 raw{};
    int *d_tile = nullptr, *d_count = nullptr, *d_fill = nullptr;
    long long *d_off = nullptr;
    HIPTRY(alloc_f64(raw, n, tagged));
    if (tagged && n > 0) HIPTRY(hipMemcpyAsync(raw.tag, tag, n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIPTRY(dalloc(&d_tile, n));
    HIPTRY(dalloc(&d_count, ng));
    HIPTRY(dalloc(&d_fill, ng));
    HIPTRY(dalloc(&d_off, ng + 1));
    const double *hs[7] = {q, x, y, z, u, v, w};
    double *ds[7] = {raw.q, raw.x, raw.y, raw.z, raw.u, raw.v, raw.w};
    for (int i = 0; i < 7 && n > 0; i++)
        HIPTRY(hipMemcpyAsync(ds[i], hs[i], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (n > 0) HIPTRY(hipMemcpyAsync(raw.flv, flv, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemsetAsync(d_count, 0, ng * sizeof(int), c->stream));
    HIPTRY(hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream));
    launch_raw_tiles_f64(c->g, c->ph, tp, raw, n, c->form == PICTOR_FORM_FP32_SPLIT ? (weighted ? 2 : 1) : 0, d_tile,
                         d_count, c->d_err, c->stream);
    HIPTRY(hipGetLastError());
    std::vector<int> cnt(ng);
    HIPTRY(hipMemcpyAsync(cnt.data(), d_count, ng * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipMemcpyAsync(c->h_flags + 1, c->d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    if (c->h_flags[1]) {
        hipMemsetAsync(c->d_err, 0, sizeof(int), c->stream);
        free_f64(raw); dfree(d_tile); dfree(d_count); dfree(d_fill); dfree(d_off);
        return fail(PICTOR_ERR_ARG, "particle outside the domain (x,y in [2, m-1)), invalid flavour, or (split form) a charge of the wrong sign for its flavour");
    }
    std::vector<long long> counts(ng);
    long long np = 0;
    for (size_t i = 0; i < ng; i++) { counts[i] = cnt[i]; np += cnt[i]; }
    compute_layout(c, counts);
    rc = ensure_buffer(c, 0, c->layout_cap);
    if (!rc) rc = ensure_buffer(c, 1, c->layout_cap);
    if (!rc) rc = ensure_spill(c, (int)std::min<long long>(std::max<long long>(65536, np / 64), 1 << 30));
    if (rc) return rc;
    std::vector<long long> off = group_offsets(c, false, counts);
    HIPTRY(hipMemcpyAsync(d_off, off.data(), (ng + 1) * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    HIPTRY(hipMemsetAsync(d_fill, 0, ng * sizeof(int), c->stream));
    c->cur = 0;
    if (c->form == PICTOR_FORM_FP64)
        launch_raw_scatter_f64(c->g, tp, raw, n, d_tile, d_off, d_fill, d_count, c->pf[0], c->stream);
    else
        launch_raw_scatter_split(c->g, c->ph, tp, raw, n, d_tile, d_off, d_fill, d_count, c->ps[0], c->stream);
    HIPTRY(hipGetLastError());
    rc = upload_layout(c);
    if (!rc) rc = upload_center_counts(c, 0, counts);
    if (rc) return rc;
    HIPTRY(hipStreamSynchronize(c->stream));
    free_f64(raw);
    dfree(d_tile); dfree(d_count); dfree(d_fill); dfree(d_off);
    c->np = np;
    return PICTOR_OK;
}

int pictor_upload_particles(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                            const double *z, const double *u, const double *v, const double *w, const int32_t *flv)
{
    return upload_particles_impl(c, n, q, x, y, z, u, v, w, flv, nullptr);
}

int pictor_upload_particles_tagged(pictor_ctx *c, int64_t n, const double *q, const double *x, const double *y,
                                   const double *z, const double *u, const double *v, const double *w,
                                   const int32_t *flv, const int32_t *tag)
{
    if (n > 0 && !tag) return fail(PICTOR_ERR_ARG, "null tag pointer");
    static const int32_t none = 0;
    return upload_particles_impl(c, n, q, x, y, z, u, v, w, flv, tag ? tag : &none);
}

int pictor_particle_count(pictor_ctx *c, int64_t *n)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->cyl) { *n = c->np; return PICTOR_OK; }
    std::vector<long long> counts;
    rc = group_counts(c, c->cur, counts);
    if (rc) return rc;
    long long a = 0;
    for (long long v : counts) a += v;
    *n = a;
    return PICTOR_OK;
}

static int download_particles_impl(pictor_ctx *c, int64_t cap, double *q, double *x, double *y, double *z,
                                   double *u, double *v, double *w, int32_t *flv, int32_t *tile, int32_t *tag,
                                   int64_t *n_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "download between move_deposit and exchange_particles");
    if (tag && !c->tagged) return fail(PICTOR_ERR_ARG, "the particles carry no tags (upload them with pictor_upload_particles_tagged)");
    if (c->cyl) {
        long long no = 0;
        rc = c->cylm ? cylm_download_particles(c, cap, q, x, y, z, u, v, w, flv, tile, &no)
                     : cyl_download_particles(c, cap, q, x, y, z, u, v, w, flv, tile, &no);
        if (n_out) *n_out = no;
        return rc;
    }
    const TileP &tp = c->tp;
    const size_t ng = (size_t)tp.ntiles * tp.nsp;
    std::vector<long long> counts;
    rc = group_counts(c, c->cur, counts);
    if (rc) return rc;
    std::vector<long long> off = group_offsets(c, true, counts);
    const long long n = off[ng];
    if (n_out) *n_out = n;
    if (n > cap) return fail(PICTOR_ERR_ARG, "output arrays too small");
    if (n == 0) return PICTOR_OK;
    long long *d_off = nullptr;
    int *d_tile = nullptr;
    PartF64 st{};
    PartSplit ss{};
    HIPTRY(dalloc(&d_off, ng + 1));
    HIPTRY(alloc_f64(st, n, c->tagged));
    HIPTRY(hipMemcpyAsync(d_off, off.data(), (ng + 1) * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    if (c->form == PICTOR_FORM_FP64) {
        launch_compact_f64(tp, c->pf[c->cur], c->d_base, c->d_cnt[c->cur], d_off, nullptr, st, c->stream);
    } else {
        HIPTRY(alloc_split(ss, n, true, c->g.ndim, c->tagged, c->weighted));
        launch_compact_split(tp, c->ps[c->cur], c->d_base, c->d_cnt[c->cur], d_off, nullptr, ss, c->stream);
        launch_split_to_f64(c->g, c->ph, ss, n, st, c->stream);
    }
    if (tile) {
        HIPTRY(dalloc(&d_tile, n));
        launch_tile_ids(tp, d_off, n, d_tile, c->stream);
    }
    HIPTRY(hipGetLastError());
    double *hd[7] = {q, x, y, z, u, v, w};
    double *dd[7] = {st.q, st.x, st.y, st.z, st.u, st.v, st.w};
    for (int i = 0; i < 7; i++)
        if (hd[i]) HIPTRY(hipMemcpyAsync(hd[i], dd[i], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (flv) HIPTRY(hipMemcpyAsync(flv, st.flv, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (tile) HIPTRY(hipMemcpyAsync(tile, d_tile, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (tag) HIPTRY(hipMemcpyAsync(tag, st.tag, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    dfree(d_off); dfree(d_tile);
    free_f64(st); free_split(ss);
    return PICTOR_OK;
}

int pictor_download_particles(pictor_ctx *c, int64_t cap, double *q, double *x, double *y, double *z, double *u,
                              double *v, double *w, int32_t *flv, int32_t *tile, int64_t *n_out)
{
    return download_particles_impl(c, cap, q, x, y, z, u, v, w, flv, tile, nullptr, n_out);
}

int pictor_download_particles_tagged(pictor_ctx *c, int64_t cap, double *q, double *x, double *y, double *z,
                                     double *u, double *v, double *w, int32_t *flv, int32_t *tile, int32_t *tag,
                                     int64_t *n_out)
{
    return download_particles_impl(c, cap, q, x, y, z, u, v, w, flv, tile, tag, n_out);
}

int pictor_update_b_half(pictor_ctx *c)
{
    int rc = check_ctx(c);
    return rc ? rc : (c->cyl ? (c->cylm ? cylm_b_half(c) : cyl_b_half(c)) : do_b_half(c));
}

int pictor_move_deposit(pictor_ctx *c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->cyl) {
        if (c->pending_swap) return fail(PICTOR_ERR_ARG, "pictor_move_deposit called twice without pictor_exchange_particles");
        return c->cylm ? cylm_push(c, false) : cyl_push(c, false);
    }
    // MoveDepositPrtl, then EnforceBC_PostMovDep (MAIN.F90:65-77): the particle faces act
    // inside the push, the current of the reflecting faces is folded back here
    rc = do_push(c, false);
    return rc ? rc : bc_after_push(c);
}

int pictor_exchange_particles(pictor_ctx *c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->pending_swap) return fail(PICTOR_ERR_ARG, "pictor_exchange_particles without a preceding pictor_move_deposit");
    return c->cyl ? (c->cylm ? cylm_exchange(c) : cyl_exchange(c)) : commit_push(c);
}

int pictor_update_e(pictor_ctx *c)
{
    int rc = check_ctx(c);
    return rc ? rc : (c->cyl ? (c->cylm ? cylm_e_phase(c, true, false) : cyl_e_phase(c, true, false)) : do_e_phase(c, true, false));
}

int pictor_current_halo_add(pictor_ctx *c)
{
    int rc = check_ctx(c);
    return rc ? rc : (c->cyl ? (c->cylm ? cylm_current_halo(c) : cyl_current_halo(c)) : do_fold(c));
}

int pictor_smoothen_current(pictor_ctx *c)
{
    int rc = check_ctx(c);
    return rc ? rc : do_smoothen(c);
}

int pictor_add_current(pictor_ctx *c)
{
    int rc = check_ctx(c);
    return rc ? rc : (c->cyl ? (c->cylm ? cylm_e_phase(c, false, true) : cyl_e_phase(c, false, true)) : do_e_phase(c, false, true));
}

int pictor_field_halo(pictor_ctx *c, int32_t which)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (which != 0 && which != 1) return fail(PICTOR_ERR_ARG, "which must be 0 (E) or 1 (B)");
    if (c->cyl) return c->cylm ? cylm_field_halo(c, which) : cyl_field_halo(c, which);
    return do_halo(c, which);
}

int pictor_reset_current(pictor_ctx *c)
{
    int rc = check_ctx(c);
    return rc ? rc : (c->cyl ? (c->cylm ? cylm_reset_current(c) : cyl_reset_current(c)) : do_reset_current(c));
}

int pictor_sort_tiles(pictor_ctx *c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "sort between move_deposit and exchange_particles");
    if (c->cylm) return cylm_sort(c);   // the modal particles by cell (holes dropped)
    if (c->cyl) return PICTOR_OK;   // the cylindrical path keeps the caller's order
    return relayout(c, 0);
}

int pictor_step(pictor_ctx *c, int32_t nsteps)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "pictor_step inside a split step");
    if (c->cyl) {
        // MAIN.F90:42-284 with the cylindrical routines
        for (int s = 0; s < nsteps; s++) {
            HIPTRY(hipEventRecord(c->ev[4], c->stream));
            if (c->cylm) {
                if ((rc = cylm_reset_current(c)) || (rc = cylm_b_half(c)) || (rc = cylm_push(c, true))) return rc;
                if ((rc = cylm_exchange(c))) return rc;
                if ((rc = cylm_b_half(c)) || (rc = cylm_e_phase(c, true, false)) || (rc = cylm_current_halo(c)) ||
                    (rc = cylm_e_phase(c, false, true)) || (rc = cylm_window_after_step(c)))
                    return rc;
            } else {
                if ((rc = cyl_reset_current(c)) || (rc = cyl_b_half(c)) || (rc = cyl_push(c, true))) return rc;
                if ((rc = cyl_exchange(c))) return rc;
                if ((rc = cyl_b_half(c)) || (rc = cyl_e_phase(c, true, false)) || (rc = cyl_current_halo(c)) ||
                    (rc = cyl_e_phase(c, false, true)))
                    return rc;
            }
            HIPTRY(hipEventRecord(c->ev[5], c->stream));
            HIPTRY(hipEventSynchronize(c->ev[5]));
            float ms = 0, f2 = 0;
            if (hipEventElapsedTime(&ms, c->ev[0], c->ev[1]) == hipSuccess) { c->t_ms[0] += ms; c->t_n[0]++; }
            if (hipEventElapsedTime(&f2, c->ev[4], c->ev[5]) == hipSuccess) { c->t_ms[2] += f2; c->t_n[2]++; }
        }
        return PICTOR_OK;
    }
    // MAIN.F90:42-284.  The host does not wait for a step's spill and error flags before
    // it enqueues the next step's ResetCurrent and first B half step (which do not touch
    // the particles): it waits for them behind those kernels, while the GPU runs them,
    // then re-lays out storage if needed and enqueues the push.  A step's events (set
    // ev[8 + 7 p .. 14 + 7 p], p alternating) are read once its flags have arrived.
    auto step_times = [&](int p) {
        hipEvent_t *E = c->ev + 8 + 7 * p;   // 0 start, 1 fields start, 2 push, 3 push end, 4 fields end, 5 end
        float ms = 0, f1 = 0, f2 = 0;
        if (hipEventElapsedTime(&ms, E[2], E[3]) == hipSuccess) { c->t_ms[0] += ms; c->t_n[0]++; }
        if (hipEventElapsedTime(&f1, E[1], E[4]) == hipSuccess) { c->t_ms[1] += f1 - ms; c->t_n[1]++; }
        if (hipEventElapsedTime(&f2, E[0], E[5]) == hipSuccess) { c->t_ms[2] += f2; c->t_n[2]++; }
    };
    int prev = -1;   // parity of the step whose flags are in flight
    for (int s = 0; s < nsteps; s++) {
        const int p = s & 1;
        hipEvent_t *E = c->ev + 8 + 7 * p;
        HIPTRY(hipEventRecord(E[0], c->stream));
        if ((rc = do_reset_current(c))) return rc;
        HIPTRY(hipEventRecord(E[1], c->stream));
        if ((rc = do_b_half(c))) return rc;
        if (prev >= 0) {
            rc = finish_push(c, c->ev[8 + 7 * prev + 6]);
            step_times(prev);
            prev = -1;
            if (rc) return rc;
        }
        if ((rc = do_push(c, true, E[2], E[3]))) return rc;
        if ((rc = bc_after_push(c))) return rc;   // EnforceBC_PostMovDep, MAIN.F90:77
        if (c->pml_on) {
            // UpdateBfieldHalf, UpdateMagFld_PML, SendRecvFlds (MAIN.F90:103-105)
            if ((rc = do_b_half(c, false)) || (rc = bc_pml(c, true)) || (rc = do_halo(c, 1))) return rc;
        } else if ((rc = do_b_half(c))) {
            return rc;
        }
        if ((rc = do_fold(c))) return rc;
        if ((rc = do_smoothen(c))) return rc;
        if (c->pml_on) {
            // UpdateEfield and UpdateElcFld_PML (MAIN.F90:108-109) before AddCurrent; the
            // halo between them only fills ghosts that the final halo fills again
            if ((rc = do_e_phase(c, true, false, false)) || (rc = bc_pml(c, false)) ||
                (rc = do_e_phase(c, false, true)))
                return rc;
        } else if ((rc = do_e_phase(c, true, true))) {
            return rc;
        }
        if ((rc = bc_final(c))) return rc;   // EnforceBC_Final, MAIN.F90:209
        HIPTRY(hipEventRecord(E[4], c->stream));
        HIPTRY(hipEventRecord(E[5], c->stream));
        if ((rc = enqueue_flags(c, E[6]))) return rc;
        prev = p;
    }
    if (prev >= 0) {
        rc = finish_push(c, c->ev[8 + 7 * prev + 6]);
        step_times(prev);
        if (rc) return rc;
    }
    return PICTOR_OK;
}

int pictor_synchronize(pictor_ctx *c)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    HIPTRY(hipStreamSynchronize(c->stream));
    return PICTOR_OK;
}

int pictor_cyl_push_scheme(pictor_ctx *c, int32_t scheme)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->cyl || c->cylm) return fail(PICTOR_ERR_ARG, "pictor_cyl_push_scheme is for the cylindrical (r, theta) grid");
    if (scheme != 0 && scheme != 1) return fail(PICTOR_ERR_ARG, "scheme must be 0 (Boris) or 1 (Strang)");
    c->cp.strang = scheme;
    return PICTOR_OK;
}

int pictor_cyl_sort_period(pictor_ctx *c, int32_t period)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->cylm) return fail(PICTOR_ERR_ARG, "pictor_cyl_sort_period is for the modal cylinder");
    if (period < 0) return fail(PICTOR_ERR_ARG, "period must be >= 0");
    c->cm.sort_period = period;
    c->cm.steps_since_sort = 0;
    return PICTOR_OK;
}

int pictor_cyl_plasma(pictor_ctx *c, int32_t nr_plasma, int32_t k0, int32_t nzc, int32_t ppc, uint64_t seed)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->cylm) return fail(PICTOR_ERR_ARG, "pictor_cyl_plasma is for the modal cylinder");
    return cylm_inject(c, nr_plasma, k0, nzc, ppc, seed, c->cm.win_shift);
}

int pictor_cyl_laser(pictor_ctx *c, double a0, double omega0, double tau, double w0, double zc)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->cylm) return fail(PICTOR_ERR_ARG, "pictor_cyl_laser is for the modal cylinder");
    if (!(omega0 > 0.0 && tau > 0.0 && w0 > 0.0)) return fail(PICTOR_ERR_ARG, "omega0, tau and w0 must be > 0");
    return cylm_laser(c, a0, omega0, tau, w0, zc);
}

int pictor_moving_window(pictor_ctx *c, int32_t on, int32_t nr_plasma, int32_t ppc, uint64_t seed)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->cylm) return fail(PICTOR_ERR_ARG, "pictor_moving_window is for the modal cylinder");
    if (on && (nr_plasma < 0 || nr_plasma > c->g.nx - 1 || ppc < 0))
        return fail(PICTOR_ERR_ARG, "nr_plasma must lie in 0..nx-1 and ppc >= 0");
    CylmState &st = c->cm;
    st.win_on = on != 0;
    c->cq.zopen = on ? 1 : 0;   // the window's box has open z ends
    st.win_nrp = nr_plasma;
    st.win_ppc = ppc;
    st.win_seed = seed;
    return PICTOR_OK;
}

int pictor_diagnostics(pictor_ctx *c, double out[9])
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "diagnostics between move_deposit and exchange_particles");
    if (c->cylm) return cylm_diagnostics(c, out);
    if (c->cyl) return cyl_diagnostics(c, out);
    const int nblk = 1024;
    std::vector<double> part((size_t)std::max(nblk * 2, c->tp.ntiles * 5));
    // field energies
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3];
        field_ptrs(c, E, B);
        launch_field_energy<double>(c->g, E, B, c->d_partial, nblk, c->stream);
    } else {
        float *E[3], *B[3];
        field_ptrs(c, E, B);
        launch_field_energy<float>(c->g, E, B, c->d_partial, nblk, c->stream, c->Elo);
    }
    HIPTRY(hipMemcpyAsync(part.data(), c->d_partial, nblk * 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    double be = 0, ee = 0;
    for (int b = 0; b < nblk; b++) { be += part[2 * b]; ee += part[2 * b + 1]; }
    out[0] = be * 0.5;
    out[1] = ee * 0.5;
    // particle energies, momentum, charge density
    launch_fill_zero<double>(c->d_rho, c->g.ncell, c->stream);
    if (c->form == PICTOR_FORM_FP64)
        launch_particle_diag_f64(c->g, c->ph, c->tp, c->pf[c->cur], c->d_base, c->d_cnt[c->cur], c->d_partial, c->d_rho,
                                 c->stream);
    else
        launch_particle_diag_split(c->g, c->ph, c->tp, c->ps[c->cur], c->d_base, c->d_cnt[c->cur], c->d_partial,
                                   c->d_rho, c->stream);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(part.data(), c->d_partial, (size_t)c->tp.ntiles * 5 * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    double acc[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < c->tp.ntiles; t++)
        for (int a = 0; a < 5; a++) acc[a] += part[(size_t)t * 5 + a];
    for (int a = 0; a < 5; a++) out[2 + a] = acc[a];
    // Gauss residual
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3];
        field_ptrs(c, E, B);
        launch_gauss<double>(c->g, E, c->d_rho, c->d_partial, nblk, c->stream);
    } else {
        float *E[3], *B[3];
        field_ptrs(c, E, B);
        launch_gauss<float>(c->g, E, c->d_rho, c->d_partial, nblk, c->stream, c->Elo);
    }
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(part.data(), c->d_partial, nblk * 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    double rm = 0, am = 0;
    for (int b = 0; b < nblk; b++) { rm = std::max(rm, part[2 * b]); am = std::max(am, part[2 * b + 1]); }
    out[7] = rm;
    out[8] = am;
    return PICTOR_OK;
}

int pictor_timer_reset(pictor_ctx *c)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    for (int i = 0; i < 3; i++) { c->t_ms[i] = 0; c->t_n[i] = 0; }
    return PICTOR_OK;
}

int pictor_timer_get(pictor_ctx *c, int32_t id, double *total_ms, int64_t *count)
{
    if (!c || id < 0 || id > 2) return fail(PICTOR_ERR_ARG, "bad timer id");
    if (total_ms) *total_ms = c->t_ms[id];
    if (count) *count = c->t_n[id];
    return PICTOR_OK;
}

int pictor_memory_bytes(pictor_ctx *c, int64_t *bytes)
{
    if (!c || !bytes) return fail(PICTOR_ERR_ARG, "null argument");
    if (c->cylm) {
        // FP64 E, B, J; the particle slots (FP64 record 60 B, split record 32 B)
        const long long per = (c->form == PICTOR_FORM_FP64) ? 60 : 32;
        *bytes = c->g.ncell * 9 * 8 + c->cm.cap * per;
        return PICTOR_OK;
    }
    const long long rs = (c->form == PICTOR_FORM_FP64) ? 8 : 4;
    long long b = c->g.ncell * (6 * rs + 3 * 8 + 8);
    const long long per = (c->form == PICTOR_FORM_FP64) ? 60 : 28;
    b += (c->buf_cap[0] + c->buf_cap[1]) * per + (long long)c->spill_cap * (per + 4);
    b += (long long)c->tp.ntiles * c->tp.NB * (8 + 4 + 8);
    *bytes = b;
    return PICTOR_OK;
}

}  // extern "C"
