// kernels_push.hip -- tile-local Boris push + zigzag deposit + binning (gfx950).
//
// One workgroup owns one tile of T x T cells.  It
//   1. stages the node-centred E and B of a (T+5)^2 window into LDS; these are
//      exactly the reference's VecE/VecB corner values (movdep_routines.F90:
//      6-223: corner nn of cell (i,j) is node (i + mod(nn-1,2), j + (nn-1)/2)),
//   2. streams its particles from HBM once (SoA, coalesced inside every bin),
//      gathers from LDS, pushes (movdep.F90:143-173), deposits the two zigzag
//      half-segments (movdep.F90:177-237) into an LDS current window,
//   3. applies the periodic exchange of the reference (loadprtlout.F90:22-97)
//      and writes each particle into the bin of its destination tile, slots
//      being reserved per wavefront with one LDS atomic per destination,
//   4. flushes the current window to the global J arrays (ghost layers
//      included; the halo fold adds them into their owners afterwards).
//
// FP64 form: positions are the reference's local cell coordinates, the push
// follows the reference expression order (bitwise with -ffp-contract=off) and
// the current window accumulates with LDS FP64 atomics.
// Split form: int32 cell + FP32 offsets, FP32 push; every current contribution
// is quantised to a multiple of 2^-30 (charge x cell) and accumulated as a
// 64-bit integer (LDS) and in FP64 (J), where all such sums are exact, so the
// current does not depend on the order of the atomics: the split form is
// deterministic.  Both species of a tile go
// through one pass, and each thread prefetches its next particle while it
// pushes the current one.
#include "pictor_common.h"
#include "push_common.h"

namespace pictor {
namespace {
using namespace pushdev;

// ---------------------------------------------------------------------------
// FP64 form, 2-D
// ---------------------------------------------------------------------------
template <int T, int BS, bool BC, bool TAG = false>
__global__ __launch_bounds__(BS) void k_push2d_f64(GridP g, PhysP ph, TileP tp, const double *__restrict__ Ex,
                                                   const double *__restrict__ Ey, const double *__restrict__ Ez,
                                                   const double *__restrict__ Bx, const double *__restrict__ By,
                                                   const double *__restrict__ Bz, double *__restrict__ Jx,
                                                   double *__restrict__ Jy, double *__restrict__ Jz, PartF64 A,
                                                   PartF64 Bo, BinsP bins, PartF64 spill, SpillP sp)
{
    constexpr int W = T + 5;
    constexpr int WW = W * W;
    __shared__ double sN[6][WW];
    __shared__ double sJ[3][WW];
    __shared__ int sStart[10];
    __shared__ long long sSrc[9];
    __shared__ long long sDstBase[9];
    __shared__ int sDstCap[9], sDstBin[9], sCnt[9];

    const int tile = tile_remap(tp, blockIdx.x);
    const int tx = tile % tp.ntx, ty = tile / tp.ntx;
    const int iw = 1 + tx * T, jw = 1 + ty * T;   // window origin = first cell - 2 (reference index)
    const int tid = threadIdx.x;
    const int mx = g.mx;

    if (tid < 9) {
        const int b = tile * tp.NB + tid;
        sSrc[tid] = bins.base[b];
        const int ddx = tid % 3 - 1, ddy = tid / 3 - 1;
        const int dt = wrap_tile(ty + ddy, tp.nty) * tp.ntx + wrap_tile(tx + ddx, tp.ntx);
        const int db = dt * tp.NB + tid;
        sDstBin[tid] = db;
        sDstBase[tid] = bins.base[db];
        sDstCap[tid] = bins.cap[db];
        sCnt[tid] = 0;
    }
    if (tid == 0) {
        int acc = 0;
        for (int b = 0; b < 9; b++) { sStart[b] = acc; acc += bins.cnt_in[tile * tp.NB + b]; }
        sStart[9] = acc;
    }
    const double bx0 = ph.bext[0], by0 = ph.bext[1], bz0 = ph.bext[2];
    for (int e = tid; e < WW; e += BS) {
        const int ii = e % W, jj = e / W;
        double nex = 0, ney = 0, nez = 0, nbx = 0, nby = 0, nbz = 0;
        if (ii >= 1 && jj >= 1) {
            const long long o = (long long)(iw + ii - 1) + (long long)mx * (jw + jj - 1);
            nex = (Ex[o - 1] + Ex[o]) * 0.5;
            ney = (Ey[o - mx] + Ey[o]) * 0.5;
            nez = Ez[o];
            nbx = (Bx[o - mx] + Bx[o]) * 0.5 + bx0;
            nby = (By[o - 1] + By[o]) * 0.5 + by0;
            nbz = (Bz[o - 1 - mx] + Bz[o - 1] + Bz[o - mx] + Bz[o]) * 0.25 + bz0;
        }
        sN[0][e] = nex; sN[1][e] = ney; sN[2][e] = nez;
        sN[3][e] = nbx; sN[4][e] = nby; sN[5][e] = nbz;
        sJ[0][e] = 0.0; sJ[1][e] = 0.0; sJ[2][e] = 0.0;
    }
    __syncthreads();

    const double c = ph.c, sqc = ph.sqc, cinv = ph.cinv, qi = ph.qi;
    const int total = sStart[9];
    for (int base = 0; base < total; base += BS) {
        const int k = base + tid;
        const bool active = k < total;
        double x0 = 0, y0 = 0, z0 = 0, un = 0, vn = 0, wn = 0, q = 0;
        double xn = 0, yn = 0, zn = 0;
        int flv = 0, d = -1, tg = 0;
        if (active) {
            int b = 0;
            while (k >= sStart[b + 1]) b++;
            const long long src = sSrc[b] + (k - sStart[b]);
            x0 = A.x[src]; y0 = A.y[src]; z0 = A.z[src];
            un = A.u[src]; vn = A.v[src]; wn = A.w[src];
            q = A.q[src]; flv = A.flv[src];
            if (TAG) tg = A.tag[src];
            const int ip0 = (int)x0, jp0 = (int)y0, kp0 = (int)z0;
            const int ii0 = ip0 - iw, jj0 = jp0 - jw;
            xn = x0; yn = y0; zn = z0;
            if (ii0 < 1 || ii0 > W - 3 || jj0 < 1 || jj0 > W - 3 || flv < 1 || flv > ph.nflvr) {
                atomicOr(sp.err, (int)kErrBounds);
                d = 4;
            } else {
                // gather, movdep.F90:88-139
                const double dx = x0 - ip0, dy = y0 - jp0;
                const double w1 = (1.0 - dx) * (1.0 - dy), w2 = dx * (1.0 - dy);
                const double w3 = (1.0 - dx) * dy, w4 = dx * dy;
                const int e0 = jj0 * W + ii0;
                double pf[6];
#pragma unroll
                for (int cc = 0; cc < 6; cc++) {
                    double a = 0.0;
                    a = a + w1 * sN[cc][e0];
                    a = a + w2 * sN[cc][e0 + 1];
                    a = a + w3 * sN[cc][e0 + W];
                    a = a + w4 * sN[cc][e0 + W + 1];
                    pf[cc] = a;
                }
                const double qm = ph.flvrqm[flv - 1] * 0.5;
                double pEx = pf[0] * qm, pEy = pf[1] * qm, pEz = pf[2] * qm;
                double pBx = pf[3] * qm, pBy = pf[4] * qm, pBz = pf[5] * qm;
                // Boris, movdep.F90:143-173
                double u0 = c * un + pEx;
                double v0 = c * vn + pEy;
                double w0 = c * wn + pEz;
                double gg = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                pBx = gg * pBx; pBy = gg * pBy; pBz = gg * pBz;
                const double f = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
                const double u1 = (u0 + v0 * pBz - w0 * pBy) * f;
                const double v1 = (v0 + w0 * pBx - u0 * pBz) * f;
                const double w1b = (w0 + u0 * pBy - v0 * pBx) * f;
                u0 = u0 + v1 * pBz - w1b * pBy + pEx;
                v0 = v0 + w1b * pBx - u1 * pBz + pEy;
                w0 = w0 + u1 * pBy - v1 * pBx + pEz;
                un = u0 * cinv; vn = v0 * cinv; wn = w0 * cinv;
                gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                xn = x0 + un * gg * ph.inv_dx;
                yn = y0 + vn * gg * ph.inv_dy;
                zn = z0 + wn * gg * ph.inv_dz;
                // zigzag deposit, movdep.F90:177-237
                const int ip = (int)xn, jp = (int)yn, kp = (int)zn;
                const double xr = dmin((double)(min(ip0, ip) + 1), dmax((double)max(ip0, ip), 0.5 * (x0 + xn)));
                const double yr = dmin((double)(min(jp0, jp) + 1), dmax((double)max(jp0, jp), 0.5 * (y0 + yn)));
                const double zr = dmin((double)(min(kp0, kp) + 1), dmax((double)max(kp0, kp), 0.5 * (z0 + zn)));
                const double qthis = q * qi;
                {
                    const double Fx = qthis * (xr - x0), Fy = qthis * (yr - y0), Fz = qthis * (zr - z0);
                    const double Wx = 0.5 * (x0 + xr) - ip0, Wy = 0.5 * (y0 + yr) - jp0;
                    const int e = e0;
                    atomicAdd(&sJ[0][e], Fx * (1.0 - Wy));
                    atomicAdd(&sJ[0][e + W], Fx * Wy);
                    atomicAdd(&sJ[1][e], Fy * (1.0 - Wx));
                    atomicAdd(&sJ[1][e + 1], Fy * Wx);
                    atomicAdd(&sJ[2][e], (Fz * (1.0 - Wx)) * (1.0 - Wy));
                    atomicAdd(&sJ[2][e + 1], (Fz * Wx) * (1.0 - Wy));
                    atomicAdd(&sJ[2][e + W], (Fz * (1.0 - Wx)) * Wy);
                    atomicAdd(&sJ[2][e + W + 1], (Fz * Wx) * Wy);
                }
                {
                    const double Fx = qthis * (xn - xr), Fy = qthis * (yn - yr), Fz = qthis * (zn - zr);
                    const double Wx = 0.5 * (xn + xr) - ip, Wy = 0.5 * (yn + yr) - jp;
                    const int e = (jp - jw) * W + (ip - iw);
                    atomicAdd(&sJ[0][e], Fx * (1.0 - Wy));
                    atomicAdd(&sJ[0][e + W], Fx * Wy);
                    atomicAdd(&sJ[1][e], Fy * (1.0 - Wx));
                    atomicAdd(&sJ[1][e + 1], Fy * Wx);
                    atomicAdd(&sJ[2][e], (Fz * (1.0 - Wx)) * (1.0 - Wy));
                    atomicAdd(&sJ[2][e + 1], (Fz * Wx) * (1.0 - Wy));
                    atomicAdd(&sJ[2][e + W], (Fz * (1.0 - Wx)) * Wy);
                    atomicAdd(&sJ[2][e + W + 1], (Fz * Wx) * Wy);
                }
                // 2-D z wrap, movdep.F90:241-249
                if (zn > ph.zmax) zn = zn - ph.zlen;
                else if (zn < ph.zmin) zn = ph.zlen + zn;
                // EnforceBC_PostMovDep (bc.F90:238-287)
                bool keep = true;
                if (BC) keep = particle_faces<double>(ph, g, xn, yn, zn, un, vn);
                // periodic exchange, loadprtlout.F90:31-76 + shifts subdomains.F90:720-740; a
                // face of a non-periodic axis sends to a null neighbour (the particle is gone)
                bool sent = false;
                if (keep) {
                    if (xn < ph.xmin) {
                        if ((yn <= (ph.ymax + ph.xmin - xn)) && (yn >= (ph.ymin - ph.xmin + xn))) {
                            if (BC && !g.per[0]) keep = false;
                            else xn = xn + ph.xlen;
                            sent = true;
                        }
                    } else if (xn > ph.xmax) {
                        if ((yn <= (ph.ymax - ph.xmax + xn)) && (yn >= (ph.ymin + ph.xmax - xn))) {
                            if (BC && !g.per[0]) keep = false;
                            else xn = xn - ph.xlen;
                            sent = true;
                        }
                    }
                }
                if (keep && !sent) {
                    if (yn < ph.ymin) {
                        if ((yn < (ph.ymin - ph.xmin + xn)) && (yn < (ph.ymin + ph.xmax - xn))) {
                            if (BC && !g.per[1]) keep = false;
                            else yn = yn + ph.ylen;
                        }
                    } else if (yn > ph.ymax) {
                        if ((yn > (ph.ymax + ph.xmin - xn)) && (yn > (ph.ymax - ph.xmax + xn))) {
                            if (BC && !g.per[1]) keep = false;
                            else yn = yn - ph.ylen;
                        }
                    }
                }
                if (keep) {
                    int cx = (int)xn - 3, cy = (int)yn - 3;
                    cx = cx < 0 ? 0 : (cx > g.nx - 1 ? g.nx - 1 : cx);
                    cy = cy < 0 ? 0 : (cy > g.ny - 1 ? g.ny - 1 : cy);
                    const int ddx = norm_dir(cx / T - tx, tp.ntx), ddy = norm_dir(cy / T - ty, tp.nty);
                    if (ddx < -1 || ddx > 1 || ddy < -1 || ddy > 1) {
                        atomicOr(sp.err, (int)kErrBounds);
                        d = 4;
                    } else {
                        d = (ddx + 1) + 3 * (ddy + 1);
                    }
                }
            }
        }
        const int slot = reserve_slot(d, sCnt);
        if (active && d >= 0) {
            if (slot < sDstCap[d]) {
                const long long o = sDstBase[d] + slot;
                Bo.x[o] = xn; Bo.y[o] = yn; Bo.z[o] = zn;
                Bo.u[o] = un; Bo.v[o] = vn; Bo.w[o] = wn;
                Bo.q[o] = q; Bo.flv[o] = flv;
                if (TAG) Bo.tag[o] = tg;
            } else {
                const int s = atomicAdd(sp.count, 1);
                if (s < sp.cap) {
                    spill.x[s] = xn; spill.y[s] = yn; spill.z[s] = zn;
                    spill.u[s] = un; spill.v[s] = vn; spill.w[s] = wn;
                    spill.q[s] = q; spill.flv[s] = flv;
                    if (TAG) spill.tag[s] = tg;
                } else {
                    atomicOr(sp.err, (int)kErrCapacity);
                }
            }
        }
    }
    __syncthreads();
    for (int e = tid; e < WW; e += BS) {
        const long long o = (long long)(iw + e % W - 1) + (long long)mx * (jw + e / W - 1);
        const double a = sJ[0][e], b = sJ[1][e], cz = sJ[2][e];
        if (a != 0.0) unsafeAtomicAdd(&Jx[o], a);
        if (b != 0.0) unsafeAtomicAdd(&Jy[o], b);
        if (cz != 0.0) unsafeAtomicAdd(&Jz[o], cz);
    }
    if (tid < 9) bins.cnt_out[sDstBin[tid]] = min(sCnt[tid], sDstCap[tid]);
}

// ---------------------------------------------------------------------------
// FP32 split form, 2-D
// ---------------------------------------------------------------------------
struct SplitRec {
    int cell, s;
    float ox, oy, u, v, w;
    int slot, id;   // input slot (TRACK) and the carried value: particle index (TRACK) or tag (TAG)
    float wq;       // charge weight (WQ)
};

// virtual bin v of a tile: the centre bins of all species first (most
// particles), then the other arrival directions
__device__ inline void vbin(int v, int nsp, int &s, int &d)
{
    if (v < nsp) { s = v; d = 4; return; }
    const int r = v - nsp;
    s = r >> 3;
    const int q = r & 7;
    d = q < 4 ? q : q + 1;
}

template <int T, int BS, bool TRACK = false, bool BC = false, bool TAG = false, bool WQ = false>
__global__ __launch_bounds__(BS) void k_push2d_split(GridP g, PhysP ph, TileP tp, const float *__restrict__ Ex,
                                                     const float *__restrict__ Ey, const float *__restrict__ Ez,
                                                     const float *__restrict__ Bx, const float *__restrict__ By,
                                                     const float *__restrict__ Bz, double *__restrict__ Jx,
                                                     double *__restrict__ Jy, double *__restrict__ Jz, PartSplit A,
                                                     PartSplit Bo, BinsP bins, PartSplit spill, SpillP sp,
                                                     int *__restrict__ dst)
{
    // FP32 form: fused multiply-add allowed (the FP64 form keeps it off for
    // bitwise parity with the reference order)
#pragma clang fp contract(fast)
    constexpr int W = T + 5;
    constexpr int WW = W * W;
    constexpr int MB = 18;   // species (<= 2) x 9 directions
    // cell-corner field layout: cell lc holds 6 float4 = (component c at corners 1..4),
    // the reference's VecE/VecB(nn, cell) (movdep_routines.F90:6-223) for the tile's cells
    __shared__ float4 sC[6 * T * T];
    __shared__ unsigned long long sJ[3][WW];
    __shared__ int sStart[MB + 1], sVs[MB];
    __shared__ unsigned sSrc[MB], sDstBase[MB];
    __shared__ int sDstCap[MB], sDstBin[MB], sCnt[MB];
    __shared__ float sQm[4], sQs[4];

    const int tile = tile_remap(tp, blockIdx.x);
    const int tx = tile % tp.ntx, ty = tile / tp.ntx;
    const int iw = 1 + tx * T, jw = 1 + ty * T;
    const int cx0 = tx * T - 2, cy0 = ty * T - 2;   // window origin as a physical cell index
    const int tid = threadIdx.x;
    const int mx = g.mx, nx = g.nx, ny = g.ny;
    const int nsp = tp.nsp, NV = nsp * 9;

    if (tid < NV) {
        int s, d;
        vbin(tid, nsp, s, d);
        sSrc[tid] = (unsigned)bins.base[tile * tp.NB + s * 9 + d];
        sVs[tid] = s;
        // destination slot tid = (species, direction) = (tid / 9, tid % 9)
        const int ds = tid / 9, dd = tid % 9;
        const int dt = wrap_tile(ty + dd / 3 - 1, tp.nty) * tp.ntx + wrap_tile(tx + dd % 3 - 1, tp.ntx);
        const int db = dt * tp.NB + ds * 9 + dd;
        sDstBin[tid] = db;
        sDstBase[tid] = (unsigned)bins.base[db];
        sDstCap[tid] = bins.cap[db];
        sCnt[tid] = 0;
    }
    if (tid == 0) {
        int acc = 0;
        for (int v = 0; v < NV; v++) {
            int s, d;
            vbin(v, nsp, s, d);
            sStart[v] = acc;
            acc += bins.cnt_in[tile * tp.NB + s * 9 + d];
        }
        sStart[NV] = acc;
        // q/m / 2 and the charge pre-scaled by 2^30 (exact) for the fixed-point current
        for (int s = 0; s < nsp; s++) { sQm[s] = (float)(ph.flvrqm[s] * 0.5); sQs[s] = (float)(ph.flvq[s] * kFixScale); }
    }
    const float bx0 = (float)ph.bext[0], by0 = (float)ph.bext[1], bz0 = (float)ph.bext[2];
    for (int lc = tid; lc < T * T; lc += BS) {
        // reference index of the cell: i = 3 + tx*T + li; its corner nodes (i..i+1, j..j+1)
        const long long o = (long long)(iw + 2 + lc % T - 1) + (long long)mx * (jw + 2 + lc / T - 1);
        float v[6][4];
#pragma unroll
        for (int cn = 0; cn < 4; cn++) {
            const long long p = o + (cn & 1) + (cn >> 1) * (long long)mx;
            v[0][cn] = (Ex[p - 1] + Ex[p]) * 0.5f;
            v[1][cn] = (Ey[p - mx] + Ey[p]) * 0.5f;
            v[2][cn] = Ez[p];
            v[3][cn] = (Bx[p - mx] + Bx[p]) * 0.5f + bx0;
            v[4][cn] = (By[p - 1] + By[p]) * 0.5f + by0;
            v[5][cn] = (Bz[p - 1 - mx] + Bz[p - 1] + Bz[p - mx] + Bz[p]) * 0.25f + bz0;
        }
#pragma unroll
        for (int cc = 0; cc < 6; cc++) sC[6 * lc + cc] = make_float4(v[cc][0], v[cc][1], v[cc][2], v[cc][3]);
    }
    for (int e = tid; e < WW; e += BS) { sJ[0][e] = 0ull; sJ[1][e] = 0ull; sJ[2][e] = 0ull; }
    __syncthreads();

    const float c = (float)ph.c, sqc = (float)ph.sqc, cinv = (float)ph.cinv;
    const float idx_ = (float)ph.inv_dx, idy_ = (float)ph.inv_dy, idz_ = (float)ph.inv_dz;
    const int total = sStart[NV];

    auto load = [&](int k, SplitRec &r) {
        int v = 0;
        while (k >= sStart[v + 1]) v++;
        const unsigned slot = sSrc[v] + (unsigned)(k - sStart[v]);
        const unsigned off = slot * 4u;
        r.cell = ldo(A.cell, off);
        r.ox = ldo(A.ox, off); r.oy = ldo(A.oy, off);
        r.u = ldo(A.u, off); r.v = ldo(A.v, off); r.w = ldo(A.w, off);
        r.s = sVs[v];
        if (TRACK) r.slot = (int)slot;
        if (TRACK || TAG) r.id = ldo(A.id, off);
        if (WQ) r.wq = ldo(A.wq, off);
    };

    SplitRec nxt{};
    if (tid < total) load(tid, nxt);
    for (int base = 0; base < total; base += BS) {
        const int k = base + tid;
        const bool active = k < total;
        const SplitRec cur = nxt;
        if (k + BS < total) load(k + BS, nxt);   // prefetch the next particle of this thread
        int d = -1, ncell = cur.cell;
        float ox = cur.ox, oy = cur.oy, un = cur.u, vn = cur.v, wn = cur.w;
        if (active) {
            const int ci = cur.cell & 0xFFFF, cj = cur.cell >> 16;
            const int ii0 = ci - cx0, jj0 = cj - cy0;
            if (ii0 < 2 || ii0 > T + 1 || jj0 < 2 || jj0 > T + 1) {
                atomicOr(sp.err, (int)kErrBounds);
                d = cur.s * 9 + 4;
            } else {
                const float x0 = cur.ox, y0 = cur.oy;
                const float qm = sQm[cur.s], qs = WQ ? sQs[cur.s] * cur.wq : sQs[cur.s];
                const float w1 = (1.0f - x0) * (1.0f - y0), w2 = x0 * (1.0f - y0);
                const float w3 = (1.0f - x0) * y0, w4 = x0 * y0;
                const int e0 = jj0 * W + ii0;
                const int lc = (jj0 - 2) * T + (ii0 - 2);
                const float4 gEx = sC[6 * lc], gEy = sC[6 * lc + 1], gEz = sC[6 * lc + 2];
                const float4 gBx = sC[6 * lc + 3], gBy = sC[6 * lc + 4], gBz = sC[6 * lc + 5];
                const float pEx = (w1 * gEx.x + w2 * gEx.y + w3 * gEx.z + w4 * gEx.w) * qm;
                const float pEy = (w1 * gEy.x + w2 * gEy.y + w3 * gEy.z + w4 * gEy.w) * qm;
                const float pEz = (w1 * gEz.x + w2 * gEz.y + w3 * gEz.z + w4 * gEz.w) * qm;
                const float hBx = (w1 * gBx.x + w2 * gBx.y + w3 * gBx.z + w4 * gBx.w) * qm;
                const float hBy = (w1 * gBy.x + w2 * gBy.y + w3 * gBy.z + w4 * gBy.w) * qm;
                const float hBz = (w1 * gBz.x + w2 * gBz.y + w3 * gBz.z + w4 * gBz.w) * qm;
                // Boris, movdep.F90:143-173 (hardware rsq/rcp: about 1 ulp, FP32 form)
                float u0 = c * un + pEx;
                float v0 = c * vn + pEy;
                float w0 = c * wn + pEz;
                float gg = __builtin_amdgcn_rsqf(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                const float pBx = gg * hBx, pBy = gg * hBy, pBz = gg * hBz;
                const float f = 2.0f * __builtin_amdgcn_rcpf(1.0f + pBx * pBx + pBy * pBy + pBz * pBz);
                const float u1 = (u0 + v0 * pBz - w0 * pBy) * f;
                const float v1 = (v0 + w0 * pBx - u0 * pBz) * f;
                const float w1b = (w0 + u0 * pBy - v0 * pBx) * f;
                u0 = u0 + v1 * pBz - w1b * pBy + pEx;
                v0 = v0 + w1b * pBx - u1 * pBz + pEy;
                w0 = w0 + u1 * pBy - v1 * pBx + pEz;
                un = u0 * cinv; vn = v0 * cinv; wn = w0 * cinv;
                gg = sqc * __builtin_amdgcn_rsqf(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                const float x1 = x0 + un * gg * idx_;
                const float y1 = y0 + vn * gg * idy_;
                const float z1 = kZMid + wn * gg * idz_;   // z is not stored (kZMid)
                int di = (int)floorf(x1), dj = (int)floorf(y1);
                // stored offsets; a position rounded onto the upper face belongs to the next cell
                ox = x1 - (float)di;
                if (ox >= 1.0f) { ox = 0.0f; di += 1; }
                oy = y1 - (float)dj;
                if (oy >= 1.0f) { oy = 0.0f; dj += 1; }
                // the two zigzag half-segments to the stored position: 8 node values
                // each, as exact integer counts of 2^-30 units
                double a[8], b[8];
                zigzag2d_split((double)qs, x0, y0, kZMid, ox, oy, z1, di, dj, a, b);
                // a particle that stays in its cell deposits both halves on the same
                // 8 nodes: add them first (exact) and halve the LDS atomics
                const bool cross = (di != 0) || (dj != 0);
                // the counts as 64-bit integers (|count| < 2^31 at the 2^-30 scale: one
                // conversion each), added with integer LDS atomics
                unsigned long long ia[8], ib[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    ia[q] = (unsigned long long)(long long)(int)a[q];
                    ib[q] = (unsigned long long)(long long)(int)b[q];
                }
                if (!cross) {
#pragma unroll
                    for (int q = 0; q < 8; q++) ia[q] += ib[q];
                }
                {
                    unsigned long long *jp = &sJ[0][e0];
                    atomicAdd(jp, ia[0]);
                    atomicAdd(jp + W, ia[1]);
                    atomicAdd(jp + WW, ia[2]);
                    atomicAdd(jp + WW + 1, ia[3]);
                    atomicAdd(jp + 2 * WW, ia[4]);
                    atomicAdd(jp + 2 * WW + 1, ia[5]);
                    atomicAdd(jp + 2 * WW + W, ia[6]);
                    atomicAdd(jp + 2 * WW + W + 1, ia[7]);
                }
                if (cross) {
                    unsigned long long *jp = &sJ[0][e0 + dj * W + di];
                    atomicAdd(jp, ib[0]);
                    atomicAdd(jp + W, ib[1]);
                    atomicAdd(jp + WW, ib[2]);
                    atomicAdd(jp + WW + 1, ib[3]);
                    atomicAdd(jp + 2 * WW, ib[4]);
                    atomicAdd(jp + 2 * WW + 1, ib[5]);
                    atomicAdd(jp + 2 * WW + W, ib[6]);
                    atomicAdd(jp + 2 * WW + W + 1, ib[7]);
                }
                int nci = ci + di, ncj = cj + dj;
                bool keep = true;
                if (BC) {
                    // the boundary faces on the FP64 position (exact for the cell and FP32
                    // offset); a reflected coordinate is split again into cell and offset
                    double X = 3.0 + (double)nci + (double)ox, Y = 3.0 + (double)ncj + (double)oy;
                    const double X0 = X, Y0 = Y;
                    keep = particle_faces<float>(ph, g, X, Y, 1.5, un, vn);
                    if (keep && X != X0) {
                        const double f = floor(X);
                        nci = (int)f - 3;
                        ox = (float)(X - f);
                        if (ox >= 1.0f) { ox = 0.0f; nci += 1; }
                    }
                    if (keep && Y != Y0) {
                        const double f = floor(Y);
                        ncj = (int)f - 3;
                        oy = (float)(Y - f);
                        if (oy >= 1.0f) { oy = 0.0f; ncj += 1; }
                    }
                    // a non-periodic axis has no image: the particle leaving it is gone
                    if ((nci < 0 || nci >= nx) && !g.per[0]) keep = false;
                    if ((ncj < 0 || ncj >= ny) && !g.per[1]) keep = false;
                }
                if (keep) {
                    nci = nci < 0 ? nci + nx : (nci >= nx ? nci - nx : nci);
                    ncj = ncj < 0 ? ncj + ny : (ncj >= ny ? ncj - ny : ncj);
                    ncell = nci | (ncj << 16);
                    const int ddx = norm_dir(nci / T - tx, tp.ntx), ddy = norm_dir(ncj / T - ty, tp.nty);
                    d = cur.s * 9 + (ddx + 1) + 3 * (ddy + 1);
                }
            }
        }
        const int slot = reserve_slot(d, sCnt);
        if (active && d >= 0) {
            if (slot < sDstCap[d]) {
                const unsigned off = (sDstBase[d] + (unsigned)slot) * 4u;
                sto(Bo.cell, off, ncell); sto(Bo.ox, off, ox); sto(Bo.oy, off, oy);
                sto(Bo.u, off, un); sto(Bo.v, off, vn); sto(Bo.w, off, wn);
                if (TRACK || TAG) sto(Bo.id, off, cur.id);
                if (WQ) sto(Bo.wq, off, cur.wq);
                if (TRACK && dst) dst[cur.slot] = (int)(sDstBase[d] + (unsigned)slot);
            } else {
                const int si = atomicAdd(sp.count, 1);
                if (si < sp.cap) {
                    spill.cell[si] = ncell; spill.ox[si] = ox; spill.oy[si] = oy;
                    spill.u[si] = un; spill.v[si] = vn; spill.w[si] = wn; spill.sp[si] = cur.s;
                    if (TAG) spill.id[si] = cur.id;
                    if (WQ) spill.wq[si] = cur.wq;
                } else {
                    atomicOr(sp.err, (int)kErrCapacity);
                }
                if (TRACK && dst) dst[cur.slot] = -1;
            }
        }
    }
    __syncthreads();
    if (tid < NV) bins.cnt_out[sDstBin[tid]] = min(sCnt[tid], sDstCap[tid]);
    constexpr double lim = kFixLimit * kFixScale;
    bool big = false;
    for (int e = tid; e < WW; e += BS) {
        const long long o = (long long)(iw + e % W - 1) + (long long)mx * (jw + e / W - 1);
        const double a = (double)(long long)sJ[0][e], b = (double)(long long)sJ[1][e];
        const double cz = (double)(long long)sJ[2][e];
        big = big || fabs(a) >= lim || fabs(b) >= lim || fabs(cz) >= lim;
        if (a != 0.0) unsafeAtomicAdd(&Jx[o], a);
        if (b != 0.0) unsafeAtomicAdd(&Jy[o], b);
        if (cz != 0.0) unsafeAtomicAdd(&Jz[o], cz);
    }
    if (big) atomicOr(sp.err, (int)kErrPrecision);
}

}  // namespace

// the kernel variant for the boundary faces (BC) and carried tags (TAG) of this context
#define PICK_BC_TAG(K, bc, tag, ...)                                                                                \
    do {                                                                                                           \
        if (bc) { if (tag) K(true, true, __VA_ARGS__); else K(true, false, __VA_ARGS__); }                           \
        else { if (tag) K(false, true, __VA_ARGS__); else K(false, false, __VA_ARGS__); }                            \
    } while (0)

// the split form's variant for the faces (BC), carried tags (TAG) and charge weights (WQ)
#define PICK_BC_TAG_WQ(K, bc, tag, wq, ...)                                                                         \
    do {                                                                                                           \
        if (wq) {                                                                                                  \
            if (bc) { if (tag) K(true, true, true, __VA_ARGS__); else K(true, false, true, __VA_ARGS__); }           \
            else { if (tag) K(false, true, true, __VA_ARGS__); else K(false, false, true, __VA_ARGS__); }            \
        } else {                                                                                                   \
            if (bc) { if (tag) K(true, true, false, __VA_ARGS__); else K(true, false, false, __VA_ARGS__); }         \
            else { if (tag) K(false, true, false, __VA_ARGS__); else K(false, false, false, __VA_ARGS__); }          \
        }                                                                                                          \
    } while (0)

void launch_push2d_f64(int T, const GridP &g, const PhysP &ph, const TileP &tp, double *const E[3],
                       double *const B[3], double *const J[3], PartF64 A, PartF64 Bo, BinsP bins, PartF64 spill,
                       SpillP sp, hipStream_t s)
{
    dim3 grid(tp.ntiles);
#define K_F64(BCV, TAGV, TT, BSZ)                                                                                  \
    hipLaunchKernelGGL((k_push2d_f64<TT, BSZ, BCV, TAGV>), grid, dim3(BSZ), 0, s, g, ph, tp, E[0], E[1], E[2], B[0], \
                       B[1], B[2], J[0], J[1], J[2], A, Bo, bins, spill, sp)
    const bool bc = ph.pbc_on != 0, tag = A.tag != nullptr;
    switch (T) {
    case 8: PICK_BC_TAG(K_F64, bc, tag, 8, 256); break;
    case 16: PICK_BC_TAG(K_F64, bc, tag, 16, 512); break;
    case 32: PICK_BC_TAG(K_F64, bc, tag, 32, 1024); break;
    default: break;
    }
#undef K_F64
}

void launch_push2d_split(int T, const GridP &g, const PhysP &ph, const TileP &tp, float *const E[3],
                         float *const B[3], double *const J[3], PartSplit A, PartSplit Bo, BinsP bins,
                         PartSplit spill, SpillP sp, hipStream_t s)
{
    dim3 grid(tp.ntiles);
#define K_SPLIT(BCV, TAGV, WQV, TT, BSZ)                                                                           \
    hipLaunchKernelGGL((k_push2d_split<TT, BSZ, false, BCV, TAGV, WQV>), grid, dim3(BSZ), 0, s, g, ph, tp, E[0],    \
                       E[1], E[2], B[0], B[1], B[2], J[0], J[1], J[2], A, Bo, bins, spill, sp, (int *)nullptr)
    const bool bc = ph.pbc_on != 0, tag = A.id != nullptr, wq = A.wq != nullptr;
    switch (T) {
    case 8: PICK_BC_TAG_WQ(K_SPLIT, bc, tag, wq, 8, 256); break;
    case 16: PICK_BC_TAG_WQ(K_SPLIT, bc, tag, wq, 16, 512); break;
    case 32: PICK_BC_TAG_WQ(K_SPLIT, bc, tag, wq, 32, 1024); break;
    default: break;
    }
#undef K_SPLIT
}

void launch_push2d_split_track(int T, const GridP &g, const PhysP &ph, const TileP &tp, float *const E[3],
                               float *const B[3], double *const J[3], PartSplit A, PartSplit Bo, BinsP bins,
                               PartSplit spill, SpillP sp, int *dst, hipStream_t s)
{
    dim3 grid(tp.ntiles);
#define PUSH_SPLIT(TT, BSZ)                                                                                         \
    hipLaunchKernelGGL((k_push2d_split<TT, BSZ, true>), grid, dim3(BSZ), 0, s, g, ph, tp, E[0], E[1], E[2], B[0], B[1], \
                       B[2], J[0], J[1], J[2], A, Bo, bins, spill, sp, dst)
    switch (T) {
    case 8: PUSH_SPLIT(8, 256); break;
    case 16: PUSH_SPLIT(16, 512); break;
    case 32: PUSH_SPLIT(32, 1024); break;
    default: break;
    }
#undef PUSH_SPLIT
}
#undef PICK_BC_TAG
#undef PICK_BC_TAG_WQ

}  // namespace pictor
