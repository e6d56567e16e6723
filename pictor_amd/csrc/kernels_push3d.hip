// kernels_push3d.hip -- 3-D tile-local Boris push + zigzag deposit + binning (gfx950).
//
// Same structure as the 2-D kernels (kernels_push.hip) on T x T x T tiles:
// node-centred E and B of the window staged in LDS (the reference's 8-corner
// VecE/VecB, movdep_routines.F90:6-223, corner nn of cell (i,j,k) is node
// (i + a, j + b, k + c), nn = 1 + a + 2b + 4c), the 12-value zigzag deposit of
// movdep.F90:187-234 into an LDS current window, the exchange of
// loadprtlout.F90:22-97 with its z rules, and 27 arrival bins per tile.
#include "pictor_common.h"
#include "push_common.h"

namespace pictor {
namespace {
using namespace pushdev;

// ---------------------------------------------------------------------------
// FP64 form, 3-D.  Gather window: nodes [i0-1, i0+T+1] (particles may sit one
// cell into a ghost layer after the exchange); current window [i0-2, i0+T+2].
// ---------------------------------------------------------------------------
template <int T, int BS, bool BC, bool TAG = false>
__global__ __launch_bounds__(BS) void k_push3d_f64(GridP g, PhysP ph, TileP tp, const double *__restrict__ Ex,
                                                   const double *__restrict__ Ey, const double *__restrict__ Ez,
                                                   const double *__restrict__ Bx, const double *__restrict__ By,
                                                   const double *__restrict__ Bz, double *__restrict__ Jx,
                                                   double *__restrict__ Jy, double *__restrict__ Jz, PartF64 A,
                                                   PartF64 Bo, BinsP bins, PartF64 spill, SpillP sp)
{
    constexpr int WN = T + 3, WN3 = WN * WN * WN;   // gather nodes
    constexpr int WJ = T + 5, WJ3 = WJ * WJ * WJ;   // current nodes
    __shared__ double sN[6][WN3];
    __shared__ double sJ[3][WJ3];
    __shared__ int sStart[28];
    __shared__ long long sSrc[27], sDstBase[27];
    __shared__ int sDstCap[27], sDstBin[27], sCnt[27];

    const int tile = tile_remap(tp, blockIdx.x);
    const int tx = tile % tp.ntx, ty = (tile / tp.ntx) % tp.nty, tz = tile / (tp.ntx * tp.nty);
    const int i0 = 3 + tx * T, j0 = 3 + ty * T, k0 = 3 + tz * T;   // first cell of the tile
    const int tid = threadIdx.x;
    const int mx = g.mx;
    const long long mxy = (long long)g.mx * g.my;

    if (tid < 27) {
        const int b = tile * tp.NB + tid;
        sSrc[tid] = bins.base[b];
        const int ddx = tid % 3 - 1, ddy = (tid / 3) % 3 - 1, ddz = tid / 9 - 1;
        const int dt = (wrap_tile(tz + ddz, tp.ntz) * tp.nty + wrap_tile(ty + ddy, tp.nty)) * tp.ntx +
                       wrap_tile(tx + ddx, tp.ntx);
        const int db = dt * tp.NB + tid;
        sDstBin[tid] = db;
        sDstBase[tid] = bins.base[db];
        sDstCap[tid] = bins.cap[db];
        sCnt[tid] = 0;
    }
    if (tid == 0) {
        int acc = 0;
        for (int b = 0; b < 27; b++) { sStart[b] = acc; acc += bins.cnt_in[tile * tp.NB + b]; }
        sStart[27] = acc;
    }
    const double bx0 = ph.bext[0], by0 = ph.bext[1], bz0 = ph.bext[2];
    for (int e = tid; e < WN3; e += BS) {
        const int ii = e % WN, jj = (e / WN) % WN, kk = e / (WN * WN);
        const long long o = fidx(g, i0 - 1 + ii, j0 - 1 + jj, k0 - 1 + kk);
        const long long oi = o - 1, oj = o - mx, ok = o - mxy;
        sN[0][e] = (Ex[oi] + Ex[o]) * 0.5;
        sN[1][e] = (Ey[oj] + Ey[o]) * 0.5;
        sN[2][e] = (Ez[ok] + Ez[o]) * 0.5;
        sN[3][e] = (Bx[oj - mxy] + Bx[ok] + Bx[oj] + Bx[o]) * 0.25 + bx0;
        sN[4][e] = (By[oi - mxy] + By[oi] + By[ok] + By[o]) * 0.25 + by0;
        sN[5][e] = (Bz[oi - mx] + Bz[oi] + Bz[oj] + Bz[o]) * 0.25 + bz0;
    }
    for (int e = tid; e < WJ3; e += BS) { sJ[0][e] = 0.0; sJ[1][e] = 0.0; sJ[2][e] = 0.0; }
    __syncthreads();

    const double c = ph.c, sqc = ph.sqc, cinv = ph.cinv, qi = ph.qi;
    const int total = sStart[27];
    for (int base = 0; base < total; base += BS) {
        const int k = base + tid;
        const bool active = k < total;
        double x0 = 0, y0 = 0, z0 = 0, un = 0, vn = 0, wn = 0, q = 0, xn = 0, yn = 0, zn = 0;
        int flv = 0, d = -1, tg = 0;
        if (active) {
            const int b = bin_of<27>(k, sStart, 27);
            const long long src = sSrc[b] + (k - sStart[b]);
            x0 = A.x[src]; y0 = A.y[src]; z0 = A.z[src];
            un = A.u[src]; vn = A.v[src]; wn = A.w[src];
            q = A.q[src]; flv = A.flv[src];
            if (TAG) tg = A.tag[src];
            xn = x0; yn = y0; zn = z0;
            const int ip0 = (int)x0, jp0 = (int)y0, kp0 = (int)z0;
            // node-window coordinates of the cell (origin i0-1)
            const int gi = ip0 - (i0 - 1), gj = jp0 - (j0 - 1), gk = kp0 - (k0 - 1);
            if (gi < 0 || gi > T + 1 || gj < 0 || gj > T + 1 || gk < 0 || gk > T + 1 || flv < 1 || flv > ph.nflvr) {
                atomicOr(sp.err, (int)kErrBounds);
                d = 13;
            } else {
                const double dx = x0 - ip0, dy = y0 - jp0, dz = z0 - kp0;
                double wt[8];
#pragma unroll
                for (int cn = 0; cn < 8; cn++) {
                    const double fx = (cn & 1) ? dx : 1.0 - dx;
                    const double fy = (cn & 2) ? dy : 1.0 - dy;
                    const double fz = (cn & 4) ? dz : 1.0 - dz;
                    wt[cn] = fx * fy * fz;
                }
                const int e0 = (gk * WN + gj) * WN + gi;
                double pf[6];
#pragma unroll
                for (int cc = 0; cc < 6; cc++) {
                    double a = 0.0;
#pragma unroll
                    for (int cn = 0; cn < 8; cn++)
                        a = a + wt[cn] * sN[cc][e0 + (cn & 1) + ((cn >> 1) & 1) * WN + (cn >> 2) * WN * WN];
                    pf[cc] = a;
                }
                const double qm = ph.flvrqm[flv - 1] * 0.5;
                double pEx = pf[0] * qm, pEy = pf[1] * qm, pEz = pf[2] * qm;
                double pBx = pf[3] * qm, pBy = pf[4] * qm, pBz = pf[5] * qm;
                double u0 = c * un + pEx;
                double v0 = c * vn + pEy;
                double w0 = c * wn + pEz;
                double gg = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                pBx = gg * pBx; pBy = gg * pBy; pBz = gg * pBz;
                const double f = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
                const double u1 = (u0 + v0 * pBz - w0 * pBy) * f;
                const double v1 = (v0 + w0 * pBx - u0 * pBz) * f;
                const double w1 = (w0 + u0 * pBy - v0 * pBx) * f;
                u0 = u0 + v1 * pBz - w1 * pBy + pEx;
                v0 = v0 + w1 * pBx - u1 * pBz + pEy;
                w0 = w0 + u1 * pBy - v1 * pBx + pEz;
                un = u0 * cinv; vn = v0 * cinv; wn = w0 * cinv;
                gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                xn = x0 + un * gg * ph.inv_dx;
                yn = y0 + vn * gg * ph.inv_dy;
                zn = z0 + wn * gg * ph.inv_dz;
                const int ip = (int)xn, jp = (int)yn, kp = (int)zn;
                const double xr = dmin((double)(min(ip0, ip) + 1), dmax((double)max(ip0, ip), 0.5 * (x0 + xn)));
                const double yr = dmin((double)(min(jp0, jp) + 1), dmax((double)max(jp0, jp), 0.5 * (y0 + yn)));
                const double zr = dmin((double)(min(kp0, kp) + 1), dmax((double)max(kp0, kp), 0.5 * (z0 + zn)));
                const double qthis = q * qi;
#pragma unroll
                for (int seg = 0; seg < 2; seg++) {
                    const double xa = seg ? xn : x0, ya = seg ? yn : y0, za = seg ? zn : z0;
                    const int ia = seg ? ip : ip0, ja = seg ? jp : jp0, ka = seg ? kp : kp0;
                    const double Fx = seg ? qthis * (xn - xr) : qthis * (xr - x0);
                    const double Fy = seg ? qthis * (yn - yr) : qthis * (yr - y0);
                    const double Fz = seg ? qthis * (zn - zr) : qthis * (zr - z0);
                    const double Wx = 0.5 * (xa + xr) - ia, Wy = 0.5 * (ya + yr) - ja, Wz = 0.5 * (za + zr) - ka;
                    const int e = ((ka - (k0 - 2)) * WJ + (ja - (j0 - 2))) * WJ + (ia - (i0 - 2));
                    constexpr int sy = WJ, sz = WJ * WJ;
                    // Jx(i, j..j+1, k..k+1), Jy(i..i+1, j, k..k+1), Jz(i..i+1, j..j+1, k)
                    atomicAdd(&sJ[0][e], (Fx * (1.0 - Wy)) * (1.0 - Wz));
                    atomicAdd(&sJ[0][e + sy], (Fx * Wy) * (1.0 - Wz));
                    atomicAdd(&sJ[0][e + sz], (Fx * (1.0 - Wy)) * Wz);
                    atomicAdd(&sJ[0][e + sy + sz], (Fx * Wy) * Wz);
                    atomicAdd(&sJ[1][e], (Fy * (1.0 - Wx)) * (1.0 - Wz));
                    atomicAdd(&sJ[1][e + 1], (Fy * Wx) * (1.0 - Wz));
                    atomicAdd(&sJ[1][e + sz], (Fy * (1.0 - Wx)) * Wz);
                    atomicAdd(&sJ[1][e + 1 + sz], (Fy * Wx) * Wz);
                    atomicAdd(&sJ[2][e], (Fz * (1.0 - Wx)) * (1.0 - Wy));
                    atomicAdd(&sJ[2][e + 1], (Fz * Wx) * (1.0 - Wy));
                    atomicAdd(&sJ[2][e + sy], (Fz * (1.0 - Wx)) * Wy);
                    atomicAdd(&sJ[2][e + 1 + sy], (Fz * Wx) * Wy);
                }
                // EnforceBC_PostMovDep (bc.F90:238-287)
                bool keep = true;
                if (BC) keep = particle_faces<double>(ph, g, xn, yn, zn, un, vn);
                // periodic exchange, loadprtlout.F90:31-96 (3-D rules); a face of a
                // non-periodic axis sends to a null neighbour (the particle is gone)
                bool sent = false;
                if (keep) {
                    if (xn < ph.xmin) {
                        if ((yn <= (ph.ymax + ph.xmin - xn)) && (yn >= (ph.ymin - ph.xmin + xn)) &&
                            (zn <= (ph.zmax + ph.xmin - xn)) && (zn >= (ph.zmin - ph.xmin + xn))) {
                            if (BC && !g.per[0]) keep = false; else xn = xn + ph.xlen;
                            sent = true;
                        }
                    } else if (xn > ph.xmax) {
                        if ((yn <= (ph.ymax - ph.xmax + xn)) && (yn >= (ph.ymin + ph.xmax - xn)) &&
                            (zn <= (ph.zmax - ph.xmax + xn)) && (zn >= (ph.zmin + ph.xmax - xn))) {
                            if (BC && !g.per[0]) keep = false; else xn = xn - ph.xlen;
                            sent = true;
                        }
                    }
                }
                if (keep && !sent) {
                    if (yn < ph.ymin) {
                        if ((yn < (ph.ymin - ph.xmin + xn)) && (yn < (ph.ymin + ph.xmax - xn)) &&
                            (zn >= (ph.zmin - ph.ymin + yn)) && (zn <= (ph.zmax + ph.ymin - yn))) {
                            if (BC && !g.per[1]) keep = false; else yn = yn + ph.ylen;
                            sent = true;
                        }
                    } else if (yn > ph.ymax) {
                        if ((yn > (ph.ymax + ph.xmin - xn)) && (yn > (ph.ymax - ph.xmax + xn)) &&
                            (zn >= (ph.zmin + ph.ymax - yn)) && (zn <= (ph.zmax - ph.ymax + yn))) {
                            if (BC && !g.per[1]) keep = false; else yn = yn - ph.ylen;
                            sent = true;
                        }
                    }
                }
                if (keep && !sent) {
                    if (zn < ph.zmin) {
                        if ((zn < (ph.zmin - ph.xmin + xn)) && (zn < (ph.zmin + ph.xmax - xn)) &&
                            (zn < (ph.zmin - ph.ymin + yn)) && (zn < (ph.zmin + ph.ymax - yn))) {
                            if (BC && !g.per[2]) keep = false; else zn = zn + ph.zlen;
                        }
                    } else if (zn > ph.zmax) {
                        if ((zn > (ph.zmax + ph.xmin - xn)) && (zn > (ph.zmax - ph.xmax + xn)) &&
                            (zn > (ph.zmax + ph.ymin - yn)) && (zn > (ph.zmax - ph.ymax + yn))) {
                            if (BC && !g.per[2]) keep = false; else zn = zn - ph.zlen;
                        }
                    }
                }
                if (keep) {
                    int cx = (int)xn - 3, cy = (int)yn - 3, cz = (int)zn - 3;
                    cx = cx < 0 ? 0 : (cx > g.nx - 1 ? g.nx - 1 : cx);
                    cy = cy < 0 ? 0 : (cy > g.ny - 1 ? g.ny - 1 : cy);
                    cz = cz < 0 ? 0 : (cz > g.nz - 1 ? g.nz - 1 : cz);
                    const int ddx = norm_dir(cx / T - tx, tp.ntx), ddy = norm_dir(cy / T - ty, tp.nty);
                    const int ddz = norm_dir(cz / T - tz, tp.ntz);
                    if (ddx < -1 || ddx > 1 || ddy < -1 || ddy > 1 || ddz < -1 || ddz > 1) {
                        atomicOr(sp.err, (int)kErrBounds);
                        d = 13;
                    } else {
                        d = (ddx + 1) + 3 * (ddy + 1) + 9 * (ddz + 1);
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
    for (int e = tid; e < WJ3; e += BS) {
        const long long o = fidx(g, i0 - 2 + e % WJ, j0 - 2 + (e / WJ) % WJ, k0 - 2 + e / (WJ * WJ));
        const double a = sJ[0][e], b = sJ[1][e], cz = sJ[2][e];
        if (a != 0.0) unsafeAtomicAdd(&Jx[o], a);
        if (b != 0.0) unsafeAtomicAdd(&Jy[o], b);
        if (cz != 0.0) unsafeAtomicAdd(&Jz[o], cz);
    }
    if (tid < 27) bins.cnt_out[sDstBin[tid]] = min(sCnt[tid], sDstCap[tid]);
}

// ---------------------------------------------------------------------------
// FP32 split form, 3-D.  Particles always sit inside their tile: gather nodes
// [i0, i0+T], current window [i0-1, i0+T+1].
// ---------------------------------------------------------------------------
struct SplitRec3 {
    int cell, s;
    float ox, oy, oz, u, v, w;
    int id;     // the caller's tag (TAG only)
    float wq;   // charge weight (WQ only)
};

template <int T, int BS, bool BC, bool TAG = false, bool WQ = false>
__global__ __launch_bounds__(BS) void k_push3d_split(GridP g, PhysP ph, TileP tp, const float *__restrict__ Ex,
                                                     const float *__restrict__ Ey, const float *__restrict__ Ez,
                                                     const float *__restrict__ Bx, const float *__restrict__ By,
                                                     const float *__restrict__ Bz, double *__restrict__ Jx,
                                                     double *__restrict__ Jy, double *__restrict__ Jz, PartSplit A,
                                                     PartSplit Bo, BinsP bins, PartSplit spill, SpillP sp)
{
#pragma clang fp contract(fast)
    constexpr int WN = T + 1, WN3 = WN * WN * WN;
    constexpr int WJ = T + 3, WJ3 = WJ * WJ * WJ;
    constexpr int MB = 54;   // species (<= 2) x 27 directions
    // node e: sNa[e] = (Ex, Ey, Ez, Bx), sNb[e] = (By, Bz); 24 B per node keeps the LDS of a
    // tile-8 workgroup under a third of the CU's
    __shared__ float4 sNa[WN3];
    __shared__ float2 sNb[WN3];
    __shared__ unsigned long long sJ[3][WJ3];
    __shared__ int sStart[MB + 1], sVs[MB];
    __shared__ unsigned sSrc[MB], sDstBase[MB];
    __shared__ int sDstCap[MB], sDstBin[MB], sCnt[MB];
    __shared__ float sQm[4], sQs[4];

    const int tile = tile_remap(tp, blockIdx.x);
    const int tx = tile % tp.ntx, ty = (tile / tp.ntx) % tp.nty, tz = tile / (tp.ntx * tp.nty);
    const int i0 = 3 + tx * T, j0 = 3 + ty * T, k0 = 3 + tz * T;
    const int tid = threadIdx.x;
    const int mx = g.mx, nx = g.nx, ny = g.ny, nz = g.nz;
    const long long mxy = (long long)g.mx * g.my;
    const int nsp = tp.nsp, NV = nsp * 27;

    if (tid < NV) {
        // virtual bins: centre bins of every species first, then the 26 others
        int s, d;
        if (tid < nsp) { s = tid; d = 13; }
        else { const int r = tid - nsp; s = r / 26; const int q = r % 26; d = q < 13 ? q : q + 1; }
        sSrc[tid] = (unsigned)bins.base[tile * tp.NB + s * 27 + d];
        sVs[tid] = s;
        const int ds = tid / 27, dd = tid % 27;
        const int ddx = dd % 3 - 1, ddy = (dd / 3) % 3 - 1, ddz = dd / 9 - 1;
        const int dt = (wrap_tile(tz + ddz, tp.ntz) * tp.nty + wrap_tile(ty + ddy, tp.nty)) * tp.ntx +
                       wrap_tile(tx + ddx, tp.ntx);
        const int db = dt * tp.NB + ds * 27 + dd;
        sDstBin[tid] = db;
        sDstBase[tid] = (unsigned)bins.base[db];
        sDstCap[tid] = bins.cap[db];
        sCnt[tid] = 0;
    }
    if (tid == 0) {
        int acc = 0;
        for (int v = 0; v < NV; v++) {
            int s, d;
            if (v < nsp) { s = v; d = 13; }
            else { const int r = v - nsp; s = r / 26; const int q = r % 26; d = q < 13 ? q : q + 1; }
            sStart[v] = acc;
            acc += bins.cnt_in[tile * tp.NB + s * 27 + d];
        }
        sStart[NV] = acc;
        for (int s = 0; s < nsp; s++) { sQm[s] = (float)(ph.flvrqm[s] * 0.5); sQs[s] = (float)(ph.flvq[s] * kFixScale); }
    }
    const float bx0 = (float)ph.bext[0], by0 = (float)ph.bext[1], bz0 = (float)ph.bext[2];
    for (int e = tid; e < WN3; e += BS) {
        const int ii = e % WN, jj = (e / WN) % WN, kk = e / (WN * WN);
        const long long o = fidx(g, i0 + ii, j0 + jj, k0 + kk);
        const long long oi = o - 1, oj = o - mx, ok = o - mxy;
        sNa[e] = make_float4((Ex[oi] + Ex[o]) * 0.5f, (Ey[oj] + Ey[o]) * 0.5f, (Ez[ok] + Ez[o]) * 0.5f,
                             (Bx[oj - mxy] + Bx[ok] + Bx[oj] + Bx[o]) * 0.25f + bx0);
        sNb[e] = make_float2((By[oi - mxy] + By[oi] + By[ok] + By[o]) * 0.25f + by0,
                             (Bz[oi - mx] + Bz[oi] + Bz[oj] + Bz[o]) * 0.25f + bz0);
    }
    for (int e = tid; e < WJ3; e += BS) { sJ[0][e] = 0ull; sJ[1][e] = 0ull; sJ[2][e] = 0ull; }
    __syncthreads();

    const float c = (float)ph.c, sqc = (float)ph.sqc, cinv = (float)ph.cinv;
    const float idx_ = (float)ph.inv_dx, idy_ = (float)ph.inv_dy, idz_ = (float)ph.inv_dz;
    const int total = sStart[NV];

    auto load = [&](int k, SplitRec3 &r) {
        const int v = bin_of<MB>(k, sStart, NV);
        const unsigned off = (sSrc[v] + (unsigned)(k - sStart[v])) * 4u;
        r.cell = ldo(A.cell, off);
        r.ox = ldo(A.ox, off); r.oy = ldo(A.oy, off); r.oz = ldo(A.oz, off);
        r.u = ldo(A.u, off); r.v = ldo(A.v, off); r.w = ldo(A.w, off);
        r.s = sVs[v];
        if (TAG) r.id = ldo(A.id, off);
        if (WQ) r.wq = ldo(A.wq, off);
    };

    SplitRec3 nxt{};
    if (tid < total) load(tid, nxt);
    for (int base = 0; base < total; base += BS) {
        const int k = base + tid;
        const bool active = k < total;
        const SplitRec3 cur = nxt;
        if (k + BS < total) load(k + BS, nxt);
        int d = -1, ncell = cur.cell;
        float ox = cur.ox, oy = cur.oy, oz = cur.oz, un = cur.u, vn = cur.v, wn = cur.w;
        if (active) {
            int ci, cj, ck;
            cell_unpack(3, cur.cell, ci, cj, ck);
            const int gi = ci - (i0 - 3), gj = cj - (j0 - 3), gk = ck - (k0 - 3);   // 0..T-1 inside the tile
            if (gi < 0 || gi >= T || gj < 0 || gj >= T || gk < 0 || gk >= T) {
                atomicOr(sp.err, (int)kErrBounds);
                d = cur.s * 27 + 13;
            } else {
                const float x0 = cur.ox, y0 = cur.oy, z0 = cur.oz;
                const float qm = sQm[cur.s], qs = WQ ? sQs[cur.s] * cur.wq : sQs[cur.s];
                float wt[8];
#pragma unroll
                for (int cn = 0; cn < 8; cn++)
                    wt[cn] = ((cn & 1) ? x0 : 1.0f - x0) * ((cn & 2) ? y0 : 1.0f - y0) * ((cn & 4) ? z0 : 1.0f - z0);
                const int e0 = (gk * WN + gj) * WN + gi;
                float pE[3] = {0.f, 0.f, 0.f}, hB[3] = {0.f, 0.f, 0.f};
#pragma unroll
                for (int cn = 0; cn < 8; cn++) {
                    const int e = e0 + (cn & 1) + ((cn >> 1) & 1) * WN + (cn >> 2) * WN * WN;
                    const float4 a = sNa[e];
                    const float2 b = sNb[e];
                    pE[0] += wt[cn] * a.x; pE[1] += wt[cn] * a.y; pE[2] += wt[cn] * a.z;
                    hB[0] += wt[cn] * a.w; hB[1] += wt[cn] * b.x; hB[2] += wt[cn] * b.y;
                }
                const float pEx = pE[0] * qm, pEy = pE[1] * qm, pEz = pE[2] * qm;
                float u0 = c * un + pEx;
                float v0 = c * vn + pEy;
                float w0 = c * wn + pEz;
                float gg = __builtin_amdgcn_rsqf(sqc + u0 * u0 + v0 * v0 + w0 * w0);
                const float pBx = gg * hB[0] * qm, pBy = gg * hB[1] * qm, pBz = gg * hB[2] * qm;
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
                const float z1 = z0 + wn * gg * idz_;
                int di = (int)floorf(x1), dj = (int)floorf(y1), dk = (int)floorf(z1);
                const float xr = fminf((float)(min(0, di) + 1), fmaxf((float)max(0, di), 0.5f * (x0 + x1)));
                const float yr = fminf((float)(min(0, dj) + 1), fmaxf((float)max(0, dj), 0.5f * (y0 + y1)));
                const float zr = fminf((float)(min(0, dk) + 1), fmaxf((float)max(0, dk), 0.5f * (z0 + z1)));
                long long a[12], bq[12];
#pragma unroll
                for (int seg = 0; seg < 2; seg++) {
                    const float Fx = seg ? qs * (x1 - xr) : qs * (xr - x0);
                    const float Fy = seg ? qs * (y1 - yr) : qs * (yr - y0);
                    const float Fz = seg ? qs * (z1 - zr) : qs * (zr - z0);
                    const float Wx = seg ? 0.5f * (x1 + xr) - (float)di : 0.5f * (x0 + xr);
                    const float Wy = seg ? 0.5f * (y1 + yr) - (float)dj : 0.5f * (y0 + yr);
                    const float Wz = seg ? 0.5f * (z1 + zr) - (float)dk : 0.5f * (z0 + zr);
                    long long *t = seg ? bq : a;
                    t[0] = quanti((Fx * (1.0f - Wy)) * (1.0f - Wz)); t[1] = quanti((Fx * Wy) * (1.0f - Wz));
                    t[2] = quanti((Fx * (1.0f - Wy)) * Wz);         t[3] = quanti((Fx * Wy) * Wz);
                    t[4] = quanti((Fy * (1.0f - Wx)) * (1.0f - Wz)); t[5] = quanti((Fy * Wx) * (1.0f - Wz));
                    t[6] = quanti((Fy * (1.0f - Wx)) * Wz);         t[7] = quanti((Fy * Wx) * Wz);
                    t[8] = quanti((Fz * (1.0f - Wx)) * (1.0f - Wy)); t[9] = quanti((Fz * Wx) * (1.0f - Wy));
                    t[10] = quanti((Fz * (1.0f - Wx)) * Wy);        t[11] = quanti((Fz * Wx) * Wy);
                }
                const bool cross = (di != 0) || (dj != 0) || (dk != 0);
                if (!cross) {
#pragma unroll
                    for (int q = 0; q < 12; q++) a[q] += bq[q];
                }
                constexpr int sy = WJ, sz = WJ * WJ;
                {
                    unsigned long long *jp = &sJ[0][((gk + 1) * WJ + (gj + 1)) * WJ + (gi + 1)];
                    atomicAdd(jp, (unsigned long long)a[0]); atomicAdd(jp + sy, (unsigned long long)a[1]); atomicAdd(jp + sz, (unsigned long long)a[2]); atomicAdd(jp + sy + sz, (unsigned long long)a[3]);
                    jp += WJ3;
                    atomicAdd(jp, (unsigned long long)a[4]); atomicAdd(jp + 1, (unsigned long long)a[5]); atomicAdd(jp + sz, (unsigned long long)a[6]); atomicAdd(jp + 1 + sz, (unsigned long long)a[7]);
                    jp += WJ3;
                    atomicAdd(jp, (unsigned long long)a[8]); atomicAdd(jp + 1, (unsigned long long)a[9]); atomicAdd(jp + sy, (unsigned long long)a[10]); atomicAdd(jp + 1 + sy, (unsigned long long)a[11]);
                }
                if (cross) {
                    unsigned long long *jp = &sJ[0][((gk + 1 + dk) * WJ + (gj + 1 + dj)) * WJ + (gi + 1 + di)];
                    atomicAdd(jp, (unsigned long long)bq[0]); atomicAdd(jp + sy, (unsigned long long)bq[1]); atomicAdd(jp + sz, (unsigned long long)bq[2]); atomicAdd(jp + sy + sz, (unsigned long long)bq[3]);
                    jp += WJ3;
                    atomicAdd(jp, (unsigned long long)bq[4]); atomicAdd(jp + 1, (unsigned long long)bq[5]); atomicAdd(jp + sz, (unsigned long long)bq[6]); atomicAdd(jp + 1 + sz, (unsigned long long)bq[7]);
                    jp += WJ3;
                    atomicAdd(jp, (unsigned long long)bq[8]); atomicAdd(jp + 1, (unsigned long long)bq[9]); atomicAdd(jp + sy, (unsigned long long)bq[10]); atomicAdd(jp + 1 + sy, (unsigned long long)bq[11]);
                }
                ox = x1 - (float)di;
                if (ox >= 1.0f) { ox = 0.0f; di += 1; }
                oy = y1 - (float)dj;
                if (oy >= 1.0f) { oy = 0.0f; dj += 1; }
                oz = z1 - (float)dk;
                if (oz >= 1.0f) { oz = 0.0f; dk += 1; }
                int nci = ci + di, ncj = cj + dj, nck = ck + dk;
                bool keep = true;
                if (BC) {
                    // the boundary faces on the FP64 position; a reflected coordinate is
                    // split again into cell and offset
                    double X = 3.0 + (double)nci + (double)ox, Y = 3.0 + (double)ncj + (double)oy;
                    const double Z = 3.0 + (double)nck + (double)oz;
                    const double X0 = X, Y0 = Y;
                    keep = particle_faces<float>(ph, g, X, Y, Z, un, vn);
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
                    if ((nci < 0 || nci >= nx) && !g.per[0]) keep = false;
                    if ((ncj < 0 || ncj >= ny) && !g.per[1]) keep = false;
                    if ((nck < 0 || nck >= nz) && !g.per[2]) keep = false;
                }
                if (keep) {
                    nci = nci < 0 ? nci + nx : (nci >= nx ? nci - nx : nci);
                    ncj = ncj < 0 ? ncj + ny : (ncj >= ny ? ncj - ny : ncj);
                    nck = nck < 0 ? nck + nz : (nck >= nz ? nck - nz : nck);
                    ncell = cell_pack(3, nci, ncj, nck);
                    const int ddx = norm_dir(nci / T - tx, tp.ntx), ddy = norm_dir(ncj / T - ty, tp.nty);
                    const int ddz = norm_dir(nck / T - tz, tp.ntz);
                    d = cur.s * 27 + (ddx + 1) + 3 * (ddy + 1) + 9 * (ddz + 1);
                }
            }
        }
        const int slot = reserve_slot(d, sCnt);
        if (active && d >= 0) {
            if (slot < sDstCap[d]) {
                const unsigned off = (sDstBase[d] + (unsigned)slot) * 4u;
                sto(Bo.cell, off, ncell); sto(Bo.ox, off, ox); sto(Bo.oy, off, oy); sto(Bo.oz, off, oz);
                sto(Bo.u, off, un); sto(Bo.v, off, vn); sto(Bo.w, off, wn);
                if (TAG) sto(Bo.id, off, cur.id);
                if (WQ) sto(Bo.wq, off, cur.wq);
            } else {
                const int si = atomicAdd(sp.count, 1);
                if (si < sp.cap) {
                    spill.cell[si] = ncell; spill.ox[si] = ox; spill.oy[si] = oy; spill.oz[si] = oz;
                    spill.u[si] = un; spill.v[si] = vn; spill.w[si] = wn; spill.sp[si] = cur.s;
                    if (TAG) spill.id[si] = cur.id;
                    if (WQ) spill.wq[si] = cur.wq;
                } else {
                    atomicOr(sp.err, (int)kErrCapacity);
                }
            }
        }
    }
    __syncthreads();
    if (tid < NV) bins.cnt_out[sDstBin[tid]] = min(sCnt[tid], sDstCap[tid]);
    constexpr double lim = kFixLimit * kFixScale;
    bool big = false;
    for (int e = tid; e < WJ3; e += BS) {
        const long long o = fidx(g, i0 - 1 + e % WJ, j0 - 1 + (e / WJ) % WJ, k0 - 1 + e / (WJ * WJ));
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

void launch_push3d_f64(int T, const GridP &g, const PhysP &ph, const TileP &tp, double *const E[3],
                       double *const B[3], double *const J[3], PartF64 A, PartF64 Bo, BinsP bins, PartF64 spill,
                       SpillP sp, hipStream_t s)
{
    dim3 grid(tp.ntiles);
#define K3_F64(BCV, TAGV, TT, BSZ)                                                                                 \
    hipLaunchKernelGGL((k_push3d_f64<TT, BSZ, BCV, TAGV>), grid, dim3(BSZ), 0, s, g, ph, tp, E[0], E[1], E[2], B[0], \
                       B[1], B[2], J[0], J[1], J[2], A, Bo, bins, spill, sp)
    const bool bc = ph.pbc_on != 0, tag = A.tag != nullptr;
    switch (T) {
    case 4: PICK_BC_TAG(K3_F64, bc, tag, 4, 256); break;
    case 8: PICK_BC_TAG(K3_F64, bc, tag, 8, 1024); break;
    default: break;
    }
#undef K3_F64
}

void launch_push3d_split(int T, const GridP &g, const PhysP &ph, const TileP &tp, float *const E[3],
                         float *const B[3], double *const J[3], PartSplit A, PartSplit Bo, BinsP bins,
                         PartSplit spill, SpillP sp, hipStream_t s)
{
    dim3 grid(tp.ntiles);
#define K3_SPLIT(BCV, TAGV, WQV, TT, BSZ)                                                                          \
    hipLaunchKernelGGL((k_push3d_split<TT, BSZ, BCV, TAGV, WQV>), grid, dim3(BSZ), 0, s, g, ph, tp, E[0], E[1],      \
                       E[2], B[0], B[1], B[2], J[0], J[1], J[2], A, Bo, bins, spill, sp)
    const bool bc = ph.pbc_on != 0, tag = A.id != nullptr, wq = A.wq != nullptr;
    switch (T) {
    case 4: PICK_BC_TAG_WQ(K3_SPLIT, bc, tag, wq, 4, 256); break;
    case 8: PICK_BC_TAG_WQ(K3_SPLIT, bc, tag, wq, 8, 512); break;
    default: break;
    }
#undef K3_SPLIT
}
#undef PICK_BC_TAG
#undef PICK_BC_TAG_WQ

}  // namespace pictor
