// cyl.hip -- the cylindrical geometry on the rank that holds the axis (gfx950, FP64).
//
// The reference's -Dcyl build: x = r (grid index i at r = i - 3.5, the axis
// at x = 3.5), y = theta (ny cells over 2 pi), z = z (3-D).  Field kernels
// follow cyl_em_update.F90 (and cyl_em_update_lehe.F90 in 3-D) expression
// for expression, so with fused multiply-add off they agree with the
// reference order bit for bit.  The push is one thread per particle: gather
// of node-averaged corner values, Boris in the local (r, theta, z) frame,
// the move as a Cartesian straight line rotated back to (r, theta), the
// zigzag deposit with global FP64 atomics, and, for particles that start
// within one cell of the axis, the reference's adaptive sub-steps with a
// bounded loop.  Particles stay in the caller's order.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <vector>

#include "ctx.h"
#include "cyl.h"

using namespace pictor;

namespace pictor {

void cyl_params(const GridP &g, const PhysP &ph, int fdtd, CylP &cp)
{
    const double pi = 3.14159265358979323846264;   // vars.F90:224
    cp = CylP{};
    cp.dtheta = 2.0 * pi / g.ny;
    cp.ax_perm_area = 4.0 / g.ny;
    cp.inv_dtheta = 1.0 / cp.dtheta;
    cp.rmin = 0.0 + (-0.5);             // xborders(0) + rshift
    cp.r_proc = cp.rmin - ph.xmin;      // xborders(0) + rshift - xmin
    const double bxy = 0.0, bxz = 0.125, byx = 0.0, byz = 0.125;
    const double bzx = 0.125 * (1.0 / 1.0) * (1.0 / 1.0);
    const double bzy = 0.125 * (1.0 / (1.0 * cp.dtheta)) * (1.0 / (1.0 * cp.dtheta));
    const double dx = 0.0, dy = 0.0;
    const double sn = std::sin(0.5 * pi * ph.c);
    const double dz = 0.25 * (1.0 - (ph.cinv * ph.cinv) * (sn * sn));
    double *L = cp.lehe;
    L[0] = 1.0 - 2.0 * bxy - 2.0 * bxz - 3.0 * dx;
    L[1] = 1.0 - 2.0 * byx - 2.0 * byz - 3.0 * dy;
    L[2] = 1.0 - 2.0 * bzx - 2.0 * bzy - 3.0 * dz;
    L[3] = bxy; L[4] = bxz; L[5] = byx; L[6] = byz; L[7] = bzx; L[8] = bzy;
    L[9] = dx; L[10] = dy; L[11] = dz;
    cp.lehe_on = (fdtd == PICTOR_FDTD_LEHE && g.ndim == 3) ? 1 : 0;
    cp.substep_limit = 100000;
}

namespace {

inline int nblocks(long long n, int bs) { return (int)std::max<long long>(1, (n + bs - 1) / bs); }

__device__ inline long long cx(const GridP &g, int i, int j, int k) { return fidx(g, i, j, k); }

// j + ny/2 folded into the physical range (the grid point across the axis)
__device__ inline int jpp_of(const GridP &g, int j)
{
    int jpp = j + g.ny / 2;
    if (jpp > g.my - 3) jpp = jpp - g.ny;
    return jpp;
}

// ---------------- field updates ----------------

// UpdateEfield, cyl_em_update.F90:7-46 (i from 4 on the axis rank)
template <bool THREE>
__global__ void k_cyl_update_e(GridP g, PhysP ph, CylP cp, double *__restrict__ Ex, double *__restrict__ Ey,
                               double *__restrict__ Ez, const double *__restrict__ Bx, const double *__restrict__ By,
                               const double *__restrict__ Bz)
{
    const int ni = g.mx - 6, nj = g.my - 5, nk = THREE ? g.mz - 5 : 1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int j = 3 + (int)(t % nj);
    const int k = THREE ? 3 + (int)(t / nj) : 1;
    const double fldc = ph.fldc, dtheta = cp.dtheta, gdx = 1.0, gidx = ph.inv_dx, gidz = ph.inv_dz;
    const double r = ((i - 3.0) + cp.rmin) * gdx;
    const double rp_half = r + 0.5 * gdx;
    const double rm_half = r - 0.5 * gdx;
    const long long o = cx(g, i, j, k), sj = g.mx, sk = (long long)g.mx * g.my;
    if (THREE) {
        Ex[o] = Ex[o] + fldc * (Bz[o] - Bz[o - sj]) / (rp_half * dtheta) - fldc * gidz * (By[o] - By[o - sk]);
        Ey[o] = Ey[o] + fldc * gidz * (Bx[o] - Bx[o - sk]) - fldc * gidx * (Bz[o] - Bz[o - 1]);
        Ez[o] = Ez[o] + fldc * gidx * (rp_half * By[o] - rm_half * By[o - 1]) / r - fldc * (Bx[o] - Bx[o - sj]) / (r * dtheta);
    } else {
        Ex[o] = Ex[o] + fldc * (Bz[o] - Bz[o - sj]) / (rp_half * dtheta);
        Ey[o] = Ey[o] - fldc * gidx * (Bz[o] - Bz[o - 1]);
        Ez[o] = Ez[o] + fldc * (gidx * (rp_half * By[o] - rm_half * By[o - 1]) - (Bx[o] - Bx[o - sj]) / dtheta) / r;
    }
}

// UpdateEFldAxis, cyl_em_update.F90:185-225
__global__ void k_cyl_e_axis(GridP g, double *__restrict__ Ex, double *__restrict__ Ey, double *__restrict__ Ez)
{
    const int nk = g.ndim == 3 ? g.mz : 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.my * nk) return;
    const int j = 1 + (int)(t % g.my), k = 1 + (int)(t / g.my);
    const int jpp = jpp_of(g, j);
    Ex[cx(g, 3, j, k)] = 0.5 * (Ex[cx(g, 4, j, k)] - Ex[cx(g, 4, jpp, k)]);
    Ex[cx(g, 2, j, k)] = -Ex[cx(g, 4, jpp, k)];
    Ey[cx(g, 3, j, k)] = -Ey[cx(g, 4, jpp, k)];
    Ez[cx(g, 3, j, k)] = Ez[cx(g, 4, jpp, k)];
}

// gradX, gradX1, gradY, gradZ, cyl_em_update_lehe.F90:255-289 (the missing
// continuation terms have zero coefficients)
__device__ inline double lgx(const CylP &cp, const double *F, long long o, long long sk)
{
    const double *L = cp.lehe;
    return L[0] * (F[o + 1] - F[o]) + L[4] * (F[o + 1 + sk] - F[o + sk]) + L[4] * (F[o + 1 - sk] - F[o - sk]);
}
__device__ inline double lgx1(const CylP &cp, const double *F, long long o, long long sk, double rp, double r)
{
    const double *L = cp.lehe;
    return L[0] * (rp * F[o + 1] - r * F[o]) + L[4] * (rp * F[o + 1 + sk] - r * F[o + sk]) +
           L[4] * (rp * F[o + 1 - sk] - r * F[o - sk]);
}
__device__ inline double lgy(const CylP &cp, const PhysP &ph, const double *F, long long o, long long sj, long long sk)
{
    const double *L = cp.lehe;
    const double gidy = ph.inv_dy;
    return L[1] * gidy * (F[o + sj] - F[o]) + L[6] * gidy * (F[o + sj + sk] - F[o + sk]) +
           L[6] * gidy * (F[o + sj - sk] - F[o - sk]);
}
__device__ inline double lgz(const CylP &cp, const double *F, long long o, long long sj, long long sk)
{
    const double *L = cp.lehe;
    return L[2] * (F[o + sk] - F[o]) + L[7] * (F[o + 1 + sk] - F[o + 1]) + L[7] * (F[o - 1 + sk] - F[o - 1]) +
           L[8] * (F[o + sj + sk] - F[o + sj]) + L[8] * (F[o - sj + sk] - F[o - sj]) + L[11] * (F[o + 2 * sk] - F[o - sk]);
}

// UpdateBfieldHalf, cyl_em_update.F90:49-110 / cyl_em_update_lehe.F90:62-118
template <bool THREE, bool LEHE>
__global__ void k_cyl_update_b_half(GridP g, PhysP ph, CylP cp, double *__restrict__ Bx, double *__restrict__ By,
                                    double *__restrict__ Bz, const double *__restrict__ Ex,
                                    const double *__restrict__ Ey, const double *__restrict__ Ez)
{
    const int ni = g.mx - 4, nj = g.my - 1, nk = THREE ? g.mz - 1 : 1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int j = 1 + (int)(t % nj);
    const int k = 1 + (int)(t / nj);
    if (LEHE && (j < 2 || k < 2 || k > g.mz - 2)) return;   // gradZ reaches j - 1, k - 1, k + 2
    const double h = ph.fld_halfc, dtheta = cp.dtheta, gdx = 1.0, gidx = ph.inv_dx, gidz = ph.inv_dz;
    const double r = ((i - 3.0) + cp.rmin) * gdx;
    const double rp_half = r + 0.5 * gdx;
    const double rp = r + gdx;
    const long long o = cx(g, i, j, k), sj = g.mx, sk = (long long)g.mx * g.my;
    if (LEHE) {
        Bx[o] = Bx[o] - h * (lgy(cp, ph, Ez, o, sj, sk)) / (r * dtheta) + h * gidz * (lgz(cp, Ey, o, sj, sk));
        By[o] = By[o] - h * gidz * (lgz(cp, Ex, o, sj, sk)) + h * gidx * (lgx(cp, Ez, o, sk));
        Bz[o] = Bz[o] - h * (gidx * lgx1(cp, Ey, o, sk, rp, r) - lgy(cp, ph, Ex, o, sj, sk) / dtheta) / rp_half;
    } else if (THREE) {
        Bx[o] = Bx[o] - h * (Ez[o + sj] - Ez[o]) / (r * dtheta) + h * gidz * (Ey[o + sk] - Ey[o]);
        By[o] = By[o] - h * gidz * (Ex[o + sk] - Ex[o]) + h * gidx * (Ez[o + 1] - Ez[o]);
        Bz[o] = Bz[o] - h * (gidx * (rp * Ey[o + 1] - r * Ey[o]) - (Ex[o + sj] - Ex[o]) / dtheta) / rp_half;
    } else {
        Bx[o] = Bx[o] - h * (Ez[o + sj] - Ez[o]) / (r * dtheta);
        By[o] = By[o] + h * gidx * (Ez[o + 1] - Ez[o]);
        Bz[o] = Bz[o] - h * (gidx * (rp * Ey[o + 1] - r * Ey[o]) - (Ex[o + sj] - Ex[o]) / dtheta) / rp_half;
    }
}

// UpdateBzAxis, cyl_em_update.F90:262-280: one thread per z plane, the
// theta sum in the reference order; then Bz(3, 1..my-1, k) = Bz_ax(k)
__global__ void k_cyl_bz_axis(GridP g, PhysP ph, CylP cp, double *__restrict__ Bz, const double *__restrict__ Ey)
{
    const int nk = g.ndim == 3 ? g.mz - 1 : 1;
    const int k = 1 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k > nk) return;
    double ax = 0.0;
    for (int j = 3; j <= g.my - 3; j++) ax = ax - ph.fld_halfc * cp.ax_perm_area * Ey[cx(g, 4, j, k)];
    ax = Bz[cx(g, 3, 3, k)] + ax;
    for (int j = 1; j <= g.my - 1; j++) Bz[cx(g, 3, j, k)] = ax;
}

// UpdateBFldAxis, cyl_em_update.F90:227-260
__global__ void k_cyl_b_axis(GridP g, double *__restrict__ Bx, double *__restrict__ By, double *__restrict__ Bz)
{
    const int nk = g.ndim == 3 ? g.mz : 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.my * nk) return;
    const int j = 1 + (int)(t % g.my), k = 1 + (int)(t / g.my);
    const int jpp = jpp_of(g, j);
    By[cx(g, 3, j, k)] = 0.5 * (By[cx(g, 4, j, k)] - By[cx(g, 4, jpp, k)]);
    By[cx(g, 2, j, k)] = -By[cx(g, 4, jpp, k)];
    Bx[cx(g, 3, j, k)] = -Bx[cx(g, 4, jpp, k)];
    Bz[cx(g, 2, j, k)] = Bz[cx(g, 4, jpp, k)];
}

// AddCurrent, cyl_em_update.F90:114-140
template <bool THREE>
__global__ void k_cyl_add_current(GridP g, PhysP ph, CylP cp, double *__restrict__ Ex, double *__restrict__ Ey,
                                  double *__restrict__ Ez, const double *__restrict__ Jx, const double *__restrict__ Jy,
                                  const double *__restrict__ Jz)
{
    const int ni = g.mx - 6, nj = g.my - 5, nk = THREE ? g.mz - 5 : 1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * nj * nk) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int j = 3 + (int)(t % nj);
    const int k = THREE ? 3 + (int)(t / nj) : 1;
    const double dtheta = cp.dtheta, gdx = 1.0, gidx = ph.inv_dx, gidz = ph.inv_dz;
    const double r = ((i - 3.0) + cp.rmin) * gdx;
    const double rp_half = r + 0.5 * gdx;
    const long long o = cx(g, i, j, k);
    Ex[o] = Ex[o] - Jx[o] * gidz / (rp_half * dtheta);
    Ey[o] = Ey[o] - Jy[o] * gidx * gidz;
    Ez[o] = Ez[o] - Jz[o] * gidx / (r * dtheta);
}

// AxisCurrentBC, cyl_movdep.F90:451-465
__global__ void k_cyl_axis_current_bc(GridP g, double *__restrict__ Jx, double *__restrict__ Jy,
                                      double *__restrict__ Jz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.my * g.mz) return;
    const int j = 1 + (int)(t % g.my), k = 1 + (int)(t / g.my);
    Jy[cx(g, 4, j, k)] = Jy[cx(g, 4, j, k)] - Jy[cx(g, 3, j, k)];
    Jz[cx(g, 4, j, k)] = Jz[cx(g, 4, j, k)] - Jz[cx(g, 3, j, k)];
    Jx[cx(g, 3, j, k)] = 0.0;
    Jy[cx(g, 2, j, k)] = 0.0;
    Jz[cx(g, 2, j, k)] = 0.0;
}

// ---------------- halos ----------------

// theta (axis 1) or z (axis 2) faces over x = 3..mx-3 and the full other
// range: ADD = 0 copies the periodic images into the 5 ghost layers
// (1..2 <- m-4..m-3, m-2..m <- 3..5); ADD = 1 adds them into the owners
// (3..5 += m-2..m, m-4..m-3 += 1..2), as comm_fld.F90 on one periodic rank
template <bool ADD>
__global__ void k_cyl_face(GridP g, int axis, double *__restrict__ F0, double *__restrict__ F1,
                           double *__restrict__ F2)
{
    const int nx = g.mx - 5;
    const int no = axis == 1 ? g.mz : g.my;   // the other transverse extent
    const int m = axis == 1 ? g.my : g.mz;
    const long long per = (long long)nx * 5 * no;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * per) return;
    const int comp = (int)(t / per);
    t %= per;
    const int i = 3 + (int)(t % nx);
    t /= nx;
    const int l = (int)(t % 5);            // which of the 5 layers
    const int oth = 1 + (int)(t / 5);
    int dst, src;
    if (!ADD) {
        dst = l < 2 ? 1 + l : m - 4 + l;    // 1, 2, m-2, m-1, m
        src = l < 2 ? dst + (m - 5) : dst - (m - 5);
    } else {
        dst = l < 3 ? 3 + l : m - 7 + l;    // 3, 4, 5, m-4, m-3
        src = l < 3 ? dst + (m - 5) : dst - (m - 5);
    }
    double *F = comp == 0 ? F0 : (comp == 1 ? F1 : F2);
    long long od, os;
    if (axis == 1) { od = cx(g, i, dst, oth); os = cx(g, i, src, oth); }
    else { od = cx(g, i, oth, dst); os = cx(g, i, oth, src); }
    if (ADD) F[od] = F[od] + F[os];
    else F[od] = F[os];
}

// the conducting outer wall: x ghosts mx-2..mx held at zero
__global__ void k_cyl_outer_zero(GridP g, double *__restrict__ F0, double *__restrict__ F1, double *__restrict__ F2)
{
    const long long per = 3LL * g.my * g.mz;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * per) return;
    const int comp = (int)(t / per);
    const long long r = t % per;
    const int i = g.mx - 2 + (int)(r % 3);
    const long long jk = r / 3;
    const int j = 1 + (int)(jk % g.my), k = 1 + (int)(jk / g.my);
    double *F = comp == 0 ? F0 : (comp == 1 ? F1 : F2);
    F[cx(g, i, j, k)] = 0.0;
}

// ---------------- particles ----------------

// node-averaged E and B at node (i, j, k): the VecE/VecB corner values of
// GatherVecEfld/GatherVecBfld (movdep_routines.F90:6-223), same sums
template <bool THREE>
__device__ inline void node_eb(const GridP &g, const double *const *E, const double *const *B, int i, int j, int k,
                               double e[3], double b[3])
{
    const long long o = cx(g, i, j, k), sj = g.mx, sk = (long long)g.mx * g.my;
    e[0] = (E[0][o - 1] + E[0][o]) * 0.5;
    e[1] = (E[1][o - sj] + E[1][o]) * 0.5;
    if (THREE) {
        e[2] = (E[2][o - sk] + E[2][o]) * 0.5;
        b[0] = (B[0][o - sj - sk] + B[0][o - sk] + B[0][o - sj] + B[0][o]) * 0.25;
        b[1] = (B[1][o - 1 - sk] + B[1][o - 1] + B[1][o - sk] + B[1][o]) * 0.25;
    } else {
        e[2] = E[2][o];
        b[0] = (B[0][o - sj] + B[0][o]) * 0.5;
        b[1] = (B[1][o - 1] + B[1][o]) * 0.5;
    }
    b[2] = (B[2][o - 1 - sj] + B[2][o - 1] + B[2][o - sj] + B[2][o]) * 0.25;
}

// Interp_VecEM_GridPoints (interpolation.F90:434-490)
template <bool THREE>
__device__ inline void cyl_gather(const GridP &g, const double *const *E, const double *const *B, double x, double y,
                                  double z, double pe[3], double pb[3])
{
    const int ip0 = (int)x, jp0 = (int)y, kp0 = THREE ? (int)z : 1;
    const double dx = x - ip0, dy = y - jp0, dz = THREE ? z - kp0 : 0.0;
    for (int c = 0; c < 3; c++) { pe[c] = 0.0; pb[c] = 0.0; }
    constexpr int NV = THREE ? 8 : 4;
#pragma unroll
    for (int nn = 0; nn < NV; nn++) {
        const int a = nn & 1, b = (nn >> 1) & 1, c = nn >> 2;
        double wt = (a ? dx : 1.0 - dx) * (b ? dy : 1.0 - dy);
        if (THREE) wt = wt * (c ? dz : 1.0 - dz);
        double e[3], bb[3];
        node_eb<THREE>(g, E, B, ip0 + a, jp0 + b, kp0 + c, e, bb);
        pe[0] = pe[0] + wt * e[0]; pe[1] = pe[1] + wt * e[1]; pe[2] = pe[2] + wt * e[2];
        pb[0] = pb[0] + wt * bb[0]; pb[1] = pb[1] + wt * bb[1]; pb[2] = pb[2] + wt * bb[2];
    }
}

// weight tables of vars.F90:313-347 (deposit component mapping per VecJ slot)
__constant__ double cFx2[8] = {1, 1, 0, 0, 0, 0, 0, 0}, cFy2[8] = {0, 0, 1, 1, 0, 0, 0, 0},
                    cFz2[8] = {0, 0, 0, 0, 1, 1, 1, 1};
__constant__ double cJx1_2[8] = {1, 1, 1, 0, 1, 0, 1, 0}, cJx2_2[8] = {0, 0, -1, 1, -1, 1, -1, 1};
__constant__ double cJy1_2[8] = {1, 0, 1, 1, 1, 1, 0, 0}, cJy2_2[8] = {-1, 1, 0, 0, -1, -1, 1, 1};
__constant__ double cJz1_2[8] = {1, 1, 1, 1, 1, 1, 1, 1}, cJz2_2[8] = {-1, -1, -1, -1, 0, 0, 0, 0};
__constant__ double cFx3[12] = {1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0}, cFy3[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0},
                    cFz3[12] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
__constant__ double cJx1_3[12] = {1, 1, 1, 1, 1, 0, 1, 0, 1, 0, 1, 0}, cJx2_3[12] = {0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
__constant__ double cJy1_3[12] = {1, 0, 1, 0, 1, 1, 1, 1, 1, 1, 0, 0}, cJy2_3[12] = {-1, 1, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1};
__constant__ double cJz1_3[12] = {1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 1, 1}, cJz2_3[12] = {-1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0};

// one zigzag half-segment into the J nodes of cell (i, j, k): the VecJ slots
// of movdep.F90:187-234 scattered as ReduceCurrentMatrix does (movdep_routines.F90:243-311)
template <bool THREE>
__device__ inline void vecj_half(const GridP &g, double *const *J, int i, int j, int k, double Fx, double Fy,
                                 double Fz, double Wx, double Wy, double Wz)
{
    const long long o = cx(g, i, j, k), sj = g.mx, sk = (long long)g.mx * g.my;
    if (THREE) {
        const long long node[12] = {o, o + sj, o + sk, o + sj + sk, o, o + 1, o + sk, o + 1 + sk, o, o + 1, o + sj, o + 1 + sj};
#pragma unroll
        for (int nn = 0; nn < 12; nn++) {
            const double v = (cFx3[nn] * Fx + cFy3[nn] * Fy + cFz3[nn] * Fz) * (cJx1_3[nn] + cJx2_3[nn] * Wx) *
                             (cJy1_3[nn] + cJy2_3[nn] * Wy) * (cJz1_3[nn] + cJz2_3[nn] * Wz);
            unsafeAtomicAdd(&J[nn >> 2][node[nn]], v);
        }
    } else {
        const long long node[8] = {o, o + sj, o, o + 1, o, o + 1, o + sj, o + 1 + sj};
        const int comp[8] = {0, 0, 1, 1, 2, 2, 2, 2};
#pragma unroll
        for (int nn = 0; nn < 8; nn++) {
            const double v = (cFx2[nn] * Fx + cFy2[nn] * Fy + cFz2[nn] * Fz) * (cJx1_2[nn] + cJx2_2[nn] * Wx) *
                             (cJy1_2[nn] + cJy2_2[nn] * Wy) * (cJz1_2[nn] + cJz2_2[nn] * Wz);
            unsafeAtomicAdd(&J[comp[nn]][node[nn]], v);
        }
    }
}

// DepositCurrentPIC, deposit.F90:11-103
template <bool THREE>
__device__ inline void deposit_pic(const GridP &g, double qi, double *const *J, double x0, double y0, double z0,
                                   double x, double y, double z, double q)
{
    const double qthis = q * qi;
    int i1 = (int)floor(x0), i2 = (int)floor(x);
    int j1 = (int)floor(y0), j2 = (int)floor(y);
    int k1 = (int)floor(z0), k2 = (int)floor(z);
    const double xr = fmin((double)(min(i1, i2) + 1), fmax((double)max(i1, i2), 0.5 * (x0 + x)));
    const double yr = fmin((double)(min(j1, j2) + 1), fmax((double)max(j1, j2), 0.5 * (y0 + y)));
    const double zr = fmin((double)(min(k1, k2) + 1), fmax((double)max(k1, k2), 0.5 * (z0 + z)));
    const double Fx1 = qthis * (xr - x0), Fy1 = qthis * (yr - y0), Fz1 = qthis * (zr - z0);
    const double Wx1 = 0.5 * (x0 + xr) - i1, Wy1 = 0.5 * (y0 + yr) - j1;
    const double Wx2 = 0.5 * (x + xr) - i2, Wy2 = 0.5 * (y + yr) - j2;
    double Wz1, Wz2;
    if (!THREE) { k1 = 1; k2 = 1; Wz1 = 0.0; Wz2 = 0.0; }
    else { Wz1 = 0.5 * (z0 + zr) - k1; Wz2 = 0.5 * (z + zr) - k2; }
    const double Fx2 = qthis * (x - xr), Fy2 = qthis * (y - yr), Fz2 = qthis * (z - zr);
    const long long sj = g.mx, sk = (long long)g.mx * g.my;
    const long long a = cx(g, i1, j1, k1), b = cx(g, i2, j2, k2);
    double *Jx = J[0], *Jy = J[1], *Jz = J[2];
    unsafeAtomicAdd(&Jx[a], Fx1 * (1.0 - Wy1) * (1.0 - Wz1));
    unsafeAtomicAdd(&Jx[a + sj], Fx1 * Wy1 * (1.0 - Wz1));
    if (THREE) {
        unsafeAtomicAdd(&Jx[a + sk], Fx1 * (1.0 - Wy1) * Wz1);
        unsafeAtomicAdd(&Jx[a + sj + sk], Fx1 * Wy1 * Wz1);
    }
    unsafeAtomicAdd(&Jx[b], Fx2 * (1.0 - Wy2) * (1.0 - Wz2));
    unsafeAtomicAdd(&Jx[b + sj], Fx2 * Wy2 * (1.0 - Wz2));
    if (THREE) {
        unsafeAtomicAdd(&Jx[b + sk], Fx2 * (1.0 - Wy2) * Wz2);
        unsafeAtomicAdd(&Jx[b + sj + sk], Fx2 * Wy2 * Wz2);
    }
    unsafeAtomicAdd(&Jy[a], Fy1 * (1.0 - Wx1) * (1.0 - Wz1));
    unsafeAtomicAdd(&Jy[a + 1], Fy1 * Wx1 * (1.0 - Wz1));
    if (THREE) {
        unsafeAtomicAdd(&Jy[a + sk], Fy1 * (1.0 - Wx1) * Wz1);
        unsafeAtomicAdd(&Jy[a + 1 + sk], Fy1 * Wx1 * Wz1);
    }
    unsafeAtomicAdd(&Jy[b], Fy2 * (1.0 - Wx2) * (1.0 - Wz2));
    unsafeAtomicAdd(&Jy[b + 1], Fy2 * Wx2 * (1.0 - Wz2));
    if (THREE) {
        unsafeAtomicAdd(&Jy[b + sk], Fy2 * (1.0 - Wx2) * Wz2);
        unsafeAtomicAdd(&Jy[b + 1 + sk], Fy2 * Wx2 * Wz2);
    }
    unsafeAtomicAdd(&Jz[a], Fz1 * (1.0 - Wx1) * (1.0 - Wy1));
    unsafeAtomicAdd(&Jz[a + 1], Fz1 * Wx1 * (1.0 - Wy1));
    unsafeAtomicAdd(&Jz[a + sj], Fz1 * (1.0 - Wx1) * Wy1);
    unsafeAtomicAdd(&Jz[a + 1 + sj], Fz1 * Wx1 * Wy1);
    unsafeAtomicAdd(&Jz[b], Fz2 * (1.0 - Wx2) * (1.0 - Wy2));
    unsafeAtomicAdd(&Jz[b + 1], Fz2 * Wx2 * (1.0 - Wy2));
    unsafeAtomicAdd(&Jz[b + sj], Fz2 * (1.0 - Wx2) * Wy2);
    unsafeAtomicAdd(&Jz[b + 1 + sj], Fz2 * Wx2 * Wy2);
}

// MovePrtlAxis, cyl_movdep.F90:381-446: one straight sub-step of length dt
// from (x, y, z) with the particle's momenta; false when the azimuthal move
// exceeds 0.45 cells (nothing is changed then)
// Boris with half-kick factors pE, pB (times q/m and dt) on the c-scaled momentum, in
// place (cyl_movdep_strang.F90:172-185); returns the 1 / (c gamma) that scaled pB
__device__ inline double boris_cscaled(const PhysP &ph, double pEx, double pEy, double pEz, double pBx, double pBy,
                                       double pBz, double &u0, double &v0, double &w0)
{
    const double sqc = ph.sqc;
    const double U = u0 + pEx, V = v0 + pEy, W = w0 + pEz;
    const double gg = 1.0 / sqrt(sqc + U * U + V * V + W * W);
    pBx = gg * pBx; pBy = gg * pBy; pBz = gg * pBz;
    const double f = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
    const double u1 = (U + V * pBz - W * pBy) * f;
    const double v1 = (V + W * pBx - U * pBz) * f;
    const double w1 = (W + U * pBy - V * pBx) * f;
    u0 = U + v1 * pBz - w1 * pBy + pEx;
    v0 = V + w1 * pBx - u1 * pBz + pEy;
    w0 = W + u1 * pBy - v1 * pBx + pEz;
    return gg;
}

// MovePrtlAxis of cyl_movdep_strang.F90:462-566 (qm = q/m / 4): half kick with the fields
// at the start, the straight move of length dt, the rotation of (u, v) into the new
// local frame, half kick with the fields at the new point
template <bool THREE>
__device__ inline bool move_axis_strang(const GridP &g, const PhysP &ph, const CylP &cp, const double *const *E,
                                        const double *const *B, double qm, double x, double y, double z, double dt,
                                        double &px, double &py, double &pz, double &pu, double &pv, double &pw)
{
    double e[3], b[3];
    cyl_gather<THREE>(g, E, B, x, y, z, e, b);
    const double c = ph.c, sqc = ph.sqc, cinv = ph.cinv;
    double u0 = c * pu, v0 = c * pv, w0 = c * pw;
    boris_cscaled(ph, e[0] * qm * dt, e[1] * qm * dt, e[2] * qm * dt, b[0] * qm * dt, b[1] * qm * dt, b[2] * qm * dt,
                  u0, v0, w0);
    double u1 = u0 * cinv, v1 = v0 * cinv;
    const double w1 = w0 * cinv;
    const double gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
    const double r0 = (x - 3.5) * 1.0;
    const double x1 = r0 + u1 * gg * dt;
    const double y1 = v1 * gg * dt;
    const double r1 = sqrt(x1 * x1 + y1 * y1);
    double dy;
    if (x1 > 0) dy = atan(y1 / x1) * cp.inv_dtheta;
    else dy = 0.0;
    if (fabs(dy) > 0.45) return false;
    double nx;
    if (x1 > 0) {
        nx = r1 / ph.inv_dx + 3.5;
        u1 = (r0 * u0 + (u0 * u1 * gg * dt + v0 * y1)) / r1;
        v1 = (v0 * r0) / r1;
    } else {
        nx = x1 / ph.inv_dx + 3.5;
        u1 = u0;
        v1 = v0;
    }
    px = nx;
    py = py + dy;
    pz = pz + w1 * gg * dt * ph.inv_dz;
    cyl_gather<THREE>(g, E, B, px, py, pz, e, b);
    u0 = u1;
    v0 = v1;
    boris_cscaled(ph, e[0] * qm * dt, e[1] * qm * dt, e[2] * qm * dt, b[0] * qm * dt, b[1] * qm * dt, b[2] * qm * dt,
                  u0, v0, w0);
    pu = u0 * cinv;
    pv = v0 * cinv;
    pw = w0 * cinv;
    return true;
}

template <bool THREE>
__device__ inline bool move_axis(const GridP &g, const PhysP &ph, const CylP &cp, const double *const *E,
                                 const double *const *B, double qm, double x, double y, double z, double dt,
                                 double &px, double &py, double &pz, double &pu, double &pv, double &pw)
{
    double e[3], b[3];
    cyl_gather<THREE>(g, E, B, x, y, z, e, b);
    double pEx = e[0] * qm * dt, pEy = e[1] * qm * dt, pEz = e[2] * qm * dt;
    double pBx = b[0] * qm * dt, pBy = b[1] * qm * dt, pBz = b[2] * qm * dt;
    const double c = ph.c, sqc = ph.sqc, cinv = ph.cinv;
    double u0 = c * pu + pEx;
    double v0 = c * pv + pEy;
    double w0 = c * pw + pEz;
    double gg = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
    pBx = gg * pBx; pBy = gg * pBy; pBz = gg * pBz;
    const double f = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
    double u1 = (u0 + v0 * pBz - w0 * pBy) * f;
    double v1 = (v0 + w0 * pBx - u0 * pBz) * f;
    const double w1 = (w0 + u0 * pBy - v0 * pBx) * f;
    u0 = u0 + v1 * pBz - w1 * pBy + pEx;
    v0 = v0 + w1 * pBx - u1 * pBz + pEy;
    w0 = w0 + u1 * pBy - v1 * pBx + pEz;
    u1 = u0 * cinv;
    v1 = v0 * cinv;
    gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
    const double x1 = (x - 3.5) * 1.0 + u1 * gg * dt;
    const double y1 = v1 * gg * dt;
    const double r1 = sqrt(x1 * x1 + y1 * y1);
    double dy;
    if (x1 > 0) dy = atan(y1 / x1) * cp.inv_dtheta;
    else dy = 0.0;
    if (fabs(dy) > 0.45) return false;
    if (x1 > 0) {
        px = r1 * ph.inv_dx + 3.5;
        pu = (x1 * u1 + y1 * v1) / r1;
        pv = (-y1 * u1 + x1 * v1) / r1;
    } else {
        px = x1 * ph.inv_dx + 3.5;
        pu = u1;
        pv = v1;
    }
    py = py + dy;
    pw = w0 * cinv;
    pz = pz + pw * gg * dt * ph.inv_dz;
    return true;
}

// MoveDepositPrtl, cyl_movdep.F90:14-289 (STRANG: cyl_movdep_strang.F90:95-294, the
// push of the reference's CPU -Dcyl build), one thread per particle.  The second
// zigzag half-segment is deposited at the particle's new cell (the reference indexes it
// with the old cell, cyl_movdep.F90:249,252, which breaks charge conservation).
template <bool THREE, bool STRANG>
__global__ __launch_bounds__(256) void k_cyl_push(GridP g, PhysP ph, CylP cp, const double *__restrict__ Ex,
                                                  const double *__restrict__ Ey, const double *__restrict__ Ez,
                                                  const double *__restrict__ Bx, const double *__restrict__ By,
                                                  const double *__restrict__ Bz, double *__restrict__ Jx,
                                                  double *__restrict__ Jy, double *__restrict__ Jz, PartF64 P,
                                                  long long n, int *__restrict__ err,
                                                  unsigned long long *__restrict__ nsub)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int flv = P.flv[m];
    if (flv == 0) return;
    const double *E[3] = {Ex, Ey, Ez}, *B[3] = {Bx, By, Bz};
    double *J[3] = {Jx, Jy, Jz};
    const double x0 = P.x[m], y0 = P.y[m], z0 = P.z[m], q = P.q[m];
    const double ui = P.u[m], vi = P.v[m], wi = P.w[m];
    const double qm = ph.flvrqm[flv - 1] * (STRANG ? 0.25 : 0.5);
    const double c = ph.c, sqc = ph.sqc, cinv = ph.cinv;
    double xn, yn, zn, un, vn, wn;
    if (x0 < 4.5) {
        // MovDepAxis / SubStepMovDep, cyl_movdep.F90:300-349
        xn = x0; yn = y0; zn = z0; un = ui; vn = vi; wn = wi;
        double sx = x0, sy = y0, sz = z0;
        double t = 0.0, dt = 1.0;
        int tries = 0;
        while (t < 1.0) {
            if (++tries > cp.substep_limit) { atomicOr(err, (int)kCylSubsteps); break; }
            double delt = dt;
            if (t + dt > 1.0) delt = 1.0 - t;
            const bool ok = STRANG ? move_axis_strang<THREE>(g, ph, cp, E, B, qm, sx, sy, sz, delt, xn, yn, zn, un, vn, wn)
                                   : move_axis<THREE>(g, ph, cp, E, B, qm, sx, sy, sz, delt, xn, yn, zn, un, vn, wn);
            if (ok) {
                t = t + delt;
                deposit_pic<THREE>(g, ph.qi, J, sx, sy, sz, xn, yn, zn, q);
                // PrtlAxisBC, cyl_movdep.F90:366-379
                if (xn < 3.5) {
                    xn = 3.5 + (3.5 - xn);
                    yn = yn + g.ny / 2.0;
                    un = -un;
                    vn = -vn;
                }
                if (yn > ph.ymax) yn = yn - ph.ylen;
                if (yn < ph.ymin) yn = yn + ph.ylen;
                dt = dt * (STRANG ? 1.7 : 1.1);   // cyl_movdep_strang.F90:427
                sx = xn; sy = yn; sz = zn;
            } else {
                dt = dt * 0.5;
            }
        }
        atomicAdd(nsub, (unsigned long long)tries);
    } else {
      if (STRANG) {
        // cyl_movdep_strang.F90:95-294 with the gather of the second half reset and the
        // z move on the half-kicked momentum, as its MovePrtlAxis does (the main loop as
        // written carries the first half's scaled fields into the second gather's sums
        // and moves z with the momentum before the kick)
        double e[3], b[3];
        cyl_gather<THREE>(g, E, B, x0, y0, z0, e, b);
        double u0 = c * ui, v0 = c * vi, w0 = c * wi;
        boris_cscaled(ph, e[0] * qm, e[1] * qm, e[2] * qm, b[0] * qm, b[1] * qm, b[2] * qm, u0, v0, w0);
        double u1 = u0 * cinv;
        const double v1 = v0 * cinv, w1 = w0 * cinv;
        const double gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        const double r0 = (x0 + cp.r_proc) * 1.0;
        const double x1 = r0 + u1 * gg;
        const double y1 = v1 * gg;
        double r1 = sqrt(x1 * x1 + y1 * y1);
        xn = r1 * ph.inv_dx - cp.r_proc;
        yn = y0 + atan(y1 / x1) * cp.inv_dtheta;
        zn = z0 + w1 * gg * ph.inv_dz;
        r1 = 1.0 / r1;
        u1 = (r0 * u0 + (u0 * u1 * gg + v0 * y1)) * r1;
        const double v2 = (v0 * r0) * r1;
        cyl_gather<THREE>(g, E, B, xn, yn, zn, e, b);
        u0 = u1;
        v0 = v2;
        boris_cscaled(ph, e[0] * qm, e[1] * qm, e[2] * qm, b[0] * qm, b[1] * qm, b[2] * qm, u0, v0, w0);
        un = u0 * cinv;
        vn = v0 * cinv;
        wn = w0 * cinv;
      } else {
        double pe[3], pb[3];
        cyl_gather<THREE>(g, E, B, x0, y0, z0, pe, pb);
        const double pEx = pe[0] * qm, pEy = pe[1] * qm, pEz = pe[2] * qm;
        double pBx = pb[0] * qm, pBy = pb[1] * qm, pBz = pb[2] * qm;
        double u0 = c * ui + pEx;
        double v0 = c * vi + pEy;
        double w0 = c * wi + pEz;
        double gg = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        pBx = gg * pBx; pBy = gg * pBy; pBz = gg * pBz;
        const double ff = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
        double u1 = (u0 + v0 * pBz - w0 * pBy) * ff;
        double v1 = (v0 + w0 * pBx - u0 * pBz) * ff;
        const double w1 = (w0 + u0 * pBy - v0 * pBx) * ff;
        u0 = u0 + v1 * pBz - w1 * pBy + pEx;
        v0 = v0 + w1 * pBx - u1 * pBz + pEy;
        w0 = w0 + u1 * pBy - v1 * pBx + pEz;
        u1 = u0 * cinv;
        v1 = v0 * cinv;
        wn = w0 * cinv;
        gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        const double x1 = (x0 + cp.r_proc) * 1.0 + u1 * gg;
        const double y1 = v1 * gg;
        double r1 = sqrt(x1 * x1 + y1 * y1);
        xn = r1 * ph.inv_dx - cp.r_proc;
        yn = y0 + atan(y1 / x1) * cp.inv_dtheta;
        zn = z0 + wn * gg * ph.inv_dz;
        r1 = 1.0 / r1;
        un = (x1 * u1 + y1 * v1) * r1;
        vn = (-y1 * u1 + x1 * v1) * r1;
      }
        // zigzag deposit, cyl_movdep.F90:215-268
        const int ip0 = (int)x0, jp0 = (int)y0, kp0 = THREE ? (int)z0 : 1;
        const int ip = (int)xn, jp = (int)yn, kp = THREE ? (int)zn : 1;
        const double qthis = q * ph.qi;
        const double xr = fmin((double)(min(ip0, ip) + 1), fmax((double)max(ip0, ip), 0.5 * (x0 + xn)));
        const double yr = fmin((double)(min(jp0, jp) + 1), fmax((double)max(jp0, jp), 0.5 * (y0 + yn)));
        const double zr = fmin((double)(min(kp0, kp) + 1), fmax((double)max(kp0, kp), 0.5 * (z0 + zn)));
        // near the axis one step can cross several theta cells (up to ny/4): the
        // second half-segment's cell row is folded by ny into the arrays, a
        // periodic image of the same nodes
        const int jq = jp < 1 ? jp + g.ny : (jp > g.my - 1 ? jp - g.ny : jp);
        if (xn < g.mx - 3) {
            vecj_half<THREE>(g, J, ip0, jp0, kp0, qthis * (xr - x0), qthis * (yr - y0), qthis * (zr - z0),
                             0.5 * (x0 + xr) - ip0, 0.5 * (y0 + yr) - jp0, THREE ? 0.5 * (z0 + zr) - kp0 : 0.0);
            vecj_half<THREE>(g, J, ip, jq, kp, qthis * (xn - xr), qthis * (yn - yr), qthis * (zn - zr),
                             0.5 * (xn + xr) - ip, 0.5 * (yn + yr) - jp, THREE ? 0.5 * (zn + zr) - kp : 0.0);
        }
    }
    if (xn >= g.mx - 3) atomicOr(err, (int)kCylWall);
    if (!THREE) {
        if (zn > ph.zmax) zn = zn - ph.zlen;
        else if (zn < ph.zmin) zn = ph.zlen + zn;
    }
    P.x[m] = xn; P.y[m] = yn; P.z[m] = zn;
    P.u[m] = un; P.v[m] = vn; P.w[m] = wn;
}

// ExchangePrtl on the axis rank: periodic theta and (3-D) z
__global__ void k_cyl_wrap(GridP g, PhysP ph, PartF64 P, long long n)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n || P.flv[m] == 0) return;
    double y = P.y[m];
    if (y < ph.ymin) y = y + ph.ylen;
    else if (y > ph.ymax) y = y - ph.ylen;
    P.y[m] = y;
    if (g.ndim == 3) {
        double z = P.z[m];
        if (z < ph.zmin) z = z + ph.zlen;
        else if (z > ph.zmax) z = z - ph.zlen;
        P.z[m] = z;
    }
}

// CalcEnergy with -Dcyl (savedata_routines.F90:265-281): CalcFldEnergy_cyl
// (cyl/cyl_savedata_routines.F90:8-31: every component weighted by the radius of its
// stagger, r = (i - 3) + rmin or r + 1/2, times dtheta, i from 4 on the axis rank) and
// CalcPrtlEnergy (savedata_routines.F90:283-308).  Per block: B, E, flavour 1, flavour 2.
__global__ __launch_bounds__(256) void k_cyl_energy(GridP g, PhysP ph, CylP cp, const double *__restrict__ Ex,
                                                    const double *__restrict__ Ey, const double *__restrict__ Ez,
                                                    const double *__restrict__ Bx, const double *__restrict__ By,
                                                    const double *__restrict__ Bz, PartF64 P, long long n,
                                                    double *partial)
{
    __shared__ double sred[4][256];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int ni = g.mx - 6, nj = g.my - 5, nk = g.ndim == 3 ? g.mz - 5 : 1, k0 = g.ndim == 3 ? 3 : 1;
    const long long ncell = (long long)ni * nj * nk;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < ncell; t += stride) {
        const int i = 4 + (int)(t % ni);
        const long long r0 = t / ni;
        const int j = 3 + (int)(r0 % nj), k = k0 + (int)(r0 / nj);
        const long long o = cx(g, i, j, k);
        const double r = (i - 3.0) + cp.rmin, rp = r + 0.5;
        acc[0] += (r * (Bx[o] * Bx[o]) + rp * (By[o] * By[o]) + rp * (Bz[o] * Bz[o])) * cp.dtheta;
        acc[1] += (rp * (Ex[o] * Ex[o]) + r * (Ey[o] * Ey[o]) + r * (Ez[o] * Ez[o])) * cp.dtheta;
    }
    const double c2 = ph.c * ph.c;
    for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += stride) {
        const int flv = P.flv[m];
        if (flv < 1 || flv > 2) continue;
        const double u = P.u[m], v = P.v[m], w = P.w[m], q = P.q[m];
        const double gamma = sqrt(1.0 + u * u + v * v + w * w);
        double e = 0.0;
        if (q > 0) e = (gamma - 1.0) * 1.0 * (q * ph.qi / ph.flvrqm[flv - 1]) * c2;
        if (q < 0) e = (gamma - 1.0) * 1.0 * (ph.mass[1] * c2);
        acc[1 + flv] += e;
    }
    for (int a = 0; a < 4; a++) sred[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 4; a++) sred[a][threadIdx.x] += sred[a][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 4; a++) partial[4 * blockIdx.x + a] = sred[a][0];
}

}  // namespace
}  // namespace pictor

// ---------------- host side ----------------

namespace {

void fields(pictor_ctx *c, double *E[3], double *B[3], double *J[3])
{
    for (int i = 0; i < 3; i++) {
        E[i] = (double *)c->E[i];
        B[i] = (double *)c->B[i];
        J[i] = (double *)c->J[i];
    }
}

void halo(pictor_ctx *c, double *F[3])
{
    const GridP &g = c->g;
    const long long nth = 3LL * (g.mx - 5) * 5 * g.mz;
    hipLaunchKernelGGL(k_cyl_face<false>, dim3(nblocks(nth, 256)), dim3(256), 0, c->stream, g, 1, F[0], F[1], F[2]);
    if (g.ndim == 3) {
        const long long nz = 3LL * (g.mx - 5) * 5 * g.my;
        hipLaunchKernelGGL(k_cyl_face<false>, dim3(nblocks(nz, 256)), dim3(256), 0, c->stream, g, 2, F[0], F[1], F[2]);
    }
    const long long no = 9LL * g.my * g.mz;
    hipLaunchKernelGGL(k_cyl_outer_zero, dim3(nblocks(no, 256)), dim3(256), 0, c->stream, g, F[0], F[1], F[2]);
}

}  // namespace

int cyl_setup(pictor_ctx *c)
{
    if (c->g.ny % 2 != 0) return fail(PICTOR_ERR_ARG, "the cylindrical grid needs an even number of theta cells (ny)");
    if (c->form != PICTOR_FORM_FP64) return fail(PICTOR_ERR_ARG, "the cylindrical geometry runs in the FP64 form");
    const int strang = c->cp.strang;   // the push scheme chosen by pictor_cyl_push_scheme survives re-setup
    cyl_params(c->g, c->ph, c->cfg.fdtd, c->cp);
    c->cp.strang = strang;
    HIPTRY(dalloc(&c->d_cerr, 1));
    HIPTRY(dalloc(&c->d_nsub, 1));
    HIPTRY(hipMemsetAsync(c->d_cerr, 0, sizeof(int), c->stream));
    HIPTRY(hipMemsetAsync(c->d_nsub, 0, sizeof(unsigned long long), c->stream));
    return PICTOR_OK;
}

void cyl_release(pictor_ctx *c)
{
    PartF64 &p = c->pcyl;
    dfree(p.q); dfree(p.x); dfree(p.y); dfree(p.z); dfree(p.u); dfree(p.v); dfree(p.w); dfree(p.flv);
    p = PartF64{};
    dfree(c->d_cerr); dfree(c->d_nsub);
    c->d_cerr = nullptr;
    c->d_nsub = nullptr;
}

int cyl_upload_particles(pictor_ctx *c, long long n, const double *q, const double *x, const double *y,
                         const double *z, const double *u, const double *v, const double *w, const int *flv)
{
    // holes are dropped; the live particles keep the caller's order
    std::vector<long long> keep;
    keep.reserve(n);
    const GridP &g = c->g;
    for (long long i = 0; i < n; i++) {
        if (flv[i] == 0) continue;
        if (flv[i] != 1 && flv[i] != 2) return fail(PICTOR_ERR_ARG, "flavour must be 0, 1 or 2");
        if (!(x[i] >= 3.5 && x[i] < g.mx - 3 && y[i] >= 3.0 && y[i] <= g.my - 2))
            return fail(PICTOR_ERR_ARG, "particle outside the cylinder (3.5 <= x < mx-3, 3 <= y <= my-2)");
        if (g.ndim == 3 && !(z[i] >= 3.0 && z[i] <= g.mz - 2)) return fail(PICTOR_ERR_ARG, "particle outside 3 <= z <= mz-2");
        keep.push_back(i);
    }
    const long long m = (long long)keep.size();
    cyl_release(c);
    int rc = cyl_setup(c);
    if (rc) return rc;
    PartF64 &p = c->pcyl;
    double **arr[7] = {&p.q, &p.x, &p.y, &p.z, &p.u, &p.v, &p.w};
    const double *src[7] = {q, x, y, z, u, v, w};
    std::vector<double> buf(std::max<long long>(m, 1));
    for (int a = 0; a < 7; a++) {
        HIPTRY(dalloc(arr[a], m));
        for (long long e = 0; e < m; e++) buf[e] = src[a][keep[e]];
        HIPTRY(hipMemcpy(*arr[a], buf.data(), m * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<int> fb(std::max<long long>(m, 1));
    for (long long e = 0; e < m; e++) fb[e] = flv[keep[e]];
    HIPTRY(dalloc(&p.flv, m));
    HIPTRY(hipMemcpy(p.flv, fb.data(), m * sizeof(int), hipMemcpyHostToDevice));
    c->np = m;
    return PICTOR_OK;
}

int cyl_download_particles(pictor_ctx *c, long long cap, double *q, double *x, double *y, double *z, double *u,
                           double *v, double *w, int *flv, int *tile, long long *n_out)
{
    const long long n = c->np;
    if (n_out) *n_out = n;
    if (n > cap) return fail(PICTOR_ERR_ARG, "output arrays too small");
    HIPTRY(hipStreamSynchronize(c->stream));
    if (n == 0) return PICTOR_OK;
    const PartF64 &p = c->pcyl;
    double *hd[7] = {q, x, y, z, u, v, w};
    double *dd[7] = {p.q, p.x, p.y, p.z, p.u, p.v, p.w};
    for (int i = 0; i < 7; i++)
        if (hd[i]) HIPTRY(hipMemcpy(hd[i], dd[i], n * sizeof(double), hipMemcpyDeviceToHost));
    if (flv) HIPTRY(hipMemcpy(flv, p.flv, n * sizeof(int), hipMemcpyDeviceToHost));
    if (tile) std::fill(tile, tile + n, 0);
    return PICTOR_OK;
}

int cyl_b_half(pictor_ctx *c)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    const long long nb = (long long)(g.mx - 4) * (g.my - 1) * (g.ndim == 3 ? g.mz - 1 : 1);
    const dim3 gb(nblocks(nb, 256)), bs(256);
    if (g.ndim == 3 && c->cp.lehe_on)
        hipLaunchKernelGGL((k_cyl_update_b_half<true, true>), gb, bs, 0, c->stream, g, c->ph, c->cp, B[0], B[1], B[2],
                           E[0], E[1], E[2]);
    else if (g.ndim == 3)
        hipLaunchKernelGGL((k_cyl_update_b_half<true, false>), gb, bs, 0, c->stream, g, c->ph, c->cp, B[0], B[1], B[2],
                           E[0], E[1], E[2]);
    else
        hipLaunchKernelGGL((k_cyl_update_b_half<false, false>), gb, bs, 0, c->stream, g, c->ph, c->cp, B[0], B[1],
                           B[2], E[0], E[1], E[2]);
    const int nk = g.ndim == 3 ? g.mz - 1 : 1;
    hipLaunchKernelGGL(k_cyl_bz_axis, dim3(nblocks(nk, 64)), dim3(64), 0, c->stream, g, c->ph, c->cp, B[2], E[1]);
    const long long na = (long long)g.my * (g.ndim == 3 ? g.mz : 1);
    hipLaunchKernelGGL(k_cyl_b_axis, dim3(nblocks(na, 256)), dim3(256), 0, c->stream, g, B[0], B[1], B[2]);
    halo(c, B);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cyl_e_phase(pictor_ctx *c, bool curl, bool add)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    const long long ne = (long long)(g.mx - 6) * (g.my - 5) * (g.ndim == 3 ? g.mz - 5 : 1);
    const long long na = (long long)g.my * (g.ndim == 3 ? g.mz : 1);
    const dim3 ge(nblocks(ne, 256)), bs(256), ga(nblocks(na, 256));
    if (curl) {
        if (g.ndim == 3)
            hipLaunchKernelGGL(k_cyl_update_e<true>, ge, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], B[0], B[1], B[2]);
        else
            hipLaunchKernelGGL(k_cyl_update_e<false>, ge, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], B[0], B[1], B[2]);
        hipLaunchKernelGGL(k_cyl_e_axis, ga, bs, 0, c->stream, g, E[0], E[1], E[2]);
        halo(c, E);
    }
    if (add) {
        if (g.ndim == 3)
            hipLaunchKernelGGL(k_cyl_add_current<true>, ge, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], J[0],
                               J[1], J[2]);
        else
            hipLaunchKernelGGL(k_cyl_add_current<false>, ge, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], J[0],
                               J[1], J[2]);
        hipLaunchKernelGGL(k_cyl_e_axis, ga, bs, 0, c->stream, g, E[0], E[1], E[2]);
        halo(c, E);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cyl_push(pictor_ctx *c, bool timed)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    const long long n = c->np;
    if (timed) HIPTRY(hipEventRecord(c->ev[0], c->stream));
    if (n > 0) {
        const dim3 gb(nblocks(n, 256)), bs(256);
        if (g.ndim == 3 && c->cp.strang)
            hipLaunchKernelGGL((k_cyl_push<true, true>), gb, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], B[0],
                               B[1], B[2], J[0], J[1], J[2], c->pcyl, n, c->d_cerr, c->d_nsub);
        else if (g.ndim == 3)
            hipLaunchKernelGGL((k_cyl_push<true, false>), gb, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], B[0],
                               B[1], B[2], J[0], J[1], J[2], c->pcyl, n, c->d_cerr, c->d_nsub);
        else if (c->cp.strang)
            hipLaunchKernelGGL((k_cyl_push<false, true>), gb, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], B[0],
                               B[1], B[2], J[0], J[1], J[2], c->pcyl, n, c->d_cerr, c->d_nsub);
        else
            hipLaunchKernelGGL((k_cyl_push<false, false>), gb, bs, 0, c->stream, g, c->ph, c->cp, E[0], E[1], E[2], B[0],
                               B[1], B[2], J[0], J[1], J[2], c->pcyl, n, c->d_cerr, c->d_nsub);
    }
    if (timed) HIPTRY(hipEventRecord(c->ev[1], c->stream));
    const long long nj = (long long)g.my * g.mz;
    hipLaunchKernelGGL(k_cyl_axis_current_bc, dim3(nblocks(nj, 256)), dim3(256), 0, c->stream, g, J[0], J[1], J[2]);
    HIPTRY(hipGetLastError());
    c->pending_swap = true;
    return PICTOR_OK;
}

int cyl_exchange(pictor_ctx *c)
{
    if (c->np > 0)
        hipLaunchKernelGGL(k_cyl_wrap, dim3(nblocks(c->np, 256)), dim3(256), 0, c->stream, c->g, c->ph, c->pcyl, c->np);
    HIPTRY(hipGetLastError());
    int err = 0;
    HIPTRY(hipMemcpyAsync(&err, c->d_cerr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    c->pending_swap = false;
    if (err & kCylWall) return fail(PICTOR_ERR_BOUNDS, "a particle reached the outer wall region (x >= mx - 3)");
    if (err & kCylSubsteps) return fail(PICTOR_ERR_BOUNDS, "the axis sub-steps of a particle did not finish");
    return PICTOR_OK;
}

int cyl_current_halo(pictor_ctx *c)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    const long long nth = 3LL * (g.mx - 5) * 5 * g.mz;
    hipLaunchKernelGGL(k_cyl_face<true>, dim3(nblocks(nth, 256)), dim3(256), 0, c->stream, g, 1, J[0], J[1], J[2]);
    if (g.ndim == 3) {
        const long long nz = 3LL * (g.mx - 5) * 5 * g.my;
        hipLaunchKernelGGL(k_cyl_face<true>, dim3(nblocks(nz, 256)), dim3(256), 0, c->stream, g, 2, J[0], J[1], J[2]);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cyl_field_halo(pictor_ctx *c, int which)
{
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    halo(c, which == 0 ? E : B);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cyl_reset_current(pictor_ctx *c)
{
    for (int i = 0; i < 3; i++) launch_fill_zero<double>((double *)c->J[i], c->g.ncell, c->stream);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cyl_diagnostics(pictor_ctx *c, double out[9])
{
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    const int nb = 1024;
    if (c->partial_len < 4 * nb) return fail(PICTOR_ERR_HIP, "diagnostic buffer too small");
    hipLaunchKernelGGL(k_cyl_energy, dim3(nb), dim3(256), 0, c->stream, c->g, c->ph, c->cp, E[0], E[1], E[2], B[0], B[1],
                       B[2], c->pcyl, c->np, c->d_partial);
    HIPTRY(hipGetLastError());
    std::vector<double> h(4 * nb);
    HIPTRY(hipMemcpyAsync(h.data(), c->d_partial, 4 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < nb; b++)
        for (int k = 0; k < 4; k++) a[k] += h[4 * b + k];
    for (int i = 0; i < 9; i++) out[i] = 0.0;
    out[0] = 0.5 * a[0];
    out[1] = 0.5 * a[1];
    out[2] = a[2];
    out[3] = a[3];
    return PICTOR_OK;
}
