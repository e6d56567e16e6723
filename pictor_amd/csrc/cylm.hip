// cylm.hip -- azimuthal-mode (m = 0, 1) cylindrical geometry, FP64.
//
// The reference's cylindrical build is a full r-theta-z grid (cyl.hip here);
// it has no modal form, no laser and no moving window.  This form keeps that
// grid's r and z staggering, metric factors, axis rules and local-frame push,
// and replaces theta by F = F0 + F1c cos(theta) + F1s sin(theta):
//   * theta differences (F(j) - F(j-1)) / dtheta of cyl_em_update.F90 become
//     the exact derivative, (dF/dtheta)_c = F1s, (dF/dtheta)_s = -F1c;
//   * the axis fills map theta to theta + pi: each mode by its parity (-1)^m;
//     the Bz axis integral acts on m = 0 only;
//   * AddCurrent divides by the circle, 2 pi for m = 0 and pi for m = 1;
//   * the deposit splits the move into the r-z zigzag at the old angle
//     (every slot weighted by its phase) and the angular move at the new
//     (r, z), which keeps each mode's discrete continuity equation exact.
// Arrays (mx, 3, mz): slot 0 = m 0, 1 = m 1 cos, 2 = m 1 sin; element
// (i, s, k), 1-based i and k, at (i - 1) + mx (s + 3 (k - 1)).  Lengths in
// axial cells (dz = 1); the radial cell is dr.  Particles: x = radial grid
// coordinate (r = (x - 3.5) dr), y = theta in radians, z = axial coordinate,
// momenta in the local (r, theta, z) frame; flavour 2 = electrons.
#include <climits>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "ctx.h"

namespace pictor {
namespace {

constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ long long mi(const GridP &g, int i, int s, int k)
{
    return (long long)(i - 1) + (long long)g.mx * (s + 3LL * (k - 1));
}

__device__ __forceinline__ double rad(const CylmP &q, double i) { return (i - 3.5) * q.dr; }

__device__ __forceinline__ double parity(int s) { return s == 0 ? 1.0 : -1.0; }

// ---- the split record: (i, k) cell word, FP32 offsets, theta in 2^-32 turns ----
constexpr double kTurnRad = 2.0 * kPi / 4294967296.0;   // radians per 2^-32 turn
constexpr double kJq = 68719476736.0;                    // 2^36: current counts per charge cell
constexpr double kJqInv = 1.0 / 68719476736.0;

__device__ __host__ inline uint32_t pack_cell(int i, int k, int flv)
{
    return (flv == 1 ? 0x80000000u : 0u) | ((uint32_t)k << 12) | (uint32_t)i;
}
__device__ __host__ inline int cell_i(uint32_t w) { return (int)(w & 0xFFFu); }
__device__ __host__ inline int cell_k(uint32_t w) { return (int)((w >> 12) & 0x7FFFFu); }
__device__ __host__ inline int cell_flv(uint32_t w) { return (w >> 31) ? 1 : 2; }

// a coordinate as (cell, offset in [0, 1)) with an FP32 offset
__device__ __host__ inline void split_coord(double x, int &i, float &o)
{
    i = (int)floor(x);
    o = (float)(x - i);
    if (o >= 1.0f) { i += 1; o = 0.0f; }
}
__device__ __host__ inline uint32_t theta_turns(double th)
{
    return (uint32_t)(unsigned long long)(long long)llrint(th * (4294967296.0 / (2.0 * kPi)));
}


// UpdateEfield (cyl_em_update.F90:7-46, 3-D): i = 4..mx-3, k = 3..mz-3, every slot
__global__ void k_cylm_update_e(GridP g, CylmP q, double *__restrict__ Ex, double *__restrict__ Ey,
                                double *__restrict__ Ez, const double *__restrict__ Bx, const double *__restrict__ By,
                                const double *__restrict__ Bz)
{
    const int ni = g.mx - 6, nk = g.mz - 5;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * 3 * nk) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int s = (int)(t % 3);
    const int k = 3 + (int)(t / 3);
    const long long o = mi(g, i, s, k), sk = 3LL * g.mx, ss = g.mx;
    const double fldc = q.fldc, idr = q.inv_dr, idz = q.inv_dz;
    const double r = rad(q, i), rp = r + 0.5 * q.dr, rm = r - 0.5 * q.dr;
    const double dBz = s == 0 ? 0.0 : (s == 1 ? Bz[o + ss] : -Bz[o - ss]);
    const double dBr = s == 0 ? 0.0 : (s == 1 ? Bx[o + ss] : -Bx[o - ss]);
    Ex[o] = Ex[o] + fldc * dBz / rp - fldc * idz * (By[o] - By[o - sk]);
    Ey[o] = Ey[o] + fldc * idz * (Bx[o] - Bx[o - sk]) - fldc * idr * (Bz[o] - Bz[o - 1]);
    Ez[o] = Ez[o] + fldc * idr * (rp * By[o] - rm * By[o - 1]) / r - fldc * dBr / r;
}

// AddCurrent (cyl_em_update.F90:114-140) over the circle of each mode
__global__ void k_cylm_add_current(GridP g, CylmP q, double *__restrict__ Ex, double *__restrict__ Ey,
                                   double *__restrict__ Ez, const double *__restrict__ Jx,
                                   const double *__restrict__ Jy, const double *__restrict__ Jz)
{
    const int ni = g.mx - 6, nk = g.mz - 5;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * 3 * nk) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int s = (int)(t % 3);
    const int k = 3 + (int)(t / 3);
    const long long o = mi(g, i, s, k);
    const double th = s == 0 ? 2.0 * kPi : kPi;
    const double r = rad(q, i), rp = r + 0.5 * q.dr;
    Ex[o] = Ex[o] - Jx[o] * q.inv_dz / (rp * th);
    Ey[o] = Ey[o] - Jy[o] * q.inv_dr * q.inv_dz / th;
    Ez[o] = Ez[o] - Jz[o] * q.inv_dr / (r * th);
}

// E axis fills (cyl_em_update.F90:143-180), theta + pi as the mode parity
__global__ void k_cylm_e_axis(GridP g, double *__restrict__ Ex, double *__restrict__ Ey, double *__restrict__ Ez)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * g.mz) return;
    const int s = (int)(t % 3), k = 1 + (int)(t / 3);
    const double p = parity(s);
    Ex[mi(g, 3, s, k)] = 0.5 * (Ex[mi(g, 4, s, k)] - p * Ex[mi(g, 4, s, k)]);
    Ex[mi(g, 2, s, k)] = -(p * Ex[mi(g, 4, s, k)]);
    Ey[mi(g, 3, s, k)] = -(p * Ey[mi(g, 4, s, k)]);
    Ez[mi(g, 3, s, k)] = p * Ez[mi(g, 4, s, k)];
}

// UpdateBfieldHalf (cyl_em_update.F90:49-110, 3-D): i = 4..mx-1, k = 1..mz-1
__global__ void k_cylm_update_b(GridP g, CylmP q, double *__restrict__ Bx, double *__restrict__ By,
                                double *__restrict__ Bz, const double *__restrict__ Ex, const double *__restrict__ Ey,
                                const double *__restrict__ Ez)
{
    const int ni = g.mx - 4, nk = g.mz - 1;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * 3 * nk) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int s = (int)(t % 3);
    const int k = 1 + (int)(t / 3);
    const long long o = mi(g, i, s, k), sk = 3LL * g.mx, ss = g.mx;
    const double h = q.fld_halfc, idr = q.inv_dr, idz = q.inv_dz;
    const double r = rad(q, i), rph = r + 0.5 * q.dr, rp = r + q.dr;
    const double dEz = s == 0 ? 0.0 : (s == 1 ? Ez[o + ss] : -Ez[o - ss]);
    const double dEr = s == 0 ? 0.0 : (s == 1 ? Ex[o + ss] : -Ex[o - ss]);
    Bx[o] = Bx[o] - h * dEz / r + h * idz * (Ey[o + sk] - Ey[o]);
    By[o] = By[o] - h * idz * (Ex[o + sk] - Ex[o]) + h * idr * (Ez[o + 1] - Ez[o]);
    Bz[o] = Bz[o] - h * (idr * (rp * Ey[o + 1] - r * Ey[o]) - dEr) / rph;
}

// the Lehe differences at o = (i, s, k) (cyl_em_update_lehe.F90:255-289 per mode, term
// order kept): gradX, gradX1 and the exact theta derivative, each with its z pair;
// gradZ with the wide z pair.  The reference's r pair and theta pair of gradZ are left
// out: the r pair reads the axis fills and makes the fields next to the axis grow at
// the z Nyquist wavenumber, and the theta pair (the Cartesian coefficient on the theta
// index) would lower every m = 1 z difference by 1/8 at any radius.
__device__ __forceinline__ double dth(const GridP &g, const double *F, long long o, int s)
{
    return s == 0 ? 0.0 : (s == 1 ? F[o + g.mx] : -F[o - g.mx]);
}

// The Lehe B half step on i = 4..mx-1, k = 2..mz-2 (gradZ reads k - 1 and k + 2; the z halo rewrites
// k = 1 and mz - 1):
//   Bx -= h gradT(Ez) / r - h/dz gradZ(Ey),  By -= h/dz gradZ(Ex) - h/dr gradX(Ez),
//   Bz -= h (1/dr gradX1(Ey) - gradT(Ex)) / (r + dr/2),
// with the differences above in the reference's term order.  Each thread walks KC consecutive k of
// one (i, s): the E rows k - 1 .. k + 2 it needs stay in registers and each further k reads one new
// row (one thread per (i, s, k) read every row three or four times through the caches).
template <int KC>
__global__ __launch_bounds__(256) void k_cylm_update_b_lehe_col(GridP g, CylmP q, double *__restrict__ Bx,
                                                                double *__restrict__ By, double *__restrict__ Bz,
                                                                const double *__restrict__ Ex,
                                                                const double *__restrict__ Ey,
                                                                const double *__restrict__ Ez)
{
    const int ni = g.mx - 4, nk = g.mz - 3, nkc = (nk + KC - 1) / KC;
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ni * 3 * nkc) return;
    const int i = 4 + (int)(t % ni);
    t /= ni;
    const int s = (int)(t % 3);
    const int kb = 2 + KC * (int)(t / 3), ke = min(kb + KC, 2 + nk);   // k = kb .. ke - 1, ke <= mz - 1
    const long long sk = 3LL * g.mx;
    const double h = q.fld_halfc, idr = q.inv_dr, idz = q.inv_dz;
    const double r = rad(q, i), rph = r + 0.5 * q.dr, rp = r + q.dr;
    // row m of (i, s): Ex, Ey at i (rows k - 1 .. k + 2), Ey, Ez at i + 1 and Ez at i, and the
    // theta differences of Ex, Ez (rows k - 1 .. k + 1)
    double xa[4], ya[4], yb[3], za[3], zb[3], xt[3], zt[3];
    auto row3 = [&](int m, int j) {
        const long long o = mi(g, i, s, m);
        yb[j] = Ey[o + 1]; za[j] = Ez[o]; zb[j] = Ez[o + 1];
        xt[j] = dth(g, Ex, o, s); zt[j] = dth(g, Ez, o, s);
    };
    auto row4 = [&](int m, int j) {
        const long long o = mi(g, i, s, m);
        xa[j] = Ex[o]; ya[j] = Ey[o];
    };
    for (int j = 0; j < 3; j++) row3(kb - 1 + j, j);
    for (int j = 0; j < 4; j++) row4(kb - 1 + j, j);
    for (int k = kb; k < ke; k++) {
        const long long o = mi(g, i, s, k);
        // index 0: k - 1, 1: k, 2: k + 1, 3: k + 2
        const double gtz = q.l_tt * zt[1] + q.l_tz * zt[2] + q.l_tz * zt[0];
        const double gzy = q.l_zz * (ya[2] - ya[1]) + q.l_zd * (ya[3] - ya[0]);
        const double gzx = q.l_zz * (xa[2] - xa[1]) + q.l_zd * (xa[3] - xa[0]);
        const double grz = q.l_rr * (zb[1] - za[1]) + q.l_rz * (zb[2] - za[2]) + q.l_rz * (zb[0] - za[0]);
        const double gr1y = q.l_rr * (rp * yb[1] - r * ya[1]) + q.l_rz * (rp * yb[2] - r * ya[2])
                            + q.l_rz * (rp * yb[0] - r * ya[0]);
        const double gtx = q.l_tt * xt[1] + q.l_tz * xt[2] + q.l_tz * xt[0];
        Bx[o] = Bx[o] - h * gtz / r + h * idz * gzy;
        By[o] = By[o] - h * idz * gzx + h * idr * grz;
        Bz[o] = Bz[o] - h * (idr * gr1y - gtx) / rph;
        if (k + 1 < ke) {
            for (int j = 0; j < 2; j++) {
                yb[j] = yb[j + 1]; za[j] = za[j + 1]; zb[j] = zb[j + 1]; xt[j] = xt[j + 1]; zt[j] = zt[j + 1];
            }
            for (int j = 0; j < 3; j++) { xa[j] = xa[j + 1]; ya[j] = ya[j + 1]; }
            row3(k + 2, 2);
            row4(k + 3, 3);
        }
    }
}

// Bz on the axis (cyl_em_update.F90:231-250, the m = 0 circulation of E_theta at
// r = dr/2 over the disk, 4/dr), then the fills of :182-229 by parity
__global__ void k_cylm_b_axis(GridP g, CylmP q, double *__restrict__ Bx, double *__restrict__ By,
                              double *__restrict__ Bz, const double *__restrict__ Ey)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * g.mz) return;
    const int s = (int)(t % 3), k = 1 + (int)(t / 3);
    if (s == 0 && k <= g.mz - 1) {
        const long long a = mi(g, 3, 0, k);
        Bz[a] = Bz[a] - q.fld_halfc * (4.0 * q.inv_dr) * Ey[mi(g, 4, 0, k)];
    }
    const double p = parity(s);
    By[mi(g, 3, s, k)] = 0.5 * (By[mi(g, 4, s, k)] - p * By[mi(g, 4, s, k)]);
    By[mi(g, 2, s, k)] = -(p * By[mi(g, 4, s, k)]);
    Bx[mi(g, 3, s, k)] = -(p * Bx[mi(g, 4, s, k)]);
    Bz[mi(g, 2, s, k)] = p * Bz[mi(g, 4, s, k)];
}

// periodic z ghosts over x = 3..mx-3 (zero ghosts for open z), then the wall (i >= mx-2
// zero); three arrays
__global__ void k_cylm_zhalo(GridP g, int zopen, double *__restrict__ F0, double *__restrict__ F1,
                             double *__restrict__ F2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.mx * 3 * 5) return;
    const int i = 1 + (int)(t % g.mx);
    const long long r = t / g.mx;
    const int s = (int)(r % 3), l = (int)(r / 3);   // l: 0,1 low ghosts, 2..4 high ghosts
    if (i < 3 || i > g.mx - 3) return;
    const int k = l < 2 ? 1 + l : g.mz - 2 + (l - 2);
    const int src = l < 2 ? k + g.nz : k - g.nz;
    double *F[3] = {F0, F1, F2};
#pragma unroll
    for (int c = 0; c < 3; c++) F[c][mi(g, i, s, k)] = zopen ? 0.0 : F[c][mi(g, i, s, src)];
}

__global__ void k_cylm_wall(GridP g, double *__restrict__ F0, double *__restrict__ F1, double *__restrict__ F2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * 3 * g.mz) return;
    const int i = g.mx - 2 + (int)(t % 3);
    const long long r = t / 3;
    const int s = (int)(r % 3), k = 1 + (int)(r / 3);
    F0[mi(g, i, s, k)] = 0.0;
    F1[mi(g, i, s, k)] = 0.0;
    F2[mi(g, i, s, k)] = 0.0;
}

// AxisCurrentBC (cyl_movdep.F90:451-465) at fixed theta: on every slot alike
__global__ void k_cylm_axis_current(GridP g, double *__restrict__ Jx, double *__restrict__ Jy,
                                    double *__restrict__ Jz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3LL * g.mz) return;
    const int s = (int)(t % 3), k = 1 + (int)(t / 3);
    Jy[mi(g, 4, s, k)] = Jy[mi(g, 4, s, k)] - Jy[mi(g, 3, s, k)];
    Jz[mi(g, 4, s, k)] = Jz[mi(g, 4, s, k)] - Jz[mi(g, 3, s, k)];
    Jx[mi(g, 3, s, k)] = 0.0;
    Jy[mi(g, 2, s, k)] = 0.0;
    Jz[mi(g, 2, s, k)] = 0.0;
}

// periodic z fold of J: 3..5 += mz-2..mz, then mz-4..mz-3 += 1..2 (two launches)
__global__ void k_cylm_zfold(GridP g, int high, double *__restrict__ J0, double *__restrict__ J1,
                             double *__restrict__ J2)
{
    const int nl = high ? 2 : 3;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.mx * 3 * nl) return;
    const int i = 1 + (int)(t % g.mx);
    const long long r = t / g.mx;
    const int s = (int)(r % 3), l = (int)(r / 3);
    const int k = high ? g.mz - 4 + l : 3 + l;
    const int src = high ? k - g.nz : k + g.nz;
    double *J[3] = {J0, J1, J2};
#pragma unroll
    for (int c = 0; c < 3; c++) J[c][mi(g, i, s, k)] = J[c][mi(g, i, s, k)] + J[c][mi(g, i, s, src)];
}

// node values of the staggered fields (no theta averaging: the modes are exact in theta)
__device__ inline void node_vals(const GridP &g, const double *const *E, const double *const *B, int i, int s, int k,
                                 double e[3], double b[3])
{
    const long long o = mi(g, i, s, k), sk = 3LL * g.mx;
    e[0] = (E[0][o - 1] + E[0][o]) * 0.5;
    e[1] = E[1][o];
    e[2] = (E[2][o - sk] + E[2][o]) * 0.5;
    b[0] = (B[0][o - sk] + B[0][o]) * 0.5;
    b[1] = (B[1][o - 1 - sk] + B[1][o - 1] + B[1][o - sk] + B[1][o]) * 0.25;
    b[2] = (B[2][o - 1] + B[2][o]) * 0.5;
}

__device__ inline void gather(const GridP &g, const double *const *E, const double *const *B, double x, double th,
                              double z, double pe[3], double pb[3])
{
    const int ip = (int)x, kp = (int)z;
    const double dx = x - ip, dz = z - kp;
    const double ph[3] = {1.0, cos(th), sin(th)};
    for (int c = 0; c < 3; c++) { pe[c] = 0.0; pb[c] = 0.0; }
    for (int s = 0; s < 3; s++) {
        double se[3] = {0.0, 0.0, 0.0}, sb[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int nn = 0; nn < 4; nn++) {
            const int a = nn & 1, b = nn >> 1;
            const double wt = (a ? dx : 1.0 - dx) * (b ? dz : 1.0 - dz);
            double e[3], bb[3];
            node_vals(g, E, B, ip + a, s, kp + b, e, bb);
            for (int c = 0; c < 3; c++) { se[c] = se[c] + wt * e[c]; sb[c] = sb[c] + wt * bb[c]; }
        }
        for (int c = 0; c < 3; c++) { pe[c] = pe[c] + ph[s] * se[c]; pb[c] = pb[c] + ph[s] * sb[c]; }
    }
}

// the modal deposit of one particle (cyl_movdep.F90's zigzag split, per mode): the r-z
// zigzag at the old angle, each slot weighted by its phase, then the angular move at
// the new (r, z); add(component, slot, i, k, value) receives every contribution
struct DepMove {
    int i1, i2, k1, k2;
    double Fx1, Fz1, Fx2, Fz2, Wx1, Wz1, Wx2, Wz2, wx, wz;
    double pa[3], pt[3];
};

template <class Add>
__device__ inline void modal_deposit(const DepMove &d, Add add)
{
#pragma unroll
    for (int s = 0; s < 3; s++) {
        add(0, s, d.i1, d.k1, d.pa[s] * d.Fx1 * (1.0 - d.Wz1));
        add(0, s, d.i1, d.k1 + 1, d.pa[s] * d.Fx1 * d.Wz1);
        add(2, s, d.i1, d.k1, d.pa[s] * d.Fz1 * (1.0 - d.Wx1));
        add(2, s, d.i1 + 1, d.k1, d.pa[s] * d.Fz1 * d.Wx1);
        add(0, s, d.i2, d.k2, d.pa[s] * d.Fx2 * (1.0 - d.Wz2));
        add(0, s, d.i2, d.k2 + 1, d.pa[s] * d.Fx2 * d.Wz2);
        add(2, s, d.i2, d.k2, d.pa[s] * d.Fz2 * (1.0 - d.Wx2));
        add(2, s, d.i2 + 1, d.k2, d.pa[s] * d.Fz2 * d.Wx2);
    }
#pragma unroll
    for (int s = 0; s < 3; s++) {
        add(1, s, d.i2, d.k2, d.pt[s] * (1.0 - d.wx) * (1.0 - d.wz));
        add(1, s, d.i2 + 1, d.k2, d.pt[s] * d.wx * (1.0 - d.wz));
        add(1, s, d.i2, d.k2 + 1, d.pt[s] * (1.0 - d.wx) * d.wz);
        add(1, s, d.i2 + 1, d.k2 + 1, d.pt[s] * d.wx * d.wz);
    }
}

__device__ inline int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// mode cells of the block's deposit window held in LDS: 9 accumulators (3 components x
// 3 slots) per (i, k) cell, FP64.  A window of at most kDepCells / 4 cells keeps 4
// interleaved copies of each accumulator (lane l adds into copy l % 4): the particles of
// one cell sit in adjacent lanes and would otherwise serialise on the same LDS words
constexpr int kDepCells = 448;
constexpr int kDepCopies = 4;
// when the block's particles spread wider than kDepCells cells (electrons expelled,
// crossing the axis or trapped in the wake, far from their array neighbours), the window
// is a box of kBoxI x kBoxK cells around the block's first depositing particle; the
// particles outside it add to J directly
constexpr int kBoxI = 28, kBoxK = 4;

// the block's deposit window [bil, bil + wi) x [bkl, bkl + wk); uniform over the block
__device__ inline void deposit_window(bool dep, const DepMove &d, int (*sred)[4], int *sbox, int &bil, int &bkl,
                                      int &wi, int &wk, bool &any)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int ilo = wave_min(dep ? min(d.i1, d.i2) : INT_MAX), ihi = wave_max(dep ? max(d.i1, d.i2) : INT_MIN);
    const int klo = wave_min(dep ? min(d.k1, d.k2) : INT_MAX), khi = wave_max(dep ? max(d.k1, d.k2) : INT_MIN);
    const int first = wave_min(dep ? (int)threadIdx.x : INT_MAX);
    if (lane == 0) { sred[wv][0] = ilo; sred[wv][1] = ihi; sred[wv][2] = klo; sred[wv][3] = khi; sbox[wv] = first; }
    __syncthreads();
    int bih = INT_MIN, bkh = INT_MIN, f = INT_MAX;
    bil = INT_MAX; bkl = INT_MAX;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        bil = min(bil, sred[w][0]); bih = max(bih, sred[w][1]);
        bkl = min(bkl, sred[w][2]); bkh = max(bkh, sred[w][3]);
        f = min(f, sbox[w]);
    }
    any = bil <= bih;
    if (!any) return;
    wi = bih + 2 - bil;
    wk = bkh + 2 - bkl;
    if ((long long)wi * wk <= kDepCells) return;
    __syncthreads();   // every thread has read sbox
    if ((int)threadIdx.x == f) { sbox[0] = min(d.i1, d.i2); sbox[1] = min(d.k1, d.k2); }
    __syncthreads();
    bil = sbox[0] - 4;
    bkl = sbox[1] - 1;
    wi = kBoxI;
    wk = kBoxK;
}

// MoveDepositPrtl: gather, Boris, the local-frame move and rotation (cyl_movdep.F90:14-289),
// then the modal deposit.  The particles keep the loader's cell order, so the 256 of a
// block deposit into a small (i, k) window: it is accumulated in LDS and added to J once
// per cell and accumulator; a block whose window exceeds kDepCells cells adds every
// contribution to J directly.
__global__ __launch_bounds__(256) void k_cylm_push(GridP g, CylmP q, const double *__restrict__ Ex,
                                                   const double *__restrict__ Ey, const double *__restrict__ Ez,
                                                   const double *__restrict__ Bx, const double *__restrict__ By,
                                                   const double *__restrict__ Bz, double *__restrict__ Jx,
                                                   double *__restrict__ Jy, double *__restrict__ Jz, PartF64 P,
                                                   long long n, int *__restrict__ err)
{
    __shared__ double sJ[9 * kDepCells];
    __shared__ int sred[4][4], sbox[4];
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int flv = m < n ? P.flv[m] : 0;
    bool dep = false;
    DepMove d;
    if (flv != 0) {
        const double *E[3] = {Ex, Ey, Ez}, *B[3] = {Bx, By, Bz};
        const double x0 = P.x[m], th0 = P.y[m], z0 = P.z[m];
        const double qm = q.flvrqm[flv - 1] * 0.5;
        const double c = q.c, sqc = q.sqc, cinv = q.cinv;
        double pe[3], pb[3];
        gather(g, E, B, x0, th0, z0, pe, pb);
        const double pEx = pe[0] * qm, pEy = pe[1] * qm, pEz = pe[2] * qm;
        double pBx = pb[0] * qm, pBy = pb[1] * qm, pBz = pb[2] * qm;
        double u0 = c * P.u[m] + pEx;
        double v0 = c * P.v[m] + pEy;
        double w0 = c * P.w[m] + pEz;
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
        const double wn = w0 * cinv;
        gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        const double X = rad(q, x0) + u1 * gg, Y = v1 * gg;
        const double r1 = sqrt(X * X + Y * Y);
        const double x1 = r1 * q.inv_dr + 3.5;
        const double th1 = th0 + atan2(Y, X);
        const double z1 = z0 + wn * gg * q.inv_dz;
        const double ir = 1.0 / r1;
        const double un = (X * u1 + Y * v1) * ir;
        const double vn = (-Y * u1 + X * v1) * ir;
        if (x1 >= g.mx - 3) {
            atomicOr(err, 1);
        } else {
            dep = true;
            const double qthis = P.q[m] * q.qi;
            d.i1 = (int)floor(x0); d.i2 = (int)floor(x1); d.k1 = (int)floor(z0); d.k2 = (int)floor(z1);
            const double xr = fmin((double)(min(d.i1, d.i2) + 1), fmax((double)max(d.i1, d.i2), 0.5 * (x0 + x1)));
            const double zr = fmin((double)(min(d.k1, d.k2) + 1), fmax((double)max(d.k1, d.k2), 0.5 * (z0 + z1)));
            d.Fx1 = qthis * (xr - x0); d.Fz1 = qthis * (zr - z0);
            d.Fx2 = qthis * (x1 - xr); d.Fz2 = qthis * (z1 - zr);
            d.Wx1 = 0.5 * (x0 + xr) - d.i1; d.Wz1 = 0.5 * (z0 + zr) - d.k1;
            d.Wx2 = 0.5 * (x1 + xr) - d.i2; d.Wz2 = 0.5 * (z1 + zr) - d.k2;
            d.wx = x1 - d.i2; d.wz = z1 - d.k2;
            d.pa[0] = 1.0; d.pa[1] = cos(th0); d.pa[2] = sin(th0);
            d.pt[0] = qthis * (th1 - th0);
            d.pt[1] = qthis * (sin(th1) - sin(th0));
            d.pt[2] = -(qthis * (cos(th1) - cos(th0)));
        }
        double zn = z1, tn = th1;
        if (q.zopen) {
            if (zn >= q.zmax || zn < q.zmin) P.flv[m] = 0;   // open z: left the box
        } else if (zn > q.zmax) {
            zn = zn - q.zlen;
        } else if (zn < q.zmin) {
            zn = q.zlen + zn;
        }
        if (tn >= 2.0 * kPi) tn = tn - 2.0 * kPi;
        else if (tn < 0.0) tn = tn + 2.0 * kPi;
        P.x[m] = x1; P.y[m] = tn; P.z[m] = zn;
        P.u[m] = un; P.v[m] = vn; P.w[m] = wn;
    }
    // the block's window: i in [bil, bil + wi), k in [bkl, bkl + wk)
    const int lane = threadIdx.x & 63;
    int bil, bkl, wi, wk;
    bool any;
    deposit_window(dep, d, sred, sbox, bil, bkl, wi, wk, any);
    if (!any) return;                           // nothing to deposit in this block (uniform)
    double *J[3] = {Jx, Jy, Jz};
    const bool inside = dep && min(d.i1, d.i2) >= bil && max(d.i1, d.i2) + 1 < bil + wi && min(d.k1, d.k2) >= bkl &&
                        max(d.k1, d.k2) + 1 < bkl + wk;
    if (dep && !inside)
        modal_deposit(d, [&](int cc, int s, int i, int k, double v) { unsafeAtomicAdd(&J[cc][mi(g, i, s, k)], v); });
    const int cells = wi * wk;
    int R = kDepCopies;                         // copies per accumulator: cells * R <= kDepCells
    while (R > 1 && cells * R > kDepCells) R >>= 1;
    const int rl = lane & (R - 1);
    for (int e = threadIdx.x; e < 9 * cells * R; e += 256) sJ[e] = 0.0;
    __syncthreads();
    if (inside)
        modal_deposit(d, [&](int cc, int s, int i, int k, double v) {
            atomicAdd(&sJ[((cc * 3 + s) * cells + (k - bkl) * wi + (i - bil)) * R + rl], v);
        });
    __syncthreads();
    for (int e = threadIdx.x; e < 9 * cells; e += 256) {
        double v = 0.0;
        for (int r = 0; r < R; r++) v += sJ[e * R + r];
        if (v == 0.0) continue;
        const int cs = e / cells, cell = e - cs * cells;
        const int i = bil + cell % wi, k = bkl + cell / wi;
        unsafeAtomicAdd(&J[cs / 3][mi(g, i, cs % 3, k)], v);
    }
}

// The node-centred FP32 copy of E and B that the split push gathers from, rebuilt before
// every split push: node (i, k) holds its 18 values v[6 s + c] (slot s = 0..2; c = 0..5:
// Ex, Ey, Ez, Bx, By, Bz, computed in FP64 as node_vals does and rounded once to FP32),
// v[0..15] in 4 float4 and v[16..17] in 1 float2, so the gather makes 4 x 5 vector loads
// per particle instead of 156 scalar FP64 loads.  Node index (k - 1) * mx + (i - 1);
// nodes with i = 1 or k = 1 (no lower neighbour) hold zero.
__global__ __launch_bounds__(256) void k_cylm_nodes_f32(GridP g, const double *__restrict__ Ex,
                                                        const double *__restrict__ Ey, const double *__restrict__ Ez,
                                                        const double *__restrict__ Bx, const double *__restrict__ By,
                                                        const double *__restrict__ Bz, float4 *__restrict__ nf4,
                                                        float2 *__restrict__ nf2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.mx * g.mz) return;
    const int i = 1 + (int)(t % g.mx), k = 1 + (int)(t / g.mx);
    float v[18];
#pragma unroll
    for (int e = 0; e < 18; e++) v[e] = 0.0f;
    if (i >= 2 && k >= 2) {
        const double *E[3] = {Ex, Ey, Ez}, *B[3] = {Bx, By, Bz};
#pragma unroll
        for (int s = 0; s < 3; s++) {
            double e[3], b[3];
            node_vals(g, E, B, i, s, k, e, b);
#pragma unroll
            for (int c = 0; c < 3; c++) { v[6 * s + c] = (float)e[c]; v[6 * s + 3 + c] = (float)b[c]; }
        }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) nf4[4 * t + j] = make_float4(v[4 * j], v[4 * j + 1], v[4 * j + 2], v[4 * j + 3]);
    nf2[t] = make_float2(v[16], v[17]);
}

// the FP32 gather of the split push from the node copy: per (slot, component) the
// bilinear sum over the cell's nodes nn = 0..3 (weights (1 - dx or dx) (1 - dz or dz)),
// then the slots combined with the phases (1, cos theta, sin theta) in order 0..2
__device__ inline void gather_nodes(const GridP &g, const float4 *__restrict__ nf4, const float2 *__restrict__ nf2,
                                    int ip, int kp, float dx, float dz, float c1, float s1, float pe[3], float pb[3])
{
    float se[18];
#pragma unroll
    for (int e = 0; e < 18; e++) se[e] = 0.0f;
    const long long n0 = (long long)(kp - 1) * g.mx + (ip - 1);
#pragma unroll
    for (int nn = 0; nn < 4; nn++) {
        const int a = nn & 1, b = nn >> 1;
        const float wt = (a ? dx : 1.0f - dx) * (b ? dz : 1.0f - dz);
        const long long n = n0 + a + (long long)b * g.mx;
        float v[18];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float4 f = nf4[4 * n + j];
            v[4 * j] = f.x; v[4 * j + 1] = f.y; v[4 * j + 2] = f.z; v[4 * j + 3] = f.w;
        }
        const float2 f2 = nf2[n];
        v[16] = f2.x; v[17] = f2.y;
#pragma unroll
        for (int e = 0; e < 18; e++) se[e] += wt * v[e];
    }
    const float ph[3] = {1.0f, c1, s1};
    for (int c = 0; c < 3; c++) { pe[c] = 0.0f; pb[c] = 0.0f; }
#pragma unroll
    for (int s = 0; s < 3; s++)
#pragma unroll
        for (int c = 0; c < 3; c++) { pe[c] += ph[s] * se[6 * s + c]; pb[c] += ph[s] * se[6 * s + 3 + c]; }
}

// modal_deposit for a particle that stays in its cell (i1 = i2, k1 = k2): the two
// r-z half-segments land on the same nodes, so their fixed-point counts are added
// first (integers below 2^53: exact) and each node takes one add instead of two
template <class Count, class Add>
__device__ inline void modal_deposit_same_cell(const DepMove &d, Count count, Add add)
{
#pragma unroll
    for (int s = 0; s < 3; s++) {
        add(0, s, d.i1, d.k1, count(d.pa[s] * d.Fx1 * (1.0 - d.Wz1)) + count(d.pa[s] * d.Fx2 * (1.0 - d.Wz2)));
        add(0, s, d.i1, d.k1 + 1, count(d.pa[s] * d.Fx1 * d.Wz1) + count(d.pa[s] * d.Fx2 * d.Wz2));
        add(2, s, d.i1, d.k1, count(d.pa[s] * d.Fz1 * (1.0 - d.Wx1)) + count(d.pa[s] * d.Fz2 * (1.0 - d.Wx2)));
        add(2, s, d.i1 + 1, d.k1, count(d.pa[s] * d.Fz1 * d.Wx1) + count(d.pa[s] * d.Fz2 * d.Wx2));
    }
#pragma unroll
    for (int s = 0; s < 3; s++) {
        add(1, s, d.i2, d.k2, count(d.pt[s] * (1.0 - d.wx) * (1.0 - d.wz)));
        add(1, s, d.i2 + 1, d.k2, count(d.pt[s] * d.wx * (1.0 - d.wz)));
        add(1, s, d.i2, d.k2 + 1, count(d.pt[s] * (1.0 - d.wx) * d.wz));
        add(1, s, d.i2 + 1, d.k2 + 1, count(d.pt[s] * d.wx * d.wz));
    }
}

// MoveDepositPrtl of the split form: the push in FP32 on (cell, offset) positions, theta
// in 2^-32 turns; the deposit in FP64 from the stored old and new positions (exact in
// FP64), each contribution rounded to a multiple of 2^-36 and summed as a 64-bit integer
// count in the LDS window: every sum is exact, so J does not depend on the order of the atomics.
// Flags: 1 = a particle reached the wall, 2 = a contribution above 64 charge cells (the
// fixed-point range).
__global__ __launch_bounds__(256) void k_cylm_push_split(GridP g, CylmP q, const float4 *__restrict__ nf4,
                                                         const float2 *__restrict__ nf2, double *__restrict__ Jx,
                                                         double *__restrict__ Jy, double *__restrict__ Jz, PartCylS P,
                                                         long long n, int *__restrict__ err)
{
    __shared__ unsigned long long sJ[9 * kDepCells];
    __shared__ int sred[4][4], sbox[4];
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t cw = m < n ? P.cell[m] : 0u;
    bool dep = false;
    DepMove d;
    if (cw != 0u) {
        const int i0 = cell_i(cw), k0 = cell_k(cw), flv = cell_flv(cw);
        const float ox = P.ox[m], oz = P.oz[m];
        const uint32_t t0 = P.th[m];
        float s0, c0;
        sincosf((float)((double)t0 * kTurnRad), &s0, &c0);
        float pe[3], pb[3];
        gather_nodes(g, nf4, nf2, i0, k0, ox, oz, c0, s0, pe, pb);
        const float qm = (float)(q.flvrqm[flv - 1] * 0.5);
        const float c = (float)q.c, sqc = (float)q.sqc, cinv = (float)q.cinv;
        const float pEx = pe[0] * qm, pEy = pe[1] * qm, pEz = pe[2] * qm;
        float pBx = pb[0] * qm, pBy = pb[1] * qm, pBz = pb[2] * qm;
        float u0 = c * P.u[m] + pEx;
        float v0 = c * P.v[m] + pEy;
        float w0 = c * P.w[m] + pEz;
        float gg = rsqrtf(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        pBx *= gg; pBy *= gg; pBz *= gg;
        const float ff = 2.0f / (1.0f + pBx * pBx + pBy * pBy + pBz * pBz);
        float u1 = (u0 + v0 * pBz - w0 * pBy) * ff;
        float v1 = (v0 + w0 * pBx - u0 * pBz) * ff;
        const float w1 = (w0 + u0 * pBy - v0 * pBx) * ff;
        u0 = u0 + v1 * pBz - w1 * pBy + pEx;
        v0 = v0 + w1 * pBx - u1 * pBz + pEy;
        w0 = w0 + u1 * pBy - v1 * pBx + pEz;
        u1 = u0 * cinv;
        v1 = v0 * cinv;
        const float wn = w0 * cinv;
        gg = sqc * rsqrtf(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        // the move in the local frame: X = r0 + dX, Y; the new radius as r0 + dR with
        // dR = (dX (2 r0 + dX) + Y^2) / (r0 + r1), so the offset keeps FP32 precision
        const float dr = (float)q.dr;
        const float r0 = ((float)(i0 - 3.5) + ox) * dr;
        const float dX = u1 * gg, Y = v1 * gg, X = r0 + dX;
        const float r1 = sqrtf(X * X + Y * Y);
        const float rs = r0 + r1;
        const float dR = rs > 0.0f ? (dX * (2.0f * r0 + dX) + Y * Y) / rs : r1;
        const float xo = ox + dR * (float)q.inv_dr, zo = oz + wn * gg * (float)q.inv_dz;
        const float fx = floorf(xo), fz = floorf(zo);
        int i2 = i0 + (int)fx, k2 = k0 + (int)fz;
        float ox2 = xo - fx, oz2 = zo - fz;
        if (ox2 >= 1.0f) { i2 += 1; ox2 = 0.0f; }
        if (oz2 >= 1.0f) { k2 += 1; oz2 = 0.0f; }
        const float ir = r1 > 0.0f ? 1.0f / r1 : 0.0f;
        const float un = r1 > 0.0f ? (X * u1 + Y * v1) * ir : u1;
        const float vn = r1 > 0.0f ? (-Y * u1 + X * v1) * ir : v1;
        const float dth = atan2f(Y, X);
        const uint32_t t1 = t0 + (uint32_t)(unsigned long long)(long long)llrintf(dth * (float)(4294967296.0 / (2.0 * kPi)));
        if (i2 >= g.mx - 3) {
            atomicOr(err, 1);
        } else {
            dep = true;
            const double qthis = (double)P.q[m] * q.qi;
            const double x0 = i0 + (double)ox, z0 = k0 + (double)oz;
            const double x1 = i2 + (double)ox2, z1 = k2 + (double)oz2;
            d.i1 = i0; d.i2 = i2; d.k1 = k0; d.k2 = k2;
            const double xr = fmin((double)(min(d.i1, d.i2) + 1), fmax((double)max(d.i1, d.i2), 0.5 * (x0 + x1)));
            const double zr = fmin((double)(min(d.k1, d.k2) + 1), fmax((double)max(d.k1, d.k2), 0.5 * (z0 + z1)));
            d.Fx1 = qthis * (xr - x0); d.Fz1 = qthis * (zr - z0);
            d.Fx2 = qthis * (x1 - xr); d.Fz2 = qthis * (z1 - zr);
            d.Wx1 = 0.5 * (x0 + xr) - d.i1; d.Wz1 = 0.5 * (z0 + zr) - d.k1;
            d.Wx2 = 0.5 * (x1 + xr) - d.i2; d.Wz2 = 0.5 * (z1 + zr) - d.k2;
            d.wx = x1 - d.i2; d.wz = z1 - d.k2;
            // the phases of the stored angles, as the next step's gather sees them
            float s1, c1;
            sincosf((float)((double)t1 * kTurnRad), &s1, &c1);
            d.pa[0] = 1.0; d.pa[1] = c0; d.pa[2] = s0;
            d.pt[0] = qthis * ((double)(int32_t)(t1 - t0) * kTurnRad);
            d.pt[1] = qthis * ((double)s1 - (double)s0);
            d.pt[2] = -(qthis * ((double)c1 - (double)c0));
        }
        // periodic z (open z: a particle that left the box becomes a hole); the stored record
        const bool gone = q.zopen && (k2 >= (int)q.zmax || k2 < (int)q.zmin);
        if (k2 >= (int)q.zmax) k2 -= (int)q.zlen;
        else if (k2 < (int)q.zmin) k2 += (int)q.zlen;
        P.cell[m] = gone ? 0u : pack_cell(i2, k2, flv);
        P.ox[m] = ox2; P.oz[m] = oz2; P.th[m] = t1;
        P.u[m] = un; P.v[m] = vn; P.w[m] = wn;
    }
    const int lane = threadIdx.x & 63;
    int bil, bkl, wi, wk;
    bool any;
    deposit_window(dep, d, sred, sbox, bil, bkl, wi, wk, any);
    if (!any) return;
    double *J[3] = {Jx, Jy, Jz};
    bool big = false;
    // the count rint(v 2^36) as a 64-bit integer: |v| <= 64 keeps v 2^36 below 2^51, where
    // adding 1.5 2^52 rounds it to an integer in the low mantissa bits (ties to even, as rint)
    auto count = [&](double v) -> long long {
        big = big || !(fabs(v) <= 64.0);
        return __double_as_longlong(v * kJq + 6755399441055744.0) - 0x4338000000000000LL;
    };
    const bool inside = dep && min(d.i1, d.i2) >= bil && max(d.i1, d.i2) + 1 < bil + wi && min(d.k1, d.k2) >= bkl &&
                        max(d.k1, d.k2) + 1 < bkl + wk;
    if (dep && !inside)
        modal_deposit(d, [&](int cc, int s, int i, int k, double v) {
            unsafeAtomicAdd(&J[cc][mi(g, i, s, k)], (double)count(v) * kJqInv);
        });
    const int cells = wi * wk;
    int R = kDepCopies;
    while (R > 1 && cells * R > kDepCells) R >>= 1;
    const int rl = lane & (R - 1);
    for (int e = threadIdx.x; e < 9 * cells * R; e += 256) sJ[e] = 0ull;
    __syncthreads();
    if (inside) {
        auto add = [&](int cc, int s, int i, int k, long long cnt) {
            atomicAdd(&sJ[((cc * 3 + s) * cells + (k - bkl) * wi + (i - bil)) * R + rl], (unsigned long long)cnt);
        };
        if (d.i1 == d.i2 && d.k1 == d.k2)
            modal_deposit_same_cell(d, count, add);
        else
            modal_deposit(d, [&](int cc, int s, int i, int k, double v) { add(cc, s, i, k, count(v)); });
    }
    if (big) atomicOr(err, 2);
    __syncthreads();
    for (int e = threadIdx.x; e < 9 * cells; e += 256) {
        long long v = 0;
        for (int r = 0; r < R; r++) v += (long long)sJ[e * R + r];
        if (v == 0) continue;
        const int cs = e / cells, cell = e - cs * cells;
        const int i = bil + cell % wi, k = bkl + cell / wi;
        unsafeAtomicAdd(&J[cs / 3][mi(g, i, cs % 3, k)], (double)v * kJqInv);
    }
}

__device__ inline uint64_t mix64d(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ inline double u01d(uint64_t seed, uint64_t idx, uint64_t stream)
{
    const uint64_t h = mix64d(mix64d(seed ^ (idx * 0xD1342543DE82EF95ULL)) + stream * 0x9E3779B97F4A7C15ULL);
    return (double)(h >> 24) * (1.0 / 1099511627776.0);
}

// uniform cold electrons: ppc per cell of i = 4..4+nrp-1, k = k0..k0+nzc-1, (r, z)
// uniform in the cell, ppc evenly spaced angles with a per-cell random offset,
// weight epc 2 pi r dr / ppc; counter-based, so a cell (shifted by `shift`
// window moves) always gets the same particles.
__device__ inline void plasma_particle(const GridP &g, const CylmP &q, long long t, int nrp, int k0, int ppc, int epc,
                                       unsigned long long seed, long long shift, double &x, double &th, double &z,
                                       double &qq)
{
    const int p = (int)(t % ppc);
    const long long cidx = t / ppc;
    const int i = 4 + (int)(cidx % nrp), k = k0 + (int)(cidx / nrp);
    const uint64_t cell = (uint64_t)((k - 3 + shift) * (long long)g.nx + (i - 4));
    const double off = u01d(seed, cell, 7);
    const uint64_t id = cell * (uint64_t)ppc + (uint64_t)p;
    x = i + u01d(seed, id, 1);
    z = k + u01d(seed, id, 2);
    th = 2.0 * kPi * ((p + off) / ppc);
    qq = -(epc * 2.0 * kPi * rad(q, x) * q.dr / ppc);
}

// appended at base + index
__global__ void k_cylm_inject(GridP g, CylmP q, PartF64 P, long long base, int nrp, int k0, int nzc, int ppc,
                              int epc, unsigned long long seed, long long shift)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nrp * nzc * ppc;
    if (t >= total) return;
    double x, th, z, qq;
    plasma_particle(g, q, t, nrp, k0, ppc, epc, seed, shift, x, th, z, qq);
    const long long n = base + t;
    P.x[n] = x; P.y[n] = th; P.z[n] = z;
    P.u[n] = 0.0; P.v[n] = 0.0; P.w[n] = 0.0;
    P.q[n] = qq;
    P.flv[n] = 2;
}

__global__ void k_cylm_inject_split(GridP g, CylmP q, PartCylS P, long long base, int nrp, int k0, int nzc, int ppc,
                                    int epc, unsigned long long seed, long long shift)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nrp * nzc * ppc;
    if (t >= total) return;
    double x, th, z, qq;
    plasma_particle(g, q, t, nrp, k0, ppc, epc, seed, shift, x, th, z, qq);
    const long long n = base + t;
    int i, k;
    float ox, oz;
    split_coord(x, i, ox);
    split_coord(z, k, oz);
    P.cell[n] = pack_cell(i, k, 2);
    P.ox[n] = ox; P.oz[n] = oz;
    P.th[n] = theta_turns(th);
    P.u[n] = 0.0f; P.v[n] = 0.0f; P.w[n] = 0.0f;
    P.q[n] = (float)qq;
}

// a linearly polarised (x) Gaussian pulse moving along +z, at focus (no curvature):
// E_x = E0 exp(-r^2/w0^2 - (z - zc)^2/L^2) cos(k0 (z - zc)), B_y = E_x at B's
// stagger, added to the m = 1 slots: E_r,c = E_x, E_theta,s = -E_x, B_r,s = B_y,
// B_theta,c = B_y.  E0 = a0 c w0dt (the quiver amplitude of gamma beta is a0).
__global__ void k_cylm_laser(GridP g, CylmP q, double E0, double k0, double w0, double L, double zc,
                             double *__restrict__ Ex, double *__restrict__ Ey, double *__restrict__ Bx,
                             double *__restrict__ By)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.mx * g.mz) return;
    const int i = 1 + (int)(t % g.mx), k = 1 + (int)(t / g.mx);
    if (i < 4 || i > g.mx - 3) return;
    auto prof = [&](double r, double z) {
        const double dz = z - zc;
        return E0 * exp(-(r * r) / (w0 * w0) - (dz * dz) / (L * L)) * cos(k0 * dz);
    };
    // staggers: E_r at r(i) + dr/2, z(k); E_theta at r(i), z(k); B_r at r(i), z(k) + 1/2;
    // B_theta at r(i) + dr/2, z(k) + 1/2 (the z centring of the Yee cell).  Between steps
    // B is half a step behind E (the E update uses B after both half pushes), so B is the
    // pulse half a step earlier, c/2 cells further back: a purely forward wave
    const double r = rad(q, i), rh = r + 0.5 * q.dr, z = (double)k, zh = k + 0.5 + 0.5 * q.c;
    Ex[mi(g, i, 1, k)] += prof(rh, z);
    Ey[mi(g, i, 2, k)] += -prof(r, z);
    Bx[mi(g, i, 2, k)] += prof(r, zh);
    By[mi(g, i, 1, k)] += prof(rh, zh);
}

// the moving window: every slot and z row moves one cell back, dst(k) = src(k + 1)
// from a copy; the front row k = mz - 3 and the ghosts above it restart at zero
__global__ void k_cylm_shift_fields(GridP g, Ptrs6<double> P)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncell) return;
    const double *__restrict__ src = P.p[blockIdx.y];
    double *__restrict__ dst = P.p[3 + blockIdx.y];
    const int k = 1 + (int)(t / (3LL * g.mx));
    dst[t] = (k + 1 <= g.mz - 3) ? src[t + 3LL * g.mx] : 0.0;
}

// particles: z -= 1; those below the window's start are dropped (flavour 0)
__global__ void k_cylm_shift_particles(CylmP q, PartF64 P, long long n)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n || P.flv[m] == 0) return;
    const double z = P.z[m] - 1.0;
    P.z[m] = z;
    if (z < q.zmin) P.flv[m] = 0;
}

// the same on the split record: k -= 1; a particle with k - 1 < zmin (z < zmin) is dropped
__global__ void k_cylm_shift_split(CylmP q, PartCylS P, long long n)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const uint32_t w = P.cell[m];
    if (w == 0u) return;
    const int k = cell_k(w) - 1;
    P.cell[m] = k < (int)q.zmin ? 0u : pack_cell(cell_i(w), k, cell_flv(w));
}

// 1/2 sum B^2, E^2 with the volume 2 pi r dr (pi for the m = 1 slots) of each stagger
__global__ void k_cylm_energy(GridP g, CylmP q, const double *__restrict__ Ex, const double *__restrict__ Ey,
                              const double *__restrict__ Ez, const double *__restrict__ Bx,
                              const double *__restrict__ By, const double *__restrict__ Bz, double *partial)
{
    __shared__ double sb[256], se[256];
    const int ni = g.mx - 6, nk = g.mz - 5;
    const long long total = (long long)ni * 3 * nk;
    double be = 0.0, ee = 0.0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        long long r0 = t;
        const int i = 4 + (int)(r0 % ni);
        r0 /= ni;
        const int s = (int)(r0 % 3), k = 3 + (int)(r0 / 3);
        const long long o = mi(g, i, s, k);
        const double wm = (s == 0 ? 2.0 * kPi : kPi) * q.dr;
        const double r = rad(q, i), rh = r + 0.5 * q.dr;
        ee += wm * (rh * Ex[o] * Ex[o] + r * Ey[o] * Ey[o] + r * Ez[o] * Ez[o]);
        be += wm * (r * Bx[o] * Bx[o] + rh * By[o] * By[o] + rh * Bz[o] * Bz[o]);
    }
    sb[threadIdx.x] = be;
    se[threadIdx.x] = ee;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) { sb[threadIdx.x] += sb[threadIdx.x + w]; se[threadIdx.x] += se[threadIdx.x + w]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = 0.5 * sb[0]; partial[2 * blockIdx.x + 1] = 0.5 * se[0]; }
}

// the two particle records behind one interface: liveness, flavour, momenta, weight, copy
struct RecF64 {
    PartF64 P;
    __device__ bool live(long long m) const { return P.flv[m] != 0; }
    __device__ uint32_t sort_key(long long m, int ib) const
    {
        return ((uint32_t)floor(P.z[m]) << ib) | (uint32_t)floor(P.x[m]);
    }
    __device__ int flv(long long m) const { return P.flv[m]; }
    __device__ void mom(long long m, double &u, double &v, double &w, double &q) const
    {
        u = P.u[m]; v = P.v[m]; w = P.w[m]; q = P.q[m];
    }
    __device__ void copy_to(long long m, const RecF64 &O, long long d) const
    {
        O.P.q[d] = P.q[m]; O.P.x[d] = P.x[m]; O.P.y[d] = P.y[m]; O.P.z[d] = P.z[m];
        O.P.u[d] = P.u[m]; O.P.v[d] = P.v[m]; O.P.w[d] = P.w[m]; O.P.flv[d] = P.flv[m];
    }
};
struct RecSplit {
    PartCylS P;
    __device__ bool live(long long m) const { return P.cell[m] != 0u; }
    __device__ uint32_t sort_key(long long m, int ib) const
    {
        const uint32_t w = P.cell[m];
        return ((uint32_t)cell_k(w) << ib) | (uint32_t)cell_i(w);
    }
    __device__ int flv(long long m) const { return cell_flv(P.cell[m]); }
    __device__ void mom(long long m, double &u, double &v, double &w, double &q) const
    {
        u = P.u[m]; v = P.v[m]; w = P.w[m]; q = P.q[m];
    }
    __device__ void copy_to(long long m, const RecSplit &O, long long d) const
    {
        O.P.cell[d] = P.cell[m]; O.P.ox[d] = P.ox[m]; O.P.oz[d] = P.oz[m]; O.P.th[d] = P.th[m];
        O.P.u[d] = P.u[m]; O.P.v[d] = P.v[m]; O.P.w[d] = P.w[m]; O.P.q[d] = P.q[m];
    }
};

// kinetic energy sum (gamma - 1) |q| per flavour (CalcPrtlEnergy, savedata_routines.F90:283-308
// without the mass factor) and the live count
template <class R>
__global__ void k_cylm_kinetic(R P, long long n, double *partial)
{
    __shared__ double s1[256], s2[256], sn[256];
    double a = 0.0, b = 0.0, cnt = 0.0;
    for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (long long)gridDim.x * blockDim.x) {
        if (!P.live(m)) continue;
        double u, v, w, qq;
        P.mom(m, u, v, w, qq);
        const double e = (sqrt(1.0 + u * u + v * v + w * w) - 1.0) * fabs(qq);
        if (P.flv(m) == 1) a += e; else b += e;
        cnt += 1.0;
    }
    s1[threadIdx.x] = a; s2[threadIdx.x] = b; sn[threadIdx.x] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w]; s2[threadIdx.x] += s2[threadIdx.x + w]; sn[threadIdx.x] += sn[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x] = s1[0]; partial[3 * blockIdx.x + 1] = s2[0]; partial[3 * blockIdx.x + 2] = sn[0];
    }
}

// order-keeping compaction of the live particles: the deposit window relies on the
// loader's cell order.  Pass 1 counts the live particles of each span of kSpan; the host
// scans the counts; pass 2 writes each span's particles at its offset, in order.
constexpr int kSpan = 1024;

template <class R>
__global__ void k_cylm_span_count(R P, long long n, int *cnt)
{
    __shared__ int sw[4];
    const long long base = (long long)blockIdx.x * kSpan;
    int c = 0;
    for (int j = threadIdx.x; j < kSpan; j += 256) {
        const long long m = base + j;
        if (m < n && P.live(m)) c++;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = sw[0] + sw[1] + sw[2] + sw[3];
}

template <class R>
__global__ void k_cylm_compact_spans(R P, long long n, const long long *__restrict__ off, R O)
{
    // thread t owns the 4 consecutive particles base + 4t .. 4t + 3
    __shared__ int ss[256];
    const long long base = (long long)blockIdx.x * kSpan + 4LL * threadIdx.x;
    int f[4], c = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const long long m = base + j;
        f[j] = (m < n && P.live(m)) ? 1 : 0;
        c += f[j];
    }
    ss[threadIdx.x] = c;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {     // inclusive scan (Hillis-Steele)
        const int v = (int)threadIdx.x >= o ? ss[threadIdx.x - o] : 0;
        __syncthreads();
        ss[threadIdx.x] += v;
        __syncthreads();
    }
    long long d = off[blockIdx.x] + ss[threadIdx.x] - c;
#pragma unroll
    for (int j = 0; j < 4; j++)
        if (f[j]) P.copy_to(base + j, O, d++);
}

void fields(pictor_ctx *c, double *E[3], double *B[3], double *J[3])
{
    for (int i = 0; i < 3; i++) {
        E[i] = (double *)c->E[i];
        B[i] = (double *)c->B[i];
        J[i] = (double *)c->J[i];
    }
}

inline int nblocks(long long n, int bs) { return (int)std::max<long long>(1, (n + bs - 1) / bs); }
dim3 blocks_of(long long n) { return dim3(nblocks(n, 256)); }

void halo(pictor_ctx *c, double *const F[3])
{
    const GridP &g = c->g;
    hipLaunchKernelGGL(k_cylm_zhalo, blocks_of((long long)g.mx * 15), dim3(256), 0, c->stream, g, c->cq.zopen, F[0],
                       F[1], F[2]);
    hipLaunchKernelGGL(k_cylm_wall, blocks_of(9LL * g.mz), dim3(256), 0, c->stream, g, F[0], F[1], F[2]);
}

void free_parts(PartF64 &p)
{
    dfree(p.q); dfree(p.x); dfree(p.y); dfree(p.z); dfree(p.u); dfree(p.v); dfree(p.w); dfree(p.flv);
    p = PartF64{};
}

int alloc_parts(PartF64 &p, long long cap)
{
    double **a[7] = {&p.q, &p.x, &p.y, &p.z, &p.u, &p.v, &p.w};
    for (auto x : a) HIPTRY(dalloc(x, cap));
    HIPTRY(dalloc(&p.flv, cap));
    return PICTOR_OK;
}

void free_parts(PartCylS &p)
{
    dfree(p.cell); dfree(p.ox); dfree(p.oz); dfree(p.th); dfree(p.u); dfree(p.v); dfree(p.w); dfree(p.q);
    p = PartCylS{};
}

int alloc_parts(PartCylS &p, long long cap)
{
    float **a[6] = {&p.ox, &p.oz, &p.u, &p.v, &p.w, &p.q};
    for (auto x : a) HIPTRY(dalloc(x, cap));
    HIPTRY(dalloc(&p.cell, cap));
    HIPTRY(dalloc(&p.th, cap));
    return PICTOR_OK;
}

bool split_form(const pictor_ctx *c) { return c->form == PICTOR_FORM_FP32_SPLIT; }

void free_particles(pictor_ctx *c)
{
    free_parts(c->pcyl);
    free_parts(c->cm.ps);
}

// order-keeping compaction of `rec` (n slots) into `out`; returns the live count
template <class R>
int compact_into(pictor_ctx *c, const R &rec, long long n, const R &out, long long &live)
{
    const long long nb = (n + kSpan - 1) / kSpan;
    int *d_cnt = nullptr;
    long long *d_off = nullptr;
    HIPTRY(dalloc(&d_cnt, nb));
    HIPTRY(dalloc(&d_off, nb));
    hipLaunchKernelGGL(k_cylm_span_count<R>, dim3((unsigned)nb), dim3(256), 0, c->stream, rec, n, d_cnt);
    std::vector<int> h(nb);
    HIPTRY(hipMemcpyAsync(h.data(), d_cnt, nb * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    std::vector<long long> off(nb);
    long long a = 0;
    for (long long b = 0; b < nb; b++) { off[b] = a; a += h[b]; }
    HIPTRY(hipMemcpyAsync(d_off, off.data(), nb * sizeof(long long), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_cylm_compact_spans<R>, dim3((unsigned)nb), dim3(256), 0, c->stream, rec, n, d_off, out);
    HIPTRY(hipGetLastError());
    HIPTRY(hipStreamSynchronize(c->stream));
    dfree(d_cnt);
    dfree(d_off);
    live = a;
    return PICTOR_OK;
}

// room for `extra` more particles: compact first (order kept), then grow (x1.25) if short
template <class P>
int ensure_room_rec(pictor_ctx *c, P &cur, long long extra)
{
    CylmState &st = c->cm;
    if (c->np + extra <= st.cap) return PICTOR_OK;
    const long long cap2 = std::max<long long>(st.cap, (long long)((c->np + extra) * 1.25) + 1024);
    P O{};
    int rc = alloc_parts(O, cap2);
    if (rc) return rc;
    long long live = 0;
    if (c->np > 0) {
        using R = typename std::conditional<std::is_same<P, PartF64>::value, RecF64, RecSplit>::type;
        if ((rc = compact_into(c, R{cur}, c->np, R{O}, live))) { free_parts(O); return rc; }
    }
    free_parts(cur);
    cur = O;
    st.cap = cap2;
    c->np = live;
    return PICTOR_OK;
}

int ensure_room(pictor_ctx *c, long long extra)
{
    return split_form(c) ? ensure_room_rec(c, c->cm.ps, extra) : ensure_room_rec(c, c->pcyl, extra);
}

// ReorderPrtlArr (mem_prtl.F90:166-301) for the modal particles: keys k * 4096 + i of the
// cell (k-major, i fastest: the loader's order), all ones for holes; a stable radix sort
// of (key, index) pairs, then every array gathered through the permutation
template <class R>
__global__ void k_cylm_sort_keys(R P, long long n, int ib, uint32_t *key, uint32_t *idx)
{
    const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) {
        key[m] = P.live(m) ? P.sort_key(m, ib) : 0xFFFFFFFFu;
        idx[m] = (uint32_t)m;
    }
}

// the live count: the holes' keys (all ones) sort last; one thread finds the first
__global__ void k_cylm_first_hole(const uint32_t *__restrict__ key, long long n, unsigned long long *out)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = lo + (hi - lo) / 2;
        if (key[mid] == 0xFFFFFFFFu) hi = mid;
        else lo = mid + 1;
    }
    *out = (unsigned long long)lo;
}

template <class R>
__global__ void k_cylm_permute(R P, const uint32_t *__restrict__ idx, long long live, R O)
{
    const long long d = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (d < live) P.copy_to(idx[d], O, d);
}

PartF64 &spare_of(CylmState &st, PartF64 *) { return st.spare64; }
PartCylS &spare_of(CylmState &st, PartCylS *) { return st.spare32; }
template <class P>
P &spare_of(CylmState &st) { return spare_of(st, (P *)nullptr); }

template <class P>
int sort_rec(pictor_ctx *c, P &cur)
{
    using R = typename std::conditional<std::is_same<P, PartF64>::value, RecF64, RecSplit>::type;
    CylmState &st = c->cm;
    const long long n = c->np;
    if (n == 0) return PICTOR_OK;
    if (n > 0x7FFFFFFFLL) return fail(PICTOR_ERR_ARG, "modal sort: more than 2^31 particle slots");
    // key = k << ib | i over ib = ceil(log2 mx) bits; holes (all ones) sort last with end_bit
    int ib = 1;
    while ((1 << ib) < c->g.mx) ib++;
    int end_bit = ib + 1;
    while ((1LL << end_bit) <= ((long long)c->g.mz << ib)) end_bit++;
    if (end_bit > 32) return fail(PICTOR_ERR_ARG, "modal sort: grid too large for 32-bit keys");
    end_bit = std::min(32, end_bit + 1);
    if (st.sbuf_cap < n) {
        for (int a = 0; a < 2; a++) { dfree(st.skey[a]); dfree(st.sidx[a]); st.skey[a] = st.sidx[a] = nullptr; }
        st.sbuf_cap = 0;
        const long long cap = std::max<long long>(n, st.cap);
        for (int a = 0; a < 2; a++)
            if (dalloc(&st.skey[a], cap) != hipSuccess || dalloc(&st.sidx[a], cap) != hipSuccess)
                return fail(PICTOR_ERR_HIP, "modal sort: key allocation failed");
        st.sbuf_cap = cap;
    }
    size_t need = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, need, st.skey[0], st.skey[1], st.sidx[0], st.sidx[1], (int)n, 0,
                                           end_bit, c->stream) != hipSuccess)
        return fail(PICTOR_ERR_HIP, "modal sort: radix sort sizing failed");
    if (st.stmp_bytes < need) {
        dfree(st.stmp);
        st.stmp = nullptr;
        st.stmp_bytes = 0;
        if (hipMalloc(&st.stmp, need) != hipSuccess) return fail(PICTOR_ERR_HIP, "modal sort: scratch allocation failed");
        st.stmp_bytes = need;
    }
    P &O = spare_of<P>(st);
    if (st.spare_cap != st.cap) {
        free_parts(O);
        st.spare_cap = 0;
        int rc = alloc_parts(O, st.cap);
        if (rc) return rc;
        st.spare_cap = st.cap;
    }
    hipLaunchKernelGGL(k_cylm_sort_keys<R>, blocks_of(n), dim3(256), 0, c->stream, R{cur}, n, ib, st.skey[0],
                       st.sidx[0]);
    size_t tb = st.stmp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(st.stmp, tb, st.skey[0], st.skey[1], st.sidx[0], st.sidx[1], (int)n, 0,
                                           end_bit, c->stream) != hipSuccess)
        return fail(PICTOR_ERR_HIP, "modal sort: radix sort failed");
    hipLaunchKernelGGL(k_cylm_first_hole, dim3(1), dim3(64), 0, c->stream, st.skey[1], n, st.d_count);
    unsigned long long live = 0;
    HIPTRY(hipMemcpyAsync(&live, st.d_count, sizeof(live), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    if (live > 0)
        hipLaunchKernelGGL(k_cylm_permute<R>, blocks_of((long long)live), dim3(256), 0, c->stream, R{cur}, st.sidx[1],
                           (long long)live, R{O});
    HIPTRY(hipGetLastError());
    std::swap(cur, O);   // the old record becomes the spare (same capacity)
    c->np = (long long)live;
    st.steps_since_sort = 0;
    return PICTOR_OK;
}

}  // namespace

void cylm_params(const pictor_config &cfg, const PhysP &ph, CylmP &q)
{
    q.dr = cfg.dr > 0.0 ? cfg.dr : 1.0;
    q.inv_dr = 1.0 / q.dr;
    q.inv_dz = 1.0;
    q.c = ph.c; q.fldc = ph.fldc; q.fld_halfc = ph.fld_halfc; q.cinv = ph.cinv; q.sqc = ph.sqc;
    q.qi = ph.qi;
    q.flvrqm[0] = ph.flvrqm[0];
    q.flvrqm[1] = ph.flvrqm[1];
    q.zmin = 3.0; q.zmax = 3.0 + cfg.nz; q.zlen = cfg.nz;
    q.zopen = 0;
    // Lehe coefficients (cyl_em_update_lehe.F90:6-17): beta_rz = beta_thetaz = 1/8,
    // d_z = (1 - sin^2(pi c / 2) / c^2) / 4; no r pair, no theta pair in gradZ
    q.lehe = cfg.fdtd == PICTOR_FDTD_LEHE ? 1 : 0;
    const double brz = 0.125, btz = 0.125;
    const double sn = std::sin(0.5 * kPi * ph.c);
    const double dz = 0.25 * (1.0 - (ph.cinv * ph.cinv) * (sn * sn));
    q.l_rr = 1.0 - 2.0 * brz; q.l_rz = brz;
    q.l_tt = 1.0 - 2.0 * btz; q.l_tz = btz;
    q.l_zz = 1.0 - 3.0 * dz; q.l_zd = dz;
}

}  // namespace pictor

using namespace pictor;

int cylm_setup(pictor_ctx *c)
{
    CylmState &st = c->cm;
    if (!st.d_err) {
        HIPTRY(dalloc(&st.d_err, 1));
        HIPTRY(dalloc(&st.d_count, 1));
        HIPTRY(dalloc(&st.d_partial, 3 * 1024));
        HIPTRY(hipMemsetAsync(st.d_err, 0, sizeof(int), c->stream));
    }
    return PICTOR_OK;
}

void cylm_release(pictor_ctx *c)
{
    CylmState &st = c->cm;
    free_particles(c);
    free_parts(st.spare64);
    free_parts(st.spare32);
    for (int a = 0; a < 2; a++) { dfree(st.skey[a]); dfree(st.sidx[a]); }
    dfree(st.stmp);
    dfree(st.d_err); dfree(st.d_count); dfree(st.d_partial);
    dfree(st.nf4); dfree(st.nf2);
    st = CylmState{};
    c->np = 0;
}

int cylm_upload_particles(pictor_ctx *c, long long n, const double *q, const double *x, const double *y,
                          const double *z, const double *u, const double *v, const double *w, const int *flv)
{
    const GridP &g = c->g;
    std::vector<long long> keep;
    keep.reserve(n);
    for (long long i = 0; i < n; i++) {
        if (flv[i] == 0) continue;
        if (flv[i] != 1 && flv[i] != 2) return fail(PICTOR_ERR_ARG, "flavour must be 0, 1 or 2");
        if (!(x[i] >= 3.5 && x[i] < g.mx - 3)) return fail(PICTOR_ERR_ARG, "particle outside 3.5 <= x < mx-3");
        if (!(y[i] >= 0.0 && y[i] < 2.0 * kPi)) return fail(PICTOR_ERR_ARG, "theta (y) must lie in [0, 2 pi)");
        if (!(z[i] >= 3.0 && z[i] <= g.mz - 2)) return fail(PICTOR_ERR_ARG, "particle outside 3 <= z <= mz-2");
        keep.push_back(i);
    }
    const long long m = (long long)keep.size();
    free_particles(c);
    c->cm.cap = 0;
    c->np = 0;
    int rc = cylm_setup(c);
    if (rc) return rc;
    if ((rc = ensure_room(c, m))) return rc;
    if (split_form(c)) {
        // (cell, FP32 offset) per coordinate, theta in 2^-32 turns, FP32 momenta and weight
        PartCylS &p = c->cm.ps;
        std::vector<uint32_t> cw(std::max<long long>(m, 1)), tw(std::max<long long>(m, 1));
        std::vector<float> f[6];
        for (auto &v : f) v.resize(std::max<long long>(m, 1));
        for (long long e = 0; e < m; e++) {
            const long long j = keep[e];
            int ii, kk;
            split_coord(x[j], ii, f[0][e]);
            split_coord(z[j], kk, f[1][e]);
            cw[e] = pack_cell(ii, kk, flv[j]);
            tw[e] = theta_turns(y[j]);
            f[2][e] = (float)u[j]; f[3][e] = (float)v[j]; f[4][e] = (float)w[j]; f[5][e] = (float)q[j];
        }
        float *dst[6] = {p.ox, p.oz, p.u, p.v, p.w, p.q};
        for (int a = 0; a < 6; a++) HIPTRY(hipMemcpy(dst[a], f[a].data(), m * sizeof(float), hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(p.cell, cw.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIPTRY(hipMemcpy(p.th, tw.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
        c->np = m;
        return PICTOR_OK;
    }
    PartF64 &p = c->pcyl;
    double *arr[7] = {p.q, p.x, p.y, p.z, p.u, p.v, p.w};
    const double *src[7] = {q, x, y, z, u, v, w};
    std::vector<double> buf(std::max<long long>(m, 1));
    for (int a = 0; a < 7; a++) {
        for (long long e = 0; e < m; e++) buf[e] = src[a][keep[e]];
        HIPTRY(hipMemcpy(arr[a], buf.data(), m * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<int> fb(std::max<long long>(m, 1));
    for (long long e = 0; e < m; e++) fb[e] = flv[keep[e]];
    HIPTRY(hipMemcpy(p.flv, fb.data(), m * sizeof(int), hipMemcpyHostToDevice));
    c->np = m;
    return PICTOR_OK;
}

int cylm_b_half(pictor_ctx *c)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    if (c->cq.lehe) {
        constexpr int KC = 8;
        const long long nb = (long long)(g.mx - 4) * 3 * ((g.mz - 3 + KC - 1) / KC);
        hipLaunchKernelGGL((k_cylm_update_b_lehe_col<KC>), blocks_of(nb), dim3(256), 0, c->stream, g, c->cq, B[0],
                           B[1], B[2], E[0], E[1], E[2]);
    } else {
        const long long nb = (long long)(g.mx - 4) * 3 * (g.mz - 1);
        hipLaunchKernelGGL(k_cylm_update_b, blocks_of(nb), dim3(256), 0, c->stream, g, c->cq, B[0], B[1], B[2], E[0],
                           E[1], E[2]);
    }
    hipLaunchKernelGGL(k_cylm_b_axis, blocks_of(3LL * g.mz), dim3(256), 0, c->stream, g, c->cq, B[0], B[1], B[2], E[1]);
    halo(c, B);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cylm_e_phase(pictor_ctx *c, bool curl, bool add)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    const long long ne = (long long)(g.mx - 6) * 3 * (g.mz - 5);
    if (curl)
        hipLaunchKernelGGL(k_cylm_update_e, blocks_of(ne), dim3(256), 0, c->stream, g, c->cq, E[0], E[1], E[2], B[0],
                           B[1], B[2]);
    if (add)
        hipLaunchKernelGGL(k_cylm_add_current, blocks_of(ne), dim3(256), 0, c->stream, g, c->cq, E[0], E[1], E[2],
                           J[0], J[1], J[2]);
    hipLaunchKernelGGL(k_cylm_e_axis, blocks_of(3LL * g.mz), dim3(256), 0, c->stream, g, E[0], E[1], E[2]);
    halo(c, E);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cylm_current_halo(pictor_ctx *c)
{
    const GridP &g = c->g;
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    hipLaunchKernelGGL(k_cylm_axis_current, blocks_of(3LL * g.mz), dim3(256), 0, c->stream, g, J[0], J[1], J[2]);
    if (!c->cq.zopen) {   // open z: the ghost rows' current leaves the box with its particles
        hipLaunchKernelGGL(k_cylm_zfold, blocks_of(9LL * g.mx), dim3(256), 0, c->stream, g, 0, J[0], J[1], J[2]);
        hipLaunchKernelGGL(k_cylm_zfold, blocks_of(6LL * g.mx), dim3(256), 0, c->stream, g, 1, J[0], J[1], J[2]);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cylm_field_halo(pictor_ctx *c, int which)
{
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    if (which == 0) {
        hipLaunchKernelGGL(k_cylm_e_axis, blocks_of(3LL * c->g.mz), dim3(256), 0, c->stream, c->g, E[0], E[1], E[2]);
        halo(c, E);
    } else {
        halo(c, B);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cylm_reset_current(pictor_ctx *c)
{
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    launch_fill_zero_n<double>(J, 3, c->g.ncell, c->stream);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

int cylm_push(pictor_ctx *c, bool timed)
{
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    if (timed) HIPTRY(hipEventRecord(c->ev[0], c->stream));
    if (c->np > 0 && split_form(c)) {
        CylmState &st = c->cm;
        const long long nodes = (long long)c->g.mx * c->g.mz;
        if (!st.nf4) {
            HIPTRY(dalloc(&st.nf4, 4 * nodes));
            HIPTRY(dalloc(&st.nf2, nodes));
        }
        hipLaunchKernelGGL(k_cylm_nodes_f32, blocks_of(nodes), dim3(256), 0, c->stream, c->g, E[0], E[1], E[2], B[0],
                           B[1], B[2], st.nf4, st.nf2);
        hipLaunchKernelGGL(k_cylm_push_split, blocks_of(c->np), dim3(256), 0, c->stream, c->g, c->cq, st.nf4, st.nf2,
                           J[0], J[1], J[2], st.ps, c->np, st.d_err);
    } else if (c->np > 0)
        hipLaunchKernelGGL(k_cylm_push, blocks_of(c->np), dim3(256), 0, c->stream, c->g, c->cq, E[0], E[1], E[2], B[0],
                           B[1], B[2], J[0], J[1], J[2], c->pcyl, c->np, c->cm.d_err);
    if (timed) HIPTRY(hipEventRecord(c->ev[1], c->stream));
    HIPTRY(hipGetLastError());
    c->pending_swap = true;
    return PICTOR_OK;
}

int cylm_exchange(pictor_ctx *c)
{
    int err = 0;
    HIPTRY(hipMemcpyAsync(&err, c->cm.d_err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    c->pending_swap = false;
    // the flags report this push only: clear them for the next one
    if (err) HIPTRY(hipMemsetAsync(c->cm.d_err, 0, sizeof(int), c->stream));
    if (err & 1) return fail(PICTOR_ERR_BOUNDS, "modal cylinder: a particle reached the outer wall");
    if (err & 2)
        return fail(PICTOR_ERR_CAPACITY, "modal cylinder, split form: a current contribution above 64 charge cells "
                                         "(the fixed-point range); use the FP64 form");
    return PICTOR_OK;
}

int cylm_inject(pictor_ctx *c, int nrp, int k0, int nzc, int ppc, unsigned long long seed, long long shift)
{
    const GridP &g = c->g;
    if (nrp < 1 || nrp > g.nx - 1 || ppc < 1 || nzc < 1 || k0 < 3 || k0 + nzc - 1 > g.mz - 3)
        return fail(PICTOR_ERR_ARG, "plasma cells outside the grid (4 <= i < 4 + nr_plasma <= mx - 3, 3 <= k <= mz - 3)");
    int rc = cylm_setup(c);
    if (rc) return rc;
    const long long add = (long long)nrp * nzc * ppc;
    if ((rc = ensure_room(c, add))) return rc;
    if (split_form(c))
        hipLaunchKernelGGL(k_cylm_inject_split, blocks_of(add), dim3(256), 0, c->stream, g, c->cq, c->cm.ps, c->np, nrp,
                           k0, nzc, ppc, c->cfg.epc, seed, shift);
    else
        hipLaunchKernelGGL(k_cylm_inject, blocks_of(add), dim3(256), 0, c->stream, g, c->cq, c->pcyl, c->np, nrp, k0,
                           nzc, ppc, c->cfg.epc, seed, shift);
    HIPTRY(hipGetLastError());
    c->np += add;
    return PICTOR_OK;
}

int cylm_laser(pictor_ctx *c, double a0, double omega0, double tau, double w0, double zc)
{
    // code units: one axial cell = dz; omega_p dt = ompe, c cells per step
    const double cc = c->ph.c, ompe = cc / c->cfg.c_ompe;
    const double w0dt = omega0 * ompe;                 // laser phase advance per step
    const double E0 = a0 * cc * w0dt;
    const double k0 = w0dt / cc;                       // rad per axial cell
    const double cells_per_skin = cc / ompe;           // c / omega_p in axial cells
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    hipLaunchKernelGGL(k_cylm_laser, blocks_of((long long)c->g.mx * c->g.mz), dim3(256), 0, c->stream, c->g, c->cq, E0,
                       k0, w0 * cells_per_skin, tau * cells_per_skin, zc, E[0], E[1], B[0], B[1]);
    HIPTRY(hipGetLastError());
    int rc = cylm_field_halo(c, 0);
    return rc ? rc : cylm_field_halo(c, 1);
}

// after a step: advance the window by whole cells at c (one cell per 1/c steps)
// after a step: the window's shifts, then the sort by cell when one is due
static int cylm_window_shifts(pictor_ctx *c);

int cylm_window_after_step(pictor_ctx *c)
{
    CylmState &st = c->cm;
    int rc = cylm_window_shifts(c);
    if (!rc && st.sort_period > 0 && ++st.steps_since_sort >= st.sort_period) rc = cylm_sort(c);
    return rc;
}

static int cylm_window_shifts(pictor_ctx *c)
{
    CylmState &st = c->cm;
    if (!st.win_on) return PICTOR_OK;
    st.win_steps += 1;
    while ((double)st.win_steps * c->ph.c >= (double)(st.win_shift + 1)) {
        // E, then B, is shifted into the J buffers, which then swap with it; J ends up
        // holding the old B buffers and is cleared (the next step's ResetCurrent clears it too)
        for (int f = 0; f < 2; f++) {
            void **F = f == 0 ? c->E : c->B;
            Ptrs6<double> P{{(double *)F[0], (double *)F[1], (double *)F[2], (double *)c->J[0], (double *)c->J[1],
                             (double *)c->J[2]}};
            hipLaunchKernelGGL(k_cylm_shift_fields, dim3(nblocks(c->g.ncell, 256), 3), dim3(256), 0, c->stream, c->g, P);
            for (int i = 0; i < 3; i++) std::swap(F[i], c->J[i]);
        }
        double *E[3], *B[3], *J[3];
        fields(c, E, B, J);
        launch_fill_zero_n<double>(J, 3, c->g.ncell, c->stream);
        if (c->np > 0 && split_form(c))
            hipLaunchKernelGGL(k_cylm_shift_split, blocks_of(c->np), dim3(256), 0, c->stream, c->cq, c->cm.ps, c->np);
        else if (c->np > 0)
            hipLaunchKernelGGL(k_cylm_shift_particles, blocks_of(c->np), dim3(256), 0, c->stream, c->cq, c->pcyl, c->np);
        st.win_shift += 1;
        int rc = PICTOR_OK;
        if (st.win_nrp > 0 && st.win_ppc > 0 &&
            (rc = cylm_inject(c, st.win_nrp, c->g.mz - 3, 1, st.win_ppc, st.win_seed, st.win_shift)))
            return rc;
        if ((rc = cylm_field_halo(c, 0)) || (rc = cylm_field_halo(c, 1))) return rc;
    }
    return PICTOR_OK;
}

int cylm_diagnostics(pictor_ctx *c, double out[9])
{
    double *E[3], *B[3], *J[3];
    fields(c, E, B, J);
    CylmState &st = c->cm;
    int rc = cylm_setup(c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cylm_energy, dim3(1024), dim3(256), 0, c->stream, c->g, c->cq, E[0], E[1], E[2], B[0], B[1],
                       B[2], st.d_partial);
    std::vector<double> h(3 * 1024);
    HIPTRY(hipMemcpyAsync(h.data(), st.d_partial, 2 * 1024 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    double be = 0, ee = 0;
    for (int b = 0; b < 1024; b++) { be += h[2 * b]; ee += h[2 * b + 1]; }
    double k1 = 0, k2 = 0, cnt = 0;
    if (c->np > 0) {
        if (split_form(c))
            hipLaunchKernelGGL(k_cylm_kinetic<RecSplit>, dim3(1024), dim3(256), 0, c->stream, RecSplit{c->cm.ps}, c->np,
                               st.d_partial);
        else
            hipLaunchKernelGGL(k_cylm_kinetic<RecF64>, dim3(1024), dim3(256), 0, c->stream, RecF64{c->pcyl}, c->np,
                               st.d_partial);
        HIPTRY(hipMemcpyAsync(h.data(), st.d_partial, 3 * 1024 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));
        for (int b = 0; b < 1024; b++) { k1 += h[3 * b]; k2 += h[3 * b + 1]; cnt += h[3 * b + 2]; }
    }
    for (int i = 0; i < 9; i++) out[i] = 0.0;
    out[0] = be; out[1] = ee; out[2] = k1; out[3] = k2;
    (void)cnt;
    return PICTOR_OK;
}

int cylm_download_particles(pictor_ctx *c, long long cap, double *q, double *x, double *y, double *z, double *u,
                            double *v, double *w, int *flv, int *tile, long long *n_out)
{
    const long long n = c->np;
    if (n_out) *n_out = n;
    if (n > cap) return fail(PICTOR_ERR_ARG, "output arrays too small");
    HIPTRY(hipStreamSynchronize(c->stream));
    if (n == 0) return PICTOR_OK;
    if (tile) std::fill(tile, tile + n, 0);
    if (!split_form(c)) {
        const PartF64 &p = c->pcyl;
        double *hd[7] = {q, x, y, z, u, v, w};
        double *dd[7] = {p.q, p.x, p.y, p.z, p.u, p.v, p.w};
        for (int i = 0; i < 7; i++)
            if (hd[i]) HIPTRY(hipMemcpy(hd[i], dd[i], n * sizeof(double), hipMemcpyDeviceToHost));
        if (flv) HIPTRY(hipMemcpy(flv, p.flv, n * sizeof(int), hipMemcpyDeviceToHost));
        return PICTOR_OK;
    }
    // the split record in the FP64 layout: x = i + ox, z = k + oz (exact), theta = turns * 2 pi / 2^32
    const PartCylS &p = c->cm.ps;
    std::vector<uint32_t> cw(n), tw(n);
    std::vector<float> f[6];
    const float *src[6] = {p.ox, p.oz, p.u, p.v, p.w, p.q};
    for (int a = 0; a < 6; a++) {
        f[a].resize(n);
        HIPTRY(hipMemcpy(f[a].data(), src[a], n * sizeof(float), hipMemcpyDeviceToHost));
    }
    HIPTRY(hipMemcpy(cw.data(), p.cell, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPTRY(hipMemcpy(tw.data(), p.th, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (long long e = 0; e < n; e++) {
        const uint32_t wd = cw[e];
        const bool live = wd != 0u;
        if (x) x[e] = live ? cell_i(wd) + (double)f[0][e] : 0.0;
        if (z) z[e] = live ? cell_k(wd) + (double)f[1][e] : 0.0;
        if (y) y[e] = (double)tw[e] * kTurnRad;
        if (u) u[e] = f[2][e];
        if (v) v[e] = f[3][e];
        if (w) w[e] = f[4][e];
        if (q) q[e] = f[5][e];
        if (flv) flv[e] = live ? cell_flv(wd) : 0;
    }
    return PICTOR_OK;
}

int cylm_sort(pictor_ctx *c)
{
    int rc = cylm_setup(c);
    if (rc) return rc;
    return split_form(c) ? sort_rec(c, c->cm.ps) : sort_rec(c, c->pcyl);
}
