// push_common.h -- device helpers shared by the 2-D and 3-D push kernels.
#pragma once
#include "pictor_common.h"

namespace pictor {
namespace pushdev {

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double dmax(double a, double b) { return (b > a) ? b : a; }

__device__ inline unsigned long long lanemask_lt()
{
    unsigned lane = __lane_id();
    return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Bin of position k (0 <= k < sStart[n]) among n bins laid end to end: the last b < n with
// sStart[b] <= k (an empty bin shares its start with the next).  A search of fixed depth (5 or 6
// LDS reads) for the 27-direction bins of the 3-D layout, where a linear walk from the first bin
// takes up to 53 steps; the 2-D kernels, whose centre bins come first, keep the walk.
template <int NMAX>
__device__ inline int bin_of(int k, const int *sStart, int n)
{
    constexpr int top = NMAX > 32 ? 32 : NMAX > 16 ? 16 : NMAX > 8 ? 8 : NMAX > 4 ? 4 : NMAX > 2 ? 2 : 1;
    static_assert(NMAX <= 64, "bin search depth");
    int b = 0;
#pragma unroll
    for (int step = top; step >= 1; step >>= 1) {
        const int c = b + step;
        const int sv = sStart[c < n ? c : n - 1];
        b = ((c < n) & (sv <= k)) ? c : b;
    }
    return b;
}

// Per-wave reservation of consecutive slots in the LDS counter of each
// destination; every lane of the wave must call it (d = -1 for idle lanes).
__device__ inline int reserve_slot(int d, int *sCnt)
{
    const bool act = d >= 0;
    unsigned long long pending = __ballot(act);
    int slot = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int ld = __shfl(d, leader);
        const unsigned long long m = __ballot(act && d == ld);
        int b = 0;
        if ((int)__lane_id() == leader) b = atomicAdd(&sCnt[ld], __popcll(m));
        b = __shfl(b, leader);
        if (act && d == ld) slot = b + __popcll(m & lanemask_lt());
        pending &= ~m;
    }
    return slot;
}

__device__ inline int wrap_tile(int t, int n)
{
    return (t < 0) ? t + n : (t >= n ? t - n : t);
}

__device__ inline int norm_dir(int dd, int n)
{
    if (dd > 1) dd -= n;
    else if (dd < -1) dd += n;
    return dd;
}

// FP32 -> exact integer count of 2^-30 units (the argument is already scaled by 2^30;
// |v| < 2^31, so one conversion)
__device__ inline long long quanti(float v) { return (long long)(int)__builtin_rintf(v); }

// The two zigzag half-segments of the 2-D split form (movdep.F90:177-237)
// from the old offsets (x0, y0, z0) to the new position as stored: offsets
// (ox, oy) in the cell displaced by (di, dj), z advanced to z1.  Evaluated in
// FP64, every relay point, flux F and weight W is exact for FP32 inputs and
// each node value is rounded once before its rounding to a 2^-30 count, so
// the deposited Jx, Jy carry exactly the charge moved between the stored
// positions up to that last rounding.  qs: charge times 2^30.  a: the first
// half-segment (nodes of the old cell), b: the second (nodes of the new cell),
// each Jx (i,j), (i,j+1); Jy (i,j), (i+1,j); Jz (i,j), (i+1,j), (i,j+1), (i+1,j+1).
__device__ inline void zigzag2d_split(double qs, float x0, float y0, float z0, float ox, float oy, float z1, int di,
                                      int dj, double a[8], double b[8])
{
    const double X0 = x0, Y0 = y0, Z0 = z0;
    const double X1 = (double)ox + (double)di, Y1 = (double)oy + (double)dj, Z1 = z1;
    const int dk = (int)floor(Z1);
    const double xr = dmin((double)(min(0, di) + 1), dmax((double)max(0, di), 0.5 * (X0 + X1)));
    const double yr = dmin((double)(min(0, dj) + 1), dmax((double)max(0, dj), 0.5 * (Y0 + Y1)));
    const double zr = dmin((double)(min(0, dk) + 1), dmax((double)max(0, dk), 0.5 * (Z0 + Z1)));
    {
        const double Fx = qs * (xr - X0), Fy = qs * (yr - Y0), Fz = qs * (zr - Z0);
        const double Wx = 0.5 * (X0 + xr), Wy = 0.5 * (Y0 + yr);
        const double zl = Fz * (1.0 - Wx), zh = Fz * Wx;
        a[0] = rint(Fx * (1.0 - Wy)); a[1] = rint(Fx * Wy);
        a[2] = rint(Fy * (1.0 - Wx)); a[3] = rint(Fy * Wx);
        a[4] = rint(zl * (1.0 - Wy)); a[5] = rint(zh * (1.0 - Wy));
        a[6] = rint(zl * Wy); a[7] = rint(zh * Wy);
    }
    {
        const double Fx = qs * (X1 - xr), Fy = qs * (Y1 - yr), Fz = qs * (Z1 - zr);
        const double Wx = 0.5 * (X1 + xr) - (double)di, Wy = 0.5 * (Y1 + yr) - (double)dj;
        const double zl = Fz * (1.0 - Wx), zh = Fz * Wx;
        b[0] = rint(Fx * (1.0 - Wy)); b[1] = rint(Fx * Wy);
        b[2] = rint(Fy * (1.0 - Wx)); b[3] = rint(Fy * Wx);
        b[4] = rint(zl * (1.0 - Wy)); b[5] = rint(zh * (1.0 - Wy));
        b[6] = rint(zl * Wy); b[7] = rint(zh * Wy);
    }
}

// EnforceBC_PostMovDep (bc.F90:238-287) on one pushed particle, in the
// reference's order: the reflecting faces top, bottom, right, left
// (RefBC_Prtl_*, bc_reflect.F90:76-186: specular image 2 x_w - x with the
// normal momentum reversed, only where the wall lies inside the array), then
// every face again in order x lo .. z hi, reflecting once more (which changes
// nothing then) or removing the particle beyond a 'remv' face (bc_open.F90).
// Positions are the reference's local coordinates.  Returns false when the
// particle is removed.  The current of a reflected particle is folded back by
// the face kernels after the push.
template <typename R>
__device__ inline bool particle_faces(const PhysP &ph, const GridP &g, double &x, double &y, double z, R &u, R &v)
{
    const int *t = ph.pbc;
    const double *p = ph.pbc_pos;
    auto top = [&]() { if (p[3] < g.my - 1 && y > p[3]) { v = -v; y = 2.0 * p[3] - y; } };
    auto bottom = [&]() { if (p[2] > 1 && y < p[2]) { v = -v; y = 2.0 * p[2] - y; } };
    auto right = [&]() { if (p[1] < g.mx - 1 && x > p[1]) { u = -u; x = 2.0 * p[1] - x; } };
    auto left = [&]() { if (p[0] > 1 && x < p[0]) { u = -u; x = 2.0 * p[0] - x; } };
    if (t[3] == kBcRefl) top();
    if (t[2] == kBcRefl) bottom();
    if (t[1] == kBcRefl) right();
    if (t[0] == kBcRefl) left();
    if (t[0] == kBcRefl) left();
    else if (t[0] == kBcRemv && x < p[0]) return false;
    if (t[1] == kBcRefl) right();
    else if (t[1] == kBcRemv && x > p[1]) return false;
    if (t[2] == kBcRefl) bottom();
    else if (t[2] == kBcRemv && y < p[2]) return false;
    if (t[3] == kBcRefl) top();
    else if (t[3] == kBcRemv && y > p[3]) return false;
    if (t[4] == kBcRemv && z < p[4]) return false;
    if (t[5] == kBcRemv && z > p[5]) return false;
    return true;
}

// byte-offset accessors: a uniform base plus a 32-bit lane offset lets the
// compiler use SGPR-base addressing (one VALU op per access)
template <typename T>
__device__ inline T ldo(const T *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ inline void sto(T *base, unsigned off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}


}  // namespace pushdev
}  // namespace pictor
