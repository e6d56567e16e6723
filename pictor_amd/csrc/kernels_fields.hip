// kernels_fields.hip -- Yee FDTD, periodic halos and field reductions (gfx950).
//
// All kernels are HBM-streaming stencils: one thread per cell, 64-wide rows
// along x (the contiguous index of the reference layout) so that each wave
// reads whole 512-byte (FP64) or 256-byte (FP32) row segments.
//
// Expression order follows src/maxwell/em_update.F90:9-102 exactly; with
// -ffp-contract=off the FP64 results are bitwise those of the reference
// CPU loops.  The update kernels write physical cells 3..m-3 only and the
// ghost layers are then copied from their periodic images, which is what the
// reference's loop over 2..m-1 followed by SendRecvFlds leaves in memory.
#include "pictor_common.h"

namespace pictor {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 phys_grid(const GridP &g)
{
    int nz = (g.ndim == 3) ? g.nz : 1;
    return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, nz);
}

// blocks over the cells yee_lo..yee_hi of every axis
inline dim3 yee_grid(const GridP &g)
{
    const int ex = yee_hi(g, 0) - yee_lo(g, 0) + 1, ey = yee_hi(g, 1) - yee_lo(g, 1) + 1;
    const int ez = (g.ndim == 3) ? yee_hi(g, 2) - yee_lo(g, 2) + 1 : 1;
    return dim3((ex + BX - 1) / BX, (ey + BY - 1) / BY, ez);
}

// em_update.F90:41-77 (B half step): the physical cells along periodic axes,
// 2..m-1 along non-periodic ones (yee_lo..yee_hi)
template <typename R>
__global__ __launch_bounds__(BX *BY) void k_update_b_half(GridP g, R hc, R idx_, R idy_, R idz_, const R *__restrict__ Ex,
                                                          const R *__restrict__ Ey, const R *__restrict__ Ez,
                                                          R *__restrict__ Bx, R *__restrict__ By, R *__restrict__ Bz)
{
    int i = yee_lo(g, 0) + blockIdx.x * BX + threadIdx.x;
    int j = yee_lo(g, 1) + blockIdx.y * BY + threadIdx.y;
    int k = (g.ndim == 3) ? yee_lo(g, 2) + (int)blockIdx.z : 1;
    if (i > yee_hi(g, 0) || j > yee_hi(g, 1)) return;
    long long o = fidx(g, i, j, k);
    long long oi = o + 1, oj = o + g.mx, ok = o + (long long)g.mx * g.my;
    if (g.ndim == 2) {
        Bx[o] = Bx[o] - hc * idy_ * (Ez[oj] - Ez[o]);
        By[o] = By[o] + hc * idx_ * (Ez[oi] - Ez[o]);
        Bz[o] = Bz[o] - hc * idx_ * (Ey[oi] - Ey[o]) + hc * idy_ * (Ex[oj] - Ex[o]);
    } else {
        Bx[o] = Bx[o] - hc * idy_ * (Ez[oj] - Ez[o]) + hc * idz_ * (Ey[ok] - Ey[o]);
        By[o] = By[o] - hc * idz_ * (Ex[ok] - Ex[o]) + hc * idx_ * (Ez[oi] - Ez[o]);
        Bz[o] = Bz[o] - hc * idx_ * (Ey[oi] - Ey[o]) + hc * idy_ * (Ex[oj] - Ex[o]);
    }
}

// Lehe B half step, em_update_lehe.F90:88-155: extended x, y (and z)
// gradients with the reference's coefficients and term order; cells 3..m-3.
template <typename R>
struct LeheC {
    R ax, ay, az, bxy, bxz, byx, byz, bzx, bzy, dx, dy, dz;
};

template <typename R>
__device__ inline R lgx(const GridP &g, const LeheC<R> &L, R d, const R *F, long long o)
{
    const long long sy = g.mx, sz = (long long)g.mx * g.my;
    R v = L.ax * d * (F[o + 1] - F[o]) + L.bxy * d * (F[o + 1 + sy] - F[o + sy]) +
          L.bxy * d * (F[o + 1 - sy] - F[o - sy]);
    if (g.ndim == 3) v = v + L.bxz * d * (F[o + 1 + sz] - F[o + sz]) + L.bxz * d * (F[o + 1 - sz] - F[o - sz]);
    return v + L.dx * d * (F[o + 2] - F[o - 1]);
}
template <typename R>
__device__ inline R lgy(const GridP &g, const LeheC<R> &L, R d, const R *F, long long o)
{
    const long long sy = g.mx, sz = (long long)g.mx * g.my;
    R v = L.ay * d * (F[o + sy] - F[o]) + L.byx * d * (F[o + 1 + sy] - F[o + 1]) +
          L.byx * d * (F[o - 1 + sy] - F[o - 1]);
    if (g.ndim == 3) v = v + L.byz * d * (F[o + sy + sz] - F[o + sz]) + L.byz * d * (F[o + sy - sz] - F[o - sz]);
    return v + L.dy * d * (F[o + 2 * sy] - F[o - sy]);
}
template <typename R>
__device__ inline R lgz(const GridP &g, const LeheC<R> &L, R d, const R *F, long long o)
{
    const long long sy = g.mx, sz = (long long)g.mx * g.my;
    return L.az * d * (F[o + sz] - F[o]) + L.bzx * d * (F[o + 1 + sz] - F[o + 1]) +
           L.bzx * d * (F[o - 1 + sz] - F[o - 1]) + L.bzy * d * (F[o + sy + sz] - F[o + sy]) +
           L.bzy * d * (F[o - sy + sz] - F[o - sy]) + L.dz * d * (F[o + 2 * sz] - F[o - sz]);
}

template <typename R>
__global__ __launch_bounds__(BX *BY) void k_update_b_half_lehe(GridP g, R hc, R idx_, R idy_, R idz_, LeheC<R> L,
                                                               const R *__restrict__ Ex, const R *__restrict__ Ey,
                                                               const R *__restrict__ Ez, R *__restrict__ Bx,
                                                               R *__restrict__ By, R *__restrict__ Bz)
{
    int i = 3 + blockIdx.x * BX + threadIdx.x;
    int j = 3 + blockIdx.y * BY + threadIdx.y;
    int k = (g.ndim == 3) ? 3 + (int)blockIdx.z : 1;
    if (i > g.mx - 3 || j > g.my - 3) return;
    long long o = fidx(g, i, j, k);
    if (g.ndim == 2) {
        Bx[o] = Bx[o] - hc * lgy(g, L, idy_, Ez, o);
        By[o] = By[o] + hc * lgx(g, L, idx_, Ez, o);
        Bz[o] = Bz[o] - hc * lgx(g, L, idx_, Ey, o) + hc * lgy(g, L, idy_, Ex, o);
    } else {
        Bx[o] = Bx[o] - hc * lgy(g, L, idy_, Ez, o) + hc * lgz(g, L, idz_, Ey, o);
        By[o] = By[o] - hc * lgz(g, L, idz_, Ex, o) + hc * lgx(g, L, idx_, Ez, o);
        Bz[o] = Bz[o] - hc * lgx(g, L, idx_, Ey, o) + hc * lgy(g, L, idy_, Ex, o);
    }
}

template <typename R, typename JT>
__device__ inline R jvalue(const JT *J, long long o, double scale);
template <>
__device__ inline double jvalue<double, double>(const double *J, long long o, double)
{
    return J[o];
}
// split form: J holds exact sums of contributions quantized to 2^-30 (charge x cell)
template <>
__device__ inline float jvalue<float, double>(const double *J, long long o, double scale)
{
    return (float)(J[o] * scale);
}

// em_update.F90:9-39 (E full step); with ADD, AddCurrent (em_update.F90:79-102)
// follows in the same thread: (E + c curl B) - J, the reference's two roundings.
// wide: the Yee range yee_lo..yee_hi; otherwise the physical cells (the Lehe
// build's UpdateEfield, em_update_lehe.F90:43-84, and AddCurrent alone)
template <typename R, typename JT, bool CURL, bool ADD>
__global__ __launch_bounds__(BX *BY) void k_update_e(GridP g, R fc, R idx_, R idy_, R idz_, const R *__restrict__ Bx,
                                                     const R *__restrict__ By, const R *__restrict__ Bz,
                                                     R *__restrict__ Ex, R *__restrict__ Ey, R *__restrict__ Ez,
                                                     const JT *__restrict__ Jx, const JT *__restrict__ Jy,
                                                     const JT *__restrict__ Jz, double jscale, int wide)
{
    const int lx = wide ? yee_lo(g, 0) : 3, ly = wide ? yee_lo(g, 1) : 3, lz = wide ? yee_lo(g, 2) : 3;
    const int ux = wide ? yee_hi(g, 0) : g.mx - 3, uy = wide ? yee_hi(g, 1) : g.my - 3;
    int i = lx + blockIdx.x * BX + threadIdx.x;
    int j = ly + blockIdx.y * BY + threadIdx.y;
    int k = (g.ndim == 3) ? lz + (int)blockIdx.z : 1;
    if (i > ux || j > uy) return;
    long long o = fidx(g, i, j, k);
    long long oi = o - 1, oj = o - g.mx, ok = o - (long long)g.mx * g.my;
    R ex = Ex[o], ey = Ey[o], ez = Ez[o];
    if (CURL) {
        if (g.ndim == 2) {
            ex = ex + fc * idy_ * (Bz[o] - Bz[oj]);
            ey = ey - fc * idx_ * (Bz[o] - Bz[oi]);
            ez = ez + fc * idx_ * (By[o] - By[oi]) - fc * idy_ * (Bx[o] - Bx[oj]);
        } else {
            ex = ex + fc * idy_ * (Bz[o] - Bz[oj]) - fc * idz_ * (By[o] - By[ok]);
            ey = ey + fc * idz_ * (Bx[o] - Bx[ok]) - fc * idx_ * (Bz[o] - Bz[oi]);
            ez = ez + fc * idx_ * (By[o] - By[oi]) - fc * idy_ * (Bx[o] - Bx[oj]);
        }
    }
    if (ADD) {
        const bool inner = i >= 3 && i <= g.mx - 3 && j >= 3 && j <= g.my - 3 &&
                           ((g.ndim == 2) || (k >= 3 && k <= g.mz - 3));
        if (inner) {
            ex = ex - jvalue<R, JT>(Jx, o, jscale) * idy_ * idz_;
            ey = ey - jvalue<R, JT>(Jy, o, jscale) * idx_ * idz_;
            ez = ez - jvalue<R, JT>(Jz, o, jscale) * idy_ * idx_;
        }
    }
    Ex[o] = ex; Ey[o] = ey; Ez[o] = ez;
}

// FP32 split form: E is kept as a pair of FP32 arrays (hi, lo) whose sum
// carries the field to about 48 bits.  The update runs in FP64: the curl
// differences of the FP32 B values are exact, so the discrete divergence of
// the curl terms cancels to FP64 rounding, and the fixed-point current
// enters exactly scaled.  Everything that reads E for the push or the B
// update reads hi, an ordinary FP32 field.
// the value a double takes after being stored as an FP32 (hi, lo) pair and read back
__device__ __forceinline__ double pair_round(double x)
{
    const float h = (float)x;
    const float l = (float)(x - (double)h);
    return (double)h + (double)l;
}

template <bool CURL, bool ADD>
__global__ __launch_bounds__(BX *BY) void k_update_e_pair(GridP g, double fc, double idx_, double idy_, double idz_,
                                                          const float *__restrict__ Bx, const float *__restrict__ By,
                                                          const float *__restrict__ Bz, float *__restrict__ Ex,
                                                          float *__restrict__ Ey, float *__restrict__ Ez,
                                                          float *__restrict__ Lx, float *__restrict__ Ly,
                                                          float *__restrict__ Lz, const double *__restrict__ Jx,
                                                          const double *__restrict__ Jy, const double *__restrict__ Jz,
                                                          double jscale, int wide)
{
    const int lx = wide ? yee_lo(g, 0) : 3, ly = wide ? yee_lo(g, 1) : 3, lz = wide ? yee_lo(g, 2) : 3;
    const int ux = wide ? yee_hi(g, 0) : g.mx - 3, uy = wide ? yee_hi(g, 1) : g.my - 3;
    int i = lx + blockIdx.x * BX + threadIdx.x;
    int j = ly + blockIdx.y * BY + threadIdx.y;
    int k = (g.ndim == 3) ? lz + (int)blockIdx.z : 1;
    if (i > ux || j > uy) return;
    long long o = fidx(g, i, j, k);
    long long oi = o - 1, oj = o - g.mx, ok = o - (long long)g.mx * g.my;
    double ex = (double)Ex[o] + (double)Lx[o], ey = (double)Ey[o] + (double)Ly[o], ez = (double)Ez[o] + (double)Lz[o];
    if (CURL) {
        const double bxo = Bx[o], byo = By[o], bzo = Bz[o];
        if (g.ndim == 2) {
            ex = ex + fc * idy_ * (bzo - (double)Bz[oj]);
            ey = ey - fc * idx_ * (bzo - (double)Bz[oi]);
            ez = ez + fc * idx_ * (byo - (double)By[oi]) - fc * idy_ * (bxo - (double)Bx[oj]);
        } else {
            ex = ex + fc * idy_ * (bzo - (double)Bz[oj]) - fc * idz_ * (byo - (double)By[ok]);
            ey = ey + fc * idz_ * (bxo - (double)Bx[ok]) - fc * idx_ * (bzo - (double)Bz[oi]);
            ez = ez + fc * idx_ * (byo - (double)By[oi]) - fc * idy_ * (bxo - (double)Bx[oj]);
        }
    }
    if (CURL && ADD) {
        // round through the stored pair between the curl and the current, as the
        // separate UpdateEfield / AddCurrent phases do: the fused step and the
        // phase-by-phase host then give the same bits
        ex = pair_round(ex); ey = pair_round(ey); ez = pair_round(ez);
    }
    if (ADD) {
        const bool inner = i >= 3 && i <= g.mx - 3 && j >= 3 && j <= g.my - 3 &&
                           ((g.ndim == 2) || (k >= 3 && k <= g.mz - 3));
        if (inner) {
            ex = ex - (Jx[o] * jscale) * idy_ * idz_;
            ey = ey - (Jy[o] * jscale) * idx_ * idz_;
            ez = ez - (Jz[o] * jscale) * idy_ * idx_;
        }
    }
    const float hx = (float)ex, hy = (float)ey, hz = (float)ez;
    Ex[o] = hx; Ey[o] = hy; Ez[o] = hz;
    Lx[o] = (float)(ex - (double)hx); Ly[o] = (float)(ey - (double)hy); Lz[o] = (float)(ez - (double)hz);
}

// FP64 field <-> FP32 (hi, lo) pair
__global__ void k_pair_from_f64(const double *__restrict__ src, float *__restrict__ hi, float *__restrict__ lo,
                                long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double v = src[t];
    const float h = (float)v;
    hi[t] = h;
    lo[t] = (float)(v - (double)h);
}

__global__ void k_pair_to_f64(const float *__restrict__ hi, const float *__restrict__ lo, double *__restrict__ dst,
                              long long n)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    dst[t] = (double)hi[t] + (double)lo[t];
}

// Ghost layers 1..2 and m-2..m take the value of their periodic image
// (comm_fld.F90:75-130 on one rank).  The exchanges run y, (z,) then x; the y
// and z faces span x = 3..mx-3, the x faces the whole cross-section, so a
// ghost is copied from the image reached by shifting x first (if x is
// periodic), then y and z while the x index is physical; all periodic: corners
// end as the doubly shifted image.  A non-periodic axis exchanges nothing.
template <typename R>
__global__ __launch_bounds__(256) void k_ghost_copy(GridP g, Ptrs6<R> F)
{
    R *__restrict__ F0 = F.p[3 * blockIdx.y];
    R *__restrict__ F1 = F.p[3 * blockIdx.y + 1];
    R *__restrict__ F2 = F.p[3 * blockIdx.y + 2];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncell) return;
    int i = (int)(t % g.mx) + 1;
    long long r = t / g.mx;
    int j = (int)(r % g.my) + 1;
    int k = (int)(r / g.my) + 1;
    int si = i, sj = j, sk = k;
    if (g.per[0]) { if (i <= 2) si = i + g.nx; else if (i >= g.mx - 2) si = i - g.nx; }
    const bool xphys = si >= 3 && si <= g.mx - 3;
    if (g.per[1] && xphys) { if (j <= 2) sj = j + g.ny; else if (j >= g.my - 2) sj = j - g.ny; }
    if (g.ndim == 3 && g.per[2] && xphys) { if (k <= 2) sk = k + g.nz; else if (k >= g.mz - 2) sk = k - g.nz; }
    if (si == i && sj == j && sk == k) return;
    long long s = fidx(g, si, sj, sk);
    F0[t] = F0[s];
    F1[t] = F1[s];
    F2[t] = F2[s];
}

// UpdateCurrentAllFaces for one axis: 3..5 += m-2..m and m-4..m-3 += 1..2.
// Axis 0 spans the full cross-section, axes 1 and 2 the x range 3..mx-3.
template <typename JT>
__global__ __launch_bounds__(256) void k_current_fold(GridP g, int axis, Ptrs6<JT> J)
{
    JT *__restrict__ J0 = J.p[3 * blockIdx.y];
    JT *__restrict__ J1 = J.p[3 * blockIdx.y + 1];
    JT *__restrict__ J2 = J.p[3 * blockIdx.y + 2];
    // target layers: 5 planes normal to `axis`
    int m[3] = {g.mx, g.my, g.mz};
    int n[3] = {g.nx, g.ny, g.nz};
    int ext[3] = {g.mx, g.my, g.mz};
    int lo0 = 1;
    if (axis != 0) { ext[0] = g.mx - 5; lo0 = 3; }
    ext[axis] = 5;
    long long total = (long long)ext[0] * ext[1] * ext[2];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int c[3];
    c[0] = (int)(t % ext[0]);
    long long r = t / ext[0];
    c[1] = (int)(r % ext[1]);
    c[2] = (int)(r / ext[1]);
    int p[3] = {c[0] + 1, c[1] + 1, c[2] + 1};
    if (axis != 0) p[0] = c[0] + lo0;
    int layer = c[axis];
    int shift;
    if (layer < 3) { p[axis] = 3 + layer; shift = n[axis]; }
    else { p[axis] = m[axis] - 4 + (layer - 3); shift = -n[axis]; }
    int q[3] = {p[0], p[1], p[2]};
    q[axis] += shift;
    long long o = fidx(g, p[0], p[1], p[2]);
    long long s = fidx(g, q[0], q[1], q[2]);
    J0[o] = J0[o] + J0[s];
    J1[o] = J1[o] + J1[s];
    J2[o] = J2[o] + J2[s];
}

// One axis pass of the 1-2-1 current filter (MovingAverageX1/Y1/Z1,
// current.F90:90-163, weights vars.F90:208-210) from src to dst over the
// physical cells, the neighbours across the periodic boundary read at their
// images (what SendRecvFldsYZ/ZX/XY put in the ghosts on one rank).  Every
// other cell is copied, so dst is a complete array.  blockIdx.y selects the
// component.
__global__ __launch_bounds__(256) void k_filter_axis(GridP g, int axis, Ptrs6<double> P)
{
    const double *__restrict__ src = P.p[blockIdx.y];
    double *__restrict__ dst = P.p[3 + blockIdx.y];
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncell) return;
    const int i = (int)(t % g.mx) + 1;
    const long long r = t / g.mx;
    const int j = (int)(r % g.my) + 1;
    const int k = (int)(r / g.my) + 1;
    const bool phys = i >= 3 && i <= g.mx - 3 && j >= 3 && j <= g.my - 3 &&
                      (g.ndim == 2 || (k >= 3 && k <= g.mz - 3));
    if (!phys) {
        dst[t] = src[t];
        return;
    }
    int c = i, n = g.nx;
    long long stride = 1;
    if (axis == 1) { c = j; n = g.ny; stride = g.mx; }
    else if (axis == 2) { c = k; n = g.nz; stride = (long long)g.mx * g.my; }
    const long long lo = (c == 3) ? t + (long long)(n - 1) * stride : t - stride;
    const long long hi = (c == 2 + n) ? t - (long long)(n - 1) * stride : t + stride;
    const double wtm1 = 0.25, wt0 = 0.5, wtp1 = 0.25;
    dst[t] = wtm1 * src[lo] + wt0 * src[t] + wtp1 * src[hi];
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill_zero(Ptrs6<T> P, long long n)
{
    T *p = P.p[blockIdx.y];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) p[t] = T(0);
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_n(Ptrs6<T> D, Ptrs6<T> S, long long n)
{
    T *d = D.p[blockIdx.y];
    const T *src = S.p[blockIdx.y];
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) d[t] = src[t];
}

// CalcFldEnergy, savedata_routines.F90:310-327: per-block FP64 partial sums
template <typename R>
__global__ __launch_bounds__(256) void k_field_energy(GridP g, const R *Ex, const R *Ey, const R *Ez, const R *Bx,
                                                      const R *By, const R *Bz, const R *Lx, const R *Ly,
                                                      const R *Lz, double *partial)
{
    __shared__ double sb[256], se[256];
    long long nphys = (long long)g.nx * g.ny * ((g.ndim == 3) ? g.nz : 1);
    double be = 0, ee = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nphys;
         t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % g.nx) + 3;
        long long r = t / g.nx;
        int j = (int)(r % g.ny) + 3;
        int k = (g.ndim == 3) ? (int)(r / g.ny) + 3 : 1;
        long long o = fidx(g, i, j, k);
        double bx = Bx[o], by = By[o], bz = Bz[o], ex = Ex[o], ey = Ey[o], ez = Ez[o];
        if (Lx) { ex = ex + (double)Lx[o]; ey = ey + (double)Ly[o]; ez = ez + (double)Lz[o]; }
        be += (bx * bx + by * by + bz * bz);
        ee += (ex * ex + ey * ey + ez * ez);
    }
    sb[threadIdx.x] = be; se[threadIdx.x] = ee;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) { sb[threadIdx.x] += sb[threadIdx.x + s]; se[threadIdx.x] += se[threadIdx.x + s]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = sb[0]; partial[2 * blockIdx.x + 1] = se[0]; }
}

// max |div E - rho| and max |rho| over physical nodes (CalcDivE stencil,
// savedata_routines.F90:329-355)
template <typename R>
__global__ __launch_bounds__(256) void k_gauss(GridP g, const R *Ex, const R *Ey, const R *Ez, const R *Lx,
                                               const R *Ly, const R *Lz, const double *rho, double *partial)
{
    __shared__ double sr[256], sa[256];
    long long nphys = (long long)g.nx * g.ny * ((g.ndim == 3) ? g.nz : 1);
    double rm = 0, am = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nphys;
         t += (long long)gridDim.x * blockDim.x) {
        int i = (int)(t % g.nx) + 3;
        long long r = t / g.nx;
        int j = (int)(r % g.ny) + 3;
        int k = (g.ndim == 3) ? (int)(r / g.ny) + 3 : 1;
        long long o = fidx(g, i, j, k);
        const long long ok = o - (long long)g.mx * g.my;
        double d;
        if (Lx) {
            // (hi + lo) of each value first: the pair is one field value
            d = ((double)Ex[o] + (double)Lx[o]) - ((double)Ex[o - 1] + (double)Lx[o - 1]) +
                ((double)Ey[o] + (double)Ly[o]) - ((double)Ey[o - g.mx] + (double)Ly[o - g.mx]);
            if (g.ndim == 3) d = d + ((double)Ez[o] + (double)Lz[o]) - ((double)Ez[ok] + (double)Lz[ok]);
        } else {
            d = (double)Ex[o] - (double)Ex[o - 1] + (double)Ey[o] - (double)Ey[o - g.mx];
            if (g.ndim == 3) d = d + (double)Ez[o] - (double)Ez[ok];
        }
        double res = fabs(d - rho[o]);
        rm = fmax(rm, res);
        am = fmax(am, fabs(rho[o]));
    }
    sr[threadIdx.x] = rm; sa[threadIdx.x] = am;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sr[threadIdx.x] = fmax(sr[threadIdx.x], sr[threadIdx.x + s]);
            sa[threadIdx.x] = fmax(sa[threadIdx.x], sa[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { partial[2 * blockIdx.x] = sr[0]; partial[2 * blockIdx.x + 1] = sa[0]; }
}

__global__ void k_cvt_f32(const double *s, float *d, long long n)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = (float)s[t];
}
__global__ void k_cvt_f64(const float *s, double *d, long long n)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = (double)s[t];
}
__global__ void k_cvt_jfix(const double *s, double *d, long long n, double scale)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = (double)s[t] * scale;
}

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

template <typename R>
void launch_update_b_half(const GridP &g, const PhysP &ph, R *const E[3], R *const B[3], hipStream_t s)
{
    hipLaunchKernelGGL(k_update_b_half<R>, yee_grid(g), dim3(BX, BY), 0, s, g, (R)ph.fld_halfc, (R)ph.inv_dx,
                       (R)ph.inv_dy, (R)ph.inv_dz, E[0], E[1], E[2], B[0], B[1], B[2]);
}

template <typename R>
void launch_update_b_half_lehe(const GridP &g, const PhysP &ph, const double lehe[12], R *const E[3], R *const B[3],
                               hipStream_t s)
{
    LeheC<R> L{(R)lehe[0], (R)lehe[1], (R)lehe[2], (R)lehe[3], (R)lehe[4],  (R)lehe[5],
               (R)lehe[6], (R)lehe[7], (R)lehe[8], (R)lehe[9], (R)lehe[10], (R)lehe[11]};
    hipLaunchKernelGGL(k_update_b_half_lehe<R>, phys_grid(g), dim3(BX, BY), 0, s, g, (R)ph.fld_halfc, (R)ph.inv_dx,
                       (R)ph.inv_dy, (R)ph.inv_dz, L, E[0], E[1], E[2], B[0], B[1], B[2]);
}

template <typename R>
void launch_update_e(const GridP &g, const PhysP &ph, R *const E[3], R *const B[3], hipStream_t s)
{
    const int wide = ph.fdtd == 0;
    hipLaunchKernelGGL((k_update_e<R, double, true, false>), wide ? yee_grid(g) : phys_grid(g), dim3(BX, BY), 0, s,
                       g, (R)ph.fldc, (R)ph.inv_dx, (R)ph.inv_dy, (R)ph.inv_dz, B[0], B[1], B[2], E[0], E[1], E[2],
                       (const double *)nullptr, (const double *)nullptr, (const double *)nullptr, 0.0, wide);
}

template <typename R, typename JT>
void launch_update_e_add_current(const GridP &g, const PhysP &ph, R *const E[3], R *const B[3], JT *const J[3],
                                 double jscale, hipStream_t s)
{
    const int wide = ph.fdtd == 0;
    hipLaunchKernelGGL((k_update_e<R, JT, true, true>), wide ? yee_grid(g) : phys_grid(g), dim3(BX, BY), 0, s, g,
                       (R)ph.fldc, (R)ph.inv_dx, (R)ph.inv_dy, (R)ph.inv_dz, B[0], B[1], B[2], E[0], E[1], E[2], J[0],
                       J[1], J[2], jscale, wide);
}

template <typename R, typename JT>
void launch_add_current(const GridP &g, const PhysP &ph, R *const E[3], JT *const J[3], double jscale,
                        hipStream_t s)
{
    hipLaunchKernelGGL((k_update_e<R, JT, false, true>), phys_grid(g), dim3(BX, BY), 0, s, g, (R)ph.fldc,
                       (R)ph.inv_dx, (R)ph.inv_dy, (R)ph.inv_dz, (const R *)nullptr, (const R *)nullptr,
                       (const R *)nullptr, E[0], E[1], E[2], J[0], J[1], J[2], jscale, 0);
}

template <typename R>
void launch_ghost_copy(const GridP &g, R *const F[3], hipStream_t s, R *const G[3])
{
    Ptrs6<R> P{{F[0], F[1], F[2], G ? G[0] : nullptr, G ? G[1] : nullptr, G ? G[2] : nullptr}};
    hipLaunchKernelGGL(k_ghost_copy<R>, dim3(nblocks(g.ncell, 256), G ? 2 : 1), dim3(256), 0, s, g, P);
}

template <typename JT>
void launch_current_fold(const GridP &g, JT *const J[3], hipStream_t s, JT *const K[3])
{
    Ptrs6<JT> P{{J[0], J[1], J[2], K ? K[0] : nullptr, K ? K[1] : nullptr, K ? K[2] : nullptr}};
    int naxes = (g.ndim == 3) ? 3 : 2;
    for (int a = 0; a < naxes; a++) {
        if (!g.per[a]) continue;   // null neighbours: the ghost current stays where it is
        long long ext[3] = {g.mx, g.my, g.mz};
        if (a != 0) ext[0] = g.mx - 5;
        ext[a] = 5;
        long long total = ext[0] * ext[1] * ext[2];
        hipLaunchKernelGGL(k_current_fold<JT>, dim3(nblocks(total, 256), K ? 2 : 1), dim3(256), 0, s, g, a, P);
    }
}

void launch_smoothen_current(const GridP &g, int passes, double *const J[3], double *const T[3], hipStream_t s)
{
    const int naxes = (g.ndim == 3) ? 3 : 2;
    Ptrs6<double> fwd{{J[0], J[1], J[2], T[0], T[1], T[2]}}, back{{T[0], T[1], T[2], J[0], J[1], J[2]}};
    bool in_t = false;   // where the current values are
    for (int ni = 0; ni < passes; ni++)
        for (int a = 0; a < naxes; a++) {
            hipLaunchKernelGGL(k_filter_axis, dim3(nblocks(g.ncell, 256), 3), dim3(256), 0, s, g, a,
                               in_t ? back : fwd);
            in_t = !in_t;
        }
    if (in_t) launch_copy_n<double>(J, T, 3, g.ncell, s);
}

template <typename T>
void launch_fill_zero_n(T *const *p, int count, long long n, hipStream_t s)
{
    unsigned nb = nblocks(n, 256);
    if (nb > 65536u) nb = 65536u;
    if (nb == 0 || count < 1) return;
    Ptrs6<T> P{};
    for (int i = 0; i < count && i < 6; i++) P.p[i] = p[i];
    hipLaunchKernelGGL(k_fill_zero<T>, dim3(nb, count < 6 ? count : 6), dim3(256), 0, s, P, n);
}

template <typename T>
void launch_fill_zero(T *p, long long n, hipStream_t s)
{
    launch_fill_zero_n<T>(&p, 1, n, s);
}

template <typename T>
void launch_copy_n(T *const *dst, T *const *src, int count, long long n, hipStream_t s)
{
    unsigned nb = nblocks(n, 256);
    if (nb > 65536u) nb = 65536u;
    if (nb == 0 || count < 1) return;
    Ptrs6<T> D{}, S{};
    for (int i = 0; i < count && i < 6; i++) { D.p[i] = dst[i]; S.p[i] = src[i]; }
    hipLaunchKernelGGL(k_copy_n<T>, dim3(nb, count < 6 ? count : 6), dim3(256), 0, s, D, S, n);
}

template <typename R>
void launch_field_energy(const GridP &g, R *const E[3], R *const B[3], double *partial, int nblk, hipStream_t s,
                         R *const *L)
{
    hipLaunchKernelGGL(k_field_energy<R>, dim3(nblk), dim3(256), 0, s, g, E[0], E[1], E[2], B[0], B[1], B[2],
                       L ? L[0] : (const R *)nullptr, L ? L[1] : (const R *)nullptr, L ? L[2] : (const R *)nullptr,
                       partial);
}

template <typename R>
void launch_gauss(const GridP &g, R *const E[3], const double *rho, double *partial, int nblk, hipStream_t s,
                  R *const *L)
{
    hipLaunchKernelGGL(k_gauss<R>, dim3(nblk), dim3(256), 0, s, g, E[0], E[1], E[2], L ? L[0] : (const R *)nullptr,
                       L ? L[1] : (const R *)nullptr, L ? L[2] : (const R *)nullptr, rho, partial);
}

void launch_update_e_pair(const GridP &g, const PhysP &ph, float *const E[3], float *const L[3], float *const B[3],
                          double *const J[3], double jscale, bool curl, bool add, hipStream_t s)
{
    const float *b0 = B ? B[0] : nullptr, *b1 = B ? B[1] : nullptr, *b2 = B ? B[2] : nullptr;
    const double *j0 = J ? J[0] : nullptr, *j1 = J ? J[1] : nullptr, *j2 = J ? J[2] : nullptr;
    const int wide = curl && ph.fdtd == 0;
    const dim3 grid = wide ? yee_grid(g) : phys_grid(g);
    if (curl && add)
        hipLaunchKernelGGL((k_update_e_pair<true, true>), grid, dim3(BX, BY), 0, s, g, ph.fldc, ph.inv_dx,
                           ph.inv_dy, ph.inv_dz, b0, b1, b2, E[0], E[1], E[2], L[0], L[1], L[2], j0, j1, j2, jscale,
                           wide);
    else if (curl)
        hipLaunchKernelGGL((k_update_e_pair<true, false>), grid, dim3(BX, BY), 0, s, g, ph.fldc, ph.inv_dx,
                           ph.inv_dy, ph.inv_dz, b0, b1, b2, E[0], E[1], E[2], L[0], L[1], L[2], j0, j1, j2, jscale,
                           wide);
    else
        hipLaunchKernelGGL((k_update_e_pair<false, true>), grid, dim3(BX, BY), 0, s, g, ph.fldc, ph.inv_dx,
                           ph.inv_dy, ph.inv_dz, b0, b1, b2, E[0], E[1], E[2], L[0], L[1], L[2], j0, j1, j2, jscale,
                           wide);
}

void launch_pair_from_f64(const double *src, float *hi, float *lo, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pair_from_f64, dim3(nblocks(n, 256)), dim3(256), 0, s, src, hi, lo, n);
}

void launch_pair_to_f64(const float *hi, const float *lo, double *dst, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_pair_to_f64, dim3(nblocks(n, 256)), dim3(256), 0, s, hi, lo, dst, n);
}

void launch_convert_fields_f32(const double *src, float *dst, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_cvt_f32, dim3(nblocks(n, 256)), dim3(256), 0, s, src, dst, n);
}
void launch_convert_fields_f64(const float *src, double *dst, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_cvt_f64, dim3(nblocks(n, 256)), dim3(256), 0, s, src, dst, n);
}
void launch_convert_jfix(const double *src, double *dst, long long n, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_cvt_jfix, dim3(nblocks(n, 256)), dim3(256), 0, s, src, dst, n, scale);
}

template void launch_update_b_half<double>(const GridP &, const PhysP &, double *const[3], double *const[3], hipStream_t);
template void launch_update_b_half<float>(const GridP &, const PhysP &, float *const[3], float *const[3], hipStream_t);
template void launch_update_b_half_lehe<double>(const GridP &, const PhysP &, const double[12], double *const[3],
                                                double *const[3], hipStream_t);
template void launch_update_b_half_lehe<float>(const GridP &, const PhysP &, const double[12], float *const[3],
                                               float *const[3], hipStream_t);
template void launch_update_e<double>(const GridP &, const PhysP &, double *const[3], double *const[3], hipStream_t);
template void launch_update_e<float>(const GridP &, const PhysP &, float *const[3], float *const[3], hipStream_t);
template void launch_update_e_add_current<double, double>(const GridP &, const PhysP &, double *const[3],
                                                          double *const[3], double *const[3], double, hipStream_t);
template void launch_update_e_add_current<float, double>(const GridP &, const PhysP &, float *const[3],
                                                         float *const[3], double *const[3], double, hipStream_t);
template void launch_add_current<double, double>(const GridP &, const PhysP &, double *const[3], double *const[3],
                                                 double, hipStream_t);
template void launch_add_current<float, double>(const GridP &, const PhysP &, float *const[3], double *const[3],
                                                double, hipStream_t);
template void launch_ghost_copy<double>(const GridP &, double *const[3], hipStream_t, double *const[3]);
template void launch_ghost_copy<float>(const GridP &, float *const[3], hipStream_t, float *const[3]);
template void launch_current_fold<double>(const GridP &, double *const[3], hipStream_t, double *const[3]);
template void launch_fill_zero_n<double>(double *const *, int, long long, hipStream_t);
template void launch_fill_zero_n<float>(float *const *, int, long long, hipStream_t);
template void launch_copy_n<float>(float *const *, float *const *, int, long long, hipStream_t);
template void launch_copy_n<double>(double *const *, double *const *, int, long long, hipStream_t);
template void launch_fill_zero<double>(double *, long long, hipStream_t);
template void launch_fill_zero<float>(float *, long long, hipStream_t);
template void launch_fill_zero<long long>(long long *, long long, hipStream_t);
template void launch_fill_zero<int>(int *, long long, hipStream_t);
template void launch_field_energy<double>(const GridP &, double *const[3], double *const[3], double *, int, hipStream_t,
                                          double *const *);
template void launch_field_energy<float>(const GridP &, float *const[3], float *const[3], double *, int, hipStream_t,
                                         float *const *);
template void launch_gauss<double>(const GridP &, double *const[3], const double *, double *, int, hipStream_t,
                                   double *const *);
template void launch_gauss<float>(const GridP &, float *const[3], const double *, double *, int, hipStream_t,
                                  float *const *);

}  // namespace pictor
