// adjoint.hip -- hand-written adjoint of the 2-D particle step (gfx950).
//
// The adjoint engine runs its own forward of the 2-D step with the particles
// kept in a fixed order (the caller's), so that a particle's adjoint state is
// found at the same index at every step.  Its forward is the split-form step
// of the production kernels in the same arithmetic (FP32: hardware rsq/rcp,
// currents as exact 2^-30 counts; FP64: IEEE sqrt/div, physical currents),
// with the field gather read from node arrays and the deposit made with
// global FP64 atomics.
//
// Per step n the forward is (MAIN.F90:51-251 order)
//   Bh = B - (c/2) curlE(E)             UpdateBfieldHalf
//   (x', u') = push(x, u; E, Bh)        MoveDepositPrtl (Boris, movdep.F90:143-173)
//   J = deposit(x, x')                  zigzag, movdep.F90:177-237
//   B' = Bh - (c/2) curlE(E)            UpdateBfieldHalf
//   E' = E + c curlB(B') - s J          UpdateEfield + AddCurrent
// and the reverse step, with curlE^T = curlB on the periodic Yee grid,
//   Bb += c curlE(Eb)                   (transpose of the E update's curl)
//   particle adjoint: reads Jb = -s Eb at the deposit stencil, back-propagates
//     through deposit, position advance, Boris and gather; scatters the
//     adjoint of the node fields (E, Bh) with atomics
//   Eb += -(c/2) curlB(Bb)              (second B half step)
//   Eb += avg^T(nodes E), Bb += avg^T(nodes B)
//   Eb += -(c/2) curlB(Bb)              (first B half step)
// The periodic halos are handled as operators on the physical cells: an
// adjoint array is copied into its ghost layers before a stencil reads it,
// and ghost-node contributions are folded into their images.
#include "adjoint.h"
#include "adjoint_math.h"
#include "push_common.h"

namespace pictor {
namespace {
using namespace pushdev;
using namespace adjm;

template <typename R>
__device__ inline void unpack(const AdjP &ap, uint32_t cs, int &ci, int &cj, int &s, long long &o)
{
    ci = (int)(cs & 0xFFFFu);
    cj = (int)((cs >> 16) & 0x7FFFu);
    s = (int)(cs >> 31);
    o = (long long)(ci + 2) + (long long)ap.mx * (cj + 2);   // node (i, j) = cell (ci + 3, cj + 3)
}

// corner values of the cell at node o in the component-major form of push_fwd
template <typename R>
__device__ inline void corners(const typename V4<R>::T *nE, const typename V4<R>::T *nB, long long o, int mx,
                               typename V4<R>::T gc[6])
{
    const typename V4<R>::T e[4] = {nE[o], nE[o + 1], nE[o + mx], nE[o + mx + 1]};
    const typename V4<R>::T b[4] = {nB[o], nB[o + 1], nB[o + mx], nB[o + mx + 1]};
    gc[0] = {e[0].x, e[1].x, e[2].x, e[3].x};
    gc[1] = {e[0].y, e[1].y, e[2].y, e[3].y};
    gc[2] = {e[0].z, e[1].z, e[2].z, e[3].z};
    gc[3] = {b[0].x, b[1].x, b[2].x, b[3].x};
    gc[4] = {b[0].y, b[1].y, b[2].y, b[3].y};
    gc[5] = {b[0].z, b[1].z, b[2].z, b[3].z};
}

// Node values of E and B (the production kernel's cell-corner staging) for
// nodes 3..m-2 of each axis.
template <typename R>
__global__ __launch_bounds__(256) void k_nodes(AdjP ap, const R *__restrict__ Ex, const R *__restrict__ Ey,
                                               const R *__restrict__ Ez, const R *__restrict__ Bx,
                                               const R *__restrict__ By, const R *__restrict__ Bz,
                                               typename V4<R>::T *__restrict__ nE, typename V4<R>::T *__restrict__ nB)
{
    const int nnx = ap.nx + 1, nny = ap.ny + 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nnx * nny) return;
    const int i = 3 + (int)(t % nnx), j = 3 + (int)(t / nnx);
    const int mx = ap.mx;
    const long long p = (long long)(i - 1) + (long long)mx * (j - 1);
    typename V4<R>::T e, b;
    e.x = (Ex[p - 1] + Ex[p]) * (R)0.5;
    e.y = (Ey[p - mx] + Ey[p]) * (R)0.5;
    e.z = Ez[p];
    e.w = (R)0;
    b.x = (Bx[p - mx] + Bx[p]) * (R)0.5;
    b.y = (By[p - 1] + By[p]) * (R)0.5;
    b.z = (Bz[p - 1 - mx] + Bz[p - 1] + Bz[p - mx] + Bz[p]) * (R)0.25;
    b.w = (R)0;
    nE[p] = e;
    nB[p] = b;
}

template <typename R>
__global__ __launch_bounds__(256) void k_eng_push(AdjP ap, const typename V4<R>::T *__restrict__ nE,
                                                  const typename V4<R>::T *__restrict__ nB, EPart<R> in,
                                                  EPart<R> out, double *__restrict__ Jx, double *__restrict__ Jy,
                                                  double *__restrict__ Jz, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int mx = ap.mx;
    int ci, cj, s;
    long long o;
    const uint32_t cs = in.cs[p];
    unpack<R>(ap, cs, ci, cj, s, o);
    typename V4<R>::T gc[6];
    corners<R>(nE, nB, o, mx, gc);
    // the split form (R = float) does not store z: every step starts at kZMid
    constexpr bool kSplit = std::is_same<R, float>::value;
    const R z0 = kSplit ? (R)kZMid : in.oz[p];
    PState<R> q;
    push_fwd<R>(ap, s, in.ox[p], in.oy[p], z0, in.u[p], in.v[p], in.w[p], gc, q);
    int di = q.d[0], dj = q.d[1];
    // stored offsets; a position rounded onto the upper face belongs to the next cell
    R ox = q.x1[0] - (R)di;
    if (ox >= (R)1) { ox = (R)0; di += 1; }
    R oy = q.x1[1] - (R)dj;
    if (oy >= (R)1) { oy = (R)0; dj += 1; }
    R oz = kSplit ? (R)kZMid : q.x1[2] - (R)q.d[2];
    if (oz >= (R)1) oz = (R)0;
    double a[8], b[8];
    int ei = di, ej = dj;   // cell of the second half-segment
    if constexpr (std::is_same<R, float>::value) {
        // the split form deposits as the production kernel: FP64 to the stored position
        pushdev::zigzag2d_split(sp2(ap.qs, s), in.ox[p], in.oy[p], kZMid, ox, oy, q.x1[2], di, dj, a, b);
    } else {
        seg_values<R>(q.F1, q.W1, a);
        seg_values<R>(q.F2, q.W2, b);
        ei = q.d[0];
        ej = q.d[1];
    }
    const bool cross = (ei != 0) || (ej != 0);
    if (!cross) {
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] += b[k];
    }
    unsafeAtomicAdd(&Jx[o], a[0]); unsafeAtomicAdd(&Jx[o + mx], a[1]);
    unsafeAtomicAdd(&Jy[o], a[2]); unsafeAtomicAdd(&Jy[o + 1], a[3]);
    unsafeAtomicAdd(&Jz[o], a[4]); unsafeAtomicAdd(&Jz[o + 1], a[5]);
    unsafeAtomicAdd(&Jz[o + mx], a[6]); unsafeAtomicAdd(&Jz[o + mx + 1], a[7]);
    if (cross) {
        const long long o2 = o + ei + (long long)ej * mx;
        unsafeAtomicAdd(&Jx[o2], b[0]); unsafeAtomicAdd(&Jx[o2 + mx], b[1]);
        unsafeAtomicAdd(&Jy[o2], b[2]); unsafeAtomicAdd(&Jy[o2 + 1], b[3]);
        unsafeAtomicAdd(&Jz[o2], b[4]); unsafeAtomicAdd(&Jz[o2 + 1], b[5]);
        unsafeAtomicAdd(&Jz[o2 + mx], b[6]); unsafeAtomicAdd(&Jz[o2 + mx + 1], b[7]);
    }
    int nci = ci + di, ncj = cj + dj;
    nci = nci < 0 ? nci + ap.nx : (nci >= ap.nx ? nci - ap.nx : nci);
    ncj = ncj < 0 ? ncj + ap.ny : (ncj >= ap.ny ? ncj - ap.ny : ncj);
    out.cs[p] = (uint32_t)nci | ((uint32_t)ncj << 16) | (cs & 0x80000000u);
    out.ox[p] = ox; out.oy[p] = oy; out.oz[p] = oz;
    out.u[p] = q.un1[0]; out.v[p] = q.un1[1]; out.w[p] = q.un1[2];
}

// One particle backwards through deposit, advance, Boris and gather.
// Eb: adjoint E field (ghosts filled); the current's adjoint is -jscale * Eb.
template <typename R>
__global__ __launch_bounds__(256) void k_eng_adjoint(AdjP ap, const typename V4<R>::T *__restrict__ nE,
                                                     const typename V4<R>::T *__restrict__ nB, EPart<R> st,
                                                     const double *__restrict__ Ebx, const double *__restrict__ Eby,
                                                     const double *__restrict__ Ebz, EPart<R> adj,
                                                     double *__restrict__ gEx, double *__restrict__ gEy,
                                                     double *__restrict__ gEz, double *__restrict__ gBx,
                                                     double *__restrict__ gBy, double *__restrict__ gBz, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int mx = ap.mx;
    int ci, cj, s;
    long long o;
    unpack<R>(ap, st.cs[p], ci, cj, s, o);
    typename V4<R>::T gc[6];
    corners<R>(nE, nB, o, mx, gc);
    constexpr bool kSplit = std::is_same<R, float>::value;   // no z in the split form (kZMid)
    PState<R> q;
    push_fwd<R>(ap, s, st.ox[p], st.oy[p], kSplit ? (R)kZMid : st.oz[p], st.u[p], st.v[p], st.w[p], gc, q);
    const R js = (R)(-ap.jscale);
    R jb1[8], jb2[8];
    jb1[0] = js * (R)Ebx[o]; jb1[1] = js * (R)Ebx[o + mx];
    jb1[2] = js * (R)Eby[o]; jb1[3] = js * (R)Eby[o + 1];
    jb1[4] = js * (R)Ebz[o]; jb1[5] = js * (R)Ebz[o + 1]; jb1[6] = js * (R)Ebz[o + mx]; jb1[7] = js * (R)Ebz[o + mx + 1];
    const long long o2 = o + q.d[0] + (long long)q.d[1] * mx;
    jb2[0] = js * (R)Ebx[o2]; jb2[1] = js * (R)Ebx[o2 + mx];
    jb2[2] = js * (R)Eby[o2]; jb2[3] = js * (R)Eby[o2 + 1];
    jb2[4] = js * (R)Ebz[o2]; jb2[5] = js * (R)Ebz[o2 + 1]; jb2[6] = js * (R)Ebz[o2 + mx]; jb2[7] = js * (R)Ebz[o2 + mx + 1];
    const R x1b[3] = {adj.ox[p], adj.oy[p], kSplit ? (R)0 : adj.oz[p]};
    const R u1b[3] = {adj.u[p], adj.v[p], adj.w[p]};
    R x0b[3], unb[3], PEb[3], PBb[3];
    reverse<R>(ap, s, q, gc, jb1, jb2, x1b, u1b, x0b, unb, PEb, PBb);
    const long long corner[4] = {o, o + 1, o + mx, o + mx + 1};
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const R w = q.w[k];
        unsafeAtomicAdd(&gEx[corner[k]], (double)(w * PEb[0]));
        unsafeAtomicAdd(&gEy[corner[k]], (double)(w * PEb[1]));
        unsafeAtomicAdd(&gEz[corner[k]], (double)(w * PEb[2]));
        unsafeAtomicAdd(&gBx[corner[k]], (double)(w * PBb[0]));
        unsafeAtomicAdd(&gBy[corner[k]], (double)(w * PBb[1]));
        unsafeAtomicAdd(&gBz[corner[k]], (double)(w * PBb[2]));
    }
    adj.ox[p] = x0b[0]; adj.oy[p] = x0b[1]; adj.oz[p] = kSplit ? (R)0 : x0b[2];
    adj.u[p] = unb[0]; adj.v[p] = unb[1]; adj.w[p] = unb[2];
}

// Eb += avg^T(node E adjoint), Bb += avg^T(node B adjoint) on the physical cells
// (node arrays folded and ghost-copied first).
__global__ __launch_bounds__(256) void k_node_to_stag(AdjP ap, const double *__restrict__ nEx,
                                                      const double *__restrict__ nEy, const double *__restrict__ nEz,
                                                      const double *__restrict__ nBx, const double *__restrict__ nBy,
                                                      const double *__restrict__ nBz, double *__restrict__ Ex,
                                                      double *__restrict__ Ey, double *__restrict__ Ez,
                                                      double *__restrict__ Bx, double *__restrict__ By,
                                                      double *__restrict__ Bz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ap.nx * ap.ny) return;
    const int i = 3 + (int)(t % ap.nx), j = 3 + (int)(t / ap.nx);
    const int mx = ap.mx;
    const long long p = (long long)(i - 1) + (long long)mx * (j - 1);
    Ex[p] += 0.5 * (nEx[p] + nEx[p + 1]);
    Ey[p] += 0.5 * (nEy[p] + nEy[p + mx]);
    Ez[p] += nEz[p];
    Bx[p] += 0.5 * (nBx[p] + nBx[p + mx]);
    By[p] += 0.5 * (nBy[p] + nBy[p + 1]);
    Bz[p] += 0.25 * (nBz[p] + nBz[p + 1] + nBz[p + mx] + nBz[p + mx + 1]);
}

// caller's FP64 SoA -> engine particles (validated on the host side)
template <typename R>
__global__ __launch_bounds__(256) void k_eng_load(AdjP ap, const double *__restrict__ x, const double *__restrict__ y,
                                                  const double *__restrict__ z, const double *__restrict__ u,
                                                  const double *__restrict__ v, const double *__restrict__ w,
                                                  const int *__restrict__ flv, EPart<R> out, long long n,
                                                  int *__restrict__ err)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double fx = floor(x[p]), fy = floor(y[p]), fz = floor(z[p]);
    int ci = (int)fx - 3, cj = (int)fy - 3;
    const int f = flv[p];
    if (ci < 0 || ci >= ap.nx || cj < 0 || cj >= ap.ny || (f != 1 && f != 2) || !(z[p] >= 1.0 && z[p] < 2.0)) {
        atomicOr(err, 1);
        ci = 0; cj = 0;
    }
    R ox = (R)(x[p] - fx), oy = (R)(y[p] - fy), oz = (R)(z[p] - fz);
    if (ox >= (R)1) ox = (R)0;   // an offset just below 1 rounded up: stay on the cell
    if (oy >= (R)1) oy = (R)0;
    if (oz >= (R)1) oz = (R)0;
    out.cs[p] = (uint32_t)ci | ((uint32_t)cj << 16) | ((uint32_t)(f == 2 ? 1 : 0) << 31);
    out.ox[p] = ox; out.oy[p] = oy; out.oz[p] = oz;
    out.u[p] = (R)u[p]; out.v[p] = (R)v[p]; out.w[p] = (R)w[p];
}

template <typename R>
__global__ __launch_bounds__(256) void k_eng_store_grad(const EPart<R> adj, double scale, double *__restrict__ du,
                                                        double *__restrict__ dv, double *__restrict__ dw, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    du[p] = scale * (double)adj.u[p];
    dv[p] = scale * (double)adj.v[p];
    dw[p] = scale * (double)adj.w[p];
}

template <typename R>
__global__ __launch_bounds__(256) void k_fill_part(EPart<R> a, long long n)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    a.ox[p] = (R)0; a.oy[p] = (R)0; a.oz[p] = (R)0;
    a.u[p] = (R)0; a.v[p] = (R)0; a.w[p] = (R)0;
}

__global__ __launch_bounds__(256) void k_seed_b(long long n, const float *__restrict__ b, double *__restrict__ d)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = (double)b[t];
}
__global__ __launch_bounds__(256) void k_seed_b(long long n, const double *__restrict__ b, double *__restrict__ d)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) d[t] = b[t];
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

template <typename R>
void launch_eng_nodes(const AdjP &ap, R *const E[3], R *const B[3], void *nE, void *nB, hipStream_t s)
{
    const long long nn = (long long)(ap.nx + 1) * (ap.ny + 1);
    hipLaunchKernelGGL(k_nodes<R>, dim3(nblk(nn)), dim3(256), 0, s, ap, E[0], E[1], E[2], B[0], B[1], B[2],
                       (typename V4<R>::T *)nE, (typename V4<R>::T *)nB);
}

template <typename R>
void launch_eng_push(const AdjP &ap, const void *nE, const void *nB, EPart<R> in, EPart<R> out, double *const J[3],
                     long long n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eng_push<R>, dim3(nblk(n)), dim3(256), 0, s, ap, (const typename V4<R>::T *)nE,
                       (const typename V4<R>::T *)nB, in, out, J[0], J[1], J[2], n);
}

template <typename R>
void launch_eng_adjoint(const AdjP &ap, const void *nE, const void *nB, EPart<R> st, double *const Eb[3],
                        EPart<R> adj, double *const gE[3], double *const gB[3], long long n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eng_adjoint<R>, dim3(nblk(n)), dim3(256), 0, s, ap, (const typename V4<R>::T *)nE,
                       (const typename V4<R>::T *)nB, st, Eb[0], Eb[1], Eb[2], adj, gE[0], gE[1], gE[2], gB[0],
                       gB[1], gB[2], n);
}

void launch_node_to_stag(const AdjP &ap, double *const gE[3], double *const gB[3], double *const Eb[3],
                         double *const Bb[3], hipStream_t s)
{
    const long long n = (long long)ap.nx * ap.ny;
    hipLaunchKernelGGL(k_node_to_stag, dim3(nblk(n)), dim3(256), 0, s, ap, gE[0], gE[1], gE[2], gB[0], gB[1], gB[2],
                       Eb[0], Eb[1], Eb[2], Bb[0], Bb[1], Bb[2]);
}

template <typename R>
void launch_eng_load(const AdjP &ap, const double *const raw[6], const int *flv, EPart<R> out, long long n, int *err,
                     hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eng_load<R>, dim3(nblk(n)), dim3(256), 0, s, ap, raw[0], raw[1], raw[2], raw[3], raw[4],
                       raw[5], flv, out, n, err);
}

template <typename R>
void launch_eng_store_grad(EPart<R> adj, double scale, double *du, double *dv, double *dw, long long n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eng_store_grad<R>, dim3(nblk(n)), dim3(256), 0, s, adj, scale, du, dv, dw, n);
}

template <typename R>
void launch_eng_zero(EPart<R> a, long long n, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_part<R>, dim3(nblk(n)), dim3(256), 0, s, a, n);
}

template <typename R>
void launch_seed_b(const R *b, double *d, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_seed_b, dim3(nblk(n)), dim3(256), 0, s, n, b, d);
}

#define INST(R)                                                                                                     \
    template void launch_eng_nodes<R>(const AdjP &, R *const[3], R *const[3], void *, void *, hipStream_t);         \
    template void launch_eng_push<R>(const AdjP &, const void *, const void *, EPart<R>, EPart<R>, double *const[3], \
                                     long long, hipStream_t);                                                       \
    template void launch_eng_adjoint<R>(const AdjP &, const void *, const void *, EPart<R>, double *const[3],      \
                                        EPart<R>, double *const[3], double *const[3], long long, hipStream_t);     \
    template void launch_eng_load<R>(const AdjP &, const double *const[6], const int *, EPart<R>, long long, int *, \
                                     hipStream_t);                                                                  \
    template void launch_eng_store_grad<R>(EPart<R>, double, double *, double *, double *, long long, hipStream_t); \
    template void launch_eng_zero<R>(EPart<R>, long long, hipStream_t);                                            \
    template void launch_seed_b<R>(const R *, double *, long long, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace pictor
