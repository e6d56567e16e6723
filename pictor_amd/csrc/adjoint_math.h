// adjoint_math.h -- per-particle forward and reverse sweep of the 2-D split push.
//
// The forward is the split-form push of the production kernel
// (kernels_push.hip): CIC gather of the cell's corner values, Boris
// (movdep.F90:143-173), position advance, zigzag relay and the two
// half-segment currents (movdep.F90:177-237), here in precision R.  The
// reverse sweep is its hand-derived transpose:
//   deposit  Jx(i, j..j+1) = F_x (1 - W_y, W_y), Jy(i..i+1, j) = F_y (1 - W_x, W_x),
//            Jz(i..i+1, j..j+1) = F_z (1 - W_x, W_x) x (1 - W_y, W_y),
//            F1 = q (x_r - x0), W1 = (x0 + x_r)/2, F2 = q (x1 - x_r), W2 = (x1 + x_r)/2 - d;
//            x_r is the midpoint unless a face was crossed (then it is that face);
//   advance  x1 = x0 + u1 G, G = c^2 / sqrt(c^2 + |U2|^2), u1 = U2 / c;
//   Boris    U2 = U0 + U1 x pB + pE, U1 = f (U0 + U0 x pB), f = 2 / (1 + |pB|^2),
//            pB = g hB, g = 1 / sqrt(c^2 + |U0|^2), U0 = c u + pE;
//   gather   pE = qm sum_k w_k E_k, hB = qm sum_k w_k B_k with bilinear w_k(x0, y0).
#pragma once
#include "adjoint.h"

namespace pictor {
namespace adjm {

template <typename R> struct V4;
template <> struct V4<float> { using T = float4; };
template <> struct V4<double> { using T = double4; };

template <typename R> __device__ inline R rsq_(R v);
template <> __device__ inline float rsq_<float>(float v) { return __builtin_amdgcn_rsqf(v); }
template <> __device__ inline double rsq_<double>(double v) { return 1.0 / sqrt(v); }
template <typename R> __device__ inline R rcp_(R v);
template <> __device__ inline float rcp_<float>(float v) { return __builtin_amdgcn_rcpf(v); }
template <> __device__ inline double rcp_<double>(double v) { return 1.0 / v; }
// current contribution: FP32 -> exact count of 2^-30 units, FP64 -> as is
template <typename R> __device__ inline double qnt(R v);
template <> __device__ inline double qnt<float>(float v) { return (double)__builtin_rintf(v); }
template <> __device__ inline double qnt<double>(double v) { return v; }
template <typename R> __device__ inline R floor_(R v);
template <> __device__ inline float floor_<float>(float v) { return floorf(v); }
template <> __device__ inline double floor_<double>(double v) { return floor(v); }

// per-species constant of a kernel argument: a select, so that a per-lane
// species does not index the argument segment (a dependent memory load)
__device__ inline double sp2(const double (&a)[2], int s) { return s == 0 ? a[0] : a[1]; }

template <typename R> __device__ inline R vmin(R a, R b) { return b < a ? b : a; }
template <typename R> __device__ inline R vmax(R a, R b) { return b > a ? b : a; }

// Forward quantities of one particle that the reverse sweep needs.
template <typename R>
struct PState {
    R x0, y0;
    R w[4];
    R pE[3], hB[3];
    R U0[3], g, pB[3], f, V[3], U1[3], U2[3], r2, G, un1[3];
    R x1[3], xr[3];
    int d[3];
    R F1[3], F2[3], W1[2], W2[2];
};

// gc[c] holds component c (Ex, Ey, Ez, Bx, By, Bz) at the 4 corners
// (i,j), (i+1,j), (i,j+1), (i+1,j+1) of the particle's cell.
template <typename R>
__device__ inline void push_fwd(const AdjP &ap, int s, R x0, R y0, R z0, R un, R vn, R wn,
                                const typename V4<R>::T gc[6], PState<R> &q)
{
#pragma clang fp contract(fast)
    const R c = (R)ap.c, sqc = (R)ap.sqc, cinv = (R)ap.cinv;
    const R qm = (R)sp2(ap.qm, s), qs = (R)sp2(ap.qs, s);
    q.x0 = x0; q.y0 = y0;
    const R w1 = ((R)1 - x0) * ((R)1 - y0), w2 = x0 * ((R)1 - y0);
    const R w3 = ((R)1 - x0) * y0, w4 = x0 * y0;
    q.w[0] = w1; q.w[1] = w2; q.w[2] = w3; q.w[3] = w4;
    R a[6];
#pragma unroll
    for (int k = 0; k < 6; k++) a[k] = (w1 * gc[k].x + w2 * gc[k].y + w3 * gc[k].z + w4 * gc[k].w) * qm;
    q.pE[0] = a[0]; q.pE[1] = a[1]; q.pE[2] = a[2];
    q.hB[0] = a[3]; q.hB[1] = a[4]; q.hB[2] = a[5];
    R *U0 = q.U0, *pB = q.pB, *U1 = q.U1, *U2 = q.U2, *V = q.V;
    U0[0] = c * un + q.pE[0];
    U0[1] = c * vn + q.pE[1];
    U0[2] = c * wn + q.pE[2];
    q.g = rsq_<R>(sqc + U0[0] * U0[0] + U0[1] * U0[1] + U0[2] * U0[2]);
    pB[0] = q.g * q.hB[0]; pB[1] = q.g * q.hB[1]; pB[2] = q.g * q.hB[2];
    q.f = (R)2 * rcp_<R>((R)1 + pB[0] * pB[0] + pB[1] * pB[1] + pB[2] * pB[2]);
    V[0] = U0[0] + U0[1] * pB[2] - U0[2] * pB[1];
    V[1] = U0[1] + U0[2] * pB[0] - U0[0] * pB[2];
    V[2] = U0[2] + U0[0] * pB[1] - U0[1] * pB[0];
    U1[0] = V[0] * q.f; U1[1] = V[1] * q.f; U1[2] = V[2] * q.f;
    U2[0] = U0[0] + U1[1] * pB[2] - U1[2] * pB[1] + q.pE[0];
    U2[1] = U0[1] + U1[2] * pB[0] - U1[0] * pB[2] + q.pE[1];
    U2[2] = U0[2] + U1[0] * pB[1] - U1[1] * pB[0] + q.pE[2];
    q.un1[0] = U2[0] * cinv; q.un1[1] = U2[1] * cinv; q.un1[2] = U2[2] * cinv;
    q.r2 = rsq_<R>(sqc + U2[0] * U2[0] + U2[1] * U2[1] + U2[2] * U2[2]);
    q.G = sqc * q.r2;
    const R xs[3] = {x0, y0, z0};
    const R idc = (R)ap.inv_d;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        q.x1[k] = xs[k] + q.un1[k] * q.G * idc;   // idc = 1: the production push's rounding of the advance
        q.d[k] = (int)floor_<R>(q.x1[k]);
        q.xr[k] = vmin<R>((R)(min(0, q.d[k]) + 1), vmax<R>((R)max(0, q.d[k]), (R)0.5 * (xs[k] + q.x1[k])));
        q.F1[k] = qs * (q.xr[k] - xs[k]);
        q.F2[k] = qs * (q.x1[k] - q.xr[k]);
    }
    q.W1[0] = (R)0.5 * (x0 + q.xr[0]);
    q.W1[1] = (R)0.5 * (y0 + q.xr[1]);
    q.W2[0] = (R)0.5 * (q.x1[0] + q.xr[0]) - (R)q.d[0];
    q.W2[1] = (R)0.5 * (q.x1[1] + q.xr[1]) - (R)q.d[1];
}

// 8 current values of one half-segment: Jx (i,j), (i,j+1); Jy (i,j), (i+1,j);
// Jz (i,j), (i+1,j), (i,j+1), (i+1,j+1)
template <typename R>
__device__ inline void seg_values(const R F[3], const R W[2], double a[8])
{
#pragma clang fp contract(fast)
    const R zl = F[2] * ((R)1 - W[0]), zh = F[2] * W[0];
    a[0] = qnt<R>(F[0] * ((R)1 - W[1])); a[1] = qnt<R>(F[0] * W[1]);
    a[2] = qnt<R>(F[1] * ((R)1 - W[0])); a[3] = qnt<R>(F[1] * W[0]);
    a[4] = qnt<R>(zl * ((R)1 - W[1])); a[5] = qnt<R>(zh * ((R)1 - W[1]));
    a[6] = qnt<R>(zl * W[1]); a[7] = qnt<R>(zh * W[1]);
}

// adjoint of one half-segment's deposit; jb: adjoint current at its 8 nodes (seg_values order)
template <typename R>
__device__ inline void seg_adjoint(const R F[3], const R W[2], const R jb[8], R Fb[3], R Wb[2])
{
    const R wx = W[0], wy = W[1];
    Fb[0] = ((R)1 - wy) * jb[0] + wy * jb[1];
    Fb[1] = ((R)1 - wx) * jb[2] + wx * jb[3];
    Fb[2] = ((R)1 - wx) * ((R)1 - wy) * jb[4] + wx * ((R)1 - wy) * jb[5] + ((R)1 - wx) * wy * jb[6] + wx * wy * jb[7];
    Wb[0] = F[1] * (jb[3] - jb[2]) + F[2] * (((R)1 - wy) * (jb[5] - jb[4]) + wy * (jb[7] - jb[6]));
    Wb[1] = F[0] * (jb[1] - jb[0]) + F[2] * (((R)1 - wx) * (jb[6] - jb[4]) + wx * (jb[7] - jb[5]));
}

template <typename R>
__device__ inline void cross3(const R a[3], const R b[3], R out[3])
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

// Reverse sweep of one particle.  In: the forward state q, the adjoint current
// at the two half-segments' nodes (jb1, jb2), the adjoint of the new position
// x1b and momentum u1b.  Out: the adjoint of the old position x0b and
// momentum unb, and the adjoint of the gathered sums (PEb, PBb): the corner
// field adjoints are w_k * PEb, w_k * PBb.
template <typename R>
__device__ inline void reverse(const AdjP &ap, int s, const PState<R> &q, const typename V4<R>::T gc[6],
                               const R jb1[8], const R jb2[8], const R x1b_in[3], const R u1b_in[3], R x0b[3],
                               R unb[3], R PEb[3], R PBb[3])
{
    const R qs = (R)sp2(ap.qs, s), qm = (R)sp2(ap.qm, s);
    const R c = (R)ap.c, sqc = (R)ap.sqc, cinv = (R)ap.cinv;
    R x1b[3] = {x1b_in[0], x1b_in[1], x1b_in[2]};
    R u1b[3] = {u1b_in[0], u1b_in[1], u1b_in[2]};
    R xrb[3] = {0, 0, 0};
    R x1d[3];
    {
        R Fb[3], Wb[2];
        seg_adjoint<R>(q.F1, q.W1, jb1, Fb, Wb);
#pragma unroll
        for (int a = 0; a < 3; a++) { xrb[a] = qs * Fb[a]; x0b[a] = x1b[a] - qs * Fb[a]; }
        x0b[0] += (R)0.5 * Wb[0]; xrb[0] += (R)0.5 * Wb[0];
        x0b[1] += (R)0.5 * Wb[1]; xrb[1] += (R)0.5 * Wb[1];
        seg_adjoint<R>(q.F2, q.W2, jb2, Fb, Wb);
#pragma unroll
        for (int a = 0; a < 3; a++) { x1d[a] = qs * Fb[a]; xrb[a] -= qs * Fb[a]; }
        x1d[0] += (R)0.5 * Wb[0]; xrb[0] += (R)0.5 * Wb[0];
        x1d[1] += (R)0.5 * Wb[1]; xrb[1] += (R)0.5 * Wb[1];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (q.d[a] == 0) { x0b[a] += (R)0.5 * xrb[a]; x1d[a] += (R)0.5 * xrb[a]; }
            x1b[a] += x1d[a];
            x0b[a] += x1d[a];   // x1 = x0 + u1 G
        }
    }
    R Gb = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) { u1b[a] += x1b[a] * q.G; Gb += x1b[a] * q.un1[a]; }
    R U2b[3];
    {
        const R k = -Gb * sqc * q.r2 * q.r2 * q.r2;
#pragma unroll
        for (int a = 0; a < 3; a++) U2b[a] = u1b[a] * cinv + k * q.U2[a];
    }
    R U0b[3] = {U2b[0], U2b[1], U2b[2]};
    R pEb[3] = {U2b[0], U2b[1], U2b[2]};
    R U1b[3], pBb[3];
    cross3<R>(q.pB, U2b, U1b);
    cross3<R>(U2b, q.U1, pBb);
    {
        R Vb[3] = {U1b[0] * q.f, U1b[1] * q.f, U1b[2] * q.f};
        const R fb = U1b[0] * q.V[0] + U1b[1] * q.V[1] + U1b[2] * q.V[2];
        R t[3];
        cross3<R>(q.pB, Vb, t);
#pragma unroll
        for (int a = 0; a < 3; a++) U0b[a] += Vb[a] + t[a];
        cross3<R>(Vb, q.U0, t);
        const R k = -fb * q.f * q.f;
#pragma unroll
        for (int a = 0; a < 3; a++) pBb[a] += t[a] + k * q.pB[a];
    }
    R hBb[3];
    {
        const R gb = pBb[0] * q.hB[0] + pBb[1] * q.hB[1] + pBb[2] * q.hB[2];
#pragma unroll
        for (int a = 0; a < 3; a++) hBb[a] = q.g * pBb[a];
        const R k = -gb * q.g * q.g * q.g;
#pragma unroll
        for (int a = 0; a < 3; a++) U0b[a] += k * q.U0[a];
    }
#pragma unroll
    for (int a = 0; a < 3; a++) { unb[a] = c * U0b[a]; pEb[a] += U0b[a]; }
#pragma unroll
    for (int a = 0; a < 3; a++) { PEb[a] = pEb[a] * qm; PBb[a] = hBb[a] * qm; }
    // bilinear weights: w = (1-x)(1-y), x(1-y), (1-x)y, xy
    R wb[4];
    const R cx[4] = {gc[0].x, gc[0].y, gc[0].z, gc[0].w};
    const R cy[4] = {gc[1].x, gc[1].y, gc[1].z, gc[1].w};
    const R cz[4] = {gc[2].x, gc[2].y, gc[2].z, gc[2].w};
    const R bx[4] = {gc[3].x, gc[3].y, gc[3].z, gc[3].w};
    const R by[4] = {gc[4].x, gc[4].y, gc[4].z, gc[4].w};
    const R bz[4] = {gc[5].x, gc[5].y, gc[5].z, gc[5].w};
#pragma unroll
    for (int k = 0; k < 4; k++)
        wb[k] = cx[k] * PEb[0] + cy[k] * PEb[1] + cz[k] * PEb[2] + bx[k] * PBb[0] + by[k] * PBb[1] + bz[k] * PBb[2];
    const R x0 = q.x0, y0 = q.y0;
    x0b[0] += -wb[0] * ((R)1 - y0) + wb[1] * ((R)1 - y0) - wb[2] * y0 + wb[3] * y0;
    x0b[1] += -wb[0] * ((R)1 - x0) - wb[1] * x0 + wb[2] * ((R)1 - x0) + wb[3] * x0;
}

}  // namespace adjm
}  // namespace pictor
