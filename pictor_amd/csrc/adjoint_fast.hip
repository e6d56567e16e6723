// adjoint_fast.hip -- reverse step of the FP32 split form on the tile bins (gfx950).
//
// The FP32 adjoint engine steps its particles with the production push
// kernel (k_push2d_split, index-carrying variant): every step's particles sit
// in the tile bins of that step, and the push records for every input slot
// the slot its particle was written to.  The reverse step below walks the
// same bins as the forward push of that step, one workgroup per tile: the
// tile's cell-corner E and B (of that step) in LDS, the adjoint current
// -s * Eb of the (T+5)^2 window in LDS, and the adjoint of the tile's corner
// fields accumulated in LDS and flushed with global atomics.  A particle's
// incoming adjoint is read at its output slot, its outgoing adjoint written
// at its input slot.
#include <type_traits>

#include "adjoint.h"
#include "adjoint_math.h"
#include "push_common.h"

namespace pictor {
namespace {
using namespace pushdev;
using namespace adjm;

// cell-corner E/B of the tile's T x T cells (the production kernel's staging)
template <int T, int BS>
__device__ inline void stage_corners(const GridP &g, int tx, int ty, const float *Ex, const float *Ey, const float *Ez,
                                     const float *Bx, const float *By, const float *Bz, float4 *sC)
{
    const int mx = g.mx;
    const int iw = 1 + tx * T, jw = 1 + ty * T;
    for (int lc = threadIdx.x; lc < T * T; lc += BS) {
        const long long o = (long long)(iw + 2 + lc % T - 1) + (long long)mx * (jw + 2 + lc / T - 1);
        float v[6][4];
#pragma unroll
        for (int cn = 0; cn < 4; cn++) {
            const long long p = o + (cn & 1) + (cn >> 1) * (long long)mx;
            v[0][cn] = (Ex[p - 1] + Ex[p]) * 0.5f;
            v[1][cn] = (Ey[p - mx] + Ey[p]) * 0.5f;
            v[2][cn] = Ez[p];
            v[3][cn] = (Bx[p - mx] + Bx[p]) * 0.5f;
            v[4][cn] = (By[p - 1] + By[p]) * 0.5f;
            v[5][cn] = (Bz[p - 1 - mx] + Bz[p - 1] + Bz[p - mx] + Bz[p]) * 0.25f;
        }
#pragma unroll
        for (int cc = 0; cc < 6; cc++) sC[6 * lc + cc] = make_float4(v[cc][0], v[cc][1], v[cc][2], v[cc][3]);
    }
}

// bins of a tile in the production kernel's order: centre bins of every species first
__device__ inline void vbin_of(int v, int nsp, int &s, int &d)
{
    if (v < nsp) { s = v; d = 4; return; }
    const int r = v - nsp;
    s = r >> 3;
    const int q = r & 7;
    d = q < 4 ? q : q + 1;
}

// one particle of the reverse sweep: its slot, species, state before the step
// and the slot it went to (where the adjoint of its state after the step is)
struct AdjRec {
    unsigned slot;
    int s, to, cell;
    float ox, oy, u, v, w;
};

__device__ inline void load_rec(const PartSplit &st, const int *dst, AdjRec &r)
{
    const unsigned sl = r.slot;
    r.to = dst[sl];
    r.cell = st.cell[sl];
    r.ox = st.ox[sl]; r.oy = st.oy[sl];
    r.u = st.u[sl]; r.v = st.v[sl]; r.w = st.w[sl];
}

template <int T, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(1024 / BS, 1024 / BS))) void k_fa_adjoint(GridP g, TileP tp, AdjP ap, const float *__restrict__ Ex,
                                                   const float *__restrict__ Ey, const float *__restrict__ Ez,
                                                   const float *__restrict__ Bx, const float *__restrict__ By,
                                                   const float *__restrict__ Bz, const double *__restrict__ Ebx,
                                                   const double *__restrict__ Eby, const double *__restrict__ Ebz,
                                                   const long long *__restrict__ base, const int *__restrict__ cnt,
                                                   PartSplit st, const int *__restrict__ dst, EPart<float> ain,
                                                   EPart<float> aout, double *__restrict__ gEx,
                                                   double *__restrict__ gEy, double *__restrict__ gEz,
                                                   double *__restrict__ gBx, double *__restrict__ gBy,
                                                   double *__restrict__ gBz)
{
    constexpr int W = T + 5, WW = W * W;
    constexpr int N1 = T + 1, NN = N1 * N1;
    constexpr int MB = 18;
    __shared__ float4 sC[6 * T * T];
    __shared__ float sJb[3 * WW];
    // FP64: an LDS FP32 atomic add runs at about 1/8 the rate of an FP64 one on
    // gfx950 (0.33 vs 2.6 lanes per CU per clock); the 32 x 32 tile keeps FP32,
    // whose FP64 window would not fit beside its corner staging in 160 KB
    using GA = typename std::conditional<(T <= 16), double, float>::type;
    __shared__ GA sG[6 * NN];
    __shared__ int sStart[MB + 1], sVs[MB];
    __shared__ unsigned sSrc[MB];
    const int tile = tile_remap(tp, blockIdx.x);
    const int tx = tile % tp.ntx, ty = tile / tp.ntx;
    const int cx0 = tx * T - 2, cy0 = ty * T - 2;
    const int mx = g.mx;
    const int iw = 1 + tx * T, jw = 1 + ty * T;
    const int nsp = tp.nsp, NV = nsp * 9;
    if ((int)threadIdx.x < NV) {
        int s, d;
        vbin_of(threadIdx.x, nsp, s, d);
        sSrc[threadIdx.x] = (unsigned)base[tile * tp.NB + s * 9 + d];
        sVs[threadIdx.x] = s;
    }
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int v = 0; v < NV; v++) {
            int s, d;
            vbin_of(v, nsp, s, d);
            sStart[v] = acc;
            acc += cnt[tile * tp.NB + s * 9 + d];
        }
        sStart[NV] = acc;
    }
    stage_corners<T, BS>(g, tx, ty, Ex, Ey, Ez, Bx, By, Bz, sC);
    const float js = (float)(-ap.jscale);
    for (int e = threadIdx.x; e < WW; e += BS) {
        const long long o = (long long)(iw + e % W - 1) + (long long)mx * (jw + e / W - 1);
        sJb[e] = js * (float)Ebx[o];
        sJb[WW + e] = js * (float)Eby[o];
        sJb[2 * WW + e] = js * (float)Ebz[o];
    }
    for (int e = threadIdx.x; e < 6 * NN; e += BS) sG[e] = (GA)0;
    __syncthreads();
    const int total = sStart[NV];
    // software pipeline: the next particle's state and output slot are loaded
    // while the current one is processed; the current one's incoming adjoint
    // (whose address is its output slot) is loaded first and used only after
    // the recomputed forward push
    // bin of position k: the last v < NV with sStart[v] <= k (an empty bin shares
    // its start with the next one).  A fixed-depth search with no loop: a loop
    // inside the particle loop makes the compiler wait for every outstanding
    // load before entering it, which would undo the pipelining.
    static_assert(MB <= 32, "bin search depth");
    auto slot_of = [&](int k, unsigned &slot, int &s) {
        int v = 0;
#pragma unroll
        for (int step = 16; step >= 1; step >>= 1) {
            const int c = v + step;
            const int sv = sStart[c < NV ? c : NV - 1];
            v = ((c < NV) & (sv <= k)) ? c : v;
        }
        slot = sSrc[v] + (unsigned)(k - sStart[v]);
        s = sVs[v];
    };
    AdjRec nxt{};
    if ((int)threadIdx.x < total) {
        slot_of(threadIdx.x, nxt.slot, nxt.s);
        load_rec(st, dst, nxt);
    }
    // the loads below are unconditional (clamped indices): a load under a branch
    // makes the compiler wait for all outstanding loads where the paths join
    for (int k = (int)threadIdx.x; k < total; k += BS) {
        const AdjRec cur = nxt;
        const int t = cur.to >= 0 ? cur.to : 0;
        // z is not stored in the 2-D split form (kZMid): no adjoint flows through it
        const float a[6] = {ain.ox[t], ain.oy[t], 0.0f, ain.u[t], ain.v[t], ain.w[t]};
        slot_of(k + BS < total ? k + BS : k, nxt.slot, nxt.s);
        load_rec(st, dst, nxt);
        const int ci = cur.cell & 0xFFFF, cj = cur.cell >> 16;
        int ii0 = ci - cx0, jj0 = cj - cy0;
        // a particle the forward checks excluded goes through the same
        // instructions on a clamped cell and contributes zeros (no branch: see above)
        const bool ok = ii0 >= 2 && ii0 <= T + 1 && jj0 >= 2 && jj0 <= T + 1 && cur.to >= 0;
        ii0 = min(max(ii0, 2), T + 1);
        jj0 = min(max(jj0, 2), T + 1);
        const int lc = (jj0 - 2) * T + (ii0 - 2);
        const float4 gc[6] = {sC[6 * lc], sC[6 * lc + 1], sC[6 * lc + 2], sC[6 * lc + 3], sC[6 * lc + 4], sC[6 * lc + 5]};
        PState<float> q;
        push_fwd<float>(ap, cur.s, cur.ox, cur.oy, kZMid, cur.u, cur.v, cur.w, gc, q);
        const int e0 = jj0 * W + ii0;
        const int e2 = e0 + min(max(q.d[1], -1), 1) * W + min(max(q.d[0], -1), 1);
        const float jb1[8] = {sJb[e0], sJb[e0 + W], sJb[WW + e0], sJb[WW + e0 + 1],
                              sJb[2 * WW + e0], sJb[2 * WW + e0 + 1], sJb[2 * WW + e0 + W], sJb[2 * WW + e0 + W + 1]};
        const float jb2[8] = {sJb[e2], sJb[e2 + W], sJb[WW + e2], sJb[WW + e2 + 1],
                              sJb[2 * WW + e2], sJb[2 * WW + e2 + 1], sJb[2 * WW + e2 + W], sJb[2 * WW + e2 + W + 1]};
        const float x1b[3] = {a[0], a[1], a[2]};
        const float u1b[3] = {a[3], a[4], a[5]};
        float x0b[3], unb[3], PEb[3], PBb[3];
        reverse<float>(ap, cur.s, q, gc, jb1, jb2, x1b, u1b, x0b, unb, PEb, PBb);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            PEb[c] = ok ? PEb[c] : 0.0f;
            PBb[c] = ok ? PBb[c] : 0.0f;
        }
        const int n0 = (jj0 - 2) * N1 + (ii0 - 2);
        const int nc[4] = {n0, n0 + 1, n0 + N1, n0 + N1 + 1};
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const float w = ok ? q.w[m] : 0.0f;
            atomicAdd(&sG[nc[m]], (GA)(w * PEb[0]));
            atomicAdd(&sG[NN + nc[m]], (GA)(w * PEb[1]));
            atomicAdd(&sG[2 * NN + nc[m]], (GA)(w * PEb[2]));
            atomicAdd(&sG[3 * NN + nc[m]], (GA)(w * PBb[0]));
            atomicAdd(&sG[4 * NN + nc[m]], (GA)(w * PBb[1]));
            atomicAdd(&sG[5 * NN + nc[m]], (GA)(w * PBb[2]));
        }
        const unsigned slot = cur.slot;
        aout.ox[slot] = ok ? x0b[0] : 0.0f; aout.oy[slot] = ok ? x0b[1] : 0.0f; aout.oz[slot] = 0.0f;
        aout.u[slot] = ok ? unb[0] : 0.0f; aout.v[slot] = ok ? unb[1] : 0.0f; aout.w[slot] = ok ? unb[2] : 0.0f;
    }
    __syncthreads();
    // node (gi, gj) of the tile is reference node (3 + tx T + gi, 3 + ty T + gj)
    double *gdst[6] = {gEx, gEy, gEz, gBx, gBy, gBz};
    for (int e = threadIdx.x; e < NN; e += BS) {
        const long long o = (long long)(2 + tx * T + e % N1) + (long long)mx * (2 + ty * T + e / N1);
#pragma unroll
        for (int cc = 0; cc < 6; cc++) {
            const double val = (double)sG[cc * NN + e];
            if (val != 0.0) unsafeAtomicAdd(&gdst[cc][o], val);
        }
    }
}

// Operations over the occupied slots of a bin layout, one workgroup per tile.
// mode 0: pos[id[slot]] = slot
// mode 1: aout[slot] = ain[pos[id[slot]]]          (adjoint to a new slot order)
// mode 2: du[id[slot]] = scale * ain.u[slot], ...  (adjoint to the caller's order)
__global__ __launch_bounds__(256) void k_fa_slots(TileP tp, int mode, const long long *__restrict__ base,
                                                  const int *__restrict__ cnt, const int *__restrict__ id,
                                                  int *__restrict__ pos, EPart<float> ain, EPart<float> aout,
                                                  double scale, double *__restrict__ du, double *__restrict__ dv,
                                                  double *__restrict__ dw)
{
    const int tile = blockIdx.x;
    for (int b = 0; b < tp.NB; b++) {
        const long long b0 = base[tile * tp.NB + b];
        const int n = cnt[tile * tp.NB + b];
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const long long slot = b0 + k;
            const int p = id[slot];
            if (mode == 0) {
                pos[p] = (int)slot;
            } else if (mode == 1) {
                const int o = pos[p];
                aout.ox[slot] = ain.ox[o]; aout.oy[slot] = ain.oy[o]; aout.oz[slot] = ain.oz[o];
                aout.u[slot] = ain.u[o]; aout.v[slot] = ain.v[o]; aout.w[slot] = ain.w[o];
            } else {
                du[p] = scale * (double)ain.u[slot];
                dv[p] = scale * (double)ain.v[slot];
                dw[p] = scale * (double)ain.w[slot];
            }
        }
    }
}

}  // namespace

void launch_fa_adjoint(const GridP &g, const TileP &tp, const AdjP &ap, float *const E[3], float *const B[3],
                       double *const Eb[3], const long long *base, const int *cnt, PartSplit st, const int *dst,
                       EPart<float> ain, EPart<float> aout, double *const gE[3], double *const gB[3], hipStream_t s)
{
#define FA_ADJ(TT, BSZ)                                                                                              \
    hipLaunchKernelGGL((k_fa_adjoint<TT, BSZ>), dim3(tp.ntiles), dim3(BSZ), 0, s, g, tp, ap, E[0], E[1], E[2],     \
                       B[0], B[1], B[2], Eb[0], Eb[1], Eb[2], base, cnt, st, dst, ain, aout, gE[0], gE[1], gE[2],  \
                       gB[0], gB[1], gB[2])
    switch (tp.T) {
    case 8: FA_ADJ(8, 256); break;
    case 16: FA_ADJ(16, 256); break;
    case 32: FA_ADJ(32, 512); break;
    default: break;
    }
#undef FA_ADJ
}

void launch_fa_slots(const TileP &tp, int mode, const long long *base, const int *cnt, const int *id, int *pos,
                     EPart<float> ain, EPart<float> aout, double scale, double *du, double *dv, double *dw,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_fa_slots, dim3(tp.ntiles), dim3(256), 0, s, tp, mode, base, cnt, id, pos, ain, aout, scale,
                       du, dv, dw);
}

}  // namespace pictor
