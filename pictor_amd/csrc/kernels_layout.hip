// kernels_layout.hip -- tile layout (counting sort into tiles), compaction,
// download conversion and particle reductions (gfx950).
//
// The counting sort plays the role of ReorderPrtlArr (src/mem/mem_prtl.F90:
// 166-301): a histogram of particles per (tile, species), a prefix sum on the
// host (a few thousand entries), and a scatter.  It runs at upload and when
// the per-step binning overflows a bin (spill), not every step: between
// layouts, the push kernel keeps particles sorted by tile on its own.
#include "pictor_common.h"

namespace pictor {
namespace {

// Bijective stride permutation of [0, n): 1000003 is prime and larger than any
// group, so it is coprime with n.  Neighbouring slots (same cell) land far apart.
__device__ inline long long perm_slot(long long s, const int *gtotal, int grp)
{
    if (!gtotal) return s;
    const long long n = gtotal[grp];
    return n > 1 ? (s * 1000003LL) % n : s;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ inline int wrapi(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

// tile of a position given in reference coordinates
__device__ inline int tile_of_ref(const GridP &g, const TileP &tp, double x, double y, double z, bool wrap)
{
    int cx = (int)floor(x) - 3, cy = (int)floor(y) - 3, cz = 0;
    if (wrap) { cx = wrapi(cx, g.nx); cy = wrapi(cy, g.ny); }
    else { cx = clampi(cx, 0, g.nx - 1); cy = clampi(cy, 0, g.ny - 1); }
    if (g.ndim == 3) {
        cz = (int)floor(z) - 3;
        cz = wrap ? wrapi(cz, g.nz) : clampi(cz, 0, g.nz - 1);
    }
    return (cz / tp.T * tp.nty + cy / tp.T) * tp.ntx + cx / tp.T;
}

// reference coordinate -> (wrapped physical cell, FP32 offset in [0,1))
__device__ inline void split_coord(double x, int n, int &cell, float &off)
{
    const double f = floor(x);
    int c = wrapi((int)f - 3, n);
    float o = (float)(x - f);
    if (o >= 1.0f) { o = 0.0f; c = (c + 1 == n) ? 0 : c + 1; }
    cell = c;
    off = o;
}

// Upload: validate, compute the tile of every particle and count per (tile, species).
// The FP64 form keeps positions as given (a particle may sit up to one cell in a
// ghost layer, as after the reference's exchange); the split form wraps them.
__global__ void k_raw_tiles(GridP g, PhysP ph, TileP tp, PartF64 raw, long long n, int split, int *tile_out,
                            int *counts, int *err)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int flv = raw.flv[i];
    if (flv == 0) { tile_out[i] = -1; return; }
    const double x = raw.x[i], y = raw.y[i], z = raw.z[i];
    bool ok = flv >= 1 && flv <= ph.nflvr;
    ok = ok && x >= 2.0 && x < (double)(g.mx - 1) && y >= 2.0 && y < (double)(g.my - 1);
    if (g.ndim == 3) ok = ok && z >= 2.0 && z < (double)(g.mz - 1);
    else ok = ok && z >= 0.0 && z < 3.0;
    // split = 1: every charge is its flavour's; split = 2 (charge weights): a nonzero
    // multiple of it with the same sign
    if (split == 1 && ok) ok = (raw.q[i] == ph.flvq[flv - 1]);
    if (split == 2 && ok) ok = (raw.q[i] / ph.flvq[flv - 1] > 0.0);
    if (!ok) { atomicOr(err, 1); tile_out[i] = -1; return; }
    int grp;
    if (split) {
        // the tile of the split representation (offset rounding may move a particle one cell)
        int cx, cy, cz = 0;
        float o;
        split_coord(x, g.nx, cx, o);
        split_coord(y, g.ny, cy, o);
        if (g.ndim == 3) split_coord(z, g.nz, cz, o);
        grp = ((cz / tp.T * tp.nty + cy / tp.T) * tp.ntx + cx / tp.T) * tp.nsp + (flv - 1);
    } else {
        grp = tile_of_ref(g, tp, x, y, z, false);
    }
    tile_out[i] = grp;
    atomicAdd(&counts[grp], 1);
}

__global__ void k_raw_scatter_f64(TileP tp, PartF64 raw, long long n, const int *tile_of, const long long *cbase,
                                  int *fill, const int *gtotal, PartF64 out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = tile_of[i];
    if (t < 0) return;
    const int slot = atomicAdd(&fill[t], 1);
    const long long o = cbase[t] + perm_slot(slot, gtotal, t);
    out.q[o] = raw.q[i]; out.x[o] = raw.x[i]; out.y[o] = raw.y[i]; out.z[o] = raw.z[i];
    out.u[o] = raw.u[i]; out.v[o] = raw.v[i]; out.w[o] = raw.w[i]; out.flv[o] = raw.flv[i];
    if (out.tag && raw.tag) out.tag[o] = raw.tag[i];
}

__global__ void k_raw_scatter_split(GridP g, PhysP ph, TileP tp, PartF64 raw, long long n, const int *tile_of,
                                    const long long *cbase, int *fill, const int *gtotal, PartSplit out)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = tile_of[i];
    if (t < 0) return;
    const int s = raw.flv[i] - 1;
    int cx, cy, cz = 0;
    float ox, oy, oz;
    split_coord(raw.x[i], g.nx, cx, ox);
    split_coord(raw.y[i], g.ny, cy, oy);
    if (g.ndim == 3) split_coord(raw.z[i], g.nz, cz, oz);
    const int grp = t;
    const int slot = atomicAdd(&fill[grp], 1);
    const long long o = cbase[grp] + perm_slot(slot, gtotal, grp);
    out.cell[o] = cell_pack(g.ndim, cx, cy, cz);
    out.ox[o] = ox; out.oy[o] = oy;
    if (g.ndim == 3) out.oz[o] = oz;
    out.u[o] = (float)raw.u[i]; out.v[o] = (float)raw.v[i]; out.w[o] = (float)raw.w[i];
    // the carried value: the caller's tag, or (adjoint engine) the caller's index
    if (out.id) out.id[o] = raw.tag ? raw.tag[i] : (int)i;
    if (out.wq) out.wq[o] = (float)(raw.q[i] / ph.flvq[raw.flv[i] - 1]);
    (void)s;
}

__global__ void k_spill_tiles_f64(GridP g, TileP tp, PartF64 sp, int n, int *tile_out, int *counts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = tile_of_ref(g, tp, sp.x[i], sp.y[i], sp.z[i], false);
    tile_out[i] = t;
    atomicAdd(&counts[t], 1);
}

__global__ void k_spill_tiles_split(GridP g, TileP tp, PartSplit sp, int n, int *tile_out, int *counts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cx, cy, cz;
    cell_unpack(g.ndim, sp.cell[i], cx, cy, cz);
    const int t = (cz / tp.T * tp.nty + cy / tp.T) * tp.ntx + cx / tp.T;
    const int grp = t * tp.nsp + sp.sp[i];
    tile_out[i] = grp;
    atomicAdd(&counts[grp], 1);
}

// one workgroup per (tile, species) group: its ND bins -> one dense run at dst_off[group]
template <typename CopyF>
__device__ inline void compact_group(const TileP &tp, const long long *base, const int *cnt, long long dst0,
                                     const int *gtotal, CopyF copy)
{
    __shared__ int sStart[28];
    __shared__ long long sBase[27];
    const int grp = blockIdx.x;
    const int tile = grp / tp.nsp, s = grp % tp.nsp;
    const int b0 = tile * tp.NB + s * tp.ND;
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int d = 0; d < tp.ND; d++) { sStart[d] = acc; sBase[d] = base[b0 + d]; acc += cnt[b0 + d]; }
        sStart[tp.ND] = acc;
    }
    __syncthreads();
    const int total = sStart[tp.ND];
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
        int d = 0;
        while (k >= sStart[d + 1]) d++;
        copy(sBase[d] + (k - sStart[d]), dst0 + perm_slot(k, gtotal, grp));
    }
}

__global__ void k_compact_f64(TileP tp, PartF64 src, const long long *base, const int *cnt, const long long *dst_off,
                              const int *gtotal, PartF64 dst)
{
    compact_group(tp, base, cnt, dst_off[blockIdx.x], gtotal, [&](long long a, long long b) {
        dst.q[b] = src.q[a]; dst.x[b] = src.x[a]; dst.y[b] = src.y[a]; dst.z[b] = src.z[a];
        dst.u[b] = src.u[a]; dst.v[b] = src.v[a]; dst.w[b] = src.w[a]; dst.flv[b] = src.flv[a];
        if (dst.tag && src.tag) dst.tag[b] = src.tag[a];
    });
}

__global__ void k_compact_split(TileP tp, PartSplit src, const long long *base, const int *cnt,
                                const long long *dst_off, const int *gtotal, PartSplit dst)
{
    const int s = blockIdx.x % tp.nsp;
    compact_group(tp, base, cnt, dst_off[blockIdx.x], gtotal, [&](long long a, long long b) {
        dst.cell[b] = src.cell[a]; dst.ox[b] = src.ox[a]; dst.oy[b] = src.oy[a];
        if (dst.oz) dst.oz[b] = src.oz[a];
        dst.u[b] = src.u[a]; dst.v[b] = src.v[a]; dst.w[b] = src.w[a];
        if (dst.sp) dst.sp[b] = s;
        if (dst.id && src.id) dst.id[b] = src.id[a];
        if (dst.wq && src.wq) dst.wq[b] = src.wq[a];
    });
}

__global__ void k_spill_scatter_f64(PartF64 sp, int n, const int *grp_of, const long long *dst_off, int *fill,
                                    const int *gtotal, PartF64 dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int grp = grp_of[i];
    const long long o = dst_off[grp] + perm_slot(atomicAdd(&fill[grp], 1), gtotal, grp);
    dst.q[o] = sp.q[i]; dst.x[o] = sp.x[i]; dst.y[o] = sp.y[i]; dst.z[o] = sp.z[i];
    dst.u[o] = sp.u[i]; dst.v[o] = sp.v[i]; dst.w[o] = sp.w[i]; dst.flv[o] = sp.flv[i];
    if (dst.tag && sp.tag) dst.tag[o] = sp.tag[i];
}

__global__ void k_spill_scatter_split(PartSplit sp, int n, const int *grp_of, const long long *dst_off, int *fill,
                                      const int *gtotal, PartSplit dst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int grp = grp_of[i];
    const long long o = dst_off[grp] + perm_slot(atomicAdd(&fill[grp], 1), gtotal, grp);
    dst.cell[o] = sp.cell[i]; dst.ox[o] = sp.ox[i]; dst.oy[o] = sp.oy[i];
    if (dst.oz) dst.oz[o] = sp.oz[i];
    dst.u[o] = sp.u[i]; dst.v[o] = sp.v[i]; dst.w[o] = sp.w[i];
    if (dst.id && sp.id) dst.id[o] = sp.id[i];
    if (dst.wq && sp.wq) dst.wq[o] = sp.wq[i];
}

// split -> reference FP64 record (x = 3 + cell + offset; 2-D z = 1 + kZMid, z is not stored)
__global__ void k_split_to_f64(GridP g, PhysP ph, PartSplit src, long long n, PartF64 dst)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int cx, cy, cz;
    cell_unpack(g.ndim, src.cell[i], cx, cy, cz);
    const int s = src.sp[i];
    dst.x[i] = (double)(3 + cx) + (double)src.ox[i];
    dst.y[i] = (double)(3 + cy) + (double)src.oy[i];
    dst.z[i] = (g.ndim == 3) ? (double)(3 + cz) + (double)src.oz[i] : 1.0 + (double)kZMid;
    dst.u[i] = src.u[i]; dst.v[i] = src.v[i]; dst.w[i] = src.w[i];
    dst.q[i] = src.wq ? ph.flvq[s] * (double)src.wq[i] : ph.flvq[s];
    dst.flv[i] = s + 1;
    if (dst.tag && src.id) dst.tag[i] = src.id[i];
}

__global__ void k_tile_ids(TileP tp, const long long *dst_off, int *tile_out)
{
    const int grp = blockIdx.x;
    const long long a = dst_off[grp], b = dst_off[grp + 1];
    for (long long k = a + threadIdx.x; k < b; k += blockDim.x) tile_out[k] = grp / tp.nsp;
}

// a node index folded onto the physical nodes of a periodic axis (per = 1);
// a non-periodic axis keeps its ghost nodes
__device__ inline int fold_node(int n, int m, int len, int per)
{
    if (!per) return n;
    if (n < 3) return n + len;
    if (n > m - 3) return n - len;
    return n;
}

// Kinetic energy (CalcPrtlEnergy, savedata_routines.F90:283-308) and momentum
// partial sums per tile, plus CIC node charge for the Gauss-law residual
// (skipped when rho is null).
template <typename Rec>
__device__ inline void particle_diag_tile(const GridP &g, const PhysP &ph, const TileP &tp, const long long *base,
                                          const int *cnt, double *partial, double *rho, Rec rec)
{
    __shared__ double red[5][256];
    const int tile = blockIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int b = 0; b < tp.NB; b++) {
        const long long b0 = base[tile * tp.NB + b];
        const int n = cnt[tile * tp.NB + b];
        const int s = b / tp.ND;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            double x, y, z, u, v, w, q;
            int flv;
            rec(b0 + k, s, x, y, z, u, v, w, q, flv);
            const double gamma = sqrt(1.0 + u * u + v * v + w * w);
            const double c2 = ph.c * ph.c;
            double e = 0.0;
            if (q > 0) e = (gamma - 1.0) * 1.0 * (q * ph.qi / ph.flvrqm[flv - 1]) * c2;
            if (q < 0) e = (gamma - 1.0) * 1.0 * (ph.mass[1] * c2);
            acc[flv - 1] += e;
            const double m = ph.mass[flv - 1];
            acc[2] += m * u; acc[3] += m * v; acc[4] += m * w;
            if (!rho) continue;
            // CIC charge at the nodes of the particle's cell
            const double qt = q * ph.qi;
            const int i = (int)floor(x), j = (int)floor(y);
            const double dx = x - i, dy = y - j;
            const double sx[2] = {1.0 - dx, dx}, sy[2] = {1.0 - dy, dy};
            if (g.ndim == 2) {
                for (int bb = 0; bb < 2; bb++)
                    for (int aa = 0; aa < 2; aa++) {
                        const int ii = fold_node(i + aa, g.mx, g.nx, g.per[0]), jj = fold_node(j + bb, g.my, g.ny, g.per[1]);
                        unsafeAtomicAdd(&rho[fidx(g, ii, jj, 1)], qt * sx[aa] * sy[bb]);
                    }
            } else {
                const int kk0 = (int)floor(z);
                const double dz = z - kk0;
                const double sz[2] = {1.0 - dz, dz};
                for (int cc = 0; cc < 2; cc++)
                    for (int bb = 0; bb < 2; bb++)
                        for (int aa = 0; aa < 2; aa++) {
                            const int ii = fold_node(i + aa, g.mx, g.nx, g.per[0]), jj = fold_node(j + bb, g.my, g.ny, g.per[1]);
                            const int kk = fold_node(kk0 + cc, g.mz, g.nz, g.per[2]);
                            unsafeAtomicAdd(&rho[fidx(g, ii, jj, kk)], qt * sx[aa] * sy[bb] * sz[cc]);
                        }
            }
        }
    }
    for (int a = 0; a < 5; a++) red[a][threadIdx.x] = acc[a];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int a = 0; a < 5; a++) red[a][threadIdx.x] += red[a][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 5) partial[tile * 5 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_particle_diag_f64(GridP g, PhysP ph, TileP tp, PartF64 p, const long long *base,
                                                           const int *cnt, double *partial, double *rho)
{
    particle_diag_tile(g, ph, tp, base, cnt, partial, rho,
                       [&](long long o, int, double &x, double &y, double &z, double &u, double &v, double &w,
                           double &q, int &flv) {
                           x = p.x[o]; y = p.y[o]; z = p.z[o]; u = p.u[o]; v = p.v[o]; w = p.w[o];
                           q = p.q[o]; flv = p.flv[o];
                       });
}

__global__ __launch_bounds__(256) void k_particle_diag_split(GridP g, PhysP ph, TileP tp, PartSplit p,
                                                             const long long *base, const int *cnt, double *partial,
                                                             double *rho)
{
    particle_diag_tile(g, ph, tp, base, cnt, partial, rho,
                       [&](long long o, int s, double &x, double &y, double &z, double &u, double &v, double &w,
                           double &q, int &flv) {
                           int cx, cy, cz;
                           cell_unpack(g.ndim, p.cell[o], cx, cy, cz);
                           x = (double)(3 + cx) + (double)p.ox[o];
                           y = (double)(3 + cy) + (double)p.oy[o];
                           z = (g.ndim == 3) ? (double)(3 + cz) + (double)p.oz[o] : 1.0 + (double)kZMid;
                           u = p.u[o]; v = p.v[o]; w = p.w[o];
                           q = p.wq ? ph.flvq[s] * (double)p.wq[o] : ph.flvq[s];
                           flv = s + 1;
                       });
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

void launch_raw_tiles_f64(const GridP &g, const PhysP &ph, const TileP &tp, PartF64 raw, long long n, int split,
                          int *tile_out, int *counts, int *err, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_raw_tiles, dim3(nblk(n, 256)), dim3(256), 0, s, g, ph, tp, raw, n, split, tile_out, counts,
                       err);
}

void launch_raw_scatter_f64(const GridP &, const TileP &tp, PartF64 raw, long long n, const int *tile_of,
                            const long long *base, int *fill, const int *gtotal, PartF64 out, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_raw_scatter_f64, dim3(nblk(n, 256)), dim3(256), 0, s, tp, raw, n, tile_of, base, fill, gtotal,
                       out);
}

void launch_raw_scatter_split(const GridP &g, const PhysP &ph, const TileP &tp, PartF64 raw, long long n,
                              const int *tile_of, const long long *base, int *fill, const int *gtotal, PartSplit out,
                              hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_raw_scatter_split, dim3(nblk(n, 256)), dim3(256), 0, s, g, ph, tp, raw, n, tile_of, base,
                       fill, gtotal, out);
}

void launch_spill_tiles_f64(const GridP &g, const TileP &tp, PartF64 spill, int n, int *tile_out, int *counts,
                            hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spill_tiles_f64, dim3(nblk(n, 256)), dim3(256), 0, s, g, tp, spill, n, tile_out, counts);
}

void launch_spill_tiles_split(const GridP &g, const TileP &tp, PartSplit spill, int n, int *tile_out, int *counts,
                              hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spill_tiles_split, dim3(nblk(n, 256)), dim3(256), 0, s, g, tp, spill, n, tile_out, counts);
}

void launch_compact_f64(const TileP &tp, PartF64 src, const long long *base, const int *cnt, const long long *dst_off,
                        const int *gtotal, PartF64 dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_compact_f64, dim3(tp.ntiles * tp.nsp), dim3(256), 0, s, tp, src, base, cnt, dst_off, gtotal,
                       dst);
}

void launch_compact_split(const TileP &tp, PartSplit src, const long long *base, const int *cnt,
                          const long long *dst_off, const int *gtotal, PartSplit dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_compact_split, dim3(tp.ntiles * tp.nsp), dim3(256), 0, s, tp, src, base, cnt, dst_off, gtotal,
                       dst);
}

void launch_spill_scatter_f64(const TileP &, PartF64 spill, int n, const int *tile_of, const long long *dst_off,
                              int *fill, const int *gtotal, PartF64 dst, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spill_scatter_f64, dim3(nblk(n, 256)), dim3(256), 0, s, spill, n, tile_of, dst_off, fill,
                       gtotal, dst);
}

void launch_spill_scatter_split(const TileP &, PartSplit spill, int n, const int *tile_of, const long long *dst_off,
                                int *fill, const int *gtotal, PartSplit dst, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_spill_scatter_split, dim3(nblk(n, 256)), dim3(256), 0, s, spill, n, tile_of, dst_off, fill,
                       gtotal, dst);
}

void launch_split_to_f64(const GridP &g, const PhysP &ph, PartSplit src, long long n, PartF64 dst, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_split_to_f64, dim3(nblk(n, 256)), dim3(256), 0, s, g, ph, src, n, dst);
}

void launch_tile_ids(const TileP &tp, const long long *dst_off, long long, int *tile_out, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_ids, dim3(tp.ntiles * tp.nsp), dim3(256), 0, s, tp, dst_off, tile_out);
}

void launch_particle_diag_f64(const GridP &g, const PhysP &ph, const TileP &tp, PartF64 p, const long long *base,
                              const int *cnt, double *partial, double *rho, hipStream_t s)
{
    hipLaunchKernelGGL(k_particle_diag_f64, dim3(tp.ntiles), dim3(256), 0, s, g, ph, tp, p, base, cnt, partial, rho);
}

void launch_particle_diag_split(const GridP &g, const PhysP &ph, const TileP &tp, PartSplit p, const long long *base,
                                const int *cnt, double *partial, double *rho, hipStream_t s)
{
    hipLaunchKernelGGL(k_particle_diag_split, dim3(tp.ntiles), dim3(256), 0, s, g, ph, tp, p, base, cnt, partial,
                       rho);
}

}  // namespace pictor
