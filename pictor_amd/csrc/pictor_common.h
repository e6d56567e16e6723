// pictor_common.h -- shared types and launchers of the gfx950 particle step.
//
// Reference layout on the device: every field is one (mx,my,mz) column-major
// array whose element (1,1,1) sits at offset 0 (src/vars.F90:23-24), so the
// host's Fortran arrays are copied without reshaping.  Particles live in
// tile bins: bin (tile, species, d) holds the particles of `tile` that
// arrived from direction d = (ddx+1) + 3*(ddy+1) [+ 9*(ddz+1)] in the last
// step, d = centre for the ones that stayed.  Every bin has exactly one
// writer (the source tile's workgroup), so binning needs no global atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pictor {

// Split-form current: every contribution is rounded to a multiple of 2^-30
// (charge x cell), below 2^31 counts for a unit charge, so one conversion makes
// it an integer; a tile sums its window with 64-bit integer LDS atomics and
// adds it to J, which holds the counts in FP64.  All partial sums stay integers
// below 2^53, so every addition is exact: the sum does not depend on the order
// of the atomics and is deterministic.  A node gets contributions from at most
// 4 (2-D) or 8 (3-D) tile windows and their ghost images; each tile's partial
// sum is checked against kFixLimit at the flush, which keeps the total below
// 16 * 2^14 = 2^18 charge cells, far inside 2^53 * 2^-30 = 2^23.
constexpr double kFixScale = 1073741824.0;           // 2^30
constexpr double kFixInv = 1.0 / 1073741824.0;
constexpr double kFixLimit = 16384.0;                // per-tile |node sum| bound (charge x cell)

enum ErrBits : int { kErrBounds = 1, kErrCapacity = 2, kErrPrecision = 4 };

struct GridP {
    int ndim, nx, ny, nz, mx, my, mz;
    long long ncell;   // mx*my*mz
    int per[3];        // 1: periodic axis; 0 once a boundary face is set on it
};

// range of cells the Yee update covers along `axis`.  The reference loops over
// 2..m-1 (em_update.F90) and lets the halo overwrite what it reaches.  In the
// periodic box the halo reaches every ghost, so the physical cells 3..m-3 are
// enough; once an axis is non-periodic, some ghosts (those of that axis, and
// where its ghosts meet another axis's, which the trimmed faces skip) keep the
// updated values, so every axis is updated over 2..m-1 as in the reference.
__host__ __device__ inline bool yee_full(const GridP &g)
{
    return !(g.per[0] && g.per[1] && (g.ndim == 2 || g.per[2]));
}
__host__ __device__ inline int yee_lo(const GridP &g, int axis)
{
    return (axis == 2 && g.ndim == 2) ? 1 : (yee_full(g) ? 2 : 3);
}
__host__ __device__ inline int yee_hi(const GridP &g, int axis)
{
    const int m = axis == 0 ? g.mx : (axis == 1 ? g.my : g.mz);
    return (axis == 2 && g.ndim == 2) ? 1 : (yee_full(g) ? m - 1 : m - 3);
}

struct PhysP {
    double c, sqc, cinv, qi, fldc, fld_halfc;
    double inv_dx, inv_dy, inv_dz;
    double xmin, xmax, ymin, ymax, zmin, zmax, xlen, ylen, zlen;
    double bext[3];
    double flvrqm[4];  // q/m per flavour (initialise.F90:84-85)
    double flvq[4];    // charge of one particle per flavour (split form)
    double mass[4];    // massi, masse (initialise.F90:49-50)
    int nflvr;
    int fdtd;          // 0 Yee, 1 Lehe
    double lehe[12];   // Lehe coefficients (fdtd 1)
    // particle faces of EnforceBC_PostMovDep: type per face (x lo, x hi, y lo,
    // y hi, z lo, z hi; 0 none, kBcRefl, kBcRemv) and local position (global + 3);
    // pbc_on: a face is set or an axis is not periodic (the push's face variant)
    int pbc[6];
    double pbc_pos[6];
    int pbc_on;
};

// particle face types (the C ABI's PICTOR_BC_REFL / PICTOR_BC_REMV)
constexpr int kBcRefl = 4, kBcRemv = 5;

struct TileP {
    int T, ntx, nty, ntz, ntiles;
    int ND;    // 9 (2-D) or 27 (3-D) arrival directions
    int nsp;   // species bins per tile (1 in the FP64 form, nflvr in the split form)
    int NB;    // nsp*ND bins per tile
    int xcd_remap;  // 1 when ntiles % 8 == 0: give each XCD a contiguous block of tiles
};

// FP64 form: the reference SoA (src/vars.F90:43-45); tag: the caller's per-particle
// tag (tagp), carried only when the particles were uploaded with tags
struct PartF64 {
    double *q = nullptr, *x = nullptr, *y = nullptr, *z = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;
    int *flv = nullptr;
    int *tag = nullptr;
};

// FP32 split form: packed physical cell coordinates + FP32 in-cell offsets.
// Cell word: 2-D  i | j << 16 (nx, ny <= 32768); 3-D  i | j << 11 | k << 22
// (nx, ny <= 2048, nz <= 512), so that no integer division is needed.
__host__ __device__ inline int cell_pack(int ndim, int i, int j, int k)
{
    return ndim == 2 ? (i | (j << 16)) : (i | (j << 11) | (k << 22));
}
__host__ __device__ inline void cell_unpack(int ndim, int c, int &i, int &j, int &k)
{
    if (ndim == 2) { i = c & 0xFFFF; j = (c >> 16) & 0xFFFF; k = 0; }
    else { i = c & 0x7FF; j = (c >> 11) & 0x7FF; k = (c >> 22) & 0x1FF; }
}

// 2-D split form: z is not stored.  A 2-D run is invariant along z; the
// reference carries z only to place the relay point of its Jz zigzag
// (movdep.F90:185), which splits Jz between the two half-segments.  The split
// form starts every particle's step at mid-cell in z (offset kZMid) and never
// stores the advanced z: with |dz| <= c dt < 1/2 cell per step the relay stays at
// the midpoint, both halves carry half the z charge flux, and the record is the
// cell word, two offsets and three momenta (24 B).  Downloads report z = 1.5,
// the centre of the reference's 2-D range [1, 2).
constexpr float kZMid = 0.5f;

struct PartSplit {
    int *cell = nullptr;
    float *ox = nullptr, *oy = nullptr, *oz = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;   // oz: 3-D only
    int *sp = nullptr;   // species id, used only by the spill / raw buffers
    // a 32-bit value carried with the particle: the caller's tag (tagp) in a context
    // whose particles were uploaded with tags, the caller's index in the adjoint
    // engine's buffers; null otherwise
    int *id = nullptr;
    // the particle's charge in units of its flavour's charge (the reference's weight,
    // q = wtp * FlvrCharge, init_prtl.F90:161), in a context whose particles do not all
    // carry their flavour's charge; null otherwise (every weight 1)
    float *wq = nullptr;
};

__host__ __device__ inline long long fidx(const GridP &g, int i, int j, int k)
{
    return (long long)(i - 1) + (long long)g.mx * ((long long)(j - 1) + (long long)g.my * (long long)(k - 1));
}

__device__ inline int tile_remap(const TileP &tp, int bid)
{
    if (!tp.xcd_remap) return bid;
    int per = tp.ntiles >> 3;
    return (bid & 7) * per + (bid >> 3);
}

// ---------------- launchers (kernels_fields.hip) ----------------
template <typename R>
void launch_update_b_half(const GridP &g, const PhysP &ph, R *const E[3], R *const B[3], hipStream_t s);
// Lehe B half step; lehe = alpha_x, alpha_y, alpha_z, beta_xy, beta_xz, beta_yx, beta_yz, beta_zx,
// beta_zy, delta_x, delta_y, delta_z (em_update_lehe.F90:8-25)
template <typename R>
void launch_update_b_half_lehe(const GridP &g, const PhysP &ph, const double lehe[12], R *const E[3], R *const B[3],
                               hipStream_t s);
template <typename R>
void launch_update_e(const GridP &g, const PhysP &ph, R *const E[3], R *const B[3], hipStream_t s);
// E update fused with AddCurrent; J is double (FP64 form) or fixed point (split form)
template <typename R, typename JT>
void launch_update_e_add_current(const GridP &g, const PhysP &ph, R *const E[3], R *const B[3],
                                 JT *const J[3], double jscale, hipStream_t s);
template <typename R, typename JT>
void launch_add_current(const GridP &g, const PhysP &ph, R *const E[3], JT *const J[3], double jscale,
                        hipStream_t s);
// up to six arrays handed to one launch (blockIdx.y selects the array or set of three)
template <typename T>
struct Ptrs6 {
    T *p[6];
};
// G / K: an optional second set of three arrays treated in the same launch
template <typename R>
void launch_ghost_copy(const GridP &g, R *const F[3], hipStream_t s, R *const G[3] = nullptr);
template <typename JT>
void launch_current_fold(const GridP &g, JT *const J[3], hipStream_t s, JT *const K[3] = nullptr);
template <typename T>
void launch_fill_zero(T *p, long long n, hipStream_t s);
template <typename T>
void launch_fill_zero_n(T *const *p, int count, long long n, hipStream_t s);   // count <= 6
template <typename T>
void launch_copy_n(T *const *dst, T *const *src, int count, long long n, hipStream_t s);   // count <= 6
// smoothen_current: `passes` rounds of the 1-2-1 filter along each axis, J in
// place, T (three arrays of ncell) as scratch
void launch_smoothen_current(const GridP &g, int passes, double *const J[3], double *const T[3], hipStream_t s);
template <typename R>
void launch_field_energy(const GridP &g, R *const E[3], R *const B[3], double *partial, int nblk, hipStream_t s,
                         R *const *L = nullptr);
template <typename R>
void launch_gauss(const GridP &g, R *const E[3], const double *rho, double *partial, int nblk, hipStream_t s,
                  R *const *L = nullptr);
// FP32 split form: E as an FP32 (hi, lo) pair updated in FP64 (L = lo); B may be null when !curl, J when !add
void launch_update_e_pair(const GridP &g, const PhysP &ph, float *const E[3], float *const L[3], float *const B[3],
                          double *const J[3], double jscale, bool curl, bool add, hipStream_t s);
void launch_pair_from_f64(const double *src, float *hi, float *lo, long long n, hipStream_t s);
void launch_pair_to_f64(const float *hi, const float *lo, double *dst, long long n, hipStream_t s);
void launch_convert_fields_f32(const double *src, float *dst, long long n, hipStream_t s);
void launch_convert_fields_f64(const float *src, double *dst, long long n, hipStream_t s);
void launch_convert_jfix(const double *src, double *dst, long long n, double scale, hipStream_t s);

// ---------------- launchers (kernels_push.hip) ----------------
struct BinsP {
    const long long *base;  // per bin
    const int *cap;         // per bin
    const int *cnt_in;      // per bin, buffer read this step
    int *cnt_out;           // per bin, buffer written this step
};

struct SpillP {
    int *count;
    int cap;
    int *err;
};

void launch_push2d_f64(int T, const GridP &g, const PhysP &ph, const TileP &tp, double *const E[3],
                       double *const B[3], double *const J[3], PartF64 A, PartF64 Bo, BinsP bins,
                       PartF64 spill, SpillP sp, hipStream_t s);
void launch_push2d_split(int T, const GridP &g, const PhysP &ph, const TileP &tp, float *const E[3],
                         float *const B[3], double *const J[3], PartSplit A, PartSplit Bo, BinsP bins,
                         PartSplit spill, SpillP sp, hipStream_t s);
// the same step carrying the particle indices (A.id -> Bo.id); dst (optional):
// the output slot of every input slot (-1 for a particle sent to the spill buffer)
void launch_push2d_split_track(int T, const GridP &g, const PhysP &ph, const TileP &tp, float *const E[3],
                               float *const B[3], double *const J[3], PartSplit A, PartSplit Bo, BinsP bins,
                               PartSplit spill, SpillP sp, int *dst, hipStream_t s);

// kernels_push3d.hip: T = 4 or 8
void launch_push3d_f64(int T, const GridP &g, const PhysP &ph, const TileP &tp, double *const E[3],
                       double *const B[3], double *const J[3], PartF64 A, PartF64 Bo, BinsP bins,
                       PartF64 spill, SpillP sp, hipStream_t s);
void launch_push3d_split(int T, const GridP &g, const PhysP &ph, const TileP &tp, float *const E[3],
                         float *const B[3], double *const J[3], PartSplit A, PartSplit Bo, BinsP bins,
                         PartSplit spill, SpillP sp, hipStream_t s);

// ---------------- launchers (kernels_layout.hip) ----------------
// raw FP64 particles (upload): tile id per particle + per (tile, species) counts
void launch_raw_tiles_f64(const GridP &g, const PhysP &ph, const TileP &tp, PartF64 raw, long long n,
                          int split, int *tile_out, int *counts, int *err, hipStream_t s);
// gtotal (per-group particle count): slots are permuted within each group so that the
// particles of one cell do not share a wavefront (same-address LDS atomics serialise)
void launch_raw_scatter_f64(const GridP &g, const TileP &tp, PartF64 raw, long long n, const int *tile_of,
                            const long long *base, int *fill, const int *gtotal, PartF64 out, hipStream_t s);
void launch_raw_scatter_split(const GridP &g, const PhysP &ph, const TileP &tp, PartF64 raw, long long n,
                              const int *tile_of, const long long *base, int *fill, const int *gtotal,
                              PartSplit out, hipStream_t s);
// spill particles: tile per spill particle + counts
void launch_spill_tiles_f64(const GridP &g, const TileP &tp, PartF64 spill, int n, int *tile_out, int *counts,
                            hipStream_t s);
void launch_spill_tiles_split(const GridP &g, const TileP &tp, PartSplit spill, int n, int *tile_out,
                              int *counts, hipStream_t s);
// compaction of every (tile, species) group of bins into one dense run
// gtotal == nullptr: dense copy in bin order (download); otherwise permuted as above
void launch_compact_f64(const TileP &tp, PartF64 src, const long long *base, const int *cnt,
                        const long long *dst_off, const int *gtotal, PartF64 dst, hipStream_t s);
void launch_compact_split(const TileP &tp, PartSplit src, const long long *base, const int *cnt,
                          const long long *dst_off, const int *gtotal, PartSplit dst, hipStream_t s);
void launch_spill_scatter_f64(const TileP &tp, PartF64 spill, int n, const int *tile_of,
                              const long long *dst_off, int *fill, const int *gtotal, PartF64 dst, hipStream_t s);
void launch_spill_scatter_split(const TileP &tp, PartSplit spill, int n, const int *tile_of,
                                const long long *dst_off, int *fill, const int *gtotal, PartSplit dst,
                                hipStream_t s);
// split -> FP64 reference coordinates, in place into an FP64 staging buffer
void launch_split_to_f64(const GridP &g, const PhysP &ph, PartSplit src, long long n, PartF64 dst,
                         hipStream_t s);
void launch_tile_ids(const TileP &tp, const long long *dst_off, long long n, int *tile_out, hipStream_t s);
// diagnostics over the particles of every bin
void launch_particle_diag_f64(const GridP &g, const PhysP &ph, const TileP &tp, PartF64 p, const long long *base,
                              const int *cnt, double *partial, double *rho, hipStream_t s);
void launch_particle_diag_split(const GridP &g, const PhysP &ph, const TileP &tp, PartSplit p,
                                const long long *base, const int *cnt, double *partial, double *rho,
                                hipStream_t s);

}  // namespace pictor
