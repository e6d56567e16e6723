// bc.hip -- boundary faces of the Cartesian step (src/bc on one rank).
//
// The reference sets faces from the setup file (SetBC_Fld / SetBC_Prtl,
// bc.F90:23-63) and enforces them at two points of its loop: after the push
// (EnforceBC_PostMovDep, bc.F90:238-287: reflecting and removing particle
// faces) and at the end of the step (EnforceBC_Final, bc.F90:289-340:
// conducting and open field faces, then updateBC_Pos).  Its PML runs after
// the second B half step and after UpdateEfield (MAIN.F90:104,109).  Here:
//   * the particle faces act inside the push kernels (particle_faces in
//     push_common.h), and the current deposited beyond a reflecting wall is
//     folded back by one small launch per face (bc_after_push);
//   * the field faces and the PML are grid kernels over the face slabs or the
//     box, launched on the context stream in the reference's order;
//   * a non-periodic axis stops its halo exchanges (GridP::per) and its ghost
//     cells evolve under the Yee update (yee_lo..yee_hi).
// Positions are global; the local coordinate on one rank is position + 3.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/pictor_hip.h"
#include "ctx.h"
#include "pictor_common.h"

using namespace pictor;

namespace {

// ---------------------------------------------------------------------------
// current folds of the reflecting faces (RefBC_Prtl_Top/Bottom, bc_reflect.F90:
// 76-125, ReflCurrentLeft/Right :195-234): per component c, J_c(dst) +=
// sign * J_c(src), then J_c(src) = 0, on the plane normal to `axis` over all
// other indices
// ---------------------------------------------------------------------------
struct FoldOp {
    int axis;          // 0: planes of constant i, 1: planes of constant j
    int dst[3], src[3], valid[3];
    double sign[3];
};

__global__ __launch_bounds__(256) void k_refl_fold(GridP g, FoldOp op, double *__restrict__ Jx,
                                                   double *__restrict__ Jy, double *__restrict__ Jz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = op.axis == 0 ? g.my : g.mx;
    if (t >= (long long)n1 * g.mz) return;
    const int a = (int)(t % n1) + 1, k = (int)(t / n1) + 1;
    double *J[3] = {Jx, Jy, Jz};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        if (!op.valid[c]) continue;
        const long long od = op.axis == 0 ? fidx(g, op.dst[c], a, k) : fidx(g, a, op.dst[c], k);
        const long long os = op.axis == 0 ? fidx(g, op.src[c], a, k) : fidx(g, a, op.src[c], k);
        if (op.sign[c] > 0) J[c][od] = J[c][od] + J[c][os];
        else J[c][od] = J[c][od] - J[c][os];
        J[c][os] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// conducting and open field faces
// ---------------------------------------------------------------------------
// F1 = F2 = 0 (with their FP32 low parts, split form) on the index range lo..hi
// along `axis`, every other index
template <typename R>
__global__ __launch_bounds__(256) void k_zero_range(GridP g, int axis, int lo, int hi, R *__restrict__ F1,
                                                    R *__restrict__ F2, float *__restrict__ L1,
                                                    float *__restrict__ L2)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ncell) return;
    const int i = (int)(t % g.mx) + 1;
    const long long r = t / g.mx;
    const int j = (int)(r % g.my) + 1;
    const int c = axis == 0 ? i : j;
    if (c < lo || c > hi) return;
    F1[t] = R(0); F2[t] = R(0);
    if (L1) { L1[t] = 0.0f; L2[t] = 0.0f; }
}

// OpenBC_Fld_Right / Left, bc.F90:410-469, one thread per (j, k)
template <typename R>
__global__ __launch_bounds__(256) void k_open_x(GridP g, int xl, int right, R *__restrict__ Ex, R *__restrict__ Ey,
                                                R *__restrict__ Ez, R *__restrict__ Bx, R *__restrict__ By,
                                                R *__restrict__ Bz, float *__restrict__ Lx, float *__restrict__ Ly,
                                                float *__restrict__ Lz)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)g.my * g.mz) return;
    const int j = (int)(t % g.my) + 1, k = (int)(t / g.my) + 1;
    const long long ol = fidx(g, xl, j, k);
    auto half = [&](R *F, float *L, long long o) {
        F[o] = R(0.5) * F[o];
        if (L) L[o] = 0.5f * L[o];
    };
    if (right) {
        for (int i = xl + 1; i <= g.mx; i++) {
            const long long o = fidx(g, i, j, k);
            half(Ey, Ly, o);
            half(Ez, Lz, o);
            Bx[o] = Bx[ol];
        }
        for (int i = xl; i <= g.mx; i++) {
            const long long o = fidx(g, i, j, k);
            half(Ex, Lx, o);
            half(By, nullptr, o);
            half(Bz, nullptr, o);
        }
    } else {
        for (int i = 1; i <= xl - 1; i++) {
            const long long o = fidx(g, i, j, k);
            half(Ey, Ly, o);
            half(Ez, Lz, o);
            Bx[o] = Bx[ol];
            half(Ex, Lx, o);
            half(By, nullptr, o);
            half(Bz, nullptr, o);
        }
    }
}

// ---------------------------------------------------------------------------
// PML, bc_pml.F90
// ---------------------------------------------------------------------------
struct PmlP {
    double pos[6];       // pml_pos (global), +-huge(1) on faces without a layer
    long long ind[6];    // pml_pos_ind
    double L[6];         // layer thickness (the default 10 on faces without a layer)
    const double *tab;   // a, b, c per (axis, E/B, index): tab[((axis * 2 + eb) * 3 + q) * mmax + i - 1]
    int mmax;
};

struct Psi12 {
    double *p[12];
};

__device__ inline bool pml_inside(const PmlP &pp, long long x, long long y, long long z)
{
    return x > pp.ind[0] && x < pp.ind[1] && y > pp.ind[2] && y < pp.ind[3] && z > pp.ind[4] && z < pp.ind[5];
}

__device__ inline double ptab(const PmlP &pp, int axis, int eb, int q, int i)
{
    return pp.tab[((axis * 2 + eb) * 3 + q) * pp.mmax + i - 1];
}

// UpdateMagFld_PML, bc_pml.F90:185-290; FP32 fields are updated in FP64
template <typename R>
__global__ __launch_bounds__(256) void k_pml_b(GridP g, PmlP pp, double fc, double idx_, double idy_, double idz_,
                                               const R *__restrict__ Ex, const R *__restrict__ Ey,
                                               const R *__restrict__ Ez, R *__restrict__ Bx, R *__restrict__ By,
                                               R *__restrict__ Bz, Psi12 ps)
{
    const int i = 2 + blockIdx.x * 64 + threadIdx.x;
    const int j = 2 + blockIdx.y * 4 + threadIdx.y;
    const int k = g.ndim == 3 ? 2 + (int)blockIdx.z : 1;
    if (i > g.mx - 1 || j > g.my - 1) return;
    if (pml_inside(pp, i - 3, j - 3, k - 3)) return;
    const double xg = (double)(i - 3) + 0.5, yg = (double)(j - 3) + 0.5, zg = (double)(k - 3) + 0.5;
    const double a_x = ptab(pp, 0, 1, 0, i), b_x = ptab(pp, 0, 1, 1, i), c_x = ptab(pp, 0, 1, 2, i);
    const double a_y = ptab(pp, 1, 1, 0, j), b_y = ptab(pp, 1, 1, 1, j), c_y = ptab(pp, 1, 1, 2, j);
    const double a_z = ptab(pp, 2, 1, 0, k), b_z = ptab(pp, 2, 1, 1, k), c_z = ptab(pp, 2, 1, 2, k);
    const bool three = g.ndim == 3;
    const long long o = fidx(g, i, j, k), oi = o + 1, oj = o + g.mx;
    const long long ok = three ? o + (long long)g.mx * g.my : o;
    double *B1x = ps.p[6], *B1y = ps.p[7], *B1z = ps.p[8], *B2x = ps.p[9], *B2y = ps.p[10], *B2z = ps.p[11];
    const double ex = Ex[o], ey = Ey[o], ez = Ez[o];
    const double exj = Ex[oj], eyi = Ey[oi], ezi = Ez[oi], ezj = Ez[oj];
    B1x[o] = a_y * B1x[o] + b_y * fc * idy_ * (ezj - ez);
    B1z[o] = a_x * B1z[o] + b_x * fc * idx_ * (eyi - ey);
    if (three) {
        const double exk = Ex[ok], eyk = Ey[ok];
        B1y[o] = a_z * B1y[o] + b_z * fc * idz_ * (exk - ex);
        B2x[o] = a_z * B2x[o] + b_z * fc * idz_ * (eyk - ey);
    }
    B2y[o] = a_x * B2y[o] + b_x * fc * idx_ * (ezi - ez);
    B2z[o] = a_y * B2z[o] + b_y * fc * idy_ * (exj - ex);
    double bx = Bx[o], by = By[o], bz = Bz[o];
    if (three) {
        const double exk = Ex[ok], eyk = Ey[ok];
        bx = bx - c_y * idy_ * (ezj - ez) + c_z * idz_ * (eyk - ey) - B1x[o] + B2x[o];
        by = by - c_z * idz_ * (exk - ex) + c_x * idx_ * (ezi - ez) - B1y[o] + B2y[o];
    } else {
        bx = bx - c_y * idy_ * (ezj - ez) - B1x[o];
        by = by + c_x * idx_ * (ezi - ez) + B2y[o];
    }
    bz = bz - c_x * idx_ * (eyi - ey) + c_y * idy_ * (exj - ex) - B1z[o] + B2z[o];
    const double *P = pp.pos, *L = pp.L;
    if (xg - 0.5 < P[0] - L[0] || xg - 0.5 > P[1] + L[1] || yg < P[2] - L[2] || yg > P[3] + L[3] || zg < P[4] - L[4] ||
        zg > P[5] + L[5])
        bx = 0.0;
    if (xg < P[0] - L[0] || xg > P[1] + L[1] || yg - 0.5 < P[2] - L[2] || yg - 0.5 > P[3] + L[3] || zg < P[4] - L[4] ||
        zg > P[5] + L[5])
        by = 0.0;
    if (xg < P[0] - L[0] || xg > P[1] + L[1] || yg < P[2] - L[2] || yg > P[3] + L[3] || zg - 0.5 < P[4] - L[4] ||
        zg - 0.5 > P[5] + L[5])
        bz = 0.0;
    Bx[o] = (R)bx; By[o] = (R)by; Bz[o] = (R)bz;
}

// UpdateElcFld_PML, bc_pml.F90:85-183; the split form's E is the FP32 pair
// (E, Lo) read and written as one FP64 value
template <typename R>
__global__ __launch_bounds__(256) void k_pml_e(GridP g, PmlP pp, double fc, double idx_, double idy_, double idz_,
                                               const R *__restrict__ Bx, const R *__restrict__ By,
                                               const R *__restrict__ Bz, R *__restrict__ Ex, R *__restrict__ Ey,
                                               R *__restrict__ Ez, float *__restrict__ Lx, float *__restrict__ Ly,
                                               float *__restrict__ Lz, Psi12 ps)
{
    const int i = 2 + blockIdx.x * 64 + threadIdx.x;
    const int j = 2 + blockIdx.y * 4 + threadIdx.y;
    const int k = g.ndim == 3 ? 2 + (int)blockIdx.z : 1;
    if (i > g.mx - 1 || j > g.my - 1) return;
    const double xg = (double)(i - 3), yg = (double)(j - 3), zg = (double)(k - 3);
    if (pml_inside(pp, i - 3, j - 3, k - 3)) return;
    const double a_x = ptab(pp, 0, 0, 0, i), b_x = ptab(pp, 0, 0, 1, i), c_x = ptab(pp, 0, 0, 2, i);
    const double a_y = ptab(pp, 1, 0, 0, j), b_y = ptab(pp, 1, 0, 1, j), c_y = ptab(pp, 1, 0, 2, j);
    const double a_z = ptab(pp, 2, 0, 0, k), b_z = ptab(pp, 2, 0, 1, k), c_z = ptab(pp, 2, 0, 2, k);
    const bool three = g.ndim == 3;
    const long long o = fidx(g, i, j, k), oi = o - 1, oj = o - g.mx;
    const long long ok = three ? o - (long long)g.mx * g.my : o;
    double *E1x = ps.p[0], *E1y = ps.p[1], *E1z = ps.p[2], *E2x = ps.p[3], *E2y = ps.p[4], *E2z = ps.p[5];
    const double bx = Bx[o], by = By[o], bz = Bz[o];
    const double bxj = Bx[oj], byi = By[oi], bzi = Bz[oi], bzj = Bz[oj];
    E1x[o] = a_y * E1x[o] + b_y * fc * idy_ * (bz - bzj);
    E1z[o] = a_x * E1z[o] + b_x * fc * idx_ * (by - byi);
    if (three) {
        const double bxk = Bx[ok], byk = By[ok];
        E1y[o] = a_z * E1y[o] + b_z * fc * idz_ * (bx - bxk);
        E2x[o] = a_z * E2x[o] + b_z * fc * idz_ * (by - byk);
    }
    E2y[o] = a_x * E2y[o] + b_x * fc * idx_ * (bz - bzi);
    E2z[o] = a_y * E2z[o] + b_y * fc * idy_ * (bx - bxj);
    double ex = Ex[o], ey = Ey[o], ez = Ez[o];
    if (Lx) { ex = ex + (double)Lx[o]; ey = ey + (double)Ly[o]; ez = ez + (double)Lz[o]; }
    if (three) {
        const double bxk = Bx[ok], byk = By[ok];
        ex = ex + c_y * idy_ * (bz - bzj) - c_z * idz_ * (by - byk) + E1x[o] - E2x[o];
        ey = ey + c_z * idz_ * (bx - bxk) - c_x * idx_ * (bz - bzi) + E1y[o] - E2y[o];
    } else {
        ex = ex + c_y * idy_ * (bz - bzj) + E1x[o];
        ey = ey - c_x * idx_ * (bz - bzi) - E2y[o];
    }
    ez = ez + c_x * idx_ * (by - byi) - c_y * idy_ * (bx - bxj) + E1z[o] - E2z[o];
    const double *P = pp.pos, *L = pp.L;
    if (xg + 0.5 < P[0] - L[0] || xg + 0.5 > P[1] + L[1] || yg < P[2] - L[2] || yg > P[3] + L[3] || zg < P[4] - L[4] ||
        zg > P[5] + L[5])
        ex = 0.0;
    if (xg < P[0] - L[0] || xg > P[1] + L[1] || yg + 0.5 < P[2] - L[2] || yg + 0.5 > P[3] + L[3] || zg < P[4] - L[4] ||
        zg > P[5] + L[5])
        ey = 0.0;
    if (xg < P[0] - L[0] || xg > P[1] + L[1] || yg < P[2] - L[2] || yg > P[3] + L[3] || zg + 0.5 < P[4] - L[4] ||
        zg + 0.5 > P[5] + L[5])
        ez = 0.0;
    if (Lx) {
        const float hx = (float)ex, hy = (float)ey, hz = (float)ez;
        Ex[o] = (R)hx; Ey[o] = (R)hy; Ez[o] = (R)hz;
        Lx[o] = (float)(ex - (double)hx); Ly[o] = (float)(ey - (double)hy); Lz[o] = (float)(ez - (double)hz);
    } else {
        Ex[o] = (R)ex; Ey[o] = (R)ey; Ez[o] = (R)ez;
    }
}

inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool face_is_pml(const pictor_ctx *c, int f) { return c->bc_fld[f] == PICTOR_BC_PML; }

// pml_params of a face: SetPMLParameters (bc_pml.F90:39-51) or the type defaults (:6-12)
struct PmlSide {
    double kappa_max, sigma_max, alpha_max, L;
    int n;
};

PmlSide pml_side(const pictor_ctx *c, int f)
{
    if (face_is_pml(c, f)) return PmlSide{1.0, 1.0, 9.0e-4, c->pml_L[f], 4};
    return PmlSide{1.0, 0.0, 1.0, 10.0, 4};
}

// real ** integer as the Fortran runtime evaluates it (repeated squaring)
double powi(double a, int n)
{
    double r = 1.0, x = a;
    unsigned u = (unsigned)n;
    for (;;) {
        if (u & 1u) r *= x;
        u >>= 1;
        if (!u) break;
        x *= x;
    }
    return r;
}

// Update_PML_Position, bc_pml.F90:54-69
void pml_position(const pictor_ctx *c, PmlP &pp, bool &none)
{
    const double huge = 2147483647.0;
    for (int f = 0; f < 6; f++) {
        const bool upper = f & 1;
        pp.pos[f] = upper ? huge : -huge;
        pp.ind[f] = upper ? 2147483647LL : -2147483647LL;
        if (face_is_pml(c, f)) {
            pp.pos[f] = c->bc_pos_fld[f];
            pp.ind[f] = upper ? (long long)std::ceil(pp.pos[f]) + 1 : (long long)std::floor(pp.pos[f]) - 1;
        }
        pp.L[f] = pml_side(c, f).L;
    }
    const GridP &g = c->g;
    none = (1 + 0 - 3 > pp.ind[0]) && (g.mx + 0 - 3 < pp.ind[1]) && (1 + 0 - 3 > pp.ind[2]) &&
           (g.my + 0 - 3 < pp.ind[3]) && (1 + 0 - 3 > pp.ind[4]) && (g.mz + 0 - 3 < pp.ind[5]);
}

// pml_ab / pml_c at global coordinate x along `axis`, bc_pml.F90:293-353
void pml_abc(const pictor_ctx *c, const PmlP &pp, double x, int axis, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    const int up = 2 * axis + 1, lo = 2 * axis;
    auto eval = [&](double d, const PmlSide &s) {
        const double alpha = s.alpha_max;
        const double sv = s.sigma_max * powi(d / s.L, s.n);
        const double sigma = (s.sigma_max < sv) ? s.sigma_max : sv;
        const double kv = 1.0 + powi(d / s.L, s.n) * (s.kappa_max - 1.0);
        const double kappa = (s.kappa_max < kv) ? s.kappa_max : kv;
        out[0] = std::exp(-(sigma / kappa) - alpha);
        out[1] = (sigma / (sigma * kappa + alpha * (kappa * kappa))) * (out[0] - 1.0);
        out[2] = c->ph.fldc * (1.0 / kappa - 1.0);
    };
    if (x > pp.pos[up]) eval(x - pp.pos[up], pml_side(c, up));
    if (x < pp.pos[lo]) eval(pp.pos[lo] - x, pml_side(c, lo));
}

// the coefficient tables of every axis for the E (integer) and B (half-integer) positions
int pml_tables(pictor_ctx *c, PmlP &pp, bool &none)
{
    pml_position(c, pp, none);
    const GridP &g = c->g;
    const int m[3] = {g.mx, g.my, g.mz};
    pp.mmax = std::max(g.mx, std::max(g.my, g.mz));
    if (c->pml_tab_stale) {
        std::vector<double> h((size_t)3 * 2 * 3 * pp.mmax, 0.0);
        for (int a = 0; a < 3; a++)
            for (int eb = 0; eb < 2; eb++)
                for (int i = 1; i <= m[a]; i++) {
                    double out[3];
                    pml_abc(c, pp, (double)(i - 3) + (eb ? 0.5 : 0.0), a, out);
                    for (int q = 0; q < 3; q++) h[(size_t)((a * 2 + eb) * 3 + q) * pp.mmax + i - 1] = out[q];
                }
        if (!c->pml_tab) HIPTRY(dalloc(&c->pml_tab, h.size()));
        HIPTRY(hipMemcpyAsync(c->pml_tab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPTRY(hipStreamSynchronize(c->stream));   // h is a local buffer
        c->pml_tab_stale = false;
    }
    pp.tab = c->pml_tab;
    return PICTOR_OK;
}

void refresh_particle_faces(pictor_ctx *c)
{
    bool any = false;
    for (int f = 0; f < 6; f++) {
        c->ph.pbc[f] = c->bc_prtl[f];
        c->ph.pbc_pos[f] = c->bc_pos_prtl[f] - 0.0 + 3.0;   // bc_pos_local, bc.F90:350-362
        any = any || c->bc_prtl[f] != PICTOR_BC_PERIODIC;
    }
    for (int a = 0; a < 3; a++) any = any || !c->g.per[a];
    c->ph.pbc_on = any ? 1 : 0;
}

int check_face(pictor_ctx *c, int axis, int side)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    HIPTRY(hipSetDevice(c->dev));
    if (c->cyl) return fail(PICTOR_ERR_ARG, "boundary faces: Cartesian geometry only");
    if (axis < 0 || axis >= c->g.ndim) return fail(PICTOR_ERR_ARG, "boundary face: axis must be 0 (x), 1 (y) or, in 3-D, 2 (z)");
    if (side != 0 && side != 1) return fail(PICTOR_ERR_ARG, "boundary face: side must be 0 (lower) or 1 (upper)");
    if (c->curr_filter > 0) return fail(PICTOR_ERR_ARG, "boundary faces with the current filter are not built (curr_filter must be 0)");
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "boundary face set between move_deposit and exchange_particles");
    return PICTOR_OK;
}

// PeriodicOff (subdomains.F90:508-547): both faces of the axis
void periodic_off(pictor_ctx *c, int axis)
{
    c->g.per[axis] = 0;
    c->bc_any = true;
}

}  // namespace

namespace pictor {

// The current folds of the reflecting faces after a push, in the order of
// EnforceBC_PostMovDep (top, bottom, right, left, then every face again).
int bc_after_push(pictor_ctx *c)
{
    bool refl = false;
    for (int f = 0; f < 4; f++) refl = refl || c->bc_prtl[f] == PICTOR_BC_REFL;
    if (!refl) return PICTOR_OK;
    const GridP &g = c->g;
    double *J[3] = {(double *)c->J[0], (double *)c->J[1], (double *)c->J[2]};
    auto launch = [&](const FoldOp &op) {
        const long long n = (long long)(op.axis == 0 ? g.my : g.mx) * g.mz;
        hipLaunchKernelGGL(k_refl_fold, dim3(nblk(n, 256)), dim3(256), 0, c->stream, g, op, J[0], J[1], J[2]);
    };
    auto face = [&](int f) {
        const double pos = c->bc_pos_prtl[f];
        FoldOp op{};
        if (f == 3) {   // RefBC_Prtl_Top
            const int W = (int)(pos - 0.0 + 3.0);
            const int ok = W >= 2 && W <= g.my - 2;
            op = FoldOp{1, {W, W - 1, W}, {W + 1, W, W + 1}, {ok, ok, ok}, {1.0, -1.0, 1.0}};
        } else if (f == 2) {   // RefBC_Prtl_Bottom
            const int W = (int)(pos - 0.0 + 3.0);
            const int ok = W >= 3 && W <= g.my - 2;
            op = FoldOp{1, {W, W, W}, {W - 1, W - 1, W - 1}, {ok, ok, ok}, {1.0, -1.0, 1.0}};
        } else if (f == 1) {   // ReflCurrentRight
            const int ind = (int)std::floor(pos - 0.0 + 3.0);
            const int ind2 = (int)std::floor(pos - 0.5 - 0.0 + 3.0);
            const int ok = ind >= 1 && ind <= g.mx - 1, ok2 = ind2 >= 1 && ind2 <= g.mx - 1;
            op = FoldOp{0, {ind2, ind, ind}, {ind2 + 1, ind + 1, ind + 1}, {ok2, ok, ok}, {-1.0, 1.0, 1.0}};
        } else {   // ReflCurrentLeft
            const int ind = (int)std::ceil(pos - 0.0 + 3.0);
            const int ind2 = (int)std::ceil(pos - 0.5 - 0.0 + 3.0);
            const int ok = ind >= 2 && ind <= g.mx, ok2 = ind2 >= 2 && ind2 <= g.mx;
            op = FoldOp{0, {ind2, ind, ind}, {ind2 - 1, ind - 1, ind - 1}, {ok2, ok, ok}, {-1.0, 1.0, 1.0}};
        }
        if (op.valid[0] || op.valid[1] || op.valid[2]) launch(op);
    };
    const int order[8] = {3, 2, 1, 0, 0, 1, 2, 3};
    for (int q = 0; q < 8; q++)
        if (c->bc_prtl[order[q]] == PICTOR_BC_REFL) face(order[q]);
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

// UpdateMagFld_PML (b = true) / UpdateElcFld_PML
int bc_pml(pictor_ctx *c, bool b)
{
    if (!c->pml_on) return PICTOR_OK;
    PmlP pp{};
    bool none = false;
    int rc = pml_tables(c, pp, none);
    if (rc) return rc;
    const GridP &g = c->g;
    if (none) {
        // DeallocatePMLArr: every auxiliary field starts again from zero
        if (b) {
            launch_fill_zero_n<double>(c->pml_psi, 6, g.ncell, c->stream);
            launch_fill_zero_n<double>(c->pml_psi + 6, 6, g.ncell, c->stream);
        }
        HIPTRY(hipGetLastError());
        return PICTOR_OK;
    }
    Psi12 ps;
    for (int q = 0; q < 12; q++) ps.p[q] = c->pml_psi[q];
    const dim3 grid((g.mx - 2 + 63) / 64, (g.my - 2 + 3) / 4, g.ndim == 3 ? g.mz - 2 : 1);
    const PhysP &ph = c->ph;
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3];
        for (int i = 0; i < 3; i++) { E[i] = (double *)c->E[i]; B[i] = (double *)c->B[i]; }
        if (b)
            hipLaunchKernelGGL(k_pml_b<double>, grid, dim3(64, 4), 0, c->stream, g, pp, ph.fldc, ph.inv_dx, ph.inv_dy,
                               ph.inv_dz, E[0], E[1], E[2], B[0], B[1], B[2], ps);
        else
            hipLaunchKernelGGL(k_pml_e<double>, grid, dim3(64, 4), 0, c->stream, g, pp, ph.fldc, ph.inv_dx, ph.inv_dy,
                               ph.inv_dz, B[0], B[1], B[2], E[0], E[1], E[2], (float *)nullptr, (float *)nullptr,
                               (float *)nullptr, ps);
    } else {
        float *E[3], *B[3];
        for (int i = 0; i < 3; i++) { E[i] = (float *)c->E[i]; B[i] = (float *)c->B[i]; }
        if (b)
            hipLaunchKernelGGL(k_pml_b<float>, grid, dim3(64, 4), 0, c->stream, g, pp, ph.fldc, ph.inv_dx, ph.inv_dy,
                               ph.inv_dz, E[0], E[1], E[2], B[0], B[1], B[2], ps);
        else
            hipLaunchKernelGGL(k_pml_e<float>, grid, dim3(64, 4), 0, c->stream, g, pp, ph.fldc, ph.inv_dx, ph.inv_dy,
                               ph.inv_dz, B[0], B[1], B[2], E[0], E[1], E[2], c->Elo[0], c->Elo[1], c->Elo[2], ps);
    }
    HIPTRY(hipGetLastError());
    return PICTOR_OK;
}

// EnforceBC_Final without the halo: the field faces, then updateBC_Pos
int bc_final(pictor_ctx *c)
{
    if (!c->bc_any) return PICTOR_OK;
    const GridP &g = c->g;
    const bool f64 = c->form == PICTOR_FORM_FP64;
    float *const *L = f64 ? nullptr : c->Elo;
    auto zero2 = [&](int axis, int lo, int hi, int c1, int c2) {
        if (f64)
            hipLaunchKernelGGL(k_zero_range<double>, dim3(nblk(g.ncell, 256)), dim3(256), 0, c->stream, g, axis, lo, hi,
                               (double *)c->E[c1], (double *)c->E[c2], (float *)nullptr, (float *)nullptr);
        else
            hipLaunchKernelGGL(k_zero_range<float>, dim3(nblk(g.ncell, 256)), dim3(256), 0, c->stream, g, axis, lo, hi,
                               (float *)c->E[c1], (float *)c->E[c2], L[c1], L[c2]);
    };
    auto open = [&](int xl, int right) {
        const long long n = (long long)g.my * g.mz;
        if (f64)
            hipLaunchKernelGGL(k_open_x<double>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, g, xl, right,
                               (double *)c->E[0], (double *)c->E[1], (double *)c->E[2], (double *)c->B[0],
                               (double *)c->B[1], (double *)c->B[2], (float *)nullptr, (float *)nullptr,
                               (float *)nullptr);
        else
            hipLaunchKernelGGL(k_open_x<float>, dim3(nblk(n, 256)), dim3(256), 0, c->stream, g, xl, right,
                               (float *)c->E[0], (float *)c->E[1], (float *)c->E[2], (float *)c->B[0],
                               (float *)c->B[1], (float *)c->B[2], L[0], L[1], L[2]);
    };
    const double *q = c->bc_pos_fld;
    // face 1 (x lower): CondBC_Fld_Left / OpenBC_Fld_Left
    if (c->bc_fld[0] == PICTOR_BC_COND) {
        int xl = (int)std::floor(q[0] - 0.0 + 3.0);
        if (xl >= 1) zero2(0, 1, std::min(g.mx, xl), 1, 2);
    } else if (c->bc_fld[0] == PICTOR_BC_OPEN) {
        const int xl = (int)std::ceil(q[0] - 0.0 + 3.0);
        if (xl >= 2 && xl <= g.mx) open(xl, 0);
    }
    // face 2 (x upper): CondBC_Fld_Right / OpenBC_Fld_Right
    if (c->bc_fld[1] == PICTOR_BC_COND) {
        int xl = (int)std::ceil(q[1] - 0.0 + 3.0);
        if (xl <= g.mx) zero2(0, std::max(1, xl), g.mx, 1, 2);
    } else if (c->bc_fld[1] == PICTOR_BC_OPEN) {
        const int xl = (int)std::ceil(q[1] - 0.0 + 3.0);
        if (xl >= 1 && xl <= g.mx - 1) open(xl, 1);
    }
    // face 3 (y lower): CondBC_Fld_Bottom
    if (c->bc_fld[2] == PICTOR_BC_COND) {
        int yl = (int)std::floor(q[2] - 0.0 + 3.0);
        if (yl >= 1) zero2(1, 1, std::min(g.my, yl), 0, 2);
    }
    // face 4 (y upper): CondBC_Fld_Top at the face's own position (bc.F90:318 passes
    // face 3's position, which would clear the whole box above the lower face)
    if (c->bc_fld[3] == PICTOR_BC_COND) {
        int yl = (int)std::ceil(q[3] - 0.0 + 3.0);
        if (yl <= g.my) zero2(1, std::max(1, yl), g.my, 0, 2);
    }
    HIPTRY(hipGetLastError());
    // updateBC_Pos, bc.F90:342-348
    bool moved_pml = false;
    for (int f = 0; f < 6; f++) {
        if (c->bc_speed[f] == 0.0) continue;
        if (c->bc_fld[f] != PICTOR_BC_PERIODIC) {
            c->bc_pos_fld[f] = c->bc_pos_fld[f] + c->bc_speed[f] * c->ph.c;
            moved_pml = moved_pml || face_is_pml(c, f);
        }
        if (c->bc_prtl[f] != PICTOR_BC_PERIODIC) c->bc_pos_prtl[f] = c->bc_pos_prtl[f] + c->bc_speed[f] * c->ph.c;
    }
    if (moved_pml) c->pml_tab_stale = true;
    refresh_particle_faces(c);
    return PICTOR_OK;
}

void bc_get(const pictor_ctx *c, BcState &s)
{
    s = BcState{};
    for (int a = 0; a < 3; a++) s.per[a] = c->g.per[a];
    for (int f = 0; f < 6; f++) {
        s.fld[f] = c->bc_fld[f]; s.prtl[f] = c->bc_prtl[f];
        s.pos_fld[f] = c->bc_pos_fld[f]; s.pos_prtl[f] = c->bc_pos_prtl[f];
        s.speed[f] = c->bc_speed[f]; s.pml_L[f] = c->pml_L[f];
    }
    s.pml_on = c->pml_on ? 1 : 0;
}

int bc_set(pictor_ctx *c, const BcState &s)
{
    bool any = false;
    for (int a = 0; a < 3; a++) { c->g.per[a] = s.per[a] ? 1 : 0; any = any || !s.per[a]; }
    for (int f = 0; f < 6; f++) {
        c->bc_fld[f] = s.fld[f]; c->bc_prtl[f] = s.prtl[f];
        c->bc_pos_fld[f] = s.pos_fld[f]; c->bc_pos_prtl[f] = s.pos_prtl[f];
        c->bc_speed[f] = s.speed[f]; c->pml_L[f] = s.pml_L[f];
        any = any || s.fld[f] != PICTOR_BC_PERIODIC || s.prtl[f] != PICTOR_BC_PERIODIC;
    }
    c->bc_any = any;
    if (s.pml_on && !c->pml_on) {
        for (int q = 0; q < 12; q++) HIPTRY(dalloc(&c->pml_psi[q], (size_t)c->g.ncell));
        c->pml_on = true;
    }
    c->pml_tab_stale = true;
    refresh_particle_faces(c);
    return PICTOR_OK;
}

void bc_release(pictor_ctx *c)
{
    for (int q = 0; q < 12; q++) { dfree(c->pml_psi[q]); c->pml_psi[q] = nullptr; }
    dfree(c->pml_tab);
    c->pml_tab = nullptr;
}

}  // namespace pictor

extern "C" {

int pictor_set_bc_fld(pictor_ctx *c, int32_t axis, int32_t side, int32_t type, double position, double thickness)
{
    int rc = check_face(c, axis, side);
    if (rc) return rc;
    const int f = 2 * axis + side;
    if (type == PICTOR_BC_COND && axis > 1) return fail(PICTOR_ERR_ARG, "conducting faces exist on x and y only (bc.F90:295-320)");
    if (type == PICTOR_BC_OPEN && axis != 0) return fail(PICTOR_ERR_ARG, "open field faces exist on x only (bc.F90:298-310)");
    if (type != PICTOR_BC_COND && type != PICTOR_BC_OPEN && type != PICTOR_BC_PML)
        return fail(PICTOR_ERR_ARG, "field face type must be PICTOR_BC_COND, PICTOR_BC_OPEN or PICTOR_BC_PML");
    if (!std::isfinite(position)) return fail(PICTOR_ERR_ARG, "boundary position must be finite");
    if (type == PICTOR_BC_PML) {
        if (!(thickness > 0.0)) return fail(PICTOR_ERR_ARG, "PML thickness must be positive");
        // CheckAttnBCPositioning, bc.F90:364-401: the layer stays inside the box
        const int n = axis == 0 ? c->g.nx : (axis == 1 ? c->g.ny : c->g.nz);
        if ((side == 0 && position - thickness < 0.0) || (side == 1 && position + thickness > n))
            return fail(PICTOR_ERR_ARG, "the PML layer extends beyond the simulation box");
        if (!c->pml_on) {
            for (int q = 0; q < 12; q++) {
                HIPTRY(dalloc(&c->pml_psi[q], (size_t)c->g.ncell));
                launch_fill_zero<double>(c->pml_psi[q], c->g.ncell, c->stream);
            }
            HIPTRY(hipGetLastError());
            c->pml_on = true;
        }
        c->pml_L[f] = thickness;
    }
    c->bc_fld[f] = type;
    c->bc_pos_fld[f] = position;
    c->pml_tab_stale = true;
    periodic_off(c, axis);
    refresh_particle_faces(c);
    return PICTOR_OK;
}

int pictor_set_bc_prtl(pictor_ctx *c, int32_t axis, int32_t side, int32_t type, double position)
{
    int rc = check_face(c, axis, side);
    if (rc) return rc;
    if (type == PICTOR_BC_REFL && axis > 1) return fail(PICTOR_ERR_ARG, "reflecting particle faces exist on x and y only (bc.F90:245-274)");
    if (type != PICTOR_BC_REFL && type != PICTOR_BC_REMV)
        return fail(PICTOR_ERR_ARG, "particle face type must be PICTOR_BC_REFL or PICTOR_BC_REMV");
    if (!std::isfinite(position)) return fail(PICTOR_ERR_ARG, "boundary position must be finite");
    const int f = 2 * axis + side;
    c->bc_prtl[f] = type;
    c->bc_pos_prtl[f] = position;
    periodic_off(c, axis);
    refresh_particle_faces(c);
    return PICTOR_OK;
}

int pictor_set_boundary_speed(pictor_ctx *c, int32_t axis, int32_t side, double speed)
{
    int rc = check_face(c, axis, side);
    if (rc) return rc;
    if (!std::isfinite(speed)) return fail(PICTOR_ERR_ARG, "boundary speed must be finite");
    c->bc_speed[2 * axis + side] = speed;
    return PICTOR_OK;
}

int pictor_pml(pictor_ctx *c, int32_t which)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    HIPTRY(hipSetDevice(c->dev));
    if (which != 0 && which != 1) return fail(PICTOR_ERR_ARG, "which must be 0 (E) or 1 (B)");
    if (c->cyl) return PICTOR_OK;   // no PML faces on the cylindrical grids
    int rc = bc_pml(c, which == 1);
    if (rc || !c->pml_on) return rc;
    // the halo of the field, as SendRecvFlds follows the PML in MAIN.F90:105,110
    return pictor_field_halo(c, which);
}

int pictor_enforce_bc_final(pictor_ctx *c)
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    HIPTRY(hipSetDevice(c->dev));
    if (c->cyl) return PICTOR_OK;
    return bc_final(c);
}

}  // extern "C"
