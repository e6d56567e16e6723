/*
 * pictor_oracle_bc.c -- boundary faces of the CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see pictor_oracle.h).
 *
 * Restates, in the reference's expression order, the non-periodic faces of
 * src/bc on one rank (xborders = yborders = zborders = 0, so a global
 * position p is the local coordinate p + 3):
 *   EnforceBC_PostMovDep   bc.F90:238-287   reflecting and removing particle faces
 *   EnforceBC_Final        bc.F90:289-340   conducting and open field faces
 *   updateBC_Pos           bc.F90:342-348   faces moving at speed * c per step
 *   UpdateMagFld_PML,
 *   UpdateElcFld_PML       bc_pml.F90:85-290, coefficients :293-353
 * Not restated: inflow ('iflw'), attenuating ('attn') and curved faces, the
 * open particle face (OpenBC_Prtl_*), ionisation.
 */
#include <math.h>
#include <string.h>

#include "pictor_oracle.h"

#define IX(p, i, j, k) ((int64_t)((i) - 1) + (int64_t)(p)->mx * ((int64_t)((j) - 1) + (int64_t)(p)->my * (int64_t)((k) - 1)))

/* ------------------------------------------------------------------ */
/* Particle faces                                                      */
/* ------------------------------------------------------------------ */

/* RefPrtlLeft / RefPrtlRight, bc_reflect.F90:160-186: specular reflection */
static void ref_prtl_left(orc_particles *s, double *x, double *u, double xmin)
{
    for (int64_t n = 0; n < s->used; n++)
        if (s->flv[n] != 0 && x[n] < xmin) {
            u[n] = -u[n];
            x[n] = 2.0 * xmin - x[n];
        }
}

static void ref_prtl_right(orc_particles *s, double *x, double *u, double xmax)
{
    for (int64_t n = 0; n < s->used; n++)
        if (s->flv[n] != 0 && x[n] > xmax) {
            u[n] = -u[n];
            x[n] = 2.0 * xmax - x[n];
        }
}

/* J(:, a, :) op= J(:, b, :) over every i and k; sign -1 subtracts */
static void row_add(const orc_params *p, double *J, int a, int b, double sign)
{
    for (int k = 1; k <= p->mz; k++)
        for (int i = 1; i <= p->mx; i++) {
            if (sign > 0) J[IX(p, i, a, k)] = J[IX(p, i, a, k)] + J[IX(p, i, b, k)];
            else J[IX(p, i, a, k)] = J[IX(p, i, a, k)] - J[IX(p, i, b, k)];
        }
}

static void row_zero(const orc_params *p, double *J, int a)
{
    for (int k = 1; k <= p->mz; k++)
        for (int i = 1; i <= p->mx; i++) J[IX(p, i, a, k)] = 0.0;
}

/* J(a, :, :) op= J(b, :, :) over every j and k */
static void col_add(const orc_params *p, double *J, int a, int b, double sign)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++) {
            if (sign > 0) J[IX(p, a, j, k)] = J[IX(p, a, j, k)] + J[IX(p, b, j, k)];
            else J[IX(p, a, j, k)] = J[IX(p, a, j, k)] - J[IX(p, b, j, k)];
        }
}

static void col_zero(const orc_params *p, double *J, int a)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++) J[IX(p, a, j, k)] = 0.0;
}

/* RefBC_Prtl_Top, bc_reflect.F90:76-101 (CPU branch) */
static void refl_top(const orc_params *p, orc_fields *f, orc_particles *s, double ymax)
{
    const double yl = ymax - 0.0 + 3.0;
    const int W = (int)yl;   /* integer assignment truncates */
    if (yl < p->my - 1) ref_prtl_right(s, s->y, s->v, yl);
    if (W >= 2 && W <= p->my - 2) {
        row_add(p, f->Jx, W, W + 1, 1.0);
        row_add(p, f->Jy, W - 1, W, -1.0);
        row_add(p, f->Jz, W, W + 1, 1.0);
        row_zero(p, f->Jx, W + 1);
        row_zero(p, f->Jy, W);
        row_zero(p, f->Jz, W + 1);
    }
}

/* RefBC_Prtl_Bottom, bc_reflect.F90:102-125 */
static void refl_bottom(const orc_params *p, orc_fields *f, orc_particles *s, double ymin)
{
    const double yl = ymin - 0.0 + 3.0;
    const int W = (int)yl;
    if (yl > 1) ref_prtl_left(s, s->y, s->v, yl);
    if (W >= 3 && W <= p->my - 2) {
        row_add(p, f->Jx, W, W - 1, 1.0);
        row_add(p, f->Jy, W, W - 1, -1.0);
        row_add(p, f->Jz, W, W - 1, 1.0);
        row_zero(p, f->Jx, W - 1);
        row_zero(p, f->Jy, W - 1);
        row_zero(p, f->Jz, W - 1);
    }
}

/* RefBC_Prtl_Right + ReflCurrentRight, bc_reflect.F90:127-137, 215-234 */
static void refl_right(const orc_params *p, orc_fields *f, orc_particles *s, double xmax)
{
    const double xl = xmax - 0.0 + 3.0;
    if (xl < p->mx - 1) ref_prtl_right(s, s->x, s->u, xl);
    int ind = (int)floor(xmax - 0.0 + 3.0);
    if (ind >= 1 && ind <= p->mx - 1) {
        col_add(p, f->Jy, ind, ind + 1, 1.0);
        col_add(p, f->Jz, ind, ind + 1, 1.0);
        col_zero(p, f->Jy, ind + 1);
        col_zero(p, f->Jz, ind + 1);
    }
    /* the reference reads bc_face(2)%pos_prtl here, the same position */
    ind = (int)floor(xmax - 0.5 - 0.0 + 3.0);
    if (ind >= 1 && ind <= p->mx - 1) {
        col_add(p, f->Jx, ind, ind + 1, -1.0);
        col_zero(p, f->Jx, ind + 1);
    }
}

/* RefBC_Prtl_Left + ReflCurrentLeft, bc_reflect.F90:139-149, 195-213 */
static void refl_left(const orc_params *p, orc_fields *f, orc_particles *s, double xmin)
{
    const double xl = xmin - 0.0 + 3.0;
    if (xl > 1) ref_prtl_left(s, s->x, s->u, xl);
    int ind = (int)ceil(xmin - 0.0 + 3.0);
    if (ind >= 2 && ind <= p->mx) {
        col_add(p, f->Jy, ind, ind - 1, 1.0);
        col_add(p, f->Jz, ind, ind - 1, 1.0);
        col_zero(p, f->Jy, ind - 1);
        col_zero(p, f->Jz, ind - 1);
    }
    ind = (int)ceil(xmin - 0.5 - 0.0 + 3.0);
    if (ind >= 2 && ind <= p->mx) {
        col_add(p, f->Jx, ind, ind - 1, -1.0);
        col_zero(p, f->Jx, ind - 1);
    }
}

/* RemovePrtl_Left / RemovePrtl_Right, bc_open.F90:8-62: the removed slot
 * becomes a hole with the reference's parking values */
static void remove_one(const orc_params *p, orc_particles *s, int64_t n)
{
    s->q[n] = 0; s->flv[n] = 0;
    s->u[n] = 0; s->v[n] = 0; s->w[n] = 0;
    s->x[n] = 3.8; s->y[n] = 3.8;
    s->z[n] = (p->ndim == 2) ? 1.5 : 3.8;
}

static void remove_below(const orc_params *p, orc_particles *s, const double *x, double pos)
{
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        if (x[n] < pos) remove_one(p, s, n);
    }
}

static void remove_above(const orc_params *p, orc_particles *s, const double *x, double pos)
{
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        if (x[n] > pos) remove_one(p, s, n);
    }
}

/* bc_pos_local, bc.F90:350-362 (borders 0) */
static double pos_local(double pos) { return pos - 0.0 + 3.0; }

void orc_bc_post_movdep(const orc_params *p, orc_fields *f, orc_particles *s)
{
    const int *t = p->bc_prtl;
    const double *q = p->bc_pos_prtl;
    /* bc.F90:245-248: the reflecting faces, top first */
    if (t[3] == ORC_BC_REFL) refl_top(p, f, s, q[3]);
    if (t[2] == ORC_BC_REFL) refl_bottom(p, f, s, q[2]);
    if (t[1] == ORC_BC_REFL) refl_right(p, f, s, q[1]);
    if (t[0] == ORC_BC_REFL) refl_left(p, f, s, q[0]);
    /* bc.F90:250-284: every face again in order 1..6 */
    if (t[0] == ORC_BC_REFL) refl_left(p, f, s, q[0]);
    else if (t[0] == ORC_BC_REMV) remove_below(p, s, s->x, pos_local(q[0]));
    if (t[1] == ORC_BC_REFL) refl_right(p, f, s, q[1]);
    else if (t[1] == ORC_BC_REMV) remove_above(p, s, s->x, pos_local(q[1]));
    if (t[2] == ORC_BC_REFL) refl_bottom(p, f, s, q[2]);
    else if (t[2] == ORC_BC_REMV) remove_below(p, s, s->y, pos_local(q[2]));
    if (t[3] == ORC_BC_REFL) refl_top(p, f, s, q[3]);
    else if (t[3] == ORC_BC_REMV) remove_above(p, s, s->y, pos_local(q[3]));
    if (t[4] == ORC_BC_REMV) remove_below(p, s, s->z, pos_local(q[4]));
    if (t[5] == ORC_BC_REMV) remove_above(p, s, s->z, pos_local(q[5]));
}

/* ------------------------------------------------------------------ */
/* Field faces                                                         */
/* ------------------------------------------------------------------ */

static void zero_x_range(const orc_params *p, double *F, int i0, int i1)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++)
            for (int i = i0; i <= i1; i++) F[IX(p, i, j, k)] = 0.0;
}

static void zero_y_range(const orc_params *p, double *F, int j0, int j1)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = j0; j <= j1; j++)
            for (int i = 1; i <= p->mx; i++) F[IX(p, i, j, k)] = 0.0;
}

/* CondBC_Fld_Left / Right / Bottom / Top, bc_reflect.F90:11-65 */
static void cond_left(const orc_params *p, orc_fields *f, double xind)
{
    int xl = (int)floor(xind - 0.0 + 3.0);
    if (xl < 1) return;
    if (xl > p->mx) xl = p->mx;
    zero_x_range(p, f->Ey, 1, xl);
    zero_x_range(p, f->Ez, 1, xl);
}

static void cond_right(const orc_params *p, orc_fields *f, double xind)
{
    int xl = (int)ceil(xind - 0.0 + 3.0);
    if (xl > p->mx) return;
    if (xl < 1) xl = 1;
    zero_x_range(p, f->Ey, xl, p->mx);
    zero_x_range(p, f->Ez, xl, p->mx);
}

static void cond_bottom(const orc_params *p, orc_fields *f, double yind)
{
    int yl = (int)floor(yind - 0.0 + 3.0);
    if (yl < 1) return;
    if (yl > p->my) yl = p->my;
    zero_y_range(p, f->Ex, 1, yl);
    zero_y_range(p, f->Ez, 1, yl);
}

static void cond_top(const orc_params *p, orc_fields *f, double yind)
{
    int yl = (int)ceil(yind - 0.0 + 3.0);
    if (yl > p->my) return;
    if (yl < 1) yl = 1;
    zero_y_range(p, f->Ex, yl, p->my);
    zero_y_range(p, f->Ez, yl, p->my);
}

static void scale_x(const orc_params *p, double *F, int i)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++) F[IX(p, i, j, k)] = 0.5 * F[IX(p, i, j, k)];
}

static void copy_x(const orc_params *p, double *F, int dst, int src)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++) F[IX(p, dst, j, k)] = F[IX(p, src, j, k)];
}

/* OpenBC_Fld_Right / Left, bc.F90:410-469 */
static void open_right(const orc_params *p, orc_fields *f, double xind)
{
    const int xl = (int)ceil(xind - 0.0 + 3.0);
    if (!(xl >= 1 && xl <= p->mx - 1)) return;
    for (int i = xl + 1; i <= p->mx; i++) {
        scale_x(p, f->Ey, i);
        scale_x(p, f->Ez, i);
        copy_x(p, f->Bx, i, xl);
    }
    for (int i = xl; i <= p->mx; i++) {
        scale_x(p, f->Ex, i);
        scale_x(p, f->By, i);
        scale_x(p, f->Bz, i);
    }
}

static void open_left(const orc_params *p, orc_fields *f, double xind)
{
    const int xl = (int)ceil(xind - 0.0 + 3.0);
    if (!(xl >= 2 && xl <= p->mx)) return;
    for (int i = 1; i <= xl - 1; i++) {
        scale_x(p, f->Ey, i);
        scale_x(p, f->Ez, i);
        copy_x(p, f->Bx, i, xl);
        scale_x(p, f->Ex, i);
        scale_x(p, f->By, i);
        scale_x(p, f->Bz, i);
    }
}

void orc_bc_final(orc_params *p, orc_fields *f)
{
    const int *t = p->bc_fld;
    const double *q = p->bc_pos_fld;
    if (t[0] == ORC_BC_COND) cond_left(p, f, q[0]);
    else if (t[0] == ORC_BC_OPEN) open_left(p, f, q[0]);
    if (t[1] == ORC_BC_COND) cond_right(p, f, q[1]);
    else if (t[1] == ORC_BC_OPEN) open_right(p, f, q[1]);
    if (t[2] == ORC_BC_COND) cond_bottom(p, f, q[2]);
    /* bc.F90:318 passes bc_face(3)%pos_fld to the top face */
    if (t[3] == ORC_BC_COND) cond_top(p, f, p->cond_top_as_written ? q[2] : q[3]);
    /* updateBC_Pos, bc.F90:342-348 */
    for (int n = 0; n < 6; n++) {
        if (p->bc_fld[n] != ORC_BC_PERIODIC) p->bc_pos_fld[n] = p->bc_pos_fld[n] + p->bc_speed[n] * p->c;
        if (p->bc_prtl[n] != ORC_BC_PERIODIC) p->bc_pos_prtl[n] = p->bc_pos_prtl[n] + p->bc_speed[n] * p->c;
    }
}

/* ------------------------------------------------------------------ */
/* PML, bc_pml.F90                                                     */
/* ------------------------------------------------------------------ */

/* pml_params defaults (bc_pml.F90:6-12) and SetPMLParameters (:39-51) */
typedef struct { double kappa_max, sigma_max, alpha_max, L; int n; } pml_side;

static pml_side side_params(const orc_params *p, int f)
{
    pml_side s = {1.0, 0.0, 1.0, 10.0, 4};
    if (p->bc_fld[f] == ORC_BC_PML) {
        s.n = 4;
        s.L = p->pml_L[f];
        s.kappa_max = 1.0;
        s.sigma_max = 1.0;
        s.alpha_max = 9.0e-4;
    }
    return s;
}

/* real ** integer by repeated squaring (the Fortran runtime's integer power) */
static double powi(double a, int n)
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

static double pml_kappa(double d, pml_side s)
{
    const double v = 1.0 + powi(d / s.L, s.n) * (s.kappa_max - 1.0);
    return (s.kappa_max < v) ? s.kappa_max : v;
}

static double pml_sigma(double d, pml_side s)
{
    const double v = s.sigma_max * powi(d / s.L, s.n);
    return (s.sigma_max < v) ? s.sigma_max : v;
}

/* Update_PML_Position, bc_pml.F90:54-69 */
typedef struct { double pos[6]; long long ind[6]; int none; pml_side side[6]; } pml_geo;

static pml_geo pml_position(const orc_params *p)
{
    pml_geo g;
    const double huge = 2147483647.0;
    for (int n = 0; n < 6; n++) {
        const int upper = n & 1;
        g.pos[n] = upper ? huge : -huge;
        g.ind[n] = upper ? 2147483647LL : -2147483647LL;
        if (p->bc_fld[n] == ORC_BC_PML) {
            g.pos[n] = p->bc_pos_fld[n];
            g.ind[n] = upper ? (long long)ceil(g.pos[n]) + 1 : (long long)floor(g.pos[n]) - 1;
        }
        g.side[n] = side_params(p, n);
    }
    g.none = (1 + 0 - 3 > g.ind[0]) && (p->mx + 0 - 3 < g.ind[1]) && (1 + 0 - 3 > g.ind[2]) &&
             (p->my + 0 - 3 < g.ind[3]) && (1 + 0 - 3 > g.ind[4]) && (p->mz + 0 - 3 < g.ind[5]);
    return g;
}

/* pml_ab and pml_c, bc_pml.F90:293-334 */
static void pml_abc(const orc_params *p, const pml_geo *g, double x, int axis, double *a, double *b, double *c)
{
    *a = 0.0; *b = 0.0; *c = 0.0;
    const int up = 2 * axis + 1, lo = 2 * axis;
    if (x > g->pos[up]) {
        const double d = x - g->pos[up];
        const pml_side s = g->side[up];
        const double alpha = s.alpha_max, sigma = pml_sigma(d, s), kappa = pml_kappa(d, s);
        *a = exp(-(sigma / kappa) - alpha);
        *b = (sigma / (sigma * kappa + alpha * (kappa * kappa))) * (*a - 1.0);
        *c = p->fldc * (1.0 / kappa - 1.0);
    }
    if (x < g->pos[lo]) {
        const double d = g->pos[lo] - x;
        const pml_side s = g->side[lo];
        const double alpha = s.alpha_max, sigma = pml_sigma(d, s), kappa = pml_kappa(d, s);
        *a = exp(-(sigma / kappa) - alpha);
        *b = (sigma / (sigma * kappa + alpha * (kappa * kappa))) * (*a - 1.0);
        *c = p->fldc * (1.0 / kappa - 1.0);
    }
}

void orc_pml_coeffs(const orc_params *p, double x, int axis, double out[3])
{
    pml_geo g = pml_position(p);
    pml_abc(p, &g, x, axis, &out[0], &out[1], &out[2]);
}

static int pml_inside(const pml_geo *g, long long xg, long long yg, long long zg)
{
    return xg > g->ind[0] && xg < g->ind[1] && yg > g->ind[2] && yg < g->ind[3] && zg > g->ind[4] && zg < g->ind[5];
}

static int pml_any(const orc_params *p)
{
    for (int n = 0; n < 6; n++)
        if (p->bc_fld[n] == ORC_BC_PML) return 1;
    return 0;
}

/* UpdateMagFld_PML, bc_pml.F90:185-290 */
void orc_pml_b(const orc_params *p, orc_fields *f)
{
    if (!pml_any(p)) return;
    const pml_geo g = pml_position(p);
    const int64_t n = (int64_t)p->mx * p->my * p->mz;
    if (g.none) {   /* DeallocatePMLArr: every psi array restarts from zero */
        for (int a = 0; a < 12; a++)
            if (f->pml[a]) memset(f->pml[a], 0, sizeof(double) * (size_t)n);
        return;
    }
    double *B1x = f->pml[6], *B1y = f->pml[7], *B1z = f->pml[8];
    double *B2x = f->pml[9], *B2y = f->pml[10], *B2z = f->pml[11];
    double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez, *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    const double fc = p->fldc, idx_ = p->inv_dx, idy_ = p->inv_dy, idz_ = p->inv_dz;
    const int three = p->ndim == 3;
    const int k0 = three ? 2 : 1, k1 = three ? p->mz - 1 : 1;
    const double *L = (const double[6]){g.side[0].L, g.side[1].L, g.side[2].L, g.side[3].L, g.side[4].L,
                                        g.side[5].L};
    for (int k = k0; k <= k1; k++)
        for (int j = 2; j <= p->my - 1; j++)
            for (int i = 2; i <= p->mx - 1; i++) {
                if (pml_inside(&g, i - 3, j - 3, k - 3)) continue;
                const double xg = i + 0 - 3 + 0.5, yg = j + 0 - 3 + 0.5, zg = k + 0 - 3 + 0.5;
                double a_x, b_x, c_x, a_y, b_y, c_y, a_z, b_z, c_z;
                pml_abc(p, &g, xg, 0, &a_x, &b_x, &c_x);
                pml_abc(p, &g, yg, 1, &a_y, &b_y, &c_y);
                pml_abc(p, &g, zg, 2, &a_z, &b_z, &c_z);
                const int64_t o = IX(p, i, j, k), oi = IX(p, i + 1, j, k), oj = IX(p, i, j + 1, k);
                const int64_t ok = three ? IX(p, i, j, k + 1) : o;
                B1x[o] = a_y * B1x[o] + b_y * fc * idy_ * (Ez[oj] - Ez[o]);
                B1z[o] = a_x * B1z[o] + b_x * fc * idx_ * (Ey[oi] - Ey[o]);
                if (three) {
                    B1y[o] = a_z * B1y[o] + b_z * fc * idz_ * (Ex[ok] - Ex[o]);
                    B2x[o] = a_z * B2x[o] + b_z * fc * idz_ * (Ey[ok] - Ey[o]);
                }
                B2y[o] = a_x * B2y[o] + b_x * fc * idx_ * (Ez[oi] - Ez[o]);
                B2z[o] = a_y * B2z[o] + b_y * fc * idy_ * (Ex[oj] - Ex[o]);
                if (three) {
                    Bx[o] = Bx[o] - c_y * idy_ * (Ez[oj] - Ez[o]) + c_z * idz_ * (Ey[ok] - Ey[o]) - B1x[o] + B2x[o];
                    By[o] = By[o] - c_z * idz_ * (Ex[ok] - Ex[o]) + c_x * idx_ * (Ez[oi] - Ez[o]) - B1y[o] + B2y[o];
                } else {
                    Bx[o] = Bx[o] - c_y * idy_ * (Ez[oj] - Ez[o]) - B1x[o];
                    By[o] = By[o] + c_x * idx_ * (Ez[oi] - Ez[o]) + B2y[o];
                }
                Bz[o] = Bz[o] - c_x * idx_ * (Ey[oi] - Ey[o]) + c_y * idy_ * (Ex[oj] - Ex[o]) - B1z[o] + B2z[o];
                if (xg - 0.5 < g.pos[0] - L[0] || xg - 0.5 > g.pos[1] + L[1] || yg < g.pos[2] - L[2] ||
                    yg > g.pos[3] + L[3] || zg < g.pos[4] - L[4] || zg > g.pos[5] + L[5])
                    Bx[o] = 0.0;
                if (xg < g.pos[0] - L[0] || xg > g.pos[1] + L[1] || yg - 0.5 < g.pos[2] - L[2] ||
                    yg - 0.5 > g.pos[3] + L[3] || zg < g.pos[4] - L[4] || zg > g.pos[5] + L[5])
                    By[o] = 0.0;
                if (xg < g.pos[0] - L[0] || xg > g.pos[1] + L[1] || yg < g.pos[2] - L[2] || yg > g.pos[3] + L[3] ||
                    zg - 0.5 < g.pos[4] - L[4] || zg - 0.5 > g.pos[5] + L[5])
                    Bz[o] = 0.0;
            }
}

/* UpdateElcFld_PML, bc_pml.F90:85-183 */
void orc_pml_e(const orc_params *p, orc_fields *f)
{
    if (!pml_any(p)) return;
    const pml_geo g = pml_position(p);
    if (g.none) return;
    double *E1x = f->pml[0], *E1y = f->pml[1], *E1z = f->pml[2];
    double *E2x = f->pml[3], *E2y = f->pml[4], *E2z = f->pml[5];
    double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez, *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    const double fc = p->fldc, idx_ = p->inv_dx, idy_ = p->inv_dy, idz_ = p->inv_dz;
    const int three = p->ndim == 3;
    const int k0 = three ? 2 : 1, k1 = three ? p->mz - 1 : 1;
    const double *L = (const double[6]){g.side[0].L, g.side[1].L, g.side[2].L, g.side[3].L, g.side[4].L,
                                        g.side[5].L};
    for (int k = k0; k <= k1; k++)
        for (int j = 2; j <= p->my - 1; j++)
            for (int i = 2; i <= p->mx - 1; i++) {
                const double xg = i + 0 - 3, yg = j + 0 - 3, zg = k + 0 - 3;
                if (pml_inside(&g, i - 3, j - 3, k - 3)) continue;
                double a_x, b_x, c_x, a_y, b_y, c_y, a_z, b_z, c_z;
                pml_abc(p, &g, xg, 0, &a_x, &b_x, &c_x);
                pml_abc(p, &g, yg, 1, &a_y, &b_y, &c_y);
                pml_abc(p, &g, zg, 2, &a_z, &b_z, &c_z);
                const int64_t o = IX(p, i, j, k), oi = IX(p, i - 1, j, k), oj = IX(p, i, j - 1, k);
                const int64_t ok = three ? IX(p, i, j, k - 1) : o;
                E1x[o] = a_y * E1x[o] + b_y * fc * idy_ * (Bz[o] - Bz[oj]);
                E1z[o] = a_x * E1z[o] + b_x * fc * idx_ * (By[o] - By[oi]);
                if (three) {
                    E1y[o] = a_z * E1y[o] + b_z * fc * idz_ * (Bx[o] - Bx[ok]);
                    E2x[o] = a_z * E2x[o] + b_z * fc * idz_ * (By[o] - By[ok]);
                }
                E2y[o] = a_x * E2y[o] + b_x * fc * idx_ * (Bz[o] - Bz[oi]);
                E2z[o] = a_y * E2z[o] + b_y * fc * idy_ * (Bx[o] - Bx[oj]);
                if (three) {
                    Ex[o] = Ex[o] + c_y * idy_ * (Bz[o] - Bz[oj]) - c_z * idz_ * (By[o] - By[ok]) + E1x[o] - E2x[o];
                    Ey[o] = Ey[o] + c_z * idz_ * (Bx[o] - Bx[ok]) - c_x * idx_ * (Bz[o] - Bz[oi]) + E1y[o] - E2y[o];
                } else {
                    Ex[o] = Ex[o] + c_y * idy_ * (Bz[o] - Bz[oj]) + E1x[o];
                    Ey[o] = Ey[o] - c_x * idx_ * (Bz[o] - Bz[oi]) - E2y[o];
                }
                Ez[o] = Ez[o] + c_x * idx_ * (By[o] - By[oi]) - c_y * idy_ * (Bx[o] - Bx[oj]) + E1z[o] - E2z[o];
                if (xg + 0.5 < g.pos[0] - L[0] || xg + 0.5 > g.pos[1] + L[1] || yg < g.pos[2] - L[2] ||
                    yg > g.pos[3] + L[3] || zg < g.pos[4] - L[4] || zg > g.pos[5] + L[5])
                    Ex[o] = 0.0;
                if (xg < g.pos[0] - L[0] || xg > g.pos[1] + L[1] || yg + 0.5 < g.pos[2] - L[2] ||
                    yg + 0.5 > g.pos[3] + L[3] || zg < g.pos[4] - L[4] || zg > g.pos[5] + L[5])
                    Ey[o] = 0.0;
                if (xg < g.pos[0] - L[0] || xg > g.pos[1] + L[1] || yg < g.pos[2] - L[2] || yg > g.pos[3] + L[3] ||
                    zg + 0.5 < g.pos[4] - L[4] || zg + 0.5 > g.pos[5] + L[5])
                    Ez[o] = 0.0;
            }
}
