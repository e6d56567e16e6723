/*
 * pictor_oracle.c -- CPU FP64 restatement of PICTOR's particle step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path (see the header).
 * Build with -O2 -ffp-contract=off so that no multiply-add is fused and the
 * rounding sequence is the one the cited Fortran expressions define.
 *
 * Reference files followed (paths relative to the PICTOR tree):
 *   src/maxwell/em_update.F90:9-102            Yee E, B half step, E -= J
 *   src/movdep/movdep.F90:12-258               Boris push + zigzag deposit
 *   src/movdep/movdep_routines.F90:6-311       VecE/VecB gather, VecJ reduce
 *   src/vars.F90:313-347                       weight tables
 *   src/comm/comm_fld.F90:15-307               halo copy / current add
 *   src/comm/subdomains/subdomains.F90:52-67,592-740  index ranges, shifts
 *   src/comm/loadprtlout.F90:13-181            outlier scan + removal
 *   src/comm/comm_prtl.F90:11-65, src/mem/mem_prtl.F90:135-160  append
 *   src/mem/mem_prtl.F90:166-301               counting sort by cell
 *   src/init/initialise.F90:44-53,84-85        charge/mass scaling
 *   src/io/savedata_routines.F90:265-355       energy, div E
 */
#include "pictor_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define VBS 32  /* VecBlockSize, src/vars.F90:36 */

struct orc_ws {
    int nthreads;
    int nv;            /* 4 (2-D) / 8 (3-D) corner values per cell */
    int nj;            /* 8 (2-D) / 12 (3-D) current entries per cell */
    int64_t ncell;     /* mx*my*mz */
    double *vec[6];    /* VecEx..VecBz (nv, ncell) */
    double *vecj;      /* VecJ (nj, ncell, nthreads) */
};

/* ---- weight tables, src/vars.F90:313-347 ---- */
static const double wtx1_2[4] = {1, 0, 1, 0}, wtx2_2[4] = {-1, 1, -1, 1};
static const double wty1_2[4] = {1, 1, 0, 0}, wty2_2[4] = {-1, -1, 1, 1};
static const double wtFx_2[8] = {1, 1, 0, 0, 0, 0, 0, 0};
static const double wtFy_2[8] = {0, 0, 1, 1, 0, 0, 0, 0};
static const double wtFz_2[8] = {0, 0, 0, 0, 1, 1, 1, 1};
static const double Jwtx1_2[8] = {1, 1, 1, 0, 1, 0, 1, 0};
static const double Jwtx2_2[8] = {0, 0, -1, 1, -1, 1, -1, 1};
static const double Jwty1_2[8] = {1, 0, 1, 1, 1, 1, 0, 0};
static const double Jwty2_2[8] = {-1, 1, 0, 0, -1, -1, 1, 1};
static const double Jwtz1_2[8] = {1, 1, 1, 1, 1, 1, 1, 1};
static const double Jwtz2_2[8] = {-1, -1, -1, -1, 0, 0, 0, 0};

static const double wtx1_3[8] = {1, 0, 1, 0, 1, 0, 1, 0}, wtx2_3[8] = {-1, 1, -1, 1, -1, 1, -1, 1};
static const double wty1_3[8] = {1, 1, 0, 0, 1, 1, 0, 0}, wty2_3[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
static const double wtz1_3[8] = {1, 1, 1, 1, 0, 0, 0, 0}, wtz2_3[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
static const double wtFx_3[12] = {1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0};
static const double wtFy_3[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0};
static const double wtFz_3[12] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
static const double Jwtx1_3[12] = {1, 1, 1, 1, 1, 0, 1, 0, 1, 0, 1, 0};
static const double Jwtx2_3[12] = {0, 0, 0, 0, -1, 1, -1, 1, -1, 1, -1, 1};
static const double Jwty1_3[12] = {1, 0, 1, 0, 1, 1, 1, 1, 1, 1, 0, 0};
static const double Jwty2_3[12] = {-1, 1, -1, 1, 0, 0, 0, 0, -1, -1, 1, 1};
static const double Jwtz1_3[12] = {1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 1, 1};
static const double Jwtz2_3[12] = {-1, -1, 1, 1, -1, -1, 1, 1, 0, 0, 0, 0};

/* 1-based (i,j,k) -> offset in a column-major (mx,my,mz) array */
#define IX(p, i, j, k) ((int64_t)((i) - 1) + (int64_t)(p)->mx * ((int64_t)((j) - 1) + (int64_t)(p)->my * (int64_t)((k) - 1)))

int orc_params_init(orc_params *p, int ndim, int nx, int ny, int nz, double c,
                    double c_ompe, double g0, double mi_me, int epc)
{
    if (ndim != 2 && ndim != 3) return ORC_ERR_ARG;
    if (nx < 6 || ny < 6 || (ndim == 3 && nz < 6) || epc < 1) return ORC_ERR_ARG;
    memset(p, 0, sizeof(*p));
    p->ndim = ndim;
    p->nx = nx; p->ny = ny; p->nz = (ndim == 2) ? 1 : nz;
    p->mx = nx + 5; p->my = ny + 5; p->mz = (ndim == 2) ? 1 : nz + 5;
    p->epc = epc;
    p->c = c; p->c_ompe = c_ompe; p->g0 = g0; p->mi_me = mi_me;
    p->me = 1.0;
    p->mi = mi_me;
    /* initialise.F90:44-53; cell_volume = 1, grid_dx = grid_dy = grid_dz = 1 */
    p->ompe = c / c_ompe;
    p->qi = (g0 * (p->ompe * p->ompe)) / (((double)epc / 1.0) * (1.0 + p->me / p->mi));
    p->qe = -p->qi;
    p->qme = -1.0;
    p->qmi = 1.0 / p->mi;
    p->masse = fabs(p->qe) * p->me;
    p->massi = p->qi * p->mi;
    /* vars.F90:220-223 (EMspeedBoost = 1) */
    p->fldc = c * 1.0;
    p->fld_halfc = c * 0.5 * 1.0;
    p->cinv = 1.0 / c;
    p->sqc = c * c;
    p->inv_dx = 1.0; p->inv_dy = 1.0; p->inv_dz = 1.0;
    /* subdomains.F90:52-67 */
    p->xmin = 3.0; p->xmax = p->mx - 2;
    p->ymin = 3.0; p->ymax = p->my - 2;
    if (ndim == 2) { p->zmin = 1.0; p->zmax = 2.0; }
    else { p->zmin = 3.0; p->zmax = p->mz - 2; }
    p->xlen = p->xmax - p->xmin;
    p->ylen = p->ymax - p->ymin;
    p->zlen = p->zmax - p->zmin;
    p->nflvr = 2;
    p->flvrqm[0] = p->qmi;   /* initialise.F90:84 */
    p->flvrqm[1] = p->qme;   /* initialise.F90:85 */
    p->reorder_period = 20;
    for (int a = 0; a < 3; a++) p->periodic[a] = 1;   /* every face periodic until a BC is set */
    for (int f = 0; f < 6; f++) p->pml_L[f] = 4.0;      /* SetBC_Fld's default PML thickness */
    return ORC_OK;
}

orc_ws *orc_ws_create(const orc_params *p, int nthreads)
{
    orc_ws *w = (orc_ws *)calloc(1, sizeof(orc_ws));
    if (!w) return NULL;
    if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
    omp_set_num_threads(nthreads);
#else
    nthreads = 1;
#endif
    w->nthreads = nthreads;
    w->nv = (p->ndim == 2) ? 4 : 8;
    w->nj = (p->ndim == 2) ? 8 : 12;
    w->ncell = (int64_t)p->mx * p->my * p->mz;
    for (int c = 0; c < 6; c++) {
        w->vec[c] = (double *)calloc((size_t)(w->nv * w->ncell), sizeof(double));
        if (!w->vec[c]) { orc_ws_destroy(w); return NULL; }
    }
    w->vecj = (double *)calloc((size_t)(w->nj * w->ncell * nthreads), sizeof(double));
    if (!w->vecj) { orc_ws_destroy(w); return NULL; }
    return w;
}

void orc_ws_destroy(orc_ws *w)
{
    if (!w) return;
    for (int c = 0; c < 6; c++) free(w->vec[c]);
    free(w->vecj);
    free(w);
}

int orc_ws_threads(const orc_ws *w) { return w->nthreads; }
const double *orc_ws_vec(const orc_ws *w, int which) { return (which >= 0 && which < 6) ? w->vec[which] : NULL; }
int orc_ws_vec_len(const orc_ws *w) { return w->nv; }

/* ------------------------------------------------------------------ */
/* Maxwell solver                                                      */
/* ------------------------------------------------------------------ */

/* em_update.F90:9-39 */
void orc_update_e(const orc_params *p, orc_fields *f)
{
    const double fc = p->fldc;
    const int mx = p->mx, my = p->my, mz = p->mz;
    double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez;
    const double *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    /* em_update.F90:9-39 loops over 2..m-1; the Lehe build's UpdateEfield
     * (em_update_lehe.F90:43-84) over 3..m-3 (the same cells once the halo ran) */
    const int e0 = (p->fdtd == 1) ? 3 : 2, e1 = (p->fdtd == 1) ? 3 : 1;
    if (p->ndim == 2) {
        #pragma omp parallel for schedule(static)
        for (int j = e0; j <= my - e1; j++)
            for (int i = e0; i <= mx - e1; i++) {
                int64_t o = IX(p, i, j, 1);
                Ex[o] = Ex[o] + fc * p->inv_dy * (Bz[o] - Bz[IX(p, i, j - 1, 1)]);
                Ey[o] = Ey[o] - fc * p->inv_dx * (Bz[o] - Bz[IX(p, i - 1, j, 1)]);
                Ez[o] = Ez[o] + fc * p->inv_dx * (By[o] - By[IX(p, i - 1, j, 1)])
                              - fc * p->inv_dy * (Bx[o] - Bx[IX(p, i, j - 1, 1)]);
            }
    } else {
        #pragma omp parallel for schedule(static)
        for (int k = e0; k <= mz - e1; k++)
            for (int j = e0; j <= my - e1; j++)
                for (int i = e0; i <= mx - e1; i++) {
                    int64_t o = IX(p, i, j, k);
                    Ex[o] = Ex[o] + fc * p->inv_dy * (Bz[o] - Bz[IX(p, i, j - 1, k)])
                                  - fc * p->inv_dz * (By[o] - By[IX(p, i, j, k - 1)]);
                    Ey[o] = Ey[o] + fc * p->inv_dz * (Bx[o] - Bx[IX(p, i, j, k - 1)])
                                  - fc * p->inv_dx * (Bz[o] - Bz[IX(p, i - 1, j, k)]);
                    Ez[o] = Ez[o] + fc * p->inv_dx * (By[o] - By[IX(p, i - 1, j, k)])
                                  - fc * p->inv_dy * (Bx[o] - Bx[IX(p, i, j - 1, k)]);
                }
    }
}

/* em_update.F90:41-77 */
/* em_update_lehe.F90:8-25 with grid_dx = grid_dy = grid_dz = 1 */
void orc_set_lehe(orc_params *p)
{
    const double pi = 3.14159265358979323846264;   /* vars.F90:224 */
    double *L = p->lehe;
    const double beta_xy = 0.125 * 1.0, beta_xz = 0.125 * 1.0, beta_yx = 0.125, beta_yz = 0.0;
    const double beta_zx = 0.125, beta_zy = 0.0;
    const double sn = sin(0.5 * pi * p->c);
    const double delta_x = 0.25 * (1.0 - (p->cinv * p->cinv) * (sn * sn));
    const double delta_y = 0.0, delta_z = 0.0;
    if (p->ndim == 3) {
        L[0] = 1.0 - 2.0 * beta_xy - 2.0 * beta_xz - 3.0 * delta_x;
        L[1] = 1.0 - 2.0 * beta_yx - 2.0 * beta_yz - 3.0 * delta_y;
        L[2] = 1.0 - 2.0 * beta_zx - 2.0 * beta_zy - 3.0 * delta_z;
    } else {
        L[0] = 1.0 - 2.0 * beta_xy - 3.0 * delta_x;
        L[1] = 1.0 - 2.0 * beta_yx - 3.0 * delta_y;
        L[2] = 1.0;
    }
    L[3] = beta_xy; L[4] = beta_xz; L[5] = beta_yx; L[6] = beta_yz; L[7] = beta_zx; L[8] = beta_zy;
    L[9] = delta_x; L[10] = delta_y; L[11] = delta_z;
    p->fdtd = 1;
}

#define F3(A, i, j, k) A[IX(p, i, j, k)]
/* gradX/gradY/gradZ, em_update_lehe.F90:120-155 (term order kept) */
static double lehe_gradx(const orc_params *p, const double *F, int i, int j, int k)
{
    const double *L = p->lehe, d = p->inv_dx;
    double g = L[0] * d * (F3(F, i + 1, j, k) - F3(F, i, j, k))
             + L[3] * d * (F3(F, i + 1, j + 1, k) - F3(F, i, j + 1, k))
             + L[3] * d * (F3(F, i + 1, j - 1, k) - F3(F, i, j - 1, k));
    if (p->ndim == 3)
        g = g + L[4] * d * (F3(F, i + 1, j, k + 1) - F3(F, i, j, k + 1))
              + L[4] * d * (F3(F, i + 1, j, k - 1) - F3(F, i, j, k - 1));
    return g + L[9] * d * (F3(F, i + 2, j, k) - F3(F, i - 1, j, k));
}
static double lehe_grady(const orc_params *p, const double *F, int i, int j, int k)
{
    const double *L = p->lehe, d = p->inv_dy;
    double g = L[1] * d * (F3(F, i, j + 1, k) - F3(F, i, j, k))
             + L[5] * d * (F3(F, i + 1, j + 1, k) - F3(F, i + 1, j, k))
             + L[5] * d * (F3(F, i - 1, j + 1, k) - F3(F, i - 1, j, k));
    if (p->ndim == 3)
        g = g + L[6] * d * (F3(F, i, j + 1, k + 1) - F3(F, i, j, k + 1))
              + L[6] * d * (F3(F, i, j + 1, k - 1) - F3(F, i, j, k - 1));
    return g + L[10] * d * (F3(F, i, j + 2, k) - F3(F, i, j - 1, k));
}
static double lehe_gradz(const orc_params *p, const double *F, int i, int j, int k)
{
    const double *L = p->lehe, d = p->inv_dz;
    return L[2] * d * (F3(F, i, j, k + 1) - F3(F, i, j, k))
         + L[7] * d * (F3(F, i + 1, j, k + 1) - F3(F, i + 1, j, k))
         + L[7] * d * (F3(F, i - 1, j, k + 1) - F3(F, i - 1, j, k))
         + L[8] * d * (F3(F, i, j + 1, k + 1) - F3(F, i, j + 1, k))
         + L[8] * d * (F3(F, i, j - 1, k + 1) - F3(F, i, j - 1, k))
         + L[11] * d * (F3(F, i, j, k + 2) - F3(F, i, j, k - 1));
}
#undef F3

/* UpdateBfieldHalf, em_update_lehe.F90:88-118 (cells 3..m-3) */
static void update_b_half_lehe(const orc_params *p, orc_fields *f)
{
    const double hc = p->fld_halfc;
    const double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez;
    double *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 3; k1 = p->mz - 3; }
    for (int k = k0; k <= k1; k++)
        #pragma omp parallel for schedule(static)
        for (int j = 3; j <= p->my - 3; j++)
            for (int i = 3; i <= p->mx - 3; i++) {
                int64_t o = IX(p, i, j, k);
                if (p->ndim == 3) {
                    Bx[o] = Bx[o] - hc * lehe_grady(p, Ez, i, j, k) + hc * lehe_gradz(p, Ey, i, j, k);
                    By[o] = By[o] - hc * lehe_gradz(p, Ex, i, j, k) + hc * lehe_gradx(p, Ez, i, j, k);
                    Bz[o] = Bz[o] - hc * lehe_gradx(p, Ey, i, j, k) + hc * lehe_grady(p, Ex, i, j, k);
                } else {
                    Bx[o] = Bx[o] - hc * lehe_grady(p, Ez, i, j, k);
                    By[o] = By[o] + hc * lehe_gradx(p, Ez, i, j, k);
                    Bz[o] = Bz[o] - hc * lehe_gradx(p, Ey, i, j, k) + hc * lehe_grady(p, Ex, i, j, k);
                }
            }
}

void orc_update_b_half(const orc_params *p, orc_fields *f)
{
    if (p->fdtd == 1) {
        update_b_half_lehe(p, f);
        return;
    }
    const double hc = p->fld_halfc;
    const int mx = p->mx, my = p->my, mz = p->mz;
    const double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez;
    double *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    if (p->ndim == 2) {
        #pragma omp parallel for schedule(static)
        for (int j = 2; j <= my - 1; j++)
            for (int i = 2; i <= mx - 1; i++) {
                int64_t o = IX(p, i, j, 1);
                Bx[o] = Bx[o] - hc * p->inv_dy * (Ez[IX(p, i, j + 1, 1)] - Ez[o]);
                By[o] = By[o] + hc * p->inv_dx * (Ez[IX(p, i + 1, j, 1)] - Ez[o]);
                Bz[o] = Bz[o] - hc * p->inv_dx * (Ey[IX(p, i + 1, j, 1)] - Ey[o])
                              + hc * p->inv_dy * (Ex[IX(p, i, j + 1, 1)] - Ex[o]);
            }
    } else {
        #pragma omp parallel for schedule(static)
        for (int k = 2; k <= mz - 1; k++)
            for (int j = 2; j <= my - 1; j++)
                for (int i = 2; i <= mx - 1; i++) {
                    int64_t o = IX(p, i, j, k);
                    Bx[o] = Bx[o] - hc * p->inv_dy * (Ez[IX(p, i, j + 1, k)] - Ez[o])
                                  + hc * p->inv_dz * (Ey[IX(p, i, j, k + 1)] - Ey[o]);
                    By[o] = By[o] - hc * p->inv_dz * (Ex[IX(p, i, j, k + 1)] - Ex[o])
                                  + hc * p->inv_dx * (Ez[IX(p, i + 1, j, k)] - Ez[o]);
                    Bz[o] = Bz[o] - hc * p->inv_dx * (Ey[IX(p, i + 1, j, k)] - Ey[o])
                                  + hc * p->inv_dy * (Ex[IX(p, i, j + 1, k)] - Ex[o]);
                }
    }
}

/* em_update.F90:79-102 */
void orc_add_current(const orc_params *p, orc_fields *f)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 3; k1 = mz - 3; }
    for (int k = k0; k <= k1; k++)
        #pragma omp parallel for schedule(static)
        for (int j = 3; j <= my - 3; j++)
            for (int i = 3; i <= mx - 3; i++) {
                int64_t o = IX(p, i, j, k);
                f->Ex[o] = f->Ex[o] - f->Jx[o] * p->inv_dy * p->inv_dz;
                f->Ey[o] = f->Ey[o] - f->Jy[o] * p->inv_dx * p->inv_dz;
                f->Ez[o] = f->Ez[o] - f->Jz[o] * p->inv_dy * p->inv_dx;
            }
}

/* current.F90:13-17 */
void orc_reset_current(const orc_params *p, orc_fields *f)
{
    size_t n = (size_t)p->mx * p->my * p->mz;
    memset(f->Jx, 0, n * sizeof(double));
    memset(f->Jy, 0, n * sizeof(double));
    memset(f->Jz, 0, n * sizeof(double));
}

/* ------------------------------------------------------------------ */
/* Periodic halos on one rank (indepLBaxis = 0, vars.F90:241).          */
/* Ranges from subdomains.F90:592-680: send face 1 = 3..5, send face 2 =  */
/* m-4..m-3, recv face 1 = 1..2, recv face 2 = m-2..m.  Faces across the  */
/* independent axis x span the full cross-section; y and z faces span    */
/* x = 3..mx-3 only (SetCommFldDomain step 2 with extended edges).        */
/* ------------------------------------------------------------------ */

typedef struct { int lo[3], hi[3]; } box3;

static void copy_box(const orc_params *p, double *F, box3 dst, int axis, int shift)
{
    for (int k = dst.lo[2]; k <= dst.hi[2]; k++)
        for (int j = dst.lo[1]; j <= dst.hi[1]; j++)
            for (int i = dst.lo[0]; i <= dst.hi[0]; i++) {
                int s[3] = {i, j, k};
                s[axis] += shift;
                F[IX(p, i, j, k)] = F[IX(p, s[0], s[1], s[2])];
            }
}

static void add_box(const orc_params *p, double *F, box3 dst, int axis, int shift)
{
    for (int k = dst.lo[2]; k <= dst.hi[2]; k++)
        for (int j = dst.lo[1]; j <= dst.hi[1]; j++)
            for (int i = dst.lo[0]; i <= dst.hi[0]; i++) {
                int s[3] = {i, j, k};
                s[axis] += shift;
                F[IX(p, i, j, k)] = F[IX(p, i, j, k)] + F[IX(p, s[0], s[1], s[2])];
            }
}

/* cross-section box of the faces normal to `axis` */
static box3 face_box(const orc_params *p, int axis, int lo, int hi)
{
    box3 b;
    int m[3] = {p->mx, p->my, p->mz};
    for (int a = 0; a < 3; a++) { b.lo[a] = 1; b.hi[a] = m[a]; }
    if (axis != 0) { b.lo[0] = 3; b.hi[0] = p->mx - 3; }
    b.lo[axis] = lo; b.hi[axis] = hi;
    return b;
}

/* SendRecvFldsYZ/ZX/XY for one axis: ghosts 1..2 <- m-4..m-3, m-2..m <- 3..5 */
static void halo_copy_axis(const orc_params *p, double *F, int axis)
{
    int m = (axis == 0) ? p->mx : (axis == 1) ? p->my : p->mz;
    copy_box(p, F, face_box(p, axis, 1, 2), axis, m - 5);
    copy_box(p, F, face_box(p, axis, m - 2, m), axis, -(m - 5));
}

/* SendRecvCurrentYZ/ZX/XY for one axis: 3..5 += m-2..m, then m-4..m-3 += 1..2 */
static void halo_add_axis(const orc_params *p, double *F, int axis)
{
    int m = (axis == 0) ? p->mx : (axis == 1) ? p->my : p->mz;
    add_box(p, F, face_box(p, axis, 3, 5), axis, m - 5);
    add_box(p, F, face_box(p, axis, m - 4, m - 3), axis, -(m - 5));
}

/* comm_fld.F90:75-91 with indepLBaxis = 0: ZX, (XY), then YZ */
void orc_field_halo(const orc_params *p, double *fx, double *fy, double *fz)
{
    double *F[3] = {fx, fy, fz};
    /* a non-periodic axis has null neighbours on one rank (PeriodicOff): no exchange */
    if (p->periodic[1])
        for (int c = 0; c < 3; c++) halo_copy_axis(p, F[c], 1);
    if (p->ndim == 3 && p->periodic[2])
        for (int c = 0; c < 3; c++) halo_copy_axis(p, F[c], 2);
    if (p->periodic[0])
        for (int c = 0; c < 3; c++) halo_copy_axis(p, F[c], 0);
}

/* comm_fld.F90:15-33 with indepLBaxis = 0: YZ, ZX, (XY) */
void orc_current_halo_add(const orc_params *p, double *jx, double *jy, double *jz)
{
    double *F[3] = {jx, jy, jz};
    if (p->periodic[0])
        for (int c = 0; c < 3; c++) halo_add_axis(p, F[c], 0);
    if (p->periodic[1])
        for (int c = 0; c < 3; c++) halo_add_axis(p, F[c], 1);
    if (p->ndim == 3 && p->periodic[2])
        for (int c = 0; c < 3; c++) halo_add_axis(p, F[c], 2);
}

/* ------------------------------------------------------------------ */
/* smoothen_current, current.F90:19-30.  Each pass is MovingAverageFilter */
/* (:55-75): SendRecvFldsYZ then MovingAverageX (:77-104) over            */
/* i = 2..mx-2, j = 2..my-2, k = k1..k2; SendRecvFldsZX then              */
/* MovingAverageY (:106-134); in 3-D SendRecvFldsXY then MovingAverageZ   */
/* (:136-163) over k = 2..mz-2 (k1..k2 = 2..mz-2 in 3-D, 1..1 in 2-D).    */
/* Every component is filtered into a scratch array over the box and     */
/* copied back (Fld_filtered).  Weights wtm1, wt0, wtp1 = 1/4, 1/2, 1/4   */
/* (vars.F90:208-210).  CurrentFilter_BC (:165-) folds currents only at   */
/* particle boundaries, which a periodic box does not have.               */
/* ------------------------------------------------------------------ */

static void moving_average_1(const orc_params *p, double *F, double *T, int axis, int k1, int k2)
{
    const double wtm1 = 0.25, wt0 = 0.5, wtp1 = 0.25;
    const int i1 = 2, i2 = p->mx - 2, j1 = 2, j2 = p->my - 2;
    for (int k = k1; k <= k2; k++)
        for (int j = j1; j <= j2; j++)
            for (int i = i1; i <= i2; i++) {
                int m[3] = {i, j, k}, q[3] = {i, j, k};
                m[axis] -= 1;
                q[axis] += 1;
                T[IX(p, i, j, k)] = wtm1 * F[IX(p, m[0], m[1], m[2])] + wt0 * F[IX(p, i, j, k)] +
                                    wtp1 * F[IX(p, q[0], q[1], q[2])];
            }
    for (int k = k1; k <= k2; k++)
        for (int j = j1; j <= j2; j++)
            for (int i = i1; i <= i2; i++) F[IX(p, i, j, k)] = T[IX(p, i, j, k)];
}

void orc_smoothen_current(const orc_params *p, orc_fields *f)
{
    if (p->curr_filter <= 0) return;
    double *T = (double *)calloc((size_t)p->mx * p->my * p->mz, sizeof(double));
    if (!T) return;
    double *J[3] = {f->Jx, f->Jy, f->Jz};
    const int k1 = (p->ndim == 3) ? 2 : 1, k2 = (p->ndim == 3) ? p->mz - 2 : 1;
    for (int ni = 1; ni <= p->curr_filter; ni++) {
        for (int c = 0; c < 3; c++) halo_copy_axis(p, J[c], 0);
        for (int c = 0; c < 3; c++) moving_average_1(p, J[c], T, 0, k1, k2);
        for (int c = 0; c < 3; c++) halo_copy_axis(p, J[c], 1);
        for (int c = 0; c < 3; c++) moving_average_1(p, J[c], T, 1, k1, k2);
        if (p->ndim == 3) {
            for (int c = 0; c < 3; c++) halo_copy_axis(p, J[c], 2);
            for (int c = 0; c < 3; c++) moving_average_1(p, J[c], T, 2, 2, p->mz - 2);
        }
    }
    free(T);
}

/* ------------------------------------------------------------------ */
/* Gather of node-averaged corner values, movdep_routines.F90:6-223     */
/* ------------------------------------------------------------------ */

void orc_gather_vec_e(const orc_params *p, orc_ws *w, const orc_fields *f)
{
    const int mx = p->mx, my = p->my;
    const int nv = w->nv;
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 2; k1 = p->mz - 1; }
    const double *Fx = f->Ex, *Fy = f->Ey, *Fz = f->Ez;
    double *Vx = w->vec[0], *Vy = w->vec[1], *Vz = w->vec[2];
#define F3(A, i, j, k) A[IX(p, i, j, k)]
    for (int k = k0; k <= k1; k++)
        #pragma omp parallel for schedule(static)
        for (int j = 2; j <= my - 1; j++)
            for (int i = 2; i <= mx - 1; i++) {
                int64_t s = (int64_t)(k - 1) * my * mx + (int64_t)(j - 1) * mx + i;  /* 1-based */
                double *v = Vx + nv * (s - 1);
                v[0] = (F3(Fx, i - 1, j, k) + F3(Fx, i, j, k)) * 0.5;
                v[1] = (F3(Fx, i, j, k) + F3(Fx, i + 1, j, k)) * 0.5;
                v[2] = (F3(Fx, i - 1, j + 1, k) + F3(Fx, i, j + 1, k)) * 0.5;
                v[3] = (F3(Fx, i, j + 1, k) + F3(Fx, i + 1, j + 1, k)) * 0.5;
                if (p->ndim == 3) {
                    v[4] = (F3(Fx, i - 1, j, k + 1) + F3(Fx, i, j, k + 1)) * 0.5;
                    v[5] = (F3(Fx, i, j, k + 1) + F3(Fx, i + 1, j, k + 1)) * 0.5;
                    v[6] = (F3(Fx, i - 1, j + 1, k + 1) + F3(Fx, i, j + 1, k + 1)) * 0.5;
                    v[7] = (F3(Fx, i, j + 1, k + 1) + F3(Fx, i + 1, j + 1, k + 1)) * 0.5;
                }
                v = Vy + nv * (s - 1);
                v[0] = (F3(Fy, i, j - 1, k) + F3(Fy, i, j, k)) * 0.5;
                v[1] = (F3(Fy, i + 1, j - 1, k) + F3(Fy, i + 1, j, k)) * 0.5;
                v[2] = (F3(Fy, i, j, k) + F3(Fy, i, j + 1, k)) * 0.5;
                v[3] = (F3(Fy, i + 1, j, k) + F3(Fy, i + 1, j + 1, k)) * 0.5;
                if (p->ndim == 3) {
                    v[4] = (F3(Fy, i, j - 1, k + 1) + F3(Fy, i, j, k + 1)) * 0.5;
                    v[5] = (F3(Fy, i + 1, j - 1, k + 1) + F3(Fy, i + 1, j, k + 1)) * 0.5;
                    v[6] = (F3(Fy, i, j, k + 1) + F3(Fy, i, j + 1, k + 1)) * 0.5;
                    v[7] = (F3(Fy, i + 1, j, k + 1) + F3(Fy, i + 1, j + 1, k + 1)) * 0.5;
                }
                v = Vz + nv * (s - 1);
                if (p->ndim == 3) {
                    v[0] = (F3(Fz, i, j, k - 1) + F3(Fz, i, j, k)) * 0.5;
                    v[1] = (F3(Fz, i + 1, j, k - 1) + F3(Fz, i + 1, j, k)) * 0.5;
                    v[2] = (F3(Fz, i, j + 1, k - 1) + F3(Fz, i, j + 1, k)) * 0.5;
                    v[3] = (F3(Fz, i + 1, j + 1, k - 1) + F3(Fz, i + 1, j + 1, k)) * 0.5;
                    v[4] = (F3(Fz, i, j, k) + F3(Fz, i, j, k + 1)) * 0.5;
                    v[5] = (F3(Fz, i + 1, j, k) + F3(Fz, i + 1, j, k + 1)) * 0.5;
                    v[6] = (F3(Fz, i, j + 1, k) + F3(Fz, i, j + 1, k + 1)) * 0.5;
                    v[7] = (F3(Fz, i + 1, j + 1, k) + F3(Fz, i + 1, j + 1, k + 1)) * 0.5;
                } else {
                    v[0] = F3(Fz, i, j, k);
                    v[1] = F3(Fz, i + 1, j, k);
                    v[2] = F3(Fz, i, j + 1, k);
                    v[3] = F3(Fz, i + 1, j + 1, k);
                }
            }
}

void orc_gather_vec_b(const orc_params *p, orc_ws *w, const orc_fields *f)
{
    const int mx = p->mx, my = p->my;
    const int nv = w->nv;
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 2; k1 = p->mz - 1; }
    const double *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    const double ex = p->Bx_ext0, ey = p->By_ext0, ez = p->Bz_ext0;
    double *Vx = w->vec[3], *Vy = w->vec[4], *Vz = w->vec[5];
    for (int k = k0; k <= k1; k++)
        #pragma omp parallel for schedule(static)
        for (int j = 2; j <= my - 1; j++)
            for (int i = 2; i <= mx - 1; i++) {
                int64_t s = (int64_t)(k - 1) * my * mx + (int64_t)(j - 1) * mx + i;
                double *v = Vx + nv * (s - 1);
                if (p->ndim == 3) {
                    v[0] = (F3(Bx, i, j - 1, k - 1) + F3(Bx, i, j, k - 1) + F3(Bx, i, j - 1, k) + F3(Bx, i, j, k)) * 0.25 + ex;
                    v[1] = (F3(Bx, i + 1, j - 1, k - 1) + F3(Bx, i + 1, j, k - 1) + F3(Bx, i + 1, j - 1, k) + F3(Bx, i + 1, j, k)) * 0.25 + ex;
                    v[2] = (F3(Bx, i, j, k - 1) + F3(Bx, i, j + 1, k - 1) + F3(Bx, i, j, k) + F3(Bx, i, j + 1, k)) * 0.25 + ex;
                    v[3] = (F3(Bx, i + 1, j, k - 1) + F3(Bx, i + 1, j + 1, k - 1) + F3(Bx, i + 1, j, k) + F3(Bx, i + 1, j + 1, k)) * 0.25 + ex;
                    v[4] = (F3(Bx, i, j - 1, k) + F3(Bx, i, j, k) + F3(Bx, i, j - 1, k + 1) + F3(Bx, i, j, k + 1)) * 0.25 + ex;
                    v[5] = (F3(Bx, i + 1, j - 1, k) + F3(Bx, i + 1, j, k) + F3(Bx, i + 1, j - 1, k + 1) + F3(Bx, i + 1, j, k + 1)) * 0.25 + ex;
                    v[6] = (F3(Bx, i, j, k) + F3(Bx, i, j + 1, k) + F3(Bx, i, j, k + 1) + F3(Bx, i, j + 1, k + 1)) * 0.25 + ex;
                    v[7] = (F3(Bx, i + 1, j, k) + F3(Bx, i + 1, j + 1, k) + F3(Bx, i + 1, j, k + 1) + F3(Bx, i + 1, j + 1, k + 1)) * 0.25 + ex;
                } else {
                    v[0] = (F3(Bx, i, j - 1, k) + F3(Bx, i, j, k)) * 0.5 + ex;
                    v[1] = (F3(Bx, i + 1, j - 1, k) + F3(Bx, i + 1, j, k)) * 0.5 + ex;
                    v[2] = (F3(Bx, i, j, k) + F3(Bx, i, j + 1, k)) * 0.5 + ex;
                    v[3] = (F3(Bx, i + 1, j, k) + F3(Bx, i + 1, j + 1, k)) * 0.5 + ex;
                }
                v = Vy + nv * (s - 1);
                if (p->ndim == 3) {
                    v[0] = (F3(By, i - 1, j, k - 1) + F3(By, i - 1, j, k) + F3(By, i, j, k - 1) + F3(By, i, j, k)) * 0.25 + ey;
                    v[1] = (F3(By, i, j, k - 1) + F3(By, i, j, k) + F3(By, i + 1, j, k - 1) + F3(By, i + 1, j, k)) * 0.25 + ey;
                    v[2] = (F3(By, i - 1, j + 1, k - 1) + F3(By, i - 1, j + 1, k) + F3(By, i, j + 1, k - 1) + F3(By, i, j + 1, k)) * 0.25 + ey;
                    v[3] = (F3(By, i, j + 1, k - 1) + F3(By, i, j + 1, k) + F3(By, i + 1, j + 1, k - 1) + F3(By, i + 1, j + 1, k)) * 0.25 + ey;
                    v[4] = (F3(By, i - 1, j, k) + F3(By, i - 1, j, k + 1) + F3(By, i, j, k) + F3(By, i, j, k + 1)) * 0.25 + ey;
                    v[5] = (F3(By, i, j, k) + F3(By, i, j, k + 1) + F3(By, i + 1, j, k) + F3(By, i + 1, j, k + 1)) * 0.25 + ey;
                    v[6] = (F3(By, i - 1, j + 1, k) + F3(By, i - 1, j + 1, k + 1) + F3(By, i, j + 1, k) + F3(By, i, j + 1, k + 1)) * 0.25 + ey;
                    v[7] = (F3(By, i, j + 1, k) + F3(By, i, j + 1, k + 1) + F3(By, i + 1, j + 1, k) + F3(By, i + 1, j + 1, k + 1)) * 0.25 + ey;
                } else {
                    v[0] = (F3(By, i - 1, j, k) + F3(By, i, j, k)) * 0.5 + ey;
                    v[1] = (F3(By, i, j, k) + F3(By, i + 1, j, k)) * 0.5 + ey;
                    v[2] = (F3(By, i - 1, j + 1, k) + F3(By, i, j + 1, k)) * 0.5 + ey;
                    v[3] = (F3(By, i, j + 1, k) + F3(By, i + 1, j + 1, k)) * 0.5 + ey;
                }
                v = Vz + nv * (s - 1);
                v[0] = (F3(Bz, i - 1, j - 1, k) + F3(Bz, i - 1, j, k) + F3(Bz, i, j - 1, k) + F3(Bz, i, j, k)) * 0.25 + ez;
                v[1] = (F3(Bz, i, j - 1, k) + F3(Bz, i, j, k) + F3(Bz, i + 1, j - 1, k) + F3(Bz, i + 1, j, k)) * 0.25 + ez;
                v[2] = (F3(Bz, i - 1, j, k) + F3(Bz, i - 1, j + 1, k) + F3(Bz, i, j, k) + F3(Bz, i, j + 1, k)) * 0.25 + ez;
                v[3] = (F3(Bz, i, j, k) + F3(Bz, i, j + 1, k) + F3(Bz, i + 1, j, k) + F3(Bz, i + 1, j + 1, k)) * 0.25 + ez;
                if (p->ndim == 3) {
                    v[4] = (F3(Bz, i - 1, j - 1, k + 1) + F3(Bz, i - 1, j, k + 1) + F3(Bz, i, j - 1, k + 1) + F3(Bz, i, j, k + 1)) * 0.25 + ez;
                    v[5] = (F3(Bz, i, j - 1, k + 1) + F3(Bz, i, j, k + 1) + F3(Bz, i + 1, j - 1, k + 1) + F3(Bz, i + 1, j, k + 1)) * 0.25 + ez;
                    v[6] = (F3(Bz, i - 1, j, k + 1) + F3(Bz, i - 1, j + 1, k + 1) + F3(Bz, i, j, k + 1) + F3(Bz, i, j + 1, k + 1)) * 0.25 + ez;
                    v[7] = (F3(Bz, i, j, k + 1) + F3(Bz, i, j + 1, k + 1) + F3(Bz, i + 1, j, k + 1) + F3(Bz, i + 1, j + 1, k + 1)) * 0.25 + ez;
                }
            }
#undef F3
}

/* ------------------------------------------------------------------ */
/* Mover: movdep.F90:12-258                                            */
/* ------------------------------------------------------------------ */

static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (b > a) ? b : a; }
static inline int imin2(int a, int b) { return a < b ? a : b; }
static inline int imax2(int a, int b) { return a > b ? a : b; }

/* ReduceCurrentMatrix, movdep_routines.F90:243-311.  The reference threads
 * the scatter over j1 (2-D) / k1 (3-D) while iteration j1 also writes row
 * j1+1, which races at chunk borders; the threaded build here runs the same
 * scatter as two passes over even and odd rows so that it is race free. */
static void reduce_current(const orc_params *p, orc_ws *w, orc_fields *f)
{
    const int nj = w->nj, nt = w->nthreads;
    const int64_t ncell = w->ncell;
    double *V = w->vecj;
    if (nt > 1) {
        #pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < ncell; c++)
            for (int t = 1; t < nt; t++)
                for (int n = 0; n < nj; n++)
                    V[n + nj * c] += V[n + nj * c + (int64_t)nj * ncell * t];
    }
    const int mx = p->mx, my = p->my, mz = p->mz;
    if (p->ndim == 2) {
        for (int par = 0; par < 2; par++) {
            #pragma omp parallel for schedule(static)
            for (int j1 = 1 + par; j1 <= my - 1; j1 += 2)
                for (int i1 = 1; i1 <= mx - 1; i1++) {
                    const double *v = V + nj * ((int64_t)(j1 - 1) * mx + i1 - 1);
                    f->Jx[IX(p, i1, j1, 1)] += v[0];
                    f->Jx[IX(p, i1, j1 + 1, 1)] += v[1];
                    f->Jy[IX(p, i1, j1, 1)] += v[2];
                    f->Jy[IX(p, i1 + 1, j1, 1)] += v[3];
                    f->Jz[IX(p, i1, j1, 1)] += v[4];
                    f->Jz[IX(p, i1 + 1, j1, 1)] += v[5];
                    f->Jz[IX(p, i1, j1 + 1, 1)] += v[6];
                    f->Jz[IX(p, i1 + 1, j1 + 1, 1)] += v[7];
                }
        }
    } else {
        for (int par = 0; par < 2; par++) {
            #pragma omp parallel for schedule(static)
            for (int k1 = 1 + par; k1 <= mz - 1; k1 += 2)
                for (int j1 = 1; j1 <= my - 1; j1++)
                    for (int i1 = 1; i1 <= mx - 1; i1++) {
                        const double *v = V + nj * ((int64_t)(k1 - 1) * my * mx + (int64_t)(j1 - 1) * mx + i1 - 1);
                        f->Jx[IX(p, i1, j1, k1)] += v[0];
                        f->Jx[IX(p, i1, j1 + 1, k1)] += v[1];
                        f->Jx[IX(p, i1, j1, k1 + 1)] += v[2];
                        f->Jx[IX(p, i1, j1 + 1, k1 + 1)] += v[3];
                        f->Jy[IX(p, i1, j1, k1)] += v[4];
                        f->Jy[IX(p, i1 + 1, j1, k1)] += v[5];
                        f->Jy[IX(p, i1, j1, k1 + 1)] += v[6];
                        f->Jy[IX(p, i1 + 1, j1, k1 + 1)] += v[7];
                        f->Jz[IX(p, i1, j1, k1)] += v[8];
                        f->Jz[IX(p, i1 + 1, j1, k1)] += v[9];
                        f->Jz[IX(p, i1, j1 + 1, k1)] += v[10];
                        f->Jz[IX(p, i1 + 1, j1 + 1, k1)] += v[11];
                    }
        }
    }
}

void orc_move_deposit(const orc_params *p, orc_ws *w, orc_fields *f, orc_particles *s)
{
    const int two = (p->ndim == 2);
    const int nv = w->nv, nj = w->nj;
    const int mx = p->mx, my = p->my;
    const int64_t used = s->used, ncell = w->ncell;
    const double c = p->c, sqc = p->sqc, cinv = p->cinv, qi = p->qi;
    const double *wtx1 = two ? wtx1_2 : wtx1_3, *wtx2 = two ? wtx2_2 : wtx2_3;
    const double *wty1 = two ? wty1_2 : wty1_3, *wty2 = two ? wty2_2 : wty2_3;
    const double *wtFx = two ? wtFx_2 : wtFx_3, *wtFy = two ? wtFy_2 : wtFy_3, *wtFz = two ? wtFz_2 : wtFz_3;
    const double *Jwtx1 = two ? Jwtx1_2 : Jwtx1_3, *Jwtx2 = two ? Jwtx2_2 : Jwtx2_3;
    const double *Jwty1 = two ? Jwty1_2 : Jwty1_3, *Jwty2 = two ? Jwty2_2 : Jwty2_3;
    const double *Jwtz1 = two ? Jwtz1_2 : Jwtz1_3, *Jwtz2 = two ? Jwtz2_2 : Jwtz2_3;

    orc_gather_vec_e(p, w, f);
    orc_gather_vec_b(p, w, f);
    memset(w->vecj, 0, sizeof(double) * (size_t)(nj * ncell * w->nthreads));   /* VecJ = 0 */

    const double *VEx = w->vec[0], *VEy = w->vec[1], *VEz = w->vec[2];
    const double *VBx = w->vec[3], *VBy = w->vec[4], *VBz = w->vec[5];

    #pragma omp parallel for schedule(static)
    for (int64_t off = 0; off < used; off += VBS) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *VJ = w->vecj + (int64_t)nj * ncell * tid;
        const int nb = (int)((used - off < VBS) ? (used - off) : VBS);
        double pEx[VBS], pEy[VBS], pEz[VBS], pBx[VBS], pBy[VBS], pBz[VBS];
        double u0[VBS], v0[VBS], w0[VBS], u1[VBS], v1[VBS], w1[VBS], g[VBS], ff[VBS];
        double x0[VBS], y0[VBS], z0[VBS], qm[VBS];
        int ip0[VBS], jp0[VBS], kp0[VBS];

        for (int n = 0; n < nb; n++) {
            x0[n] = s->x[off + n]; y0[n] = s->y[off + n]; z0[n] = s->z[off + n];
        }
        for (int n = 0; n < nb; n++)
            qm[n] = (s->flv[off + n] == 0) ? 0.0 : p->flvrqm[s->flv[off + n] - 1] * 0.5;
        for (int n = 0; n < nb; n++) {
            pEx[n] = 0.0; pEy[n] = 0.0; pEz[n] = 0.0;
            pBx[n] = 0.0; pBy[n] = 0.0; pBz[n] = 0.0;
        }
        for (int n = 0; n < nb; n++) {
            double wt[8];
            ip0[n] = (int)x0[n];        /* Fortran real -> integer truncation */
            jp0[n] = (int)y0[n];
            kp0[n] = (int)z0[n];
            double dx = x0[n] - ip0[n];
            double dy = y0[n] - jp0[n];
            int64_t sidx;
            if (two) {
                for (int nn = 0; nn < 4; nn++)
                    wt[nn] = (wtx1[nn] + wtx2[nn] * dx) * (wty1[nn] + wty2[nn] * dy);
                sidx = (int64_t)(jp0[n] - 1) * mx + ip0[n];
            } else {
                double dz = z0[n] - kp0[n];
                for (int nn = 0; nn < 8; nn++)
                    wt[nn] = (wtx1[nn] + wtx2[nn] * dx) * (wty1[nn] + wty2[nn] * dy) * (wtz1_3[nn] + wtz2_3[nn] * dz);
                sidx = (int64_t)(kp0[n] - 1) * mx * my + (int64_t)(jp0[n] - 1) * mx + ip0[n];
            }
            const int64_t b = nv * (sidx - 1);
            for (int nn = 0; nn < nv; nn++) {
                pEx[n] = pEx[n] + wt[nn] * VEx[b + nn];
                pEy[n] = pEy[n] + wt[nn] * VEy[b + nn];
                pEz[n] = pEz[n] + wt[nn] * VEz[b + nn];
                pBx[n] = pBx[n] + wt[nn] * VBx[b + nn];
                pBy[n] = pBy[n] + wt[nn] * VBy[b + nn];
                pBz[n] = pBz[n] + wt[nn] * VBz[b + nn];
            }
        }
        for (int n = 0; n < nb; n++) {
            pEx[n] = pEx[n] * qm[n]; pEy[n] = pEy[n] * qm[n]; pEz[n] = pEz[n] * qm[n];
            pBx[n] = pBx[n] * qm[n]; pBy[n] = pBy[n] * qm[n]; pBz[n] = pBz[n] * qm[n];
        }
        /* Boris pusher, movdep.F90:143-173 */
        for (int n = 0; n < nb; n++) {
            const int64_t m = off + n;
            u0[n] = c * s->u[m] + pEx[n];
            v0[n] = c * s->v[m] + pEy[n];
            w0[n] = c * s->w[m] + pEz[n];
            g[n] = 1.0 / sqrt(sqc + u0[n] * u0[n] + v0[n] * v0[n] + w0[n] * w0[n]);
            pBx[n] = g[n] * pBx[n];
            pBy[n] = g[n] * pBy[n];
            pBz[n] = g[n] * pBz[n];
            ff[n] = 2.0 / (1.0 + pBx[n] * pBx[n] + pBy[n] * pBy[n] + pBz[n] * pBz[n]);
            u1[n] = (u0[n] + v0[n] * pBz[n] - w0[n] * pBy[n]) * ff[n];
            v1[n] = (v0[n] + w0[n] * pBx[n] - u0[n] * pBz[n]) * ff[n];
            w1[n] = (w0[n] + u0[n] * pBy[n] - v0[n] * pBx[n]) * ff[n];
            u0[n] = u0[n] + v1[n] * pBz[n] - w1[n] * pBy[n] + pEx[n];
            v0[n] = v0[n] + w1[n] * pBx[n] - u1[n] * pBz[n] + pEy[n];
            w0[n] = w0[n] + u1[n] * pBy[n] - v1[n] * pBx[n] + pEz[n];
            s->u[m] = u0[n] * cinv;
            s->v[m] = v0[n] * cinv;
            s->w[m] = w0[n] * cinv;
            g[n] = sqc / sqrt(sqc + u0[n] * u0[n] + v0[n] * v0[n] + w0[n] * w0[n]);
            s->x[m] = s->x[m] + s->u[m] * g[n] * p->inv_dx;
            s->y[m] = s->y[m] + s->v[m] * g[n] * p->inv_dy;
            s->z[m] = s->z[m] + s->w[m] * g[n] * p->inv_dz;
        }
        /* zigzag deposit, movdep.F90:177-237 */
        for (int n = 0; n < nb; n++) {
            const int64_t m = off + n;
            const int ip = (int)s->x[m], jp = (int)s->y[m], kp = (int)s->z[m];
            const double xr = dmin((double)(imin2(ip0[n], ip) + 1), dmax((double)imax2(ip0[n], ip), 0.5 * (x0[n] + s->x[m])));
            const double yr = dmin((double)(imin2(jp0[n], jp) + 1), dmax((double)imax2(jp0[n], jp), 0.5 * (y0[n] + s->y[m])));
            const double zr = dmin((double)(imin2(kp0[n], kp) + 1), dmax((double)imax2(kp0[n], kp), 0.5 * (z0[n] + s->z[m])));
            const double qthis = s->q[m] * qi;
            double Fx1 = qthis * (xr - x0[n]);
            double Fy1 = qthis * (yr - y0[n]);
            double Fz1 = qthis * (zr - z0[n]);
            double Wx1 = 0.5 * (x0[n] + xr) - ip0[n];
            double Wy1 = 0.5 * (y0[n] + yr) - jp0[n];
            double Wz1;
            int64_t sidx;
            if (two) {
                Wz1 = 0.0;
                sidx = (int64_t)(jp0[n] - 1) * mx + ip0[n];
            } else {
                Wz1 = 0.5 * (z0[n] + zr) - kp0[n];
                sidx = (int64_t)(kp0[n] - 1) * mx * my + (int64_t)(jp0[n] - 1) * mx + ip0[n];
            }
            double *vj = VJ + nj * (sidx - 1);
            for (int nn = 0; nn < nj; nn++)
                vj[nn] = vj[nn] + (wtFx[nn] * Fx1 + wtFy[nn] * Fy1 + wtFz[nn] * Fz1) * (Jwtx1[nn] + Jwtx2[nn] * Wx1)
                                  * (Jwty1[nn] + Jwty2[nn] * Wy1) * (Jwtz1[nn] + Jwtz2[nn] * Wz1);

            double Fx2 = qthis * (s->x[m] - xr);
            double Fy2 = qthis * (s->y[m] - yr);
            double Fz2 = qthis * (s->z[m] - zr);
            double Wx2 = 0.5 * (s->x[m] + xr) - ip;
            double Wy2 = 0.5 * (s->y[m] + yr) - jp;
            double Wz2;
            if (two) {
                Wz2 = 0.0;
                sidx = (int64_t)(jp - 1) * mx + ip;
            } else {
                Wz2 = 0.5 * (s->z[m] + zr) - kp;
                sidx = (int64_t)(kp - 1) * mx * my + (int64_t)(jp - 1) * mx + ip;
            }
            vj = VJ + nj * (sidx - 1);
            for (int nn = 0; nn < nj; nn++)
                vj[nn] = vj[nn] + (wtFx[nn] * Fx2 + wtFy[nn] * Fy2 + wtFz[nn] * Fz2) * (Jwtx1[nn] + Jwtx2[nn] * Wx2)
                                  * (Jwty1[nn] + Jwty2[nn] * Wy2) * (Jwtz1[nn] + Jwtz2[nn] * Wz2);
        }
        /* 2-D periodic wrap in z, movdep.F90:241-249 */
        if (two)
            for (int n = 0; n < nb; n++) {
                const int64_t m = off + n;
                if (s->z[m] > p->zmax) s->z[m] = s->z[m] - p->zlen;
                else if (s->z[m] < p->zmin) s->z[m] = p->zlen + s->z[m];
            }
    }
    reduce_current(p, w, f);
}

/* ------------------------------------------------------------------ */
/* Particle exchange on one periodic rank                              */
/* ------------------------------------------------------------------ */

/* RemovePrtl, loadprtlout.F90:164-181 */
static void remove_prtl(const orc_params *p, orc_particles *s, int64_t i)
{
    s->q[i] = 0; s->flv[i] = 0;
    s->u[i] = 0; s->v[i] = 0; s->w[i] = 0;
    s->x[i] = 3.8; s->y[i] = 3.8;
    s->z[i] = (p->ndim == 2) ? 1.5 : 3.8;
}

typedef struct { int64_t n, cap; double *d; } obuf;   /* 7 doubles + flv per record */

static int obuf_push(obuf *b, const orc_particles *s, int64_t i, double xs, double ys, double zs)
{
    if (b->n == b->cap) {
        int64_t nc = b->cap ? 2 * b->cap : 1024;
        double *nd = (double *)realloc(b->d, sizeof(double) * 8 * (size_t)nc);
        if (!nd) return ORC_ERR_CAPACITY;
        b->d = nd; b->cap = nc;
    }
    double *r = b->d + 8 * b->n;
    /* CopyTransferPrtl, loadprtlout.F90:128-142: position minus the send shift */
    r[0] = s->q[i]; r[1] = s->x[i] - xs; r[2] = s->y[i] - ys; r[3] = s->z[i] - zs;
    r[4] = s->u[i]; r[5] = s->v[i]; r[6] = s->w[i]; r[7] = (double)s->flv[i];
    b->n++;
    return ORC_OK;
}

/* ExchangePrtl + AppendParticles for nSubDomains = 1 (every face is its own
 * neighbour).  Send shifts: face 2 xlen, face 4 ylen, face 6 zlen
 * (subdomains.F90:720-740); the receiving face adds its own shift
 * (mem_prtl.F90:135-160).  Receive face n gets what the opposite send face
 * sent (tag pairing subdomains.F90:595-610), and arrivals are appended in
 * receive-face order 1..6. */
int orc_exchange_particles(const orc_params *p, orc_particles *s)
{
    obuf out[6];
    memset(out, 0, sizeof(out));
    const double xmin = p->xmin, xmax = p->xmax, ymin = p->ymin, ymax = p->ymax, zmin = p->zmin, zmax = p->zmax;
    const double sendx[6] = {0, p->xlen, 0, 0, 0, 0};
    const double sendy[6] = {0, 0, 0, p->ylen, 0, 0};
    const double sendz[6] = {0, 0, 0, 0, 0, p->zlen};
    const int three = (p->ndim == 3);
    int rc = ORC_OK;
    /* LoadPrtlOutliers, loadprtlout.F90:22-97 */
    for (int64_t i = 0; i < s->used && rc == ORC_OK; i++) {
        if (s->flv[i] == 0) continue;
        double x = s->x[i], y = s->y[i], z = s->z[i];
        if (x < xmin) {
            if ((y <= (ymax + xmin - x)) && (y >= (ymin - xmin + x))) {
                if (!three || ((z <= (zmax + xmin - x)) && (z >= (zmin - xmin + x)))) {
                    rc |= obuf_push(&out[0], s, i, sendx[0], sendy[0], sendz[0]);
                    remove_prtl(p, s, i);
                }
            }
        } else if (x > xmax) {
            if ((y <= (ymax - xmax + x)) && (y >= (ymin + xmax - x))) {
                if (!three || ((z <= (zmax - xmax + x)) && (z >= (zmin + xmax - x)))) {
                    rc |= obuf_push(&out[1], s, i, sendx[1], sendy[1], sendz[1]);
                    remove_prtl(p, s, i);
                }
            }
        }
        x = s->x[i]; y = s->y[i]; z = s->z[i];
        if (y < ymin) {
            if ((y < (ymin - xmin + x)) && (y < (ymin + xmax - x))) {
                if (!three || ((z >= (zmin - ymin + y)) && (z <= (zmax + ymin - y)))) {
                    rc |= obuf_push(&out[2], s, i, sendx[2], sendy[2], sendz[2]);
                    remove_prtl(p, s, i);
                }
            }
        } else if (y > ymax) {
            if ((y > (ymax + xmin - x)) && (y > (ymax - xmax + x))) {
                if (!three || ((z >= (zmin + ymax - y)) && (z <= (zmax - ymax + y)))) {
                    rc |= obuf_push(&out[3], s, i, sendx[3], sendy[3], sendz[3]);
                    remove_prtl(p, s, i);
                }
            }
        }
        if (three) {
            x = s->x[i]; y = s->y[i]; z = s->z[i];
            if (z < zmin) {
                if ((z < (zmin - xmin + x)) && (z < (zmin + xmax - x))) {
                    if ((z < (zmin - ymin + y)) && (z < (zmin + ymax - y))) {
                        rc |= obuf_push(&out[4], s, i, sendx[4], sendy[4], sendz[4]);
                        remove_prtl(p, s, i);
                    }
                }
            } else if (z > zmax) {
                if ((z > (zmax + xmin - x)) && (z > (zmax - xmax + x))) {
                    if ((z > (zmax + ymin - y)) && (z > (zmax - ymax + y))) {
                        rc |= obuf_push(&out[5], s, i, sendx[5], sendy[5], sendz[5]);
                        remove_prtl(p, s, i);
                    }
                }
            }
        }
    }
    /* append: receive face n <- send face n^1 (0-based pairs (0,1),(2,3),(4,5)) */
    const double recvx[6] = {0, p->xlen, 0, 0, 0, 0};
    const double recvy[6] = {0, 0, 0, p->ylen, 0, 0};
    const double recvz[6] = {0, 0, 0, 0, 0, p->zlen};
    int64_t total = 0;
    for (int n = 0; n < 6; n++) total += out[n].n;
    if (rc == ORC_OK && s->used + total > s->cap) rc = ORC_ERR_CAPACITY;
    if (rc == ORC_OK) {
        int nfaces = three ? 6 : 4;
        for (int n = 0; n < nfaces; n++) {
            /* a non-periodic face sends to a null neighbour: the particle is gone */
            if (!p->periodic[n >> 1]) continue;
            obuf *b = &out[n ^ 1];
            for (int64_t r = 0; r < b->n; r++) {
                const double *d = b->d + 8 * r;
                int64_t m = s->used + r;
                s->q[m] = d[0];
                s->x[m] = d[1] + recvx[n];
                s->y[m] = d[2] + recvy[n];
                s->z[m] = d[3] + recvz[n];
                s->u[m] = d[4]; s->v[m] = d[5]; s->w[m] = d[6];
                s->flv[m] = (int32_t)d[7];
            }
            s->used += b->n;
        }
    }
    for (int n = 0; n < 6; n++) free(out[n].d);
    return rc;
}

/* ReorderPrtlArr, mem_prtl.F90:166-301 (single thread: stable by index) */
int orc_reorder(const orc_params *p, orc_particles *s)
{
    const int mx = p->mx, my = p->my, mz = p->mz;
    const int64_t ncell = (int64_t)mx * my * mz;
    int64_t *pc = (int64_t *)calloc((size_t)ncell, sizeof(int64_t));
    if (!pc) return ORC_ERR_CAPACITY;
    int64_t np = 0;
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        int i = (int)floor(s->x[n]), j = (int)floor(s->y[n]);
        int k = (p->ndim == 2) ? 1 : (int)floor(s->z[n]);
        if (i < 1 || i > mx || j < 1 || j > my || k < 1 || k > mz) { free(pc); return ORC_ERR_ARG; }
        pc[IX(p, i, j, k)]++;
        np++;
    }
    int64_t offset = 0;
    for (int64_t c = 0; c < ncell; c++) { int64_t t = pc[c]; pc[c] = offset; offset += t; }
    double *buf = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
    int32_t *ibuf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
    int64_t *dst = (int64_t *)malloc(sizeof(int64_t) * (size_t)(s->used > 0 ? s->used : 1));
    if (!buf || !ibuf || !dst) { free(pc); free(buf); free(ibuf); free(dst); return ORC_ERR_CAPACITY; }
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) { dst[n] = -1; continue; }
        int i = (int)floor(s->x[n]), j = (int)floor(s->y[n]);
        int k = (p->ndim == 2) ? 1 : (int)floor(s->z[n]);
        dst[n] = pc[IX(p, i, j, k)]++;
    }
    double *arrs[7] = {s->q, s->x, s->y, s->z, s->u, s->v, s->w};
    for (int a = 0; a < 7; a++) {
        for (int64_t n = 0; n < s->used; n++) if (dst[n] >= 0) buf[dst[n]] = arrs[a][n];
        memcpy(arrs[a], buf, sizeof(double) * (size_t)np);
    }
    for (int64_t n = 0; n < s->used; n++) if (dst[n] >= 0) ibuf[dst[n]] = s->flv[n];
    memcpy(s->flv, ibuf, sizeof(int32_t) * (size_t)np);
    for (int64_t n = np; n < s->used; n++) { s->q[n] = 0; s->flv[n] = 0; }
    s->used = np;
    free(pc); free(buf); free(ibuf); free(dst);
    return ORC_OK;
}

/* One time step in the MAIN.F90:42-284 order.  The reference zeroes J at
 * the end of a step (MAIN.F90:248); here the same reset runs at the start
 * of the next step so J stays readable after the step (nothing reads J in
 * between). */
int orc_step(orc_params *p, orc_ws *w, orc_fields *f, orc_particles *s, int t)
{
    orc_reset_current(p, f);
    orc_update_b_half(p, f);
    orc_field_halo(p, f->Bx, f->By, f->Bz);
    orc_move_deposit(p, w, f, s);
    orc_bc_post_movdep(p, f, s);   /* EnforceBC_PostMovDep, MAIN.F90:77 */
    int rc = orc_exchange_particles(p, s);
    if (rc != ORC_OK) return rc;
    orc_update_b_half(p, f);
    orc_pml_b(p, f);               /* MAIN.F90:104 */
    orc_field_halo(p, f->Bx, f->By, f->Bz);
    orc_update_e(p, f);
    orc_pml_e(p, f);               /* MAIN.F90:109 */
    orc_field_halo(p, f->Ex, f->Ey, f->Ez);
    orc_current_halo_add(p, f->Jx, f->Jy, f->Jz);
    orc_smoothen_current(p, f);
    orc_add_current(p, f);
    orc_field_halo(p, f->Ex, f->Ey, f->Ez);
    orc_bc_final(p, f);            /* EnforceBC_Final, MAIN.F90:209 */
    if (p->reorder_period > 0 && t % p->reorder_period == 0) rc = orc_reorder(p, s);
    return rc;
}

int orc_run(orc_params *p, orc_ws *w, orc_fields *f, orc_particles *s, int t0, int nsteps)
{
    for (int t = t0; t < t0 + nsteps; t++) {
        int rc = orc_step(p, w, f, s, t);
        if (rc != ORC_OK) return rc;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------ */
/* Diagnostics                                                         */
/* ------------------------------------------------------------------ */

/* CalcEnergy, savedata_routines.F90:265-327: out = {B energy, E energy,
 * particle energy flavour 1, flavour 2}; wtp = 1 (unit weights). */
void orc_energy(const orc_params *p, const orc_fields *f, const orc_particles *s, double out[4])
{
    double be = 0, ee = 0;
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 3; k1 = p->mz - 3; }
    for (int k = k0; k <= k1; k++)
        for (int j = 3; j <= p->my - 3; j++)
            for (int i = 3; i <= p->mx - 3; i++) {
                int64_t o = IX(p, i, j, k);
                be = be + (f->Bx[o] * f->Bx[o] + f->By[o] * f->By[o] + f->Bz[o] * f->Bz[o]);
                ee = ee + (f->Ex[o] * f->Ex[o] + f->Ey[o] * f->Ey[o] + f->Ez[o] * f->Ez[o]);
            }
    out[0] = be * 0.5;
    out[1] = ee * 0.5;
    out[2] = 0; out[3] = 0;
    const double mec2 = p->masse * p->c * p->c;
    for (int64_t n = 0; n < s->used; n++) {
        int fl = s->flv[n];
        if (fl < 1 || fl > 2) continue;
        double gamma = sqrt(1.0 + s->u[n] * s->u[n] + s->v[n] * s->v[n] + s->w[n] * s->w[n]);
        if (s->q[n] > 0)
            out[1 + fl] = out[1 + fl] + (gamma - 1.0) * 1.0 * (s->q[n] * p->qi / p->flvrqm[fl - 1]) * p->c * p->c;
        if (s->q[n] < 0)
            out[1 + fl] = out[1 + fl] + (gamma - 1.0) * 1.0 * mec2;
    }
}

/* CalcDivE, savedata_routines.F90:329-355 */
void orc_div_e(const orc_params *p, const orc_fields *f, double *div)
{
    memset(div, 0, sizeof(double) * (size_t)p->mx * p->my * p->mz);
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 2; k1 = p->mz - 2; }
    for (int k = k0; k <= k1; k++)
        for (int j = 2; j <= p->my - 2; j++)
            for (int i = 2; i <= p->mx - 2; i++) {
                int64_t o = IX(p, i, j, k);
                if (p->ndim == 2)
                    div[o] = f->Ex[o] - f->Ex[IX(p, i - 1, j, k)] + f->Ey[o] - f->Ey[IX(p, i, j - 1, k)];
                else
                    div[o] = f->Ex[o] - f->Ex[IX(p, i - 1, j, k)] + f->Ey[o] - f->Ey[IX(p, i, j - 1, k)]
                           + f->Ez[o] - f->Ez[IX(p, i, j, k - 1)];
            }
}

static inline int fold(int n, int lo, int hi, int len)
{
    while (n < lo) n += len;
    while (n > hi) n -= len;
    return n;
}

/* node index folded onto the physical nodes along a periodic axis; a
 * non-periodic axis keeps its ghost nodes */
static inline int pfold(const orc_params *p, int axis, int n)
{
    if (!p->periodic[axis]) return n;
    if (axis == 0) return fold(n, 3, p->mx - 3, p->nx);
    if (axis == 1) return fold(n, 3, p->my - 3, p->ny);
    return fold(n, 3, p->mz - 3, p->nz);
}

/* CIC node charge rho(i,j,k) = sum q*qi * S_x * S_y (* S_z) at the nodes of
 * the cell holding the particle, folded onto physical nodes 3..m-3.  With
 * charge-conserving deposition div E - rho is invariant in time. */
void orc_rho(const orc_params *p, const orc_particles *s, double *rho)
{
    memset(rho, 0, sizeof(double) * (size_t)p->mx * p->my * p->mz);
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        const double qt = s->q[n] * p->qi;
        const int i = (int)floor(s->x[n]), j = (int)floor(s->y[n]);
        const double dx = s->x[n] - i, dy = s->y[n] - j;
        const double sx[2] = {1.0 - dx, dx}, sy[2] = {1.0 - dy, dy};
        if (p->ndim == 2) {
            for (int b = 0; b < 2; b++)
                for (int a = 0; a < 2; a++) {
                    int ii = pfold(p, 0, i + a), jj = pfold(p, 1, j + b);
                    rho[IX(p, ii, jj, 1)] += qt * sx[a] * sy[b];
                }
        } else {
            const int k = (int)floor(s->z[n]);
            const double dz = s->z[n] - k;
            const double sz[2] = {1.0 - dz, dz};
            for (int cc = 0; cc < 2; cc++)
                for (int b = 0; b < 2; b++)
                    for (int a = 0; a < 2; a++) {
                        int ii = pfold(p, 0, i + a), jj = pfold(p, 1, j + b);
                        int kk = pfold(p, 2, k + cc);
                        rho[IX(p, ii, jj, kk)] += qt * sx[a] * sy[b] * sz[cc];
                    }
        }
    }
}

void orc_gauss(const orc_params *p, const orc_fields *f, const orc_particles *s, double out[2])
{
    size_t n = (size_t)p->mx * p->my * p->mz;
    double *div = (double *)malloc(sizeof(double) * n);
    double *rho = (double *)malloc(sizeof(double) * n);
    out[0] = out[1] = -1.0;
    if (!div || !rho) { free(div); free(rho); return; }
    orc_div_e(p, f, div);
    orc_rho(p, s, rho);
    double rmax = 0, amax = 0;
    int k0 = 1, k1 = 1;
    if (p->ndim == 3) { k0 = 3; k1 = p->mz - 3; }
    for (int k = k0; k <= k1; k++)
        for (int j = 3; j <= p->my - 3; j++)
            for (int i = 3; i <= p->mx - 3; i++) {
                int64_t o = IX(p, i, j, k);
                double r = fabs(div[o] - rho[o]);
                if (r > rmax) rmax = r;
                if (fabs(rho[o]) > amax) amax = fabs(rho[o]);
            }
    out[0] = rmax; out[1] = amax;
    free(div); free(rho);
}

/* sum of m * gamma*beta per component (m = massi for flavour 1, masse for 2) */
void orc_momentum(const orc_params *p, const orc_particles *s, double out[3])
{
    out[0] = out[1] = out[2] = 0;
    for (int64_t n = 0; n < s->used; n++) {
        int fl = s->flv[n];
        if (fl < 1 || fl > 2) continue;
        double m = (fl == 1) ? p->massi : p->masse;
        out[0] += m * s->u[n]; out[1] += m * s->v[n]; out[2] += m * s->w[n];
    }
}

/* ================================================================== */
/* Cylindrical geometry (the reference built with -Dcyl, inc_axis)     */
/* x = r, y = theta, z = z; one rank that holds the axis.              */
/* ================================================================== */

/* InitParam_cyl, cyl_common.F90:19-24, for one rank with xborders(0) = 0;
 * the Lehe-cyl coefficients of cyl_em_update_lehe.F90:6-17 with dtheta
 * evaluated at run time (grid_dx = grid_dz = 1). */
int orc_cyl_init(orc_cyl *cy, const orc_params *p)
{
    if (p->ny % 2 != 0) return ORC_ERR_ARG;   /* the axis maps j to j + ny/2 */
    memset(cy, 0, sizeof(*cy));
    const double pi = 3.14159265358979323846264;   /* vars.F90:224 */
    cy->dtheta = 2.0 * pi / p->ny;
    cy->ax_perm_area = 4.0 / p->ny;
    cy->inv_dtheta = 1.0 / cy->dtheta;
    cy->rshift = -0.5;
    cy->rmin = 0.0 + cy->rshift;               /* xborders(procxind) + rshift */
    cy->r_proc = 0.0 + cy->rshift - p->xmin;   /* cyl_movdep.F90:56 */
    cy->second_half_new_cell = 1;
    cy->substep_limit = 1000000;
    const double bxy = 0.0, bxz = 0.125, byx = 0.0, byz = 0.125;
    const double bzx = 0.125 * (1.0 / 1.0) * (1.0 / 1.0);
    const double bzy = 0.125 * (1.0 / (1.0 * cy->dtheta)) * (1.0 / (1.0 * cy->dtheta));
    const double dx = 0.0, dy = 0.0;
    const double sn = sin(0.5 * pi * p->c);
    const double dz = 0.25 * (1.0 - (p->cinv * p->cinv) * (sn * sn));
    cy->lehe[0] = 1.0 - 2.0 * bxy - 2.0 * bxz - 3.0 * dx;
    cy->lehe[1] = 1.0 - 2.0 * byx - 2.0 * byz - 3.0 * dy;
    cy->lehe[2] = 1.0 - 2.0 * bzx - 2.0 * bzy - 3.0 * dz;
    cy->lehe[3] = bxy; cy->lehe[4] = bxz; cy->lehe[5] = byx;
    cy->lehe[6] = byz; cy->lehe[7] = bzx; cy->lehe[8] = bzy;
    cy->lehe[9] = dx; cy->lehe[10] = dy; cy->lehe[11] = dz;
    return ORC_OK;
}

/* the grid point across the axis: j + ny/2 folded into the physical range */
static inline int cyl_jpp(const orc_params *p, int j)
{
    int jpp = j + p->ny / 2;
    if (jpp > p->my - 3) jpp = jpp - p->ny;
    return jpp;
}

/* UpdateEFldAxis, cyl_em_update.F90:185-225 */
static void cyl_e_axis(const orc_params *p, orc_fields *f)
{
    const int k1 = (p->ndim == 2) ? 1 : p->mz;
    for (int k = 1; k <= k1; k++)
        for (int j = 1; j <= p->my; j++) {
            const int jpp = cyl_jpp(p, j);
            f->Ex[IX(p, 3, j, k)] = 0.5 * (f->Ex[IX(p, 4, j, k)] - f->Ex[IX(p, 4, jpp, k)]);
            f->Ex[IX(p, 2, j, k)] = -f->Ex[IX(p, 4, jpp, k)];
            f->Ey[IX(p, 3, j, k)] = -f->Ey[IX(p, 4, jpp, k)];
            f->Ez[IX(p, 3, j, k)] = f->Ez[IX(p, 4, jpp, k)];
        }
}

/* UpdateBFldAxis, cyl_em_update.F90:227-260 */
static void cyl_b_axis(const orc_params *p, orc_fields *f)
{
    const int k1 = (p->ndim == 2) ? 1 : p->mz;
    for (int k = 1; k <= k1; k++)
        for (int j = 1; j <= p->my; j++) {
            const int jpp = cyl_jpp(p, j);
            f->By[IX(p, 3, j, k)] = 0.5 * (f->By[IX(p, 4, j, k)] - f->By[IX(p, 4, jpp, k)]);
            f->By[IX(p, 2, j, k)] = -f->By[IX(p, 4, jpp, k)];
            f->Bx[IX(p, 3, j, k)] = -f->Bx[IX(p, 4, jpp, k)];
            f->Bz[IX(p, 2, j, k)] = f->Bz[IX(p, 4, jpp, k)];
        }
}

/* UpdateEfield, cyl_em_update.F90:7-46 (axis rank: i from 4) */
void orc_cyl_update_e(const orc_params *p, const orc_cyl *cy, orc_fields *f)
{
    const int three = (p->ndim == 3);
    const int i1 = 4, k1 = three ? 3 : 1, k2 = three ? p->mz - 3 : 1;
    const double fldc = p->fldc, dtheta = cy->dtheta, rmin = cy->rmin;
    const double gdx = 1.0, gidx = p->inv_dx, gidz = p->inv_dz;
    double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez;
    const double *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    for (int k = k1; k <= k2; k++)
        for (int j = 3; j <= p->my - 3; j++)
            for (int i = i1; i <= p->mx - 3; i++) {
                const double r = ((i - 3.0) + rmin) * gdx;
                const double rp_half = r + 0.5 * gdx;
                const double rm_half = r - 0.5 * gdx;
                const int64_t o = IX(p, i, j, k);
                if (three) {
                    Ex[o] = Ex[o] + fldc * (Bz[o] - Bz[IX(p, i, j - 1, k)]) / (rp_half * dtheta)
                            - fldc * gidz * (By[o] - By[IX(p, i, j, k - 1)]);
                    Ey[o] = Ey[o] + fldc * gidz * (Bx[o] - Bx[IX(p, i, j, k - 1)])
                            - fldc * gidx * (Bz[o] - Bz[IX(p, i - 1, j, k)]);
                    Ez[o] = Ez[o] + fldc * gidx * (rp_half * By[o] - rm_half * By[IX(p, i - 1, j, k)]) / r
                            - fldc * (Bx[o] - Bx[IX(p, i, j - 1, k)]) / (r * dtheta);
                } else {
                    Ex[o] = Ex[o] + fldc * (Bz[o] - Bz[IX(p, i, j - 1, k)]) / (rp_half * dtheta);
                    Ey[o] = Ey[o] - fldc * gidx * (Bz[o] - Bz[IX(p, i - 1, j, k)]);
                    Ez[o] = Ez[o] + fldc * (gidx * (rp_half * By[o] - rm_half * By[IX(p, i - 1, j, k)])
                                            - (Bx[o] - Bx[IX(p, i, j - 1, k)]) / dtheta) / r;
                }
            }
    cyl_e_axis(p, f);
}

/* gradX, gradX1, gradY, gradZ of cyl_em_update_lehe.F90:255-289.  The
 * reference file ends gradX, gradX1 and gradY with a dangling continuation;
 * the terms it would add have zero coefficients there (beta_xy = delta_x =
 * beta_yx = delta_y = 0), so the sums below are complete. */
static double cyl_gradx(const orc_cyl *cy, const orc_params *p, const double *F, int i, int j, int k)
{
    const double *L = cy->lehe;
    return L[0] * (F[IX(p, i + 1, j, k)] - F[IX(p, i, j, k)])
           + L[4] * (F[IX(p, i + 1, j, k + 1)] - F[IX(p, i, j, k + 1)])
           + L[4] * (F[IX(p, i + 1, j, k - 1)] - F[IX(p, i, j, k - 1)]);
}

static double cyl_gradx1(const orc_cyl *cy, const orc_params *p, const double *F, int i, int j, int k, double rp,
                         double r)
{
    const double *L = cy->lehe;
    return L[0] * (rp * F[IX(p, i + 1, j, k)] - r * F[IX(p, i, j, k)])
           + L[4] * (rp * F[IX(p, i + 1, j, k + 1)] - r * F[IX(p, i, j, k + 1)])
           + L[4] * (rp * F[IX(p, i + 1, j, k - 1)] - r * F[IX(p, i, j, k - 1)]);
}

static double cyl_grady(const orc_cyl *cy, const orc_params *p, const double *F, int i, int j, int k)
{
    const double *L = cy->lehe;
    const double gidy = p->inv_dy;
    return L[1] * gidy * (F[IX(p, i, j + 1, k)] - F[IX(p, i, j, k)])
           + L[6] * gidy * (F[IX(p, i, j + 1, k + 1)] - F[IX(p, i, j, k + 1)])
           + L[6] * gidy * (F[IX(p, i, j + 1, k - 1)] - F[IX(p, i, j, k - 1)]);
}

static double cyl_gradz(const orc_cyl *cy, const orc_params *p, const double *F, int i, int j, int k)
{
    const double *L = cy->lehe;
    return L[2] * (F[IX(p, i, j, k + 1)] - F[IX(p, i, j, k)])
           + L[7] * (F[IX(p, i + 1, j, k + 1)] - F[IX(p, i + 1, j, k)])
           + L[7] * (F[IX(p, i - 1, j, k + 1)] - F[IX(p, i - 1, j, k)])
           + L[8] * (F[IX(p, i, j + 1, k + 1)] - F[IX(p, i, j + 1, k)])
           + L[8] * (F[IX(p, i, j - 1, k + 1)] - F[IX(p, i, j - 1, k)])
           + L[11] * (F[IX(p, i, j, k + 2)] - F[IX(p, i, j, k - 1)]);
}

/* UpdateBzAxis, cyl_em_update.F90:262-280: Faraday's law over the disk r < 1/2 */
static void cyl_bz_axis(const orc_params *p, const orc_cyl *cy, orc_fields *f)
{
    const int three = (p->ndim == 3);
    const int kt = three ? p->mz - 1 : 1;
    for (int k = 1; k <= kt; k++) {
        double ax = 0.0;
        for (int j = 3; j <= p->my - 3; j++) ax = ax - p->fld_halfc * cy->ax_perm_area * f->Ey[IX(p, 4, j, k)];
        ax = f->Bz[IX(p, 3, 3, k)] + ax;
        for (int j = 1; j <= p->my - 1; j++) f->Bz[IX(p, 3, j, k)] = ax;
    }
}

/* UpdateBfieldHalf, cyl_em_update.F90:49-110 (fdtd = 0) and
 * cyl_em_update_lehe.F90:62-118 (fdtd = 1, 3-D), axis rank */
void orc_cyl_update_b_half(const orc_params *p, const orc_cyl *cy, orc_fields *f)
{
    const int three = (p->ndim == 3);
    const int lehe = (p->fdtd == 1) && three;
    const int i1 = 4, k1 = 1, k2 = three ? p->mz - 1 : 1;
    const double h = p->fld_halfc, dtheta = cy->dtheta, rmin = cy->rmin;
    const double gdx = 1.0, gidx = p->inv_dx, gidz = p->inv_dz;
    double *Bx = f->Bx, *By = f->By, *Bz = f->Bz;
    const double *Ex = f->Ex, *Ey = f->Ey, *Ez = f->Ez;
    for (int k = k1; k <= k2; k++)
        for (int j = 1; j <= p->my - 1; j++)
            for (int i = i1; i <= p->mx - 1; i++) {
                const double r = ((i - 3.0) + rmin) * gdx;
                const double rp_half = r + 0.5 * gdx;
                const double rp = r + gdx;
                const int64_t o = IX(p, i, j, k);
                if (lehe) {
                    /* gradZ reaches j - 1, k - 1 and k + 2: the outermost ghost layers are
                     * skipped (the reference reads outside its arrays there; the halo
                     * exchange overwrites these layers) */
                    if (j < 2 || k < 2 || k > p->mz - 2) continue;
                    Bx[o] = Bx[o] - h * (cyl_grady(cy, p, Ez, i, j, k)) / (r * dtheta)
                            + h * gidz * (cyl_gradz(cy, p, Ey, i, j, k));
                    By[o] = By[o] - h * gidz * (cyl_gradz(cy, p, Ex, i, j, k)) + h * gidx * (cyl_gradx(cy, p, Ez, i, j, k));
                    Bz[o] = Bz[o] - h * (gidx * cyl_gradx1(cy, p, Ey, i, j, k, rp, r) - cyl_grady(cy, p, Ex, i, j, k) / dtheta)
                                        / rp_half;
                } else if (three) {
                    Bx[o] = Bx[o] - h * (Ez[IX(p, i, j + 1, k)] - Ez[o]) / (r * dtheta)
                            + h * gidz * (Ey[IX(p, i, j, k + 1)] - Ey[o]);
                    By[o] = By[o] - h * gidz * (Ex[IX(p, i, j, k + 1)] - Ex[o]) + h * gidx * (Ez[IX(p, i + 1, j, k)] - Ez[o]);
                    Bz[o] = Bz[o] - h * (gidx * (rp * Ey[IX(p, i + 1, j, k)] - r * Ey[o])
                                         - (Ex[IX(p, i, j + 1, k)] - Ex[o]) / dtheta) / rp_half;
                } else {
                    Bx[o] = Bx[o] - h * (Ez[IX(p, i, j + 1, k)] - Ez[o]) / (r * dtheta);
                    By[o] = By[o] + h * gidx * (Ez[IX(p, i + 1, j, k)] - Ez[o]);
                    Bz[o] = Bz[o] - h * (gidx * (rp * Ey[IX(p, i + 1, j, k)] - r * Ey[o])
                                         - (Ex[IX(p, i, j + 1, k)] - Ex[o]) / dtheta) / rp_half;
                }
            }
    cyl_bz_axis(p, cy, f);
    cyl_b_axis(p, f);
}

/* CalcEnergy with -Dcyl (savedata_routines.F90:265-281): CalcFldEnergy_cyl
 * (cyl/cyl_savedata_routines.F90:8-31, the axis rank: i from 4, r = (i - 3) + rmin
 * on each component's stagger, times dtheta), then CalcPrtlEnergy as orc_energy.
 * out: B energy, E energy, flavour 1, flavour 2. */
void orc_cyl_energy(const orc_params *p, const orc_cyl *cy, const orc_fields *f, const orc_particles *s, double out[4])
{
    double be = 0.0, ee = 0.0;
    const int k0 = p->ndim == 3 ? 3 : 1, k1 = p->ndim == 3 ? p->mz - 3 : 1;
    for (int k = k0; k <= k1; k++)
        for (int j = 3; j <= p->my - 3; j++)
            for (int i = 4; i <= p->mx - 3; i++) {
                const int64_t o = IX(p, i, j, k);
                const double r = (i - 3.0) + cy->rmin;
                const double rp = r + 0.5;
                be = be + (r * (f->Bx[o] * f->Bx[o]) + rp * (f->By[o] * f->By[o]) + rp * (f->Bz[o] * f->Bz[o])) * cy->dtheta;
                ee = ee + (rp * (f->Ex[o] * f->Ex[o]) + r * (f->Ey[o] * f->Ey[o]) + r * (f->Ez[o] * f->Ez[o])) * cy->dtheta;
            }
    double tmp[4];
    orc_energy(p, f, s, tmp);   /* its particle part is CalcPrtlEnergy */
    out[0] = be * 0.5;
    out[1] = ee * 0.5;
    out[2] = tmp[2];
    out[3] = tmp[3];
}

/* AddCurrent, cyl_em_update.F90:114-140 (RZtwoD = .false.) */
void orc_cyl_add_current(const orc_params *p, const orc_cyl *cy, orc_fields *f)
{
    const int three = (p->ndim == 3);
    const int i1 = 4, k1 = three ? 3 : 1, k2 = three ? p->mz - 3 : 1;
    const double dtheta = cy->dtheta, rmin = cy->rmin;
    const double gdx = 1.0, gidx = p->inv_dx, gidz = p->inv_dz;
    for (int k = k1; k <= k2; k++)
        for (int j = 3; j <= p->my - 3; j++)
            for (int i = i1; i <= p->mx - 3; i++) {
                const double r = ((i - 3.0) + rmin) * gdx;
                const double rp_half = r + 0.5 * gdx;
                const int64_t o = IX(p, i, j, k);
                f->Ex[o] = f->Ex[o] - f->Jx[o] * gidz / (rp_half * dtheta);
                f->Ey[o] = f->Ey[o] - f->Jy[o] * gidx * gidz;
                f->Ez[o] = f->Ez[o] - f->Jz[o] * gidx / (r * dtheta);
            }
    cyl_e_axis(p, f);
}

/* AxisCurrentBC, cyl_movdep.F90:451-465 */
void orc_cyl_axis_current_bc(const orc_params *p, orc_fields *f)
{
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++) {
            f->Jy[IX(p, 4, j, k)] = f->Jy[IX(p, 4, j, k)] - f->Jy[IX(p, 3, j, k)];
            f->Jz[IX(p, 4, j, k)] = f->Jz[IX(p, 4, j, k)] - f->Jz[IX(p, 3, j, k)];
        }
    for (int k = 1; k <= p->mz; k++)
        for (int j = 1; j <= p->my; j++) {
            f->Jx[IX(p, 3, j, k)] = 0.0;
            f->Jy[IX(p, 2, j, k)] = 0.0;
            f->Jz[IX(p, 2, j, k)] = 0.0;
        }
}

/* Field halos of the axis rank: theta (periodic) faces, then z faces (3-D),
 * as comm_fld.F90:75-91 with the x faces left out (the axis routines fill
 * the inner x ghosts); the outer x ghosts are held at zero, a conducting
 * wall at r = nx - 1/2 where the tangential E sits. */
void orc_cyl_field_halo(const orc_params *p, double *fx, double *fy, double *fz)
{
    double *F[3] = {fx, fy, fz};
    for (int c = 0; c < 3; c++) halo_copy_axis(p, F[c], 1);
    if (p->ndim == 3)
        for (int c = 0; c < 3; c++) halo_copy_axis(p, F[c], 2);
    for (int c = 0; c < 3; c++)
        for (int k = 1; k <= p->mz; k++)
            for (int j = 1; j <= p->my; j++)
                for (int i = p->mx - 2; i <= p->mx; i++) F[c][IX(p, i, j, k)] = 0.0;
}

/* Current halo of the axis rank: theta faces, then z faces (3-D) */
void orc_cyl_current_halo_add(const orc_params *p, double *jx, double *jy, double *jz)
{
    double *F[3] = {jx, jy, jz};
    for (int c = 0; c < 3; c++) halo_add_axis(p, F[c], 1);
    if (p->ndim == 3)
        for (int c = 0; c < 3; c++) halo_add_axis(p, F[c], 2);
}

/* DepositCurrentPIC, deposit.F90:11-103 */
static void cyl_deposit_pic(const orc_params *p, orc_fields *f, double x0, double y0, double z0, double x, double y,
                            double z, double q)
{
    const double qthis = q * p->qi;
    int i1 = (int)floor(x0), i2 = (int)floor(x);
    int j1 = (int)floor(y0), j2 = (int)floor(y);
    int k1 = (int)floor(z0), k2 = (int)floor(z);
    const double xr = dmin((double)(imin2(i1, i2) + 1), dmax((double)imax2(i1, i2), 0.5 * (x0 + x)));
    const double yr = dmin((double)(imin2(j1, j2) + 1), dmax((double)imax2(j1, j2), 0.5 * (y0 + y)));
    const double zr = dmin((double)(imin2(k1, k2) + 1), dmax((double)imax2(k1, k2), 0.5 * (z0 + z)));
    const double Fx1 = qthis * (xr - x0), Fy1 = qthis * (yr - y0), Fz1 = qthis * (zr - z0);
    const double Wx1 = 0.5 * (x0 + xr) - i1, Wy1 = 0.5 * (y0 + yr) - j1;
    const double Wx2 = 0.5 * (x + xr) - i2, Wy2 = 0.5 * (y + yr) - j2;
    double Wz1, Wz2;
    const int three = (p->ndim == 3);
    if (!three) { k1 = 1; k2 = 1; Wz1 = 0.0; Wz2 = 0.0; }
    else { Wz1 = 0.5 * (z0 + zr) - k1; Wz2 = 0.5 * (z + zr) - k2; }
    const double Fx2 = qthis * (x - xr), Fy2 = qthis * (y - yr), Fz2 = qthis * (z - zr);
    double *Jx = f->Jx, *Jy = f->Jy, *Jz = f->Jz;
#define JA(A, i, j, k, v) A[IX(p, i, j, k)] = A[IX(p, i, j, k)] + (v)
    JA(Jx, i1, j1, k1, Fx1 * (1.0 - Wy1) * (1.0 - Wz1));
    JA(Jx, i1, j1 + 1, k1, Fx1 * Wy1 * (1.0 - Wz1));
    if (three) {
        JA(Jx, i1, j1, k1 + 1, Fx1 * (1.0 - Wy1) * Wz1);
        JA(Jx, i1, j1 + 1, k1 + 1, Fx1 * Wy1 * Wz1);
    }
    JA(Jx, i2, j2, k2, Fx2 * (1.0 - Wy2) * (1.0 - Wz2));
    JA(Jx, i2, j2 + 1, k2, Fx2 * Wy2 * (1.0 - Wz2));
    if (three) {
        JA(Jx, i2, j2, k2 + 1, Fx2 * (1.0 - Wy2) * Wz2);
        JA(Jx, i2, j2 + 1, k2 + 1, Fx2 * Wy2 * Wz2);
    }
    JA(Jy, i1, j1, k1, Fy1 * (1.0 - Wx1) * (1.0 - Wz1));
    JA(Jy, i1 + 1, j1, k1, Fy1 * Wx1 * (1.0 - Wz1));
    if (three) {
        JA(Jy, i1, j1, k1 + 1, Fy1 * (1.0 - Wx1) * Wz1);
        JA(Jy, i1 + 1, j1, k1 + 1, Fy1 * Wx1 * Wz1);
    }
    JA(Jy, i2, j2, k2, Fy2 * (1.0 - Wx2) * (1.0 - Wz2));
    JA(Jy, i2 + 1, j2, k2, Fy2 * Wx2 * (1.0 - Wz2));
    if (three) {
        JA(Jy, i2, j2, k2 + 1, Fy2 * (1.0 - Wx2) * Wz2);
        JA(Jy, i2 + 1, j2, k2 + 1, Fy2 * Wx2 * Wz2);
    }
    JA(Jz, i1, j1, k1, Fz1 * (1.0 - Wx1) * (1.0 - Wy1));
    JA(Jz, i1 + 1, j1, k1, Fz1 * Wx1 * (1.0 - Wy1));
    JA(Jz, i1, j1 + 1, k1, Fz1 * (1.0 - Wx1) * Wy1);
    JA(Jz, i1 + 1, j1 + 1, k1, Fz1 * Wx1 * Wy1);
    JA(Jz, i2, j2, k2, Fz2 * (1.0 - Wx2) * (1.0 - Wy2));
    JA(Jz, i2 + 1, j2, k2, Fz2 * Wx2 * (1.0 - Wy2));
    JA(Jz, i2, j2 + 1, k2, Fz2 * (1.0 - Wx2) * Wy2);
    JA(Jz, i2 + 1, j2 + 1, k2, Fz2 * Wx2 * Wy2);
#undef JA
}

/* Interp_VecEM_GridPoints, interpolation.F90:434-490 */
static void cyl_interp(const orc_params *p, const orc_ws *w, double x, double y, double z, double e[3], double b[3])
{
    const int two = (p->ndim == 2);
    const int ip0 = (int)x, jp0 = (int)y;
    const double dx = x - ip0, dy = y - jp0;
    double wt[8];
    int64_t sidx;
    if (!two) {
        const int kp0 = (int)z;
        const double dz = z - kp0;
        for (int nn = 0; nn < 8; nn++)
            wt[nn] = (wtx1_3[nn] + wtx2_3[nn] * dx) * (wty1_3[nn] + wty2_3[nn] * dy) * (wtz1_3[nn] + wtz2_3[nn] * dz);
        sidx = (int64_t)(kp0 - 1) * p->mx * p->my + (int64_t)(jp0 - 1) * p->mx + ip0;
    } else {
        for (int nn = 0; nn < 4; nn++) wt[nn] = (wtx1_2[nn] + wtx2_2[nn] * dx) * (wty1_2[nn] + wty2_2[nn] * dy);
        sidx = (int64_t)(jp0 - 1) * p->mx + ip0;
    }
    const int nv = w->nv;
    const int64_t bo = nv * (sidx - 1);
    for (int c = 0; c < 3; c++) { e[c] = 0.0; b[c] = 0.0; }
    for (int nn = 0; nn < nv; nn++) {
        e[0] = e[0] + wt[nn] * w->vec[0][bo + nn];
        e[1] = e[1] + wt[nn] * w->vec[1][bo + nn];
        e[2] = e[2] + wt[nn] * w->vec[2][bo + nn];
        b[0] = b[0] + wt[nn] * w->vec[3][bo + nn];
        b[1] = b[1] + wt[nn] * w->vec[4][bo + nn];
        b[2] = b[2] + wt[nn] * w->vec[5][bo + nn];
    }
}

/* MovePrtlAxis, cyl_movdep.F90:381-446: one straight sub-step of length dt;
 * returns 0 when the azimuthal move exceeds 0.45 cells (the caller halves dt) */
static int cyl_move_axis(const orc_params *p, const orc_cyl *cy, const orc_ws *w, orc_particles *s, int64_t n,
                         double x, double y, double z, double dt)
{
    double e[3], b[3];
    cyl_interp(p, w, x, y, z, e, b);
    const double qm = (s->flv[n] == 0) ? 0.0 : p->flvrqm[s->flv[n] - 1] * 0.5;
    double pEx = e[0] * qm * dt, pEy = e[1] * qm * dt, pEz = e[2] * qm * dt;
    double pBx = b[0] * qm * dt, pBy = b[1] * qm * dt, pBz = b[2] * qm * dt;
    const double c = p->c, sqc = p->sqc, cinv = p->cinv;
    double u0 = c * s->u[n] + pEx;
    double v0 = c * s->v[n] + pEy;
    double w0 = c * s->w[n] + pEz;
    double g = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
    pBx = g * pBx; pBy = g * pBy; pBz = g * pBz;
    const double f = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
    double u1 = (u0 + v0 * pBz - w0 * pBy) * f;
    double v1 = (v0 + w0 * pBx - u0 * pBz) * f;
    const double w1 = (w0 + u0 * pBy - v0 * pBx) * f;
    u0 = u0 + v1 * pBz - w1 * pBy + pEx;
    v0 = v0 + w1 * pBx - u1 * pBz + pEy;
    w0 = w0 + u1 * pBy - v1 * pBx + pEz;
    u1 = u0 * cinv;
    v1 = v0 * cinv;
    g = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
    const double x1 = (x - 3.5) * 1.0 + u1 * g * dt;
    const double y1 = v1 * g * dt;
    const double r1 = sqrt(x1 * x1 + y1 * y1);
    double dy;
    if (x1 > 0) dy = atan(y1 / x1) * cy->inv_dtheta;
    else dy = 0.0;
    if (fabs(dy) > 0.45) return 0;
    if (x1 > 0) {
        s->x[n] = r1 * p->inv_dx + 3.5;
        s->u[n] = (x1 * u1 + y1 * v1) / r1;
        s->v[n] = (-y1 * u1 + x1 * v1) / r1;
    } else {
        s->x[n] = x1 * p->inv_dx + 3.5;
        s->u[n] = u1;
        s->v[n] = v1;
    }
    s->y[n] = s->y[n] + dy;
    s->w[n] = w0 * cinv;
    s->z[n] = s->z[n] + s->w[n] * g * dt * p->inv_dz;
    return 1;
}

/* Boris with the half-kick factors pE, pB (already times q/m and dt) on the
 * c-scaled momentum (u0, v0, w0), in place (cyl_movdep_strang.F90:172-185);
 * returns the factor 1 / (c gamma) it scaled pB by */
static double boris_cscaled(const orc_params *p, double pEx, double pEy, double pEz, double pBx, double pBy, double pBz,
                          double *u0, double *v0, double *w0)
{
    const double sqc = p->sqc;
    double U = *u0 + pEx, V = *v0 + pEy, W = *w0 + pEz;
    const double g = 1.0 / sqrt(sqc + U * U + V * V + W * W);
    pBx = g * pBx; pBy = g * pBy; pBz = g * pBz;
    const double f = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
    const double u1 = (U + V * pBz - W * pBy) * f;
    const double v1 = (V + W * pBx - U * pBz) * f;
    const double w1 = (W + U * pBy - V * pBx) * f;
    *u0 = U + v1 * pBz - w1 * pBy + pEx;
    *v0 = V + w1 * pBx - u1 * pBz + pEy;
    *w0 = W + u1 * pBy - v1 * pBx + pEz;
    return g;
}

/* MovePrtlAxis, cyl_movdep_strang.F90:462-566: half kick with the fields at the
 * start, the straight move of length dt, the rotation of (u, v) into the new local
 * frame, half kick with the fields at the new point.  push = 2 keeps the written
 * x1 <= 0 branch, which leaves (u, v) divided by c before the second kick. */
static int cyl_move_axis_strang(const orc_params *p, const orc_cyl *cy, const orc_ws *w, orc_particles *s, int64_t n,
                                double x, double y, double z, double dt)
{
    double e[3], b[3];
    cyl_interp(p, w, x, y, z, e, b);
    const double qm = (s->flv[n] == 0) ? 0.0 : p->flvrqm[s->flv[n] - 1] * 0.25;
    const double c = p->c, sqc = p->sqc, cinv = p->cinv;
    double u0 = c * s->u[n], v0 = c * s->v[n], w0 = c * s->w[n];
    boris_cscaled(p, e[0] * qm * dt, e[1] * qm * dt, e[2] * qm * dt, b[0] * qm * dt, b[1] * qm * dt, b[2] * qm * dt,
                  &u0, &v0, &w0);
    double u1 = u0 * cinv, v1 = v0 * cinv;
    const double w1 = w0 * cinv;
    const double g = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
    const double r0 = (x - 3.5) * 1.0;
    const double x1 = r0 + u1 * g * dt;
    const double y1 = v1 * g * dt;
    const double r1 = sqrt(x1 * x1 + y1 * y1);
    double dy;
    if (x1 > 0) dy = atan(y1 / x1) * cy->inv_dtheta;
    else dy = 0.0;
    if (fabs(dy) > 0.45) return 0;
    if (x1 > 0) {
        s->x[n] = r1 / p->inv_dx + 3.5;
        u1 = (r0 * u0 + (u0 * u1 * g * dt + v0 * y1)) / r1;
        v1 = (v0 * r0) / r1;
    } else {
        s->x[n] = x1 / p->inv_dx + 3.5;
        if (cy->push != 2) { u1 = u0; v1 = v0; }
    }
    s->y[n] = s->y[n] + dy;
    s->z[n] = s->z[n] + w1 * g * dt * p->inv_dz;
    cyl_interp(p, w, s->x[n], s->y[n], s->z[n], e, b);
    u0 = u1; v0 = v1;
    boris_cscaled(p, e[0] * qm * dt, e[1] * qm * dt, e[2] * qm * dt, b[0] * qm * dt, b[1] * qm * dt, b[2] * qm * dt,
                  &u0, &v0, &w0);
    s->u[n] = u0 * cinv;
    s->v[n] = v0 * cinv;
    s->w[n] = w0 * cinv;
    return 1;
}

/* PrtlAxisBC, cyl_movdep.F90:366-379 */
static void cyl_axis_bc(const orc_params *p, orc_particles *s, int64_t n)
{
    if (s->x[n] < 3.5) {
        s->x[n] = 3.5 + (3.5 - s->x[n]);
        s->y[n] = s->y[n] + p->ny / 2.0;
        s->u[n] = -s->u[n];
        s->v[n] = -s->v[n];
    }
    if (s->y[n] > p->ymax) s->y[n] = s->y[n] - p->ylen;
    if (s->y[n] < p->ymin) s->y[n] = s->y[n] + p->ylen;
}

/* SubStepMovDep, cyl_movdep.F90:315-349; returns the number of sub-steps or -1 past the limit */
static int64_t cyl_substep(const orc_params *p, const orc_cyl *cy, const orc_ws *w, orc_fields *f, orc_particles *s,
                           int64_t n, double x0, double y0, double z0, double u0, double v0, double w0)
{
    if (s->flv[n] == 0) return 0;
    s->x[n] = x0; s->y[n] = y0; s->z[n] = z0;
    s->u[n] = u0; s->v[n] = v0; s->w[n] = w0;
    double t = 0.0, dt = 1.0;
    int64_t tries = 0;
    while (t < 1.0) {
        if (++tries > cy->substep_limit) return -1;
        double delt = dt;
        if (t + dt > 1.0) delt = 1.0 - t;
        const int res = cy->push ? cyl_move_axis_strang(p, cy, w, s, n, x0, y0, z0, delt)
                                 : cyl_move_axis(p, cy, w, s, n, x0, y0, z0, delt);
        if (res == 1) {
            t = t + delt;
            cyl_deposit_pic(p, f, x0, y0, z0, s->x[n], s->y[n], s->z[n], s->q[n]);
            cyl_axis_bc(p, s, n);
            dt = dt * (cy->push ? 1.7 : 1.1);   /* cyl_movdep_strang.F90:427 */
            x0 = s->x[n]; y0 = s->y[n]; z0 = s->z[n];
        } else {
            dt = dt * 0.5;
        }
    }
    return tries;
}

/* MoveDepositPrtl, cyl_movdep.F90:14-289, one rank holding the axis.
 * second_half_new_cell = 1 deposits the second zigzag half-segment at the
 * particle's new cell (charge-conserving); 0 keeps the reference's
 * short_arr_ind of (ip0, jp0) for it (cyl_movdep.F90:249,252).
 * Returns ORC_ERR_ARG if a particle reaches the outer two cells or a
 * sub-step loop passes substep_limit. */
int orc_cyl_move_deposit(const orc_params *p, orc_cyl *cy, orc_ws *w, orc_fields *f, orc_particles *s)
{
    const int two = (p->ndim == 2);
    const int nj = w->nj;
    const int mx = p->mx, my = p->my;
    const int64_t used = s->used, ncell = w->ncell;
    const double c = p->c, sqc = p->sqc, cinv = p->cinv, qi = p->qi;
    const double *wtFx = two ? wtFx_2 : wtFx_3, *wtFy = two ? wtFy_2 : wtFy_3, *wtFz = two ? wtFz_2 : wtFz_3;
    const double *Jwtx1 = two ? Jwtx1_2 : Jwtx1_3, *Jwtx2 = two ? Jwtx2_2 : Jwtx2_3;
    const double *Jwty1 = two ? Jwty1_2 : Jwty1_3, *Jwty2 = two ? Jwty2_2 : Jwty2_3;
    const double *Jwtz1 = two ? Jwtz1_2 : Jwtz1_3, *Jwtz2 = two ? Jwtz2_2 : Jwtz2_3;
    const double r_proc = cy->r_proc;
    int rc = ORC_OK;
    cy->substeps = 0;

    orc_gather_vec_e(p, w, f);
    orc_gather_vec_b(p, w, f);
    memset(w->vecj, 0, sizeof(double) * (size_t)(nj * ncell * w->nthreads));
    double *VJ = w->vecj;   /* one thread: the sub-steps deposit into J directly */

    for (int64_t off = 0; off < used; off += VBS) {
        const int nb = (int)((used - off < VBS) ? (used - off) : VBS);
        double x0[VBS], y0[VBS], z0[VBS], qthis[VBS], ui[VBS], vi[VBS], wi[VBS];
        int ip0[VBS], jp0[VBS], kp0[VBS];
        for (int n = 0; n < nb; n++) {
            const int64_t m = off + n;
            x0[n] = s->x[m]; y0[n] = s->y[m]; z0[n] = s->z[m];
            qthis[n] = s->q[m] * qi;
            ui[n] = s->u[m]; vi[n] = s->v[m]; wi[n] = s->w[m];
        }
        for (int n = 0; n < nb && cy->push; n++) {
            /* MoveDepositPrtl of cyl_movdep_strang.F90:95-294 (Strang splitting, Delzanno
             * et al. 2013): half kick with the fields at the old point (q/m / 4), the move
             * with the half-kicked momentum, the rotation into the new local frame, half
             * kick with the fields at the new point.  As written (push = 2) the second
             * gather adds onto the first one's scaled fields and the z move uses the
             * momentum before the kick; push = 1 follows MovePrtlAxis of the same file,
             * which resets the gather and moves z with the half-kicked momentum. */
            const int64_t m = off + n;
            const double qm = (s->flv[m] == 0) ? 0.0 : p->flvrqm[s->flv[m] - 1] * 0.25;
            double e[3], b[3];
            ip0[n] = (int)x0[n];
            jp0[n] = (int)y0[n];
            kp0[n] = two ? 1 : (int)z0[n];
            cyl_interp(p, w, x0[n], y0[n], z0[n], e, b);
            const double pE1[3] = {e[0] * qm, e[1] * qm, e[2] * qm};
            const double pB1[3] = {b[0] * qm, b[1] * qm, b[2] * qm};
            double u0 = c * s->u[m], v0 = c * s->v[m], w0 = c * s->w[m];
            const double g1 = boris_cscaled(p, pE1[0], pE1[1], pE1[2], pB1[0], pB1[1], pB1[2], &u0, &v0, &w0);
            double u1 = u0 * cinv;
            const double v1 = v0 * cinv, w1 = w0 * cinv;
            const double g = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
            const double r0 = (s->x[m] + r_proc) * 1.0;
            const double x1 = r0 + u1 * g;
            const double y1 = v1 * g;
            double r1 = sqrt(x1 * x1 + y1 * y1);
            s->x[m] = r1 * p->inv_dx - r_proc;
            s->y[m] = s->y[m] + atan(y1 / x1) * cy->inv_dtheta;
            s->z[m] = s->z[m] + (cy->push == 2 ? s->w[m] : w1) * g * p->inv_dz;
            r1 = 1.0 / r1;
            u1 = (r0 * u0 + (u0 * u1 * g + v0 * y1)) * r1;
            const double v2 = (v0 * r0) * r1;
            if (x0[n] < 4.5) continue;   /* redone by the axis sub-steps */
            cyl_interp(p, w, s->x[m], s->y[m], s->z[m], e, b);
            double pE2[3] = {e[0], e[1], e[2]}, pB2[3] = {b[0], b[1], b[2]};
            if (cy->push == 2)   /* the sums start from the first half's pE and its gamma-scaled pB */
                for (int a = 0; a < 3; a++) {
                    pE2[a] = pE1[a] + e[a];
                    pB2[a] = g1 * pB1[a] + b[a];
                }
            u0 = u1; v0 = v2;
            boris_cscaled(p, pE2[0] * qm, pE2[1] * qm, pE2[2] * qm, pB2[0] * qm, pB2[1] * qm, pB2[2] * qm, &u0, &v0,
                          &w0);
            s->u[m] = u0 * cinv;
            s->v[m] = v0 * cinv;
            s->w[m] = w0 * cinv;
        }
        for (int n = 0; n < nb && !cy->push; n++) {
            const int64_t m = off + n;
            const double qm = (s->flv[m] == 0) ? 0.0 : p->flvrqm[s->flv[m] - 1] * 0.5;
            double e[3], b[3];
            ip0[n] = (int)x0[n];
            jp0[n] = (int)y0[n];
            kp0[n] = two ? 1 : (int)z0[n];
            cyl_interp(p, w, x0[n], y0[n], z0[n], e, b);
            double pEx = e[0] * qm, pEy = e[1] * qm, pEz = e[2] * qm;
            double pBx = b[0] * qm, pBy = b[1] * qm, pBz = b[2] * qm;
            double u0 = c * s->u[m] + pEx;
            double v0 = c * s->v[m] + pEy;
            double w0 = c * s->w[m] + pEz;
            double g = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
            pBx = g * pBx; pBy = g * pBy; pBz = g * pBz;
            const double ff = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
            double u1 = (u0 + v0 * pBz - w0 * pBy) * ff;
            double v1 = (v0 + w0 * pBx - u0 * pBz) * ff;
            const double w1 = (w0 + u0 * pBy - v0 * pBx) * ff;
            u0 = u0 + v1 * pBz - w1 * pBy + pEx;
            v0 = v0 + w1 * pBx - u1 * pBz + pEy;
            w0 = w0 + u1 * pBy - v1 * pBx + pEz;
            u1 = u0 * cinv;
            v1 = v0 * cinv;
            s->w[m] = w0 * cinv;
            g = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
            const double x1 = (s->x[m] + r_proc) * 1.0 + u1 * g;
            const double y1 = v1 * g;
            double r1 = sqrt(x1 * x1 + y1 * y1);
            s->x[m] = r1 * p->inv_dx - r_proc;
            s->y[m] = s->y[m] + atan(y1 / x1) * cy->inv_dtheta;
            s->z[m] = s->z[m] + s->w[m] * g * p->inv_dz;
            r1 = 1.0 / r1;
            s->u[m] = (x1 * u1 + y1 * v1) * r1;
            s->v[m] = (-y1 * u1 + x1 * v1) * r1;
        }
        /* MovDepAxis, cyl_movdep.F90:300-312 */
        for (int n = 0; n < nb && rc == ORC_OK; n++)
            if (x0[n] < 4.5) {
                const int64_t k = cyl_substep(p, cy, w, f, s, off + n, x0[n], y0[n], z0[n], ui[n], vi[n], wi[n]);
                if (k < 0) rc = ORC_ERR_ARG;
                else cy->substeps += k;
                qthis[n] = 0.0;
            }
        for (int n = 0; n < nb && rc == ORC_OK; n++) {
            const int64_t m = off + n;
            if (s->x[m] >= mx - 3) { rc = ORC_ERR_ARG; break; }   /* the outer wall region */
            const int ip = (int)s->x[m], jp = (int)s->y[m];
            /* near the axis one step can cross several theta cells (up to ny/4):
             * the second half-segment's cell row is folded by ny into the arrays,
             * a periodic image of the same nodes */
            const int jq = (jp < 1) ? jp + p->ny : (jp > my - 1 ? jp - p->ny : jp);
            const int kp = two ? 1 : (int)s->z[m];
            const double xr = dmin((double)(imin2(ip0[n], ip) + 1), dmax((double)imax2(ip0[n], ip), 0.5 * (x0[n] + s->x[m])));
            const double yr = dmin((double)(imin2(jp0[n], jp) + 1), dmax((double)imax2(jp0[n], jp), 0.5 * (y0[n] + s->y[m])));
            const double zr = dmin((double)(imin2(kp0[n], kp) + 1), dmax((double)imax2(kp0[n], kp), 0.5 * (z0[n] + s->z[m])));
            const double Fx1 = qthis[n] * (xr - x0[n]);
            const double Fy1 = qthis[n] * (yr - y0[n]);
            const double Fz1 = qthis[n] * (zr - z0[n]);
            const double Wx1 = 0.5 * (x0[n] + xr) - ip0[n];
            const double Wy1 = 0.5 * (y0[n] + yr) - jp0[n];
            const double Wz1 = two ? 0.0 : 0.5 * (z0[n] + zr) - kp0[n];
            int64_t sidx = two ? (int64_t)(jp0[n] - 1) * mx + ip0[n]
                               : (int64_t)(kp0[n] - 1) * mx * my + (int64_t)(jp0[n] - 1) * mx + ip0[n];
            double *vj = VJ + nj * (sidx - 1);
            for (int nn = 0; nn < nj; nn++)
                vj[nn] = vj[nn] + (wtFx[nn] * Fx1 + wtFy[nn] * Fy1 + wtFz[nn] * Fz1) * (Jwtx1[nn] + Jwtx2[nn] * Wx1)
                                  * (Jwty1[nn] + Jwty2[nn] * Wy1) * (Jwtz1[nn] + Jwtz2[nn] * Wz1);
            const double Fx2 = qthis[n] * (s->x[m] - xr);
            const double Fy2 = qthis[n] * (s->y[m] - yr);
            const double Fz2 = qthis[n] * (s->z[m] - zr);
            const double Wx2 = 0.5 * (s->x[m] + xr) - ip;
            const double Wy2 = 0.5 * (s->y[m] + yr) - jp;
            const double Wz2 = two ? 0.0 : 0.5 * (s->z[m] + zr) - kp;
            if (cy->second_half_new_cell)
                sidx = two ? (int64_t)(jq - 1) * mx + ip : (int64_t)(kp - 1) * mx * my + (int64_t)(jq - 1) * mx + ip;
            vj = VJ + nj * (sidx - 1);
            for (int nn = 0; nn < nj; nn++)
                vj[nn] = vj[nn] + (wtFx[nn] * Fx2 + wtFy[nn] * Fy2 + wtFz[nn] * Fz2) * (Jwtx1[nn] + Jwtx2[nn] * Wx2)
                                  * (Jwty1[nn] + Jwty2[nn] * Wy2) * (Jwtz1[nn] + Jwtz2[nn] * Wz2);
        }
        if (two)
            for (int n = 0; n < nb; n++) {
                const int64_t m = off + n;
                if (s->z[m] > p->zmax) s->z[m] = s->z[m] - p->zlen;
                else if (s->z[m] < p->zmin) s->z[m] = p->zlen + s->z[m];
            }
        if (rc != ORC_OK) break;
    }
    if (rc != ORC_OK) return rc;
    reduce_current(p, w, f);
    orc_cyl_axis_current_bc(p, f);
    return ORC_OK;
}

/* ExchangePrtl on the axis rank: periodic wrap in theta and (3-D) z, in
 * place (the reference removes and appends; the order of particles is not
 * part of the result) */
void orc_cyl_exchange_particles(const orc_params *p, orc_particles *s)
{
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        if (s->y[n] < p->ymin) s->y[n] = s->y[n] + p->ylen;
        else if (s->y[n] > p->ymax) s->y[n] = s->y[n] - p->ylen;
        if (p->ndim == 3) {
            if (s->z[n] < p->zmin) s->z[n] = s->z[n] + p->zlen;
            else if (s->z[n] > p->zmax) s->z[n] = s->z[n] - p->zlen;
        }
    }
}

/* One time step in the MAIN.F90:42-284 order, cylindrical build */
int orc_cyl_step(const orc_params *p, orc_cyl *cy, orc_ws *w, orc_fields *f, orc_particles *s)
{
    orc_reset_current(p, f);
    orc_cyl_update_b_half(p, cy, f);
    orc_cyl_field_halo(p, f->Bx, f->By, f->Bz);
    int rc = orc_cyl_move_deposit(p, cy, w, f, s);
    if (rc != ORC_OK) return rc;
    orc_cyl_exchange_particles(p, s);
    orc_cyl_update_b_half(p, cy, f);
    orc_cyl_field_halo(p, f->Bx, f->By, f->Bz);
    orc_cyl_update_e(p, cy, f);
    orc_cyl_field_halo(p, f->Ex, f->Ey, f->Ez);
    orc_cyl_current_halo_add(p, f->Jx, f->Jy, f->Jz);
    orc_cyl_add_current(p, cy, f);
    orc_cyl_field_halo(p, f->Ex, f->Ey, f->Ez);
    return ORC_OK;
}
