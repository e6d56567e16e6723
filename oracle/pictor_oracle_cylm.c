/*
 * pictor_oracle_cylm.c -- CPU FP64 statement of the azimuthal-mode (m = 0, 1)
 * cylindrical step.  TEST INFRASTRUCTURE ONLY (see pictor_oracle.h).
 *
 * The reference has no modal solver: its cylindrical build is a full r-theta-z
 * grid (src/cyl/cyl_em_update.F90, cyl_movdep.F90).  This form keeps that grid's
 * r and z staggering, metric factors, axis rules and push, and replaces the
 * theta grid by F(r, theta, z) = F0 + F1c cos(theta) + F1s sin(theta):
 *   * every theta difference (F(j) - F(j-1)) / dtheta of cyl_em_update.F90
 *     becomes the exact derivative, (dF/dtheta)_c = F1s, (dF/dtheta)_s = -F1c;
 *   * the axis fills (:143-246), which map j to j + ny/2 (theta + pi), become
 *     the parity (-1)^m of each mode; the Bz axis integral acts on m = 0 only;
 *   * AddCurrent divides by the whole circle: 2 pi for m = 0, pi for m = 1;
 *   * AxisCurrentBC (cyl_movdep.F90:451-465) acts at fixed theta, so on every
 *     mode alike;
 *   * the push is cyl_movdep.F90's local-frame Boris step and atan rotation;
 *     the deposit splits the move into the r-z zigzag at the old angle and the
 *     angular move at the new (r, z), which keeps the modal continuity
 *     equation exact (rho_m + div_m J_m node by node).
 * Lengths are in axial cells (dz = 1); the radial cell is dr.  Arrays
 * (mx, 3, mz), column-major, 1-based i and k, slot s = 0 (m = 0), 1 (cos),
 * 2 (sin): element (i, s, k) at (i - 1) + mx (s + 3 (k - 1)).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pictor_oracle.h"

#define PI_D 3.14159265358979323846

static inline long long MI(const orc_cylm *m, int i, int s, int k)
{
    return (long long)(i - 1) + (long long)m->mx * (s + 3LL * (k - 1));
}

int orc_cylm_init(orc_cylm *m, int nr, int nz, double dr, double c, double c_ompe, double g0, double mi_me, int epc)
{
    if (nr < 8 || nz < 8 || dr <= 0.0 || c <= 0.0 || epc < 1) return ORC_ERR_ARG;
    memset(m, 0, sizeof(*m));
    m->nr = nr; m->nz = nz; m->mx = nr + 5; m->mz = nz + 5;
    m->dr = dr; m->inv_dr = 1.0 / dr; m->inv_dz = 1.0;
    m->c = c; m->fldc = c; m->fld_halfc = 0.5 * c;
    m->cinv = 1.0 / c; m->sqc = c * c;
    /* InitScale, initialise.F90:44-53, with epc macro-particle weight per unit volume */
    const double me = 1.0, mi = mi_me * me;
    m->ompe = c / c_ompe;
    m->qi = g0 * m->ompe * m->ompe / ((double)epc * (1.0 + me / mi));
    m->flvrqm[0] = 1.0 / mi;   /* flavour 1: ions */
    m->flvrqm[1] = -1.0;       /* flavour 2: electrons */
    m->zmin = 3.0; m->zmax = 3.0 + nz; m->zlen = nz;
    return ORC_OK;
}

/* The Lehe B update per mode (cyl_em_update_lehe.F90:6-17, 62-118, 255-289).  The
 * reference averages the r and theta differences along z with beta_xz = beta_yz = 1/8,
 * and its z difference is
 *   (1 - 2 b_zx - 2 b_zy - 3 d_z) dF + b_zx (dF(i+1) + dF(i-1)) + b_zy (dF(j+1) + dF(j-1))
 *   + d_z (F(k+2) - F(k-1)),  b_zx = (dz/dr)^2 / 8, b_zy = 1 / (8 dtheta^2),
 *   d_z = (1 - sin^2(pi c / 2) / c^2) / 4.
 * The modal form keeps beta_rz = beta_thetaz = 1/8 and d_z and sets b_zx = 0 and the
 * theta pair to 0 (l_zr = l_zm = 0):
 *   * the r pair F(i+1) - 2 F(i) + F(i-1) reads the axis fills at i = 4 and is not
 *     symmetric in the r-weighted energy norm: with it the fields next to the axis
 *     grow without bound at the z Nyquist wavenumber, on the reference's full grid
 *     as on the modes (b_zx = 1/128 at dr = 4 dz);
 *   * b_zy (2 cos(m dtheta) - 2) tends to -m^2 / 8 at every radius (the Cartesian
 *     coefficient applied to the theta index), which would slow an m = 1 pulse.
 * l_zr and l_zm stay as fields: setting them restores the reference's terms. */
int orc_cylm_set_fdtd(orc_cylm *m, int fdtd)
{
    if (fdtd != 0 && fdtd != 1) return ORC_ERR_ARG;
    m->fdtd = fdtd;
    const double pi = 3.14159265358979323846264;
    const double brz = 0.125, btz = 0.125;
    const double sn = sin(0.5 * pi * m->c);
    const double dz = 0.25 * (1.0 - (m->cinv * m->cinv) * (sn * sn));
    m->l_rr = 1.0 - 2.0 * brz; m->l_rz = brz;
    m->l_tt = 1.0 - 2.0 * btz; m->l_tz = btz;
    m->l_zz = 1.0 - 3.0 * dz; m->l_zr = 0.0; m->l_zd = dz;
    m->l_zm = 0.0;
    return ORC_OK;
}

/* radius of grid index i: r(i) = (i - 3.5) dr (cyl_common.F90:23 with rmin = -dr/2) */
static inline double rad(const orc_cylm *m, double i) { return (i - 3.5) * m->dr; }

/* ---------------------------------------------------------------- fields */

/* z ghosts: 1..2 <- nz+1..nz+2, mz-2..mz <- 3..5 (periodic z), over x = 3..mx-3
 * as the reference's z faces (subdomains.F90:592-680, SetCommFldDomain).  Open z
 * (the moving window): the ghosts hold no field, so nothing from the back of the
 * box reappears at its front. */
static void cylm_zhalo(const orc_cylm *m, double *F)
{
    for (int s = 0; s < 3; s++)
        for (int i = 3; i <= m->mx - 3; i++) {
            for (int k = 1; k <= 2; k++) F[MI(m, i, s, k)] = m->zopen ? 0.0 : F[MI(m, i, s, k + m->nz)];
            for (int k = m->mz - 2; k <= m->mz; k++) F[MI(m, i, s, k)] = m->zopen ? 0.0 : F[MI(m, i, s, k - m->nz)];
        }
}

void orc_cylm_set_zopen(orc_cylm *m, int zopen) { m->zopen = zopen ? 1 : 0; }

/* the conducting wall: cells i >= mx - 2 hold no field */
static void cylm_outer_zero(const orc_cylm *m, double *F)
{
    for (int k = 1; k <= m->mz; k++)
        for (int s = 0; s < 3; s++)
            for (int i = m->mx - 2; i <= m->mx; i++) F[MI(m, i, s, k)] = 0.0;
}

static inline double par(int s) { return s == 0 ? 1.0 : -1.0; }   /* (-1)^m */

/* UpdateEfield, cyl_em_update.F90:7-46 (3-D branch), i = 4..mx-3, k = 3..mz-3 */
static void cylm_update_e(const orc_cylm *m, orc_fields *f)
{
    const double fldc = m->fldc, idr = m->inv_dr, idz = m->inv_dz;
    const long long sk = 3LL * m->mx, ss = m->mx;
    for (int k = 3; k <= m->mz - 3; k++)
        for (int s = 0; s < 3; s++)
            for (int i = 4; i <= m->mx - 3; i++) {
                const long long o = MI(m, i, s, k);
                const double r = rad(m, i), rp = r + 0.5 * m->dr, rm = r - 0.5 * m->dr;
                /* (dF/dtheta) of slot s: 0 for m = 0, F_sin for cos, -F_cos for sin */
                const double dBz = s == 0 ? 0.0 : (s == 1 ? f->Bz[o + ss] : -f->Bz[o - ss]);
                const double dBr = s == 0 ? 0.0 : (s == 1 ? f->Bx[o + ss] : -f->Bx[o - ss]);
                f->Ex[o] = f->Ex[o] + fldc * dBz / rp - fldc * idz * (f->By[o] - f->By[o - sk]);
                f->Ey[o] = f->Ey[o] + fldc * idz * (f->Bx[o] - f->Bx[o - sk]) - fldc * idr * (f->Bz[o] - f->Bz[o - 1]);
                f->Ez[o] = f->Ez[o] + fldc * idr * (rp * f->By[o] - rm * f->By[o - 1]) / r - fldc * dBr / r;
            }
}

/* AddCurrent, cyl_em_update.F90:114-140, divided by the circle (2 pi, pi) */
static void cylm_add_current(const orc_cylm *m, orc_fields *f)
{
    const double idr = m->inv_dr, idz = m->inv_dz;
    for (int k = 3; k <= m->mz - 3; k++)
        for (int s = 0; s < 3; s++) {
            const double th = s == 0 ? 2.0 * PI_D : PI_D;
            for (int i = 4; i <= m->mx - 3; i++) {
                const long long o = MI(m, i, s, k);
                const double r = rad(m, i), rp = r + 0.5 * m->dr;
                f->Ex[o] = f->Ex[o] - f->Jx[o] * idz / (rp * th);
                f->Ey[o] = f->Ey[o] - f->Jy[o] * idr * idz / th;
                f->Ez[o] = f->Ez[o] - f->Jz[o] * idr / (r * th);
            }
        }
}

/* E axis fills, cyl_em_update.F90:143-180, theta + pi as the parity of the mode */
static void cylm_e_axis(const orc_cylm *m, orc_fields *f)
{
    for (int k = 1; k <= m->mz; k++)
        for (int s = 0; s < 3; s++) {
            const double p = par(s);
            f->Ex[MI(m, 3, s, k)] = 0.5 * (f->Ex[MI(m, 4, s, k)] - p * f->Ex[MI(m, 4, s, k)]);
            f->Ex[MI(m, 2, s, k)] = -(p * f->Ex[MI(m, 4, s, k)]);
            f->Ey[MI(m, 3, s, k)] = -(p * f->Ey[MI(m, 4, s, k)]);
            f->Ez[MI(m, 3, s, k)] = p * f->Ez[MI(m, 4, s, k)];
        }
}

/* UpdateBfieldHalf, cyl_em_update.F90:49-110 (3-D branch), i = 4..mx-1, k = 1..mz-1 */
static void cylm_update_b_half(const orc_cylm *m, orc_fields *f)
{
    const double h = m->fld_halfc, idr = m->inv_dr, idz = m->inv_dz;
    const long long sk = 3LL * m->mx, ss = m->mx;
    for (int k = 1; k <= m->mz - 1; k++)
        for (int s = 0; s < 3; s++)
            for (int i = 4; i <= m->mx - 1; i++) {
                const long long o = MI(m, i, s, k);
                const double r = rad(m, i), rph = r + 0.5 * m->dr, rp = r + m->dr;
                const double dEz = s == 0 ? 0.0 : (s == 1 ? f->Ez[o + ss] : -f->Ez[o - ss]);
                const double dEr = s == 0 ? 0.0 : (s == 1 ? f->Ex[o + ss] : -f->Ex[o - ss]);
                f->Bx[o] = f->Bx[o] - h * dEz / r + h * idz * (f->Ey[o + sk] - f->Ey[o]);
                f->By[o] = f->By[o] - h * idz * (f->Ex[o + sk] - f->Ex[o]) + h * idr * (f->Ez[o + 1] - f->Ez[o]);
                f->Bz[o] = f->Bz[o] - h * (idr * (rp * f->Ey[o + 1] - r * f->Ey[o]) - dEr) / rph;
            }
}

/* the Lehe differences at o = (i, s, k): gradX, gradX1 with the z pair, the theta
 * derivative (exact per mode) with the z pair, gradZ with the r pair, the mode's
 * theta term and the wide z pair (term order of cyl_em_update_lehe.F90:255-289) */
static inline double dth(const orc_cylm *m, const double *F, long long o, int s)
{
    return s == 0 ? 0.0 : (s == 1 ? F[o + m->mx] : -F[o - m->mx]);
}
static double lehe_gr(const orc_cylm *m, const double *F, long long o)
{
    const long long sk = 3LL * m->mx;
    return m->l_rr * (F[o + 1] - F[o]) + m->l_rz * (F[o + 1 + sk] - F[o + sk]) + m->l_rz * (F[o + 1 - sk] - F[o - sk]);
}
static double lehe_gr1(const orc_cylm *m, const double *F, long long o, double rp, double r)
{
    const long long sk = 3LL * m->mx;
    return m->l_rr * (rp * F[o + 1] - r * F[o]) + m->l_rz * (rp * F[o + 1 + sk] - r * F[o + sk])
           + m->l_rz * (rp * F[o + 1 - sk] - r * F[o - sk]);
}
static double lehe_gt(const orc_cylm *m, const double *F, long long o, int s)
{
    const long long sk = 3LL * m->mx;
    return m->l_tt * dth(m, F, o, s) + m->l_tz * dth(m, F, o + sk, s) + m->l_tz * dth(m, F, o - sk, s);
}
static double lehe_gz(const orc_cylm *m, const double *F, long long o, int s)
{
    const long long sk = 3LL * m->mx;
    const double lzz = s == 0 ? m->l_zz : m->l_zz - m->l_zm;
    return lzz * (F[o + sk] - F[o]) + m->l_zr * (F[o + 1 + sk] - F[o + 1]) + m->l_zr * (F[o - 1 + sk] - F[o - 1])
           + m->l_zd * (F[o + 2 * sk] - F[o - sk]);
}

/* UpdateBfieldHalf of cyl_em_update_lehe.F90:62-118 per mode, i = 4..mx-1,
 * k = 2..mz-2 (gradZ reads k - 1 and k + 2; the z halo rewrites k = 1, mz - 1) */
static void cylm_update_b_half_lehe(const orc_cylm *m, orc_fields *f)
{
    const double h = m->fld_halfc, idr = m->inv_dr, idz = m->inv_dz;
    for (int k = 2; k <= m->mz - 2; k++)
        for (int s = 0; s < 3; s++)
            for (int i = 4; i <= m->mx - 1; i++) {
                const long long o = MI(m, i, s, k);
                const double r = rad(m, i), rph = r + 0.5 * m->dr, rp = r + m->dr;
                f->Bx[o] = f->Bx[o] - h * lehe_gt(m, f->Ez, o, s) / r + h * idz * lehe_gz(m, f->Ey, o, s);
                f->By[o] = f->By[o] - h * idz * lehe_gz(m, f->Ex, o, s) + h * idr * lehe_gr(m, f->Ez, o);
                f->Bz[o] = f->Bz[o] - h * (idr * lehe_gr1(m, f->Ey, o, rp, r) - lehe_gt(m, f->Ex, o, s)) / rph;
            }
}

/* B axis: the Bz integral over the disk r < dr/2 (cyl_em_update.F90:231-250: the
 * circulation of E_theta at r = dr/2 over the disk area, 4 / dr for m = 0; the
 * m = 1 circulation vanishes), then the fills of :182-229 */
static void cylm_b_axis(const orc_cylm *m, orc_fields *f)
{
    for (int k = 1; k <= m->mz - 1; k++) {
        const long long a = MI(m, 3, 0, k);
        f->Bz[a] = f->Bz[a] - m->fld_halfc * (4.0 * m->inv_dr) * f->Ey[MI(m, 4, 0, k)];
    }
    for (int k = 1; k <= m->mz; k++)
        for (int s = 0; s < 3; s++) {
            const double p = par(s);
            f->By[MI(m, 3, s, k)] = 0.5 * (f->By[MI(m, 4, s, k)] - p * f->By[MI(m, 4, s, k)]);
            f->By[MI(m, 2, s, k)] = -(p * f->By[MI(m, 4, s, k)]);
            f->Bx[MI(m, 3, s, k)] = -(p * f->Bx[MI(m, 4, s, k)]);
            f->Bz[MI(m, 2, s, k)] = p * f->Bz[MI(m, 4, s, k)];
        }
}

void orc_cylm_b_half(const orc_cylm *m, orc_fields *f)
{
    if (m->fdtd == 1)
        cylm_update_b_half_lehe(m, f);
    else
        cylm_update_b_half(m, f);
    cylm_b_axis(m, f);
    cylm_zhalo(m, f->Bx); cylm_zhalo(m, f->By); cylm_zhalo(m, f->Bz);
    cylm_outer_zero(m, f->Bx); cylm_outer_zero(m, f->By); cylm_outer_zero(m, f->Bz);
}

void orc_cylm_e_phase(const orc_cylm *m, orc_fields *f, int curl, int add)
{
    if (curl) cylm_update_e(m, f);
    if (add) cylm_add_current(m, f);
    cylm_e_axis(m, f);
    cylm_zhalo(m, f->Ex); cylm_zhalo(m, f->Ey); cylm_zhalo(m, f->Ez);
    cylm_outer_zero(m, f->Ex); cylm_outer_zero(m, f->Ey); cylm_outer_zero(m, f->Ez);
}

/* AxisCurrentBC (cyl_movdep.F90:451-465) on every mode, then the periodic z
 * fold: 3..5 += mz-2..mz, mz-4..mz-3 += 1..2 (comm_fld.F90:15-33); with open z
 * the current deposited in the ghost rows leaves the box with its particles */
void orc_cylm_current_halo(const orc_cylm *m, orc_fields *f)
{
    for (int k = 1; k <= m->mz; k++)
        for (int s = 0; s < 3; s++) {
            f->Jy[MI(m, 4, s, k)] = f->Jy[MI(m, 4, s, k)] - f->Jy[MI(m, 3, s, k)];
            f->Jz[MI(m, 4, s, k)] = f->Jz[MI(m, 4, s, k)] - f->Jz[MI(m, 3, s, k)];
            f->Jx[MI(m, 3, s, k)] = 0.0;
            f->Jy[MI(m, 2, s, k)] = 0.0;
            f->Jz[MI(m, 2, s, k)] = 0.0;
        }
    if (m->zopen) return;
    double *J[3] = {f->Jx, f->Jy, f->Jz};
    for (int c = 0; c < 3; c++)
        for (int s = 0; s < 3; s++)
            for (int i = 1; i <= m->mx; i++) {
                for (int k = 3; k <= 5; k++) J[c][MI(m, i, s, k)] = J[c][MI(m, i, s, k)] + J[c][MI(m, i, s, k + m->nz)];
                for (int k = m->mz - 4; k <= m->mz - 3; k++)
                    J[c][MI(m, i, s, k)] = J[c][MI(m, i, s, k)] + J[c][MI(m, i, s, k - m->nz)];
            }
}

void orc_cylm_reset_current(const orc_cylm *m, orc_fields *f)
{
    const size_t n = (size_t)m->mx * 3 * m->mz;
    memset(f->Jx, 0, n * sizeof(double));
    memset(f->Jy, 0, n * sizeof(double));
    memset(f->Jz, 0, n * sizeof(double));
}

/* ------------------------------------------------------------- particles */

/* node values of the staggered fields (the averaging of GatherVecE/BFld in the
 * r-z plane, movdep_routines.F90:6-223 with cyl_movdep.F90's node layout) */
static void cylm_node(const orc_cylm *m, const orc_fields *f, int i, int s, int k, double e[3], double b[3])
{
    const long long o = MI(m, i, s, k), sk = 3LL * m->mx;
    e[0] = (f->Ex[o - 1] + f->Ex[o]) * 0.5;
    e[1] = f->Ey[o];
    e[2] = (f->Ez[o - sk] + f->Ez[o]) * 0.5;
    b[0] = (f->Bx[o - sk] + f->Bx[o]) * 0.5;
    b[1] = (f->By[o - 1 - sk] + f->By[o - 1] + f->By[o - sk] + f->By[o]) * 0.25;
    b[2] = (f->Bz[o - 1] + f->Bz[o]) * 0.5;
}

/* fields at (x, theta, z): CIC in (x, z) of every slot, summed with the phases */
static void cylm_gather(const orc_cylm *m, const orc_fields *f, double x, double th, double z, double pe[3],
                        double pb[3])
{
    const int ip = (int)x, kp = (int)z;
    const double dx = x - ip, dz = z - kp;
    const double ph[3] = {1.0, cos(th), sin(th)};
    for (int c = 0; c < 3; c++) { pe[c] = 0.0; pb[c] = 0.0; }
    for (int s = 0; s < 3; s++) {
        double se[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
        for (int nn = 0; nn < 4; nn++) {
            const int a = nn & 1, b = nn >> 1;
            const double wt = (a ? dx : 1.0 - dx) * (b ? dz : 1.0 - dz);
            double e[3], bb[3];
            cylm_node(m, f, ip + a, s, kp + b, e, bb);
            for (int c = 0; c < 3; c++) { se[c] = se[c] + wt * e[c]; sb[c] = sb[c] + wt * bb[c]; }
        }
        for (int c = 0; c < 3; c++) { pe[c] = pe[c] + ph[s] * se[c]; pb[c] = pb[c] + ph[s] * sb[c]; }
    }
}

/* the deposit of one step: (a) the r-z zigzag from (x0, z0) to (x1, z1) at the
 * old angle th0, every slot weighted by its phase; (b) the angular move th0 ->
 * th1 at the node weights of (x1, z1): m = 0 carries the angle, m = 1 the
 * differences of sin and -cos (the charge of slot c changes by cos(th1) -
 * cos(th0), which the sin slot's angular current divergence removes) */
static void cylm_deposit(const orc_cylm *m, orc_fields *f, double q, double x0, double th0, double z0, double x1,
                         double th1, double z1)
{
    const double qthis = q * m->qi;
    const int i1 = (int)floor(x0), i2 = (int)floor(x1), k1 = (int)floor(z0), k2 = (int)floor(z1);
    const double xr = fmin((double)((i1 < i2 ? i1 : i2) + 1), fmax((double)(i1 > i2 ? i1 : i2), 0.5 * (x0 + x1)));
    const double zr = fmin((double)((k1 < k2 ? k1 : k2) + 1), fmax((double)(k1 > k2 ? k1 : k2), 0.5 * (z0 + z1)));
    const double Fx1 = qthis * (xr - x0), Fz1 = qthis * (zr - z0);
    const double Fx2 = qthis * (x1 - xr), Fz2 = qthis * (z1 - zr);
    const double Wx1 = 0.5 * (x0 + xr) - i1, Wz1 = 0.5 * (z0 + zr) - k1;
    const double Wx2 = 0.5 * (x1 + xr) - i2, Wz2 = 0.5 * (z1 + zr) - k2;
    const double pa[3] = {1.0, cos(th0), sin(th0)};
    const long long sk = 3LL * m->mx;
    for (int s = 0; s < 3; s++) {
        const long long a = MI(m, i1, s, k1), b = MI(m, i2, s, k2);
        f->Jx[a] = f->Jx[a] + pa[s] * Fx1 * (1.0 - Wz1);
        f->Jx[a + sk] = f->Jx[a + sk] + pa[s] * Fx1 * Wz1;
        f->Jz[a] = f->Jz[a] + pa[s] * Fz1 * (1.0 - Wx1);
        f->Jz[a + 1] = f->Jz[a + 1] + pa[s] * Fz1 * Wx1;
        f->Jx[b] = f->Jx[b] + pa[s] * Fx2 * (1.0 - Wz2);
        f->Jx[b + sk] = f->Jx[b + sk] + pa[s] * Fx2 * Wz2;
        f->Jz[b] = f->Jz[b] + pa[s] * Fz2 * (1.0 - Wx2);
        f->Jz[b + 1] = f->Jz[b + 1] + pa[s] * Fz2 * Wx2;
    }
    const double wx = x1 - i2, wz = z1 - k2;
    const double pt[3] = {qthis * (th1 - th0), qthis * (sin(th1) - sin(th0)), -(qthis * (cos(th1) - cos(th0)))};
    for (int s = 0; s < 3; s++) {
        const long long b = MI(m, i2, s, k2);
        f->Jy[b] = f->Jy[b] + pt[s] * (1.0 - wx) * (1.0 - wz);
        f->Jy[b + 1] = f->Jy[b + 1] + pt[s] * wx * (1.0 - wz);
        f->Jy[b + sk] = f->Jy[b + sk] + pt[s] * (1.0 - wx) * wz;
        f->Jy[b + 1 + sk] = f->Jy[b + 1 + sk] + pt[s] * wx * wz;
    }
}

/* MoveDepositPrtl with the cylindrical local-frame step (cyl_movdep.F90:14-289).
 * Returns ORC_ERR_BOUNDS when a particle reaches the wall (x >= mx - 3). */
int orc_cylm_move_deposit(const orc_cylm *m, orc_fields *f, orc_particles *s)
{
    int rc = ORC_OK;
    const double c = m->c, sqc = m->sqc, cinv = m->cinv;
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        const int flv = s->flv[n];
        const double qm = m->flvrqm[flv - 1] * 0.5;
        const double x0 = s->x[n], th0 = s->y[n], z0 = s->z[n];
        double pe[3], pb[3];
        cylm_gather(m, f, x0, th0, z0, pe, pb);
        const double pEx = pe[0] * qm, pEy = pe[1] * qm, pEz = pe[2] * qm;
        double pBx = pb[0] * qm, pBy = pb[1] * qm, pBz = pb[2] * qm;
        double u0 = c * s->u[n] + pEx;
        double v0 = c * s->v[n] + pEy;
        double w0 = c * s->w[n] + pEz;
        double gg = 1.0 / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        pBx = gg * pBx; pBy = gg * pBy; pBz = gg * pBz;
        const double ff = 2.0 / (1.0 + pBx * pBx + pBy * pBy + pBz * pBz);
        double u1 = (u0 + v0 * pBz - w0 * pBy) * ff;
        double v1 = (v0 + w0 * pBx - u0 * pBz) * ff;
        const double w1 = (w0 + u0 * pBy - v0 * pBx) * ff;
        u0 = u0 + v1 * pBz - w1 * pBy + pEx;
        v0 = v0 + w1 * pBx - u1 * pBz + pEy;
        w0 = w0 + u1 * pBy - v1 * pBx + pEz;
        u1 = u0 * cinv;
        v1 = v0 * cinv;
        const double wn = w0 * cinv;
        gg = sqc / sqrt(sqc + u0 * u0 + v0 * v0 + w0 * w0);
        /* the move in the particle's local frame (x along r, y along theta) */
        const double X = rad(m, x0) + u1 * gg, Y = v1 * gg;
        const double r1 = sqrt(X * X + Y * Y);
        const double x1 = r1 * m->inv_dr + 3.5;
        const double th1 = th0 + atan2(Y, X);
        const double z1 = z0 + wn * gg * m->inv_dz;
        const double ir = 1.0 / r1;
        const double un = (X * u1 + Y * v1) * ir;
        const double vn = (-Y * u1 + X * v1) * ir;
        if (x1 >= m->mx - 3) {
            rc = ORC_ERR_BOUNDS;
        } else {
            cylm_deposit(m, f, s->q[n], x0, th0, z0, x1, th1, z1);
        }
        double zn = z1, tn = th1;
        if (m->zopen) {
            /* open z: a particle that leaves through either end is dropped */
            if (zn >= m->zmax || zn < m->zmin) s->flv[n] = 0;
        } else if (zn > m->zmax) {
            zn = zn - m->zlen;
        } else if (zn < m->zmin) {
            zn = m->zlen + zn;
        }
        if (tn >= 2.0 * PI_D) tn = tn - 2.0 * PI_D;
        else if (tn < 0.0) tn = tn + 2.0 * PI_D;
        s->x[n] = x1; s->y[n] = tn; s->z[n] = zn;
        s->u[n] = un; s->v[n] = vn; s->w[n] = wn;
    }
    return rc;
}

/* one step in the MAIN.F90:42-284 order */
int orc_cylm_step(const orc_cylm *m, orc_fields *f, orc_particles *s)
{
    orc_cylm_reset_current(m, f);
    orc_cylm_b_half(m, f);
    int rc = orc_cylm_move_deposit(m, f, s);
    orc_cylm_b_half(m, f);
    orc_cylm_e_phase(m, f, 1, 0);
    orc_cylm_current_halo(m, f);
    orc_cylm_e_phase(m, f, 0, 1);
    return rc;
}

/* ------------------------------------------------------------- diagnostics */

/* 1/2 integral of B^2 and E^2 over the physical cells with the volume 2 pi r dr dz
 * of each staggered position; the m = 1 slots carry half the weight of m = 0 */
void orc_cylm_energy(const orc_cylm *m, const orc_fields *f, double out[2])
{
    double be = 0.0, ee = 0.0;
    for (int k = 3; k <= m->mz - 3; k++)
        for (int s = 0; s < 3; s++) {
            const double wm = s == 0 ? 2.0 * PI_D : PI_D;
            for (int i = 4; i <= m->mx - 3; i++) {
                const long long o = MI(m, i, s, k);
                const double r = rad(m, i), rh = r + 0.5 * m->dr;
                ee += wm * m->dr * (rh * f->Ex[o] * f->Ex[o] + r * f->Ey[o] * f->Ey[o] + r * f->Ez[o] * f->Ez[o]);
                be += wm * m->dr * (r * f->Bx[o] * f->Bx[o] + rh * f->By[o] * f->By[o] + rh * f->Bz[o] * f->Bz[o]);
            }
        }
    out[0] = 0.5 * be;
    out[1] = 0.5 * ee;
}

/* discrete Gauss residual per node (i, s, k): the flux of E out of the node's
 * cell, theta_m [dz (r+ Er(i) - r- Er(i-1)) + r dr (Ez(k) - Ez(k-1)) + dr dz
 * (dE_theta/dtheta)_s], minus the node-CIC charge (in index coordinates) of the
 * particles times their phase; written into g (zero where not evaluated:
 * i < 5, i > mx - 4, ghosts).  AddCurrent and the deposit keep it constant. */
void orc_cylm_gauss(const orc_cylm *m, const orc_fields *f, const orc_particles *s, double *g)
{
    const size_t nc = (size_t)m->mx * 3 * m->mz;
    double *rho = (double *)calloc(nc, sizeof(double));
    if (!rho) return;
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        const double x = s->x[n], z = s->z[n], th = s->y[n];
        const int i = (int)floor(x), k = (int)floor(z);
        const double wx = x - i, wz = z - k, ph[3] = {1.0, cos(th), sin(th)};
        const double qq = s->q[n] * m->qi;
        for (int sl = 0; sl < 3; sl++)
            for (int nn = 0; nn < 4; nn++) {
                const int a = nn & 1, b = nn >> 1;
                int kk = k + b;
                if (kk > m->mz - 3 && !m->zopen) kk -= m->nz;   /* periodic image */
                rho[MI(m, i + a, sl, kk)] += qq * ph[sl] * (a ? wx : 1.0 - wx) * (b ? wz : 1.0 - wz);
            }
    }
    memset(g, 0, nc * sizeof(double));
    const long long sk = 3LL * m->mx, ss = m->mx;
    for (int k = 3; k <= m->mz - 3; k++)
        for (int sl = 0; sl < 3; sl++) {
            const double th = sl == 0 ? 2.0 * PI_D : PI_D;
            for (int i = 5; i <= m->mx - 4; i++) {
                const long long o = MI(m, i, sl, k);
                const double r = rad(m, i), rp = r + 0.5 * m->dr, rm = r - 0.5 * m->dr;
                const double dEt = sl == 0 ? 0.0 : (sl == 1 ? f->Ey[o + ss] : -f->Ey[o - ss]);
                const double flux = th * ((rp * f->Ex[o] - rm * f->Ex[o - 1]) + r * m->dr * (f->Ez[o] - f->Ez[o - sk]) +
                                          m->dr * dEt);
                g[o] = flux - rho[o];
            }
        }
    free(rho);
}

/* ------------------------------------------------------------- loading */

static inline uint64_t mix64c(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static inline double u01(uint64_t seed, uint64_t idx, uint64_t stream)
{
    const uint64_t h = mix64c(mix64c(seed ^ (idx * 0xD1342543DE82EF95ULL)) + stream * 0x9E3779B97F4A7C15ULL);
    return (double)(h >> 24) * (1.0 / 1099511627776.0);
}

/* electrons (flavour 2) of uniform density in the cells i = 4..4+nr_p-1 (r < nr_p dr),
 * z cells k0..k0+nzc-1: per cell ppc particles, (r, z) uniform in the cell, angles
 * evenly spaced with a random offset, weight epc * 2 pi r dr / ppc (epc weight per
 * unit volume); cold (zero momentum).  Cell (i, k) uses counter indices of the cell
 * number (k - 3) * nr + (i - 4) and the window shift `shift`, so the same cells give
 * the same particles on every device. Appends at s->used; returns the count. */
int64_t orc_cylm_load(const orc_cylm *m, orc_particles *s, int nr_p, int k0, int nzc, int ppc, int epc,
                      uint64_t seed, int64_t shift)
{
    int64_t added = 0;
    for (int k = k0; k < k0 + nzc; k++)
        for (int i = 4; i < 4 + nr_p; i++) {
            const uint64_t cell = (uint64_t)((k - 3 + shift) * (int64_t)m->nr + (i - 4));
            const double off = u01(seed, cell, 7);
            for (int p = 0; p < ppc; p++) {
                if (s->used >= s->cap) return added;
                const uint64_t id = cell * (uint64_t)ppc + (uint64_t)p;
                const double x = i + u01(seed, id, 1), z = k + u01(seed, id, 2);
                const double th = 2.0 * PI_D * ((p + off) / ppc);
                const int64_t n = s->used++;
                s->x[n] = x; s->y[n] = th; s->z[n] = z;
                s->u[n] = 0.0; s->v[n] = 0.0; s->w[n] = 0.0;
                s->q[n] = -(epc * 2.0 * PI_D * rad(m, x) * m->dr / ppc);
                s->flv[n] = 2;
                added++;
            }
        }
    return added;
}

/* the moving window's shift by one axial cell: fields (k) <- (k + 1) with the
 * front row k = mz - 3 and the ghosts above it zero, then the z halos and the
 * wall; particles z -= 1, those below z = 3 dropped (flavour 0); fresh plasma
 * in the front row (orc_cylm_load with the window's shift count) */
void orc_cylm_shift(const orc_cylm *m, orc_fields *f, orc_particles *s, int nr_p, int ppc, int epc, uint64_t seed,
                    int64_t shift)
{
    double *F[6] = {f->Ex, f->Ey, f->Ez, f->Bx, f->By, f->Bz};
    for (int c = 0; c < 6; c++)
        for (int k = 1; k <= m->mz; k++)
            for (int sl = 0; sl < 3; sl++)
                for (int i = 1; i <= m->mx; i++)
                    F[c][MI(m, i, sl, k)] = (k + 1 <= m->mz - 3) ? F[c][MI(m, i, sl, k + 1)] : 0.0;
    for (int64_t n = 0; n < s->used; n++) {
        if (s->flv[n] == 0) continue;
        s->z[n] = s->z[n] - 1.0;
        if (s->z[n] < m->zmin) s->flv[n] = 0;
    }
    if (nr_p > 0 && ppc > 0) orc_cylm_load(m, s, nr_p, m->mz - 3, 1, ppc, epc, seed, shift);
    cylm_e_axis(m, f);
    for (int c = 0; c < 6; c++) { cylm_zhalo(m, F[c]); cylm_outer_zero(m, F[c]); }
}
