/*
 * pictor_oracle.h -- CPU FP64 restatement of PICTOR's particle step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg).  It is never linked into, loaded by or called from the
 * product library (pictor_amd/).
 *
 * Every routine follows the expression order of the cited Fortran so that a
 * build with -ffp-contract=off reproduces the reference CPU path's rounding
 * (single thread).  Arrays use the reference layout: fields are (mx,my,mz)
 * column-major with 1-based indices (element (1,1,1) at offset 0); particles
 * are SoA q,x,y,z,u,v,w (double) + flv (int32), holes marked flv=0.
 */
#ifndef PICTOR_ORACLE_H
#define PICTOR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t ndim;            /* 2 or 3 */
    int32_t nx, ny, nz;      /* physical cells (nz ignored in 2-D) */
    int32_t mx, my, mz;      /* n + 5 (mz = 1 in 2-D), subdomains.F90:42-50 */
    int32_t epc;             /* particles per cell per species */
    double c, c_ompe, g0, mi_me;
    double me, mi, ompe, qi, qe, qmi, qme, masse, massi;  /* initialise.F90:44-53 */
    double fldc, fld_halfc, cinv, sqc;                    /* vars.F90:220-223 */
    double inv_dx, inv_dy, inv_dz;                        /* vars.F90:15-17 */
    double xmin, xmax, ymin, ymax, zmin, zmax, xlen, ylen, zlen; /* subdomains.F90:52-67 */
    double Bx_ext0, By_ext0, Bz_ext0;
    int32_t nflvr;
    int32_t reorder_period;  /* parameters.F90:71 */
    double flvrqm[8];        /* initialise.F90:84-85 */
    int32_t fdtd;            /* 0 = Yee (em_update.F90), 1 = Lehe (em_update_lehe.F90) */
    int32_t curr_filter;     /* passes of smoothen_current per step, parameters.F90:22 */
    /* em_update_lehe.F90:8-25: alpha_x, alpha_y, alpha_z, beta_xy, beta_xz, beta_yx,
     * beta_yz, beta_zx, beta_zy, delta_x, delta_y, delta_z */
    double lehe[12];
    /* Boundary faces (bc.F90): face f = 0..5 is the reference's side f + 1 (x lower,
     * x upper, y lower, y upper, z lower, z upper).  periodic[a] = 0 once any face of
     * axis a has a type (PeriodicOff, subdomains.F90:508-547).  Positions are global
     * (x = local - 3 on one rank) and move by speed * c per step (updateBC_Pos).
     * pml_L: thickness of a 'pml' face (SetPMLParameters, bc_pml.F90:39-51).
     * cond_top_as_written = 1 applies CondBC_Fld_Top at face 3's position, as
     * bc.F90:318 is written; 0 (default) uses face 4's own position. */
    int32_t periodic[3];
    int32_t bc_fld[6], bc_prtl[6];
    double bc_pos_fld[6], bc_pos_prtl[6], bc_speed[6], pml_L[6];
    int32_t cond_top_as_written;
} orc_params;

/* face types (bc_face%type_fld / type_prtl) */
#define ORC_BC_PERIODIC 0   /* 'prdc' */
#define ORC_BC_COND 1       /* field 'cond': CondBC_Fld_*, bc_reflect.F90:11-65 */
#define ORC_BC_OPEN 2       /* field 'open': OpenBC_Fld_*, bc.F90:410-469 (x faces) */
#define ORC_BC_PML 3        /* field 'pml': bc_pml.F90 */
#define ORC_BC_REFL 4       /* particle 'refl': RefBC_Prtl_*, bc_reflect.F90:76-237 (x, y faces) */
#define ORC_BC_REMV 5       /* particle 'remv': RemovePrtl_*, bc_open.F90 */

typedef struct orc_fields {
    double *Ex, *Ey, *Ez, *Bx, *By, *Bz, *Jx, *Jy, *Jz;
    /* PML auxiliary fields (bc_pml.F90:21-31), each (mx,my,mz): pml_E_psi1 x,y,z,
     * pml_E_psi2 x,y,z, pml_B_psi1 x,y,z, pml_B_psi2 x,y,z; NULL without a 'pml' face */
    double *pml[12];
} orc_fields;

typedef struct orc_particles {
    int64_t cap;    /* allocated length of every array */
    int64_t used;   /* used_prtl_arr_size (vars.F90:125-127) */
    double *q, *x, *y, *z, *u, *v, *w;
    int32_t *flv;
} orc_particles;

typedef struct orc_ws orc_ws;   /* VecE/VecB/VecJ work arrays (mem_fld.F90:6-32) */

/* status codes */
#define ORC_OK 0
#define ORC_ERR_CAPACITY 12     /* same meaning as the reference's Abort(12) */
#define ORC_ERR_ARG 1
#define ORC_ERR_BOUNDS 3        /* a particle reached the wall of the cylindrical grid */

int  orc_params_init(orc_params *p, int ndim, int nx, int ny, int nz, double c,
                     double c_ompe, double g0, double mi_me, int epc);
orc_ws *orc_ws_create(const orc_params *p, int nthreads);
void orc_ws_destroy(orc_ws *w);
int  orc_ws_threads(const orc_ws *w);

/* Lehe coefficients (em_update_lehe.F90:8-25) and fdtd = 1 */
void orc_set_lehe(orc_params *p);
/* Maxwell solver, em_update.F90 (fdtd 0) / em_update_lehe.F90 B half step (fdtd 1) */
void orc_update_b_half(const orc_params *p, orc_fields *f);
void orc_update_e(const orc_params *p, orc_fields *f);
void orc_add_current(const orc_params *p, orc_fields *f);
void orc_reset_current(const orc_params *p, orc_fields *f);
/* Periodic single-rank halos, comm_fld.F90 */
void orc_field_halo(const orc_params *p, double *fx, double *fy, double *fz);
void orc_current_halo_add(const orc_params *p, double *jx, double *jy, double *jz);
/* smoothen_current, current.F90:19-30: p->curr_filter passes of the 1-2-1
 * moving average (MovingAverageFilter, :55-75) with the per-axis periodic
 * ghost exchanges of one rank */
void orc_smoothen_current(const orc_params *p, orc_fields *f);
/* Particle mover, movdep.F90 + movdep_routines.F90 */
void orc_gather_vec_e(const orc_params *p, orc_ws *w, const orc_fields *f);
void orc_gather_vec_b(const orc_params *p, orc_ws *w, const orc_fields *f);
void orc_move_deposit(const orc_params *p, orc_ws *w, orc_fields *f, orc_particles *s);
/* The VecE/VecB corner arrays of the last gather (for tests). */
const double *orc_ws_vec(const orc_ws *w, int which);   /* 0..5 = Ex,Ey,Ez,Bx,By,Bz */
int  orc_ws_vec_len(const orc_ws *w);                   /* 4 (2-D) or 8 (3-D) */
/* Particle exchange (wrap + append), loadprtlout.F90 + comm_prtl.F90 */
int  orc_exchange_particles(const orc_params *p, orc_particles *s);
/* Counting sort by cell, mem_prtl.F90:166-301 */
int  orc_reorder(const orc_params *p, orc_particles *s);
/* Boundary conditions (bc.F90:238-348, bc_reflect.F90, bc_open.F90, bc_pml.F90) */
void orc_bc_post_movdep(const orc_params *p, orc_fields *f, orc_particles *s);  /* EnforceBC_PostMovDep */
void orc_bc_final(orc_params *p, orc_fields *f);      /* EnforceBC_Final incl. updateBC_Pos */
void orc_pml_b(const orc_params *p, orc_fields *f);   /* UpdateMagFld_PML */
void orc_pml_e(const orc_params *p, orc_fields *f);   /* UpdateElcFld_PML */
/* pml_ab / pml_c of bc_pml.F90:293-353 at global coordinate x along axis (0..2):
 * out = {a, b, c}; exposed for the coefficient pins */
void orc_pml_coeffs(const orc_params *p, double x, int axis, double out[3]);
/* One MAIN.F90 time step (J reset moved to the step start). */
int  orc_step(orc_params *p, orc_ws *w, orc_fields *f, orc_particles *s, int t);
int  orc_run(orc_params *p, orc_ws *w, orc_fields *f, orc_particles *s, int t0, int nsteps);

/* Diagnostics, savedata_routines.F90:265-355 */
void orc_energy(const orc_params *p, const orc_fields *f, const orc_particles *s, double out[4]);
void orc_div_e(const orc_params *p, const orc_fields *f, double *div);
/* CIC charge density at the nodes, periodically folded onto physical nodes. */
void orc_rho(const orc_params *p, const orc_particles *s, double *rho);
/* max |div E - rho| and max |rho| over physical nodes. */
void orc_gauss(const orc_params *p, const orc_fields *f, const orc_particles *s, double out[2]);
void orc_momentum(const orc_params *p, const orc_particles *s, double out[3]);

/* ---- cylindrical geometry (-Dcyl, inc_axis, one rank holding the axis):
 * x = r, y = theta, z = z; r of grid index i is i - 3.5 (the axis at x = 3.5) */
typedef struct orc_cyl {
    double dtheta, inv_dtheta, ax_perm_area, rshift;  /* cyl_common.F90:19-24 */
    double rmin;             /* xborders(0) + rshift */
    double r_proc;           /* xborders(0) + rshift - xmin, cyl_movdep.F90:56 */
    int32_t second_half_new_cell;  /* 1: second zigzag half at the new cell; 0: cyl_movdep.F90:249,252 indexing */
    int32_t push;            /* 0: Boris (cyl_movdep.F90); 1: Strang (cyl_movdep_strang.F90, the
                                CPU -Dcyl build); 2: Strang as its main loop is written */
    int64_t substep_limit;   /* bound on the axis sub-step loop (unbounded in the reference) */
    int64_t substeps;        /* sub-steps taken by the last orc_cyl_move_deposit */
    double lehe[12];         /* cyl_em_update_lehe.F90:6-17, the order of orc_params.lehe */
} orc_cyl;

int  orc_cyl_init(orc_cyl *cy, const orc_params *p);
void orc_cyl_update_e(const orc_params *p, const orc_cyl *cy, orc_fields *f);
void orc_cyl_update_b_half(const orc_params *p, const orc_cyl *cy, orc_fields *f);
void orc_cyl_add_current(const orc_params *p, const orc_cyl *cy, orc_fields *f);
void orc_cyl_axis_current_bc(const orc_params *p, orc_fields *f);
void orc_cyl_field_halo(const orc_params *p, double *fx, double *fy, double *fz);
void orc_cyl_current_halo_add(const orc_params *p, double *jx, double *jy, double *jz);
int  orc_cyl_move_deposit(const orc_params *p, orc_cyl *cy, orc_ws *w, orc_fields *f, orc_particles *s);
void orc_cyl_exchange_particles(const orc_params *p, orc_particles *s);
void orc_cyl_energy(const orc_params *p, const orc_cyl *cy, const orc_fields *f, const orc_particles *s, double out[4]);
int  orc_cyl_step(const orc_params *p, orc_cyl *cy, orc_ws *w, orc_fields *f, orc_particles *s);

/* ---- azimuthal-mode cylindrical form (m = 0, 1), pictor_oracle_cylm.c ----
 * No reference counterpart: the reference's cylindrical grid is full r-theta-z.
 * Fields (mx, 3, mz) per component, slot 0 = m 0, 1 = m 1 cos, 2 = m 1 sin;
 * particles: x = radial grid coordinate (r = (x - 3.5) dr), y = theta in
 * radians, z = axial grid coordinate; momenta in the local (r, theta, z) frame. */
typedef struct orc_cylm {
    int32_t nr, nz, mx, mz;
    double dr, inv_dr, inv_dz;
    double c, fldc, fld_halfc, cinv, sqc;
    double ompe, qi;
    double flvrqm[2];
    double zmin, zmax, zlen;
    /* fdtd = 1: the Lehe B update (orc_cylm_set_fdtd); zopen = 1: open z ends
     * (the moving window, orc_cylm_set_zopen) instead of periodic z */
    int32_t fdtd, zopen;
    double l_rr, l_rz, l_tt, l_tz, l_zz, l_zr, l_zd, l_zm;
} orc_cylm;

int orc_cylm_init(orc_cylm *m, int nr, int nz, double dr, double c, double c_ompe, double g0, double mi_me, int epc);
int orc_cylm_set_fdtd(orc_cylm *m, int fdtd);
void orc_cylm_set_zopen(orc_cylm *m, int zopen);
void orc_cylm_b_half(const orc_cylm *m, orc_fields *f);
void orc_cylm_e_phase(const orc_cylm *m, orc_fields *f, int curl, int add);
void orc_cylm_current_halo(const orc_cylm *m, orc_fields *f);
void orc_cylm_reset_current(const orc_cylm *m, orc_fields *f);
int orc_cylm_move_deposit(const orc_cylm *m, orc_fields *f, orc_particles *s);
int orc_cylm_step(const orc_cylm *m, orc_fields *f, orc_particles *s);
void orc_cylm_energy(const orc_cylm *m, const orc_fields *f, double out[2]);
void orc_cylm_gauss(const orc_cylm *m, const orc_fields *f, const orc_particles *s, double *g);
int64_t orc_cylm_load(const orc_cylm *m, orc_particles *s, int nr_p, int k0, int nzc, int ppc, int epc,
                      uint64_t seed, int64_t shift);
void orc_cylm_shift(const orc_cylm *m, orc_fields *f, orc_particles *s, int nr_p, int ppc, int epc, uint64_t seed,
                    int64_t shift);

#ifdef __cplusplus
}
#endif
#endif
