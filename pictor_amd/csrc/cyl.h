// cyl.h -- cylindrical geometry (the reference's -Dcyl build with the axis
// included): x = r, y = theta, z = z on the rank that holds the axis.  FP64.
#pragma once
#include "pictor_common.h"

struct pictor_ctx;

namespace pictor {

// InitParam_cyl (cyl_common.F90:19-24) for xborders(0) = 0, and the Lehe-cyl
// coefficients of cyl_em_update_lehe.F90:6-17 with dtheta at run time.
struct CylP {
    double dtheta, inv_dtheta, ax_perm_area;
    double rmin;     // r of grid index i is (i - 3) + rmin = i - 3.5
    double r_proc;   // rmin - xmin (cyl_movdep.F90:56)
    double lehe[12]; // alpha_x, alpha_y, alpha_z, beta_xy, beta_xz, beta_yx, beta_yz, beta_zx, beta_zy, delta_x..z
    int lehe_on;     // fdtd = Lehe in 3-D
    int substep_limit;
    int strang;      // 1: the Strang push of cyl_movdep_strang.F90 (the CPU -Dcyl build); 0: cyl_movdep.F90
};

enum CylErr : int { kCylWall = 1, kCylSubsteps = 2 };

void cyl_params(const GridP &g, const PhysP &ph, int fdtd, CylP &cp);

}  // namespace pictor

// host side of the cylindrical phases (cyl.hip); each returns a PICTOR_* code
int cyl_setup(pictor_ctx *c);
void cyl_release(pictor_ctx *c);
int cyl_upload_particles(pictor_ctx *c, long long n, const double *q, const double *x, const double *y,
                         const double *z, const double *u, const double *v, const double *w, const int *flv);
int cyl_diagnostics(pictor_ctx *c, double out[9]);
int cyl_download_particles(pictor_ctx *c, long long cap, double *q, double *x, double *y, double *z, double *u,
                           double *v, double *w, int *flv, int *tile, long long *n_out);
int cyl_b_half(pictor_ctx *c);
int cyl_push(pictor_ctx *c, bool timed);
int cyl_exchange(pictor_ctx *c);
int cyl_e_phase(pictor_ctx *c, bool curl, bool add);
int cyl_current_halo(pictor_ctx *c);
int cyl_field_halo(pictor_ctx *c, int which);
int cyl_reset_current(pictor_ctx *c);
