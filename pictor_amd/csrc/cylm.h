// cylm.h -- azimuthal-mode (m = 0, 1) cylindrical geometry (cylm.hip).  FP64.
#pragma once
#include "pictor_common.h"

struct pictor_ctx;
struct pictor_config;

namespace pictor {

// scalars of the modal grid: lengths in axial cells, radial cell dr
struct CylmP {
    double dr, inv_dr, inv_dz;
    double c, fldc, fld_halfc, cinv, sqc;
    double qi;
    double flvrqm[2];
    double zmin, zmax, zlen;
    // open z ends while the moving window is on: the z ghosts hold no field, J is not
    // folded along z and a particle that leaves through either end becomes a hole
    int zopen;
    // the Lehe B update (fdtd = PICTOR_FDTD_LEHE): z averaging of the r and theta
    // differences (l_rr, l_rz; l_tt, l_tz) and the z difference with the wide pair (l_zz, l_zd)
    int lehe;
    double l_rr, l_rz, l_tt, l_tz, l_zz, l_zd;
};

// the split particle record of the modal path (FP32 push, fixed-point current):
// 32 B read and 28 B written per particle update
struct PartCylS {
    uint32_t *cell = nullptr;   // 0: hole; else (flavour 1: bit 31) | k << 12 | i
    float *ox = nullptr;        // radial offset in the cell, [0, 1) in index units
    float *oz = nullptr;        // axial offset in the cell, [0, 1)
    uint32_t *th = nullptr;     // theta in units of 2^-32 turns
    float *u = nullptr, *v = nullptr, *w = nullptr;   // local-frame momenta
    float *q = nullptr;         // charge weight (the FP64 record's q)
};

// particle storage and moving-window state of the modal path
struct CylmState {
    PartCylS ps;                         // the split form's particles (FP32_SPLIT); pcyl holds FP64
    int sort_period = 0;                 // pictor_step sorts by cell every sort_period steps (0: never)
    long long steps_since_sort = 0;
    // the sort's buffers, kept between sorts: keys and indices (in, out), the radix sort's
    // scratch, and a spare particle record of the current capacity
    uint32_t *skey[2] = {nullptr, nullptr}, *sidx[2] = {nullptr, nullptr};
    long long sbuf_cap = 0;
    void *stmp = nullptr;
    size_t stmp_bytes = 0;
    PartF64 spare64;
    PartCylS spare32;
    long long spare_cap = 0;
    long long cap = 0;                   // allocated particle slots (c->np slots in use, holes included)
    // the split push's node-centred FP32 copy of E and B (4 float4 + 1 float2 per
    // (i, k) node, mx * mz nodes), rebuilt before every split push
    float4 *nf4 = nullptr;
    float2 *nf2 = nullptr;
    int *d_err = nullptr;
    unsigned long long *d_count = nullptr;
    double *d_partial = nullptr;
    bool win_on = false;
    long long win_steps = 0, win_shift = 0;
    int win_nrp = 0, win_ppc = 0;
    unsigned long long win_seed = 0;
};

void cylm_params(const pictor_config &cfg, const PhysP &ph, CylmP &q);

}  // namespace pictor

int cylm_setup(pictor_ctx *c);
void cylm_release(pictor_ctx *c);
int cylm_upload_particles(pictor_ctx *c, long long n, const double *q, const double *x, const double *y,
                          const double *z, const double *u, const double *v, const double *w, const int *flv);
int cylm_sort(pictor_ctx *c);
int cylm_download_particles(pictor_ctx *c, long long cap, double *q, double *x, double *y, double *z, double *u,
                            double *v, double *w, int *flv, int *tile, long long *n_out);
int cylm_b_half(pictor_ctx *c);
int cylm_e_phase(pictor_ctx *c, bool curl, bool add);
int cylm_current_halo(pictor_ctx *c);
int cylm_field_halo(pictor_ctx *c, int which);
int cylm_reset_current(pictor_ctx *c);
int cylm_push(pictor_ctx *c, bool timed);
int cylm_exchange(pictor_ctx *c);
int cylm_inject(pictor_ctx *c, int nrp, int k0, int nzc, int ppc, unsigned long long seed, long long shift);
int cylm_laser(pictor_ctx *c, double a0, double omega0, double tau, double w0, double zc);
int cylm_window_after_step(pictor_ctx *c);
int cylm_diagnostics(pictor_ctx *c, double out[9]);
