/*
 * pictor_hip.h -- C ABI of the MI355X (gfx950) particle step of PICTOR.
 *
 * The reference calls zero-argument Fortran subroutines from MAIN.F90 in a
 * fixed order and keeps all state in module globals (src/vars.F90,
 * src/var_gpu.F90).  This ABI replaces those calls one for one; the state
 * lives in an opaque context instead of module globals.  All pointers are
 * host pointers in the reference layout:
 *
 *   fields    real(8) F(mx,my,mz), column-major, F(1,1,1) at offset 0,
 *             mx = nx+5, my = ny+5, mz = nz+5 (mz = 1 in 2-D),
 *             physical cells 3..m-3 (src/comm/subdomains/subdomains.F90:42-67)
 *   particles SoA q,x,y,z,u,v,w real(8) + flv integer(4), 1-based in Fortran,
 *             holes marked flv = 0 (src/vars.F90:43-45, 125-127)
 *
 * Every call returns 0 on success or a PICTOR_ERR_* code; the message of the
 * last failure is pictor_last_error().  One context = one HIP stream on the
 * current device; calls on one context must come from one host thread.
 */
#ifndef PICTOR_HIP_H
#define PICTOR_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PICTOR_OK 0
#define PICTOR_ERR_ARG 1         /* invalid argument or configuration        */
#define PICTOR_ERR_HIP 2         /* a HIP runtime call failed                */
#define PICTOR_ERR_BOUNDS 3      /* a particle left the allowed window       */
#define PICTOR_ERR_CAPACITY 12   /* particle storage exhausted; the reference's
                                    Abort(12), src/mem/particles_gpu.F90:21 */

#define PICTOR_FORM_FP64 0       /* reference SoA in FP64                    */
#define PICTOR_FORM_FP32_SPLIT 1 /* int32 cell + FP32 offset, FP32 push/field,
                                    fixed-point (int64) current              */

#define PICTOR_FDTD_YEE 0        /* src/maxwell/em_update.F90                */
#define PICTOR_FDTD_LEHE 1       /* src/maxwell/em_update_lehe.F90:8-155:
                                    extended B-update stencils, Yee E update */

/* Size of the static device array behind pictor_energy_noalloc, in doubles
 * (2 MiB per GPU): 2056 + 5 per particle tile. */
#define PICTOR_NOALLOC_SCRATCH_DOUBLES 262144

#define PICTOR_GEOM_CARTESIAN 0  /* periodic x, y (, z)                      */
#define PICTOR_GEOM_CYL 1        /* the -Dcyl build with the axis (inc_axis):
                                    x = r (index i at r = i - 3.5), y = theta
                                    (ny even, 2 pi periodic), z periodic; FP64
                                    form; src/cyl/cyl_em_update.F90,
                                    cyl_movdep.F90; conducting wall beyond
                                    r = nx - 1/2                              */
#define PICTOR_GEOM_CYL_MODAL 2  /* no reference counterpart: the cylinder in
                                    azimuthal modes m = 0, 1 (ndim 3, nx = Nr,
                                    ny = 2 modes, nz = Nz), the reference's r-z
                                    staggering, axis rules and push; fields
                                    (nx+5, 3, nz+5) with slot 0 = m 0, 1 = m 1
                                    cos, 2 = m 1 sin; particles y = theta in
                                    radians; radial cell cfg.dr axial cells;
                                    FP64, Yee; periodic z, or open z ends while
                                    the moving window is on */

/* Run configuration: the compile-time parameters of parameters.F90:5-87 and
 * src/vars.F90:10-18 as a runtime struct.  grid_dx = grid_dy = grid_dz = 1,
 * EMspeedBoost = 1, periodic faces on every side until pictor_set_bc_fld /
 * pictor_set_bc_prtl set a face. */
typedef struct pictor_config {
    int32_t ndim;          /* 2 (-DtwoD) or 3                                */
    int32_t nx, ny, nz;    /* physical cells; nz ignored in 2-D              */
    int32_t epc;           /* particles per cell per species (parameters.F90:17) */
    int32_t form;          /* PICTOR_FORM_*                                  */
    int32_t tile;          /* tile edge in cells (0 = choose); must divide nx, ny (nz) */
    int32_t fdtd;          /* PICTOR_FDTD_*: the B half step of em_update.F90 or em_update_lehe.F90 */
    int32_t geometry;      /* PICTOR_GEOM_* (the reference's -Dcyl switch)   */
    int32_t curr_filter;   /* passes of the 1-2-1 current filter per step (parameters.F90:22,
                              smoothen_current current.F90:19-30); Cartesian only */
    double c;              /* parameters.F90:20 */
    double c_ompe;         /* parameters.F90:19 */
    double g0;             /* parameters.F90:16 */
    double mi_me;          /* parameters.F90:18 */
    double capacity_factor;/* particle storage per tile / particles at layout time (0 = 1.75) */
    double dr;             /* PICTOR_GEOM_CYL_MODAL: radial cell in axial cells (0 = 1) */
} pictor_config;

typedef struct pictor_ctx pictor_ctx;

/* ---- life cycle: Initialisations, src/MAIN.F90:338-395; InitScale,
 *      src/init/initialise.F90:44-53; InitAll_gpu, src/init/initialise_gpu.F90:9-101 */
int pictor_init(const pictor_config *cfg, pictor_ctx **out);
int pictor_finalize(pictor_ctx *ctx);
const char *pictor_last_error(void);
/* Derived scalars: {qi, qe, masse, massi, ompe, fldc, xlen, ylen, zlen, mx, my, mz,
 * tile edge, number of tiles, re-layouts after a bin overflow since init,
 * particles those re-layouts moved from the spill buffer}. */
int pictor_get_scalars(pictor_ctx *ctx, double out[16]);

/* ---- host <-> device copies in the reference layout (the CUDA Fortran array
 *      assignments of src/mem/particles_gpu.F90:28-72, src/maxwell/fields_gpu.F90:19-56) */
int pictor_upload_fields(pictor_ctx *ctx, const double *ex, const double *ey, const double *ez,
                         const double *bx, const double *by, const double *bz);
int pictor_download_fields(pictor_ctx *ctx, double *ex, double *ey, double *ez,
                           double *bx, double *by, double *bz,
                           double *jx, double *jy, double *jz);
/* Holes (flv = 0) are dropped.  The FP32 split form keeps the charge implicit when every
 * particle carries its flavour's charge; otherwise (weighted particles, q = wtp * FlvrCharge,
 * init_prtl.F90:161) each particle record gains an FP32 charge weight q / FlvrCharge (4 B
 * read and written per particle-step more); a charge of the wrong sign is refused. */
int pictor_upload_particles(pictor_ctx *ctx, int64_t n, const double *q, const double *x,
                            const double *y, const double *z, const double *u, const double *v,
                            const double *w, const int32_t *flv);
int pictor_particle_count(pictor_ctx *ctx, int64_t *n);
/* Output in tile order; tile (optional, may be NULL) receives each particle's tile id. */
int pictor_download_particles(pictor_ctx *ctx, int64_t cap, double *q, double *x, double *y,
                              double *z, double *u, double *v, double *w, int32_t *flv,
                              int32_t *tile, int64_t *n_out);
/* The same two calls with the reference's per-particle tag (tagp, src/vars.F90:43 and the
 * particle type :54-79; set by GetTag, init_prtl.F90:160): an int32 carried unchanged with
 * every particle through pushes, re-layouts, spills and restart files, so that a caller can
 * re-associate the particles after a download in tile order (Cartesian geometry, both forms).
 * A tagged upload makes the context carry tags until an untagged upload; downloading tags
 * from an untagged context is an error.  The tag costs 4 B read and 4 B written per
 * particle-step. */
/* SetFilteredEfldGPU, src/maxwell/fields_gpu.F90:70-90 (the E field the reference's test
 * particles feel): E through `passes` rounds (nMoverEMfilter) of the 1-2-1 filter of
 * smoothen_current, then the periodic halo copy, written to fx, fy, fz (mx*my*mz each,
 * the reference layout).  Cartesian geometry without boundary faces. */
int pictor_filtered_efield(pictor_ctx *ctx, int32_t passes, double *fx, double *fy, double *fz);
int pictor_upload_particles_tagged(pictor_ctx *ctx, int64_t n, const double *q, const double *x,
                                   const double *y, const double *z, const double *u, const double *v,
                                   const double *w, const int32_t *flv, const int32_t *tag);
int pictor_download_particles_tagged(pictor_ctx *ctx, int64_t cap, double *q, double *x, double *y,
                                     double *z, double *u, double *v, double *w, int32_t *flv,
                                     int32_t *tile, int32_t *tag, int64_t *n_out);

/* ---- restart files: SaveRestartData, src/io/savedata.F90:73-141 (writers :940-1210),
 *      and the reload, src/io/reload.F90:17-338 (restart branch src/MAIN.F90:35-39).
 * One binary file (no HDF5 here) with the reference's restart content: the run
 * configuration and step t (param_<proc>_<t>), E and B in the form's own precision
 * (fld_<proc>_<t>; the split form's FP32 B and FP32 (hi, lo) pair of E, so that a
 * restarted split run continues bit for bit), and the particles in the reference SoA
 * q, x, y, z, u, v, w, flv (prtl_<proc>_<t>), which holds the split record exactly.
 * The modal cylinder's moving window and sort schedule are saved too.  J is not saved
 * (each step resets it).  The file is written to path.tmp and renamed when complete.
 * pictor_load_state needs a context of the same configuration (the tile edge may
 * differ) and returns the saved t. */
int pictor_save_state(pictor_ctx *ctx, const char *path, int64_t t);
int pictor_load_state(pictor_ctx *ctx, const char *path, int64_t *t);

/* ---- output files: save_total_energy, src/io/savedata.F90:856-890 (energy_<t>),
 *      and save_field_collective, savedata.F90:356-497 with CollectFld,
 *      savedata_routines.F90:207-261 (fld_<t>).  The reference writes HDF5; this
 * image has none, so each file is a NumPy .npz archive (uncompressed zip of .npy
 * arrays, Fortran order) with the reference's dataset names:
 *   energy: Benergy(1), Eenergy(1), PrtlEnergy(Nflvr), FP64 (CalcEnergy);
 *   fields: Ex, Ey, Ez, Bx, By, Bz (+ Jx, Jy, Jz if save_current, after the
 *           current's halo copy) as FP32 (fdatax, fdatay, fdataz): the nodes
 *           3, 3 + res, ... 3 + res * int(n / res) of each axis over the whole box,
 *           the staggered components averaged onto the node as CollectFld does.
 * The field file is written for the Cartesian geometry. */
int pictor_save_energy(pictor_ctx *ctx, const char *path);
int pictor_save_fields(pictor_ctx *ctx, const char *path, int32_t res, int32_t save_current);

/* ---- the step, one call per MAIN.F90 phase ---- */
/* UpdateBfieldHalf + SendRecvFlds(B): em_update.F90:41-77, comm_fld.F90:75-91 */
int pictor_update_b_half(pictor_ctx *ctx);
/* MoveDepositPrtl: movdep.F90:12-258 (GPU: movdep_gpu.F90:14-49).  Also bins
 * every particle into its destination tile for the next step. */
int pictor_move_deposit(pictor_ctx *ctx);
/* ExchangePrtl + AppendParticles on one periodic rank: comm_prtl.F90:11-65,
 * loadprtlout.F90:13-181.  The wrap itself happens inside pictor_move_deposit;
 * this call commits the new tile bins and re-lays out storage on overflow. */
int pictor_exchange_particles(pictor_ctx *ctx);
/* UpdateEfield + SendRecvFlds(E): em_update.F90:9-39 */
int pictor_update_e(pictor_ctx *ctx);
/* UpdateCurrentAllFaces: comm_fld.F90:15-72 (ghost current added into owners) */
int pictor_current_halo_add(pictor_ctx *ctx);
/* smoothen_current: current.F90:19-30 (MAIN.F90:158-162), cfg.curr_filter passes of
 * the 1-2-1 moving average over the physical cells of the periodic box, along x,
 * y (and z) in turn.  The physical cells match the reference bit for bit in the
 * FP64 form; J's ghost layers are left as they were (nothing reads them before
 * the next pictor_reset_current).  No-op for curr_filter = 0. */
int pictor_smoothen_current(pictor_ctx *ctx);
/* AddCurrent + SendRecvFlds(E): em_update.F90:79-102 */
int pictor_add_current(pictor_ctx *ctx);
/* SendRecvFlds on E (which = 0) or B (which = 1): comm_fld.F90:75-130 */
int pictor_field_halo(pictor_ctx *ctx, int32_t which);
/* ResetCurrent: current.F90:13-17 */
int pictor_reset_current(pictor_ctx *ctx);
/* ReorderPrtlArr: mem_prtl.F90:166-301 -- counting sort into tiles */
int pictor_sort_tiles(pictor_ctx *ctx);
/* nsteps whole time steps of MAIN.F90:42-284 (J reset at the step start),
 * with the boundary faces below where they are set. */
int pictor_step(pictor_ctx *ctx, int32_t nsteps);

/* ---- boundary faces, src/bc (Cartesian geometry, both forms, 2-D and 3-D).
 * axis 0 = x, 1 = y, 2 = z (z only in 3-D); side 0 = lower, 1 = upper;
 * position is global (0 .. n along the axis; local coordinate = position + 3).
 * Setting any face of an axis makes the axis non-periodic on both sides
 * (PeriodicOff, subdomains.F90:508-547): its halo exchanges stop, its ghost
 * cells evolve under the Yee update (cells 2..m-1, em_update.F90), and a
 * particle that leaves through its faces is dropped.  Not built: inflow
 * ('iflw'), attenuating ('attn') and curved faces, the open particle face,
 * the current filter on a non-periodic axis (curr_filter must be 0), the
 * adjoint with faces set. */
#define PICTOR_BC_PERIODIC 0     /* 'prdc' */
#define PICTOR_BC_COND 1         /* field 'cond': CondBC_Fld_*, bc_reflect.F90:11-65
                                    (x, y faces): tangential E zero beyond it */
#define PICTOR_BC_OPEN 2         /* field 'open': OpenBC_Fld_*, bc.F90:410-469 (x faces) */
#define PICTOR_BC_PML 3          /* field 'pml': UpdateMagFld_PML / UpdateElcFld_PML,
                                    bc_pml.F90 (any face; thickness = layer depth) */
#define PICTOR_BC_REFL 4         /* particle 'refl': RefBC_Prtl_*, bc_reflect.F90:76-237
                                    (x, y faces), with the current folded back */
#define PICTOR_BC_REMV 5         /* particle 'remv': RemovePrtl_*, bc_open.F90 (any face) */
/* SetBC_Fld, bc.F90:23-50 (PML: SetPMLParameters, bc_pml.F90:39-51; the layer
 * must fit inside the box, CheckAttnBCPositioning bc.F90:364-401) */
int pictor_set_bc_fld(pictor_ctx *ctx, int32_t axis, int32_t side, int32_t type, double position,
                      double thickness);
/* SetBC_Prtl, bc.F90:52-63 */
int pictor_set_bc_prtl(pictor_ctx *ctx, int32_t axis, int32_t side, int32_t type, double position);
/* SetBoundarySpeed, bc.F90:210-217: both positions of the face move by speed * c
 * at the end of every step (updateBC_Pos, bc.F90:342-348) */
int pictor_set_boundary_speed(pictor_ctx *ctx, int32_t axis, int32_t side, double speed);
/* UpdateMagFld_PML (which = 1, after the second B half step, MAIN.F90:104) or
 * UpdateElcFld_PML (which = 0, after UpdateEfield, MAIN.F90:109); no-op
 * without a PML face */
int pictor_pml(pictor_ctx *ctx, int32_t which);
/* EnforceBC_Final, bc.F90:289-340 (MAIN.F90:209): conducting and open field
 * faces, then updateBC_Pos.  The reflecting and removing particle faces of
 * EnforceBC_PostMovDep (MAIN.F90:77) run inside pictor_move_deposit. */
int pictor_enforce_bc_final(pictor_ctx *ctx);
int pictor_synchronize(pictor_ctx *ctx);

/* ---- diagnostics: CalcEnergy/CalcDivE, savedata_routines.F90:265-355 (FP64).
 * out[0] B energy, out[1] E energy, out[2..3] kinetic energy of flavours 1, 2,
 * out[4..6] sum m*u, m*v, m*w, out[7] max|div E - rho|, out[8] max|rho|.   */
int pictor_diagnostics(pictor_ctx *ctx, double out[9]);

/* ---- kernel timers (HIP events on the context stream) ----
 * id 0 = push/deposit kernel, 1 = field kernels, 2 = whole step.
 * Returns the summed milliseconds and the number of launches since reset. */
int pictor_timer_reset(pictor_ctx *ctx);
int pictor_timer_get(pictor_ctx *ctx, int32_t id, double *total_ms, int64_t *count);
/* Bytes of device memory owned by the context. */
int pictor_memory_bytes(pictor_ctx *ctx, int64_t *bytes);

/* ---- synthetic input (setup_weibel.F90:17-26 restated for the benchmark
 * configurations): epc/2 uniformly random positions per cell per InitPrtl call,
 * two calls with opposite drifts, both flavours at every position
 * (init_prtl.F90:138-221), Gaussian thermal gamma*beta of spread uth per
 * component, drift added with AddDriftVel (prob_dist.F90:223-254).
 * Arrays must hold 2*epc*nx*ny(*nz) entries; counter-based RNG (seed).      */
int pictor_setup_weibel(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, int32_t epc,
                        double drift, double uth, uint64_t seed, int64_t cap,
                        double *q, double *x, double *y, double *z, double *u, double *v,
                        double *w, int32_t *flv, int64_t *n_out);

/* ---- adjoint of the 2-D particle step (no counterpart in the reference:
 * the hand-written adjoint of push + deposit + FDTD).  Objective
 * L = 1/2 sum B^2 over the physical cells after nsteps steps, started from
 * the context's current E and B and the given particles (caller's order,
 * reference SoA, q = flavour charge).  The context's form selects the
 * arithmetic: FP32 split (hardware rsq/rcp, fixed-point current) or FP64.
 * pictor_adjoint_forward runs the steps keeping a checkpoint every
 * `segment` steps and returns L; pictor_adjoint_backward recomputes each
 * segment from its checkpoint and runs the reverse steps, writing
 * scale * dL/du, dL/dv, dL/dw of the initial momenta (any may be NULL).
 * The context's own particles and fields are not touched.
 * pictor_adjoint_stats: out[0] forward ms, out[1] recompute ms, out[2]
 * reverse-step ms, out[3] device bytes of the engine, out[4] nsteps,
 * out[5] segment, out[6] checkpoints, out[7] L, out[8] engine (0 tile-sorted,
 * 1 unsorted), out[9] forward passes repeated with larger tile bins. */
int pictor_adjoint_forward(pictor_ctx *ctx, int64_t n, const double *q, const double *x,
                           const double *y, const double *z, const double *u, const double *v,
                           const double *w, const int32_t *flv, int32_t nsteps, int32_t segment,
                           double *objective);
int pictor_adjoint_backward(pictor_ctx *ctx, double scale, double *du, double *dv, double *dw);
/* The same two calls with the particle arrays (forward) and the gradient
 * arrays (backward) in device memory of the context's GPU (for example torch
 * tensors on that device): nothing crosses PCIe but L.  The call waits for
 * its own stream only, so the caller finishes writing the inputs (and reads
 * the outputs) on its own streams around it. */
int pictor_adjoint_forward_device(pictor_ctx *ctx, int64_t n, const double *q, const double *x,
                                  const double *y, const double *z, const double *u, const double *v,
                                  const double *w, const int32_t *flv, int32_t nsteps, int32_t segment,
                                  double *objective);
int pictor_adjoint_backward_device(pictor_ctx *ctx, double scale, double *du, double *dv, double *dw);
int pictor_adjoint_stats(pictor_ctx *ctx, double out[10]);
/* Engine of the FP32 split form: 0 = tile-sorted (default; the particles in
 * the tile bins of pictor_move_deposit, every forward step the production
 * push, the reverse step over the same bins with LDS gather and LDS
 * accumulation of the field adjoint; tile edge 8, 16 or 32), 1 = unsorted
 * (one thread per particle, global atomics).  The FP64 form always uses the
 * unsorted engine.  Takes effect at the next pictor_adjoint_forward. */
int pictor_adjoint_engine(pictor_ctx *ctx, int32_t engine);
int pictor_adjoint_release(pictor_ctx *ctx);

/* Push scheme of the cylindrical grid (PICTOR_GEOM_CYL): 0 = Boris, MoveDepositPrtl of
 * src/cyl/cyl_movdep.F90:14-465 (the algorithm of the reference's GPU kernel,
 * cyl_movdep_gpu.F90:69-340; the default); 1 = Strang splitting, MoveDepositPrtl of
 * src/cyl/cyl_movdep_strang.F90:14-584, which the reference's CPU -Dcyl build links
 * (Makefile:104): half kick, move, second gather at the new point, half kick, axis
 * sub-steps growing by 1.7.  Its main loop as written keeps the first gather's scaled
 * fields in the second gather's sums and moves z before the kick; this follows its
 * MovePrtlAxis, which does neither. */
int pictor_cyl_push_scheme(pictor_ctx *ctx, int32_t scheme);

/* ---- the azimuthal-mode cylinder (PICTOR_GEOM_CYL_MODAL; no counterpart in
 * the reference, which has no modes, laser or moving window) ----
 * Cold electrons of uniform density (weight epc per unit volume): ppc per cell
 * in the radial cells 4..4+nr_plasma-1 and axial cells k0..k0+nzc-1, (r, z)
 * uniform in the cell, ppc evenly spaced angles with a random offset per cell;
 * counter-based (seed), so a cell always gets the same particles. */
int pictor_cyl_plasma(pictor_ctx *ctx, int32_t nr_plasma, int32_t k0, int32_t nzc, int32_t ppc, uint64_t seed);
/* Adds a Gaussian pulse, linearly polarised along x and moving along +z, at
 * focus: E_x = a0 c w0dt exp(-r^2/w0^2 - (z - zc)^2/(c tau)^2) cos(k0 (z - zc)),
 * B_y the same profile half a step earlier (at z + c/2: between steps B is half
 * a step behind E), into the m = 1 slots: a purely forward pulse.  omega0 in
 * omega_p, tau in 1/omega_p (c tau is the 1/e half length of the field), w0 in
 * c/omega_p, zc in axial cells. */
int pictor_cyl_laser(pictor_ctx *ctx, double a0, double omega0, double tau, double w0, double zc);
/* Sorting of the modal particles: pictor_sort_tiles (ReorderPrtlArr, mem_prtl.F90:166-301)
 * orders them by cell, k-major with i fastest (the loader's order), and drops holes, with
 * a stable radix sort; pictor_cyl_sort_period(ctx, n) makes pictor_step do that every n
 * steps (0, the default: never).  Strong wakes scatter the particles of neighbouring
 * array slots over many cells; the push's deposit window needs them in cell order. */
int pictor_cyl_sort_period(pictor_ctx *ctx, int32_t period);
/* Moving window: after every fused step the fields and particles move back by
 * one axial cell whenever the window (speed c) has advanced one; the front
 * cells restart empty and receive fresh plasma (nr_plasma, ppc, seed as
 * pictor_cyl_plasma).  While it is on, z is open instead of periodic: the z
 * ghost cells hold no field, the current deposited in them is dropped, and a
 * particle that leaves through either end becomes a hole.  on = 0 stops it
 * and makes z periodic again. */
int pictor_moving_window(pictor_ctx *ctx, int32_t on, int32_t nr_plasma, int32_t ppc, uint64_t seed);

/* ---- allocation-free diagnostics (no counterpart in the reference) ----
 * For callers that cannot allocate device memory (code under stream capture,
 * serving loops, harnesses that forbid hipMalloc): out[0..6] of
 * pictor_diagnostics (field energies, kinetic energy per flavour, momentum sums),
 * bitwise equal to it, computed with the same kernels into a static device array
 * of PICTOR_NOALLOC_SCRATCH_DOUBLES doubles per GPU instead of the context's
 * buffer; it calls no hipMalloc or hipFree.  Cartesian geometry; a grid whose
 * tiles need more than the array holds (2056 + 5 per tile) is PICTOR_ERR_ARG.
 * Every call on one GPU shares that GPU's array, so calls that overlap on one
 * GPU (two host threads, or two contexts on different streams) race. */
int pictor_energy_noalloc(pictor_ctx *ctx, double out[7]);

/* ---- measurement (no counterpart in the reference) ----
 * Atomic throughput on the current device: every lane of 4 workgroups per CU
 * adds `iters` times at pseudo-random addresses.  out[0..5]: lane-ops per
 * second of LDS FP32, LDS FP64, LDS u32, LDS u64 atomics (32 KiB window) and
 * global FP64, FP32 atomics (4 Mi-element array); out[8..13] the same per CU
 * per clock; out[6] CUs, out[7] clock (Hz). */
int pictor_probe_atomics(int32_t iters, double out[16]);
/* The stream ceiling of the particle record's access pattern on the current
 * device: six FP32 arrays of n elements read and six written (48 B per
 * element; n a multiple of 512, at most 2^28), with one 4-byte element per
 * lane (out[0]), two lane-contiguous 4-byte elements per lane (out[1]), a
 * float2 (out[2]) and a float4 (out[3]) per lane: bytes per second; out[4..7]
 * the milliseconds per pass. */
int pictor_probe_stream(int64_t n, double out[8]);

#ifdef __cplusplus
}
#endif
#endif
