// adjoint.h -- types and launchers of the 2-D adjoint engine (adjoint.hip).
#pragma once
#include "pictor_common.h"

namespace pictor {

// Engine particles in the caller's order.  cs = ci | cj << 16 | species << 31
// (0-based physical cell, nx, ny <= 32768), in-cell offsets and gamma*beta.
template <typename R>
struct EPart {
    uint32_t *cs;
    R *ox, *oy, *oz, *u, *v, *w;
};

struct AdjP {
    int nx, ny, mx, my;
    double c, sqc, cinv;
    double qm[2];    // flvrqm / 2 per species
    double qs[2];    // charge factor of the deposit per species
    double jscale;   // E -= jscale * J (J as deposited)
    double inv_d;    // 1 / cell size (= 1); a runtime factor so the advance rounds as the production push's
};

// node values of E and B: 4-vectors of R (x, y, z, 0) per node of the (mx, my) grid
template <typename R>
void launch_eng_nodes(const AdjP &ap, R *const E[3], R *const B[3], void *nE, void *nB, hipStream_t s);
template <typename R>
void launch_eng_push(const AdjP &ap, const void *nE, const void *nB, EPart<R> in, EPart<R> out, double *const J[3],
                     long long n, hipStream_t s);
template <typename R>
void launch_eng_adjoint(const AdjP &ap, const void *nE, const void *nB, EPart<R> st, double *const Eb[3],
                        EPart<R> adj, double *const gE[3], double *const gB[3], long long n, hipStream_t s);
void launch_node_to_stag(const AdjP &ap, double *const gE[3], double *const gB[3], double *const Eb[3],
                         double *const Bb[3], hipStream_t s);
template <typename R>
void launch_eng_load(const AdjP &ap, const double *const raw[6], const int *flv, EPart<R> out, long long n, int *err,
                     hipStream_t s);
template <typename R>
void launch_eng_store_grad(EPart<R> adj, double scale, double *du, double *dv, double *dw, long long n, hipStream_t s);
template <typename R>
void launch_eng_zero(EPart<R> a, long long n, hipStream_t s);
template <typename R>
void launch_seed_b(const R *b, double *d, long long n, hipStream_t s);

// FP32 engine on the tile bins (adjoint_fast.hip): reverse step of the split
// push over the bins (base, cnt) of the step's state st; dst[slot] = output
// slot of the particle (from the forward push); ain at output slots, aout at
// input slots (EPart's cs is unused); gE/gB: node adjoints (accumulated).
void launch_fa_adjoint(const GridP &g, const TileP &tp, const AdjP &ap, float *const E[3], float *const B[3],
                       double *const Eb[3], const long long *base, const int *cnt, PartSplit st, const int *dst,
                       EPart<float> ain, EPart<float> aout, double *const gE[3], double *const gB[3], hipStream_t s);
// over the occupied slots: mode 0 pos[id] = slot; mode 1 aout[slot] = ain[pos[id[slot]]];
// mode 2 du/dv/dw[id[slot]] = scale * ain.u/v/w[slot]
void launch_fa_slots(const TileP &tp, int mode, const long long *base, const int *cnt, const int *id, int *pos,
                     EPart<float> ain, EPart<float> aout, double scale, double *du, double *dv, double *dw,
                     hipStream_t s);

}  // namespace pictor
