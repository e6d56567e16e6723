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

}  // namespace pictor
