// setup_weibel.cpp -- seeded synthetic input for the Weibel configurations.
//
// Restates what setup_weibel.F90:17-26 asks of InitPrtl (src/help_setup.F90:
// 43-105) for the benchmark configurations, with a counter-based generator
// instead of the compiler runtime's random_number (whose stream is specific
// to the Fortran runtime):
//   * two InitPrtl calls per run, each placing epc/2 positions per cell;
//   * both flavours are inserted at every position (InsertPrtl_PSP,
//     init_prtl.F90:138-169): flavour 1 (q = +1) then flavour 2 (q = -1);
//   * call 0 drifts flavour 1 by -drift and flavour 2 by +drift, call 1 the
//     opposite (DriftLeft / DriftRight swapped, setup_weibel.F90:23-24);
//   * thermal gamma*beta: a Gaussian of spread uth per component;
//   * drift added with AddDriftVel (prob_dist.F90:223-254).
// Local coordinates: x = 3 + i + r (xmin = 3), 2-D z = 1 + r (zmin = 1),
// as in InitPrtlFromPSP_RandomPosition (init_prtl.F90:177-221).
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "../../include/pictor_hip.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

inline uint64_t hash3(uint64_t seed, uint64_t idx, uint64_t stream)
{
    return mix64(mix64(seed ^ (idx * 0xD1342543DE82EF95ULL)) + stream * 0x9E3779B97F4A7C15ULL);
}

// uniform in [0, 1 - 2^-40]: 40 bits so that 3 + i + r stays below 3 + i + 1 exactly
inline double upos(uint64_t h) { return (double)(h >> 24) * (1.0 / 1099511627776.0); }
// uniform in (0, 1]
inline double uopen(uint64_t h) { return ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0); }

void gauss2(uint64_t seed, uint64_t n, uint64_t s0, double &g1, double &g2)
{
    const double r = std::sqrt(-2.0 * std::log(uopen(hash3(seed, n, s0))));
    const double t = 6.283185307179586 * uopen(hash3(seed, n, s0 + 1));
    g1 = r * std::cos(t);
    g2 = r * std::sin(t);
}

// AddDriftVel, prob_dist.F90:223-254
void add_drift(double &ug, double &vg, double &wg, double vx, double vy, double vz)
{
    const double drift = std::sqrt(vx * vx + vy * vy + vz * vz);
    double dbeta, dgamma;
    if (std::fabs(drift) < 1.0) {
        dbeta = drift;
        dgamma = 1.0 / std::sqrt((1.0 - drift) * (1.0 + drift));
    } else {
        dgamma = std::fabs(drift);
        dbeta = std::sqrt((drift - 1.0) * (drift + 1.0)) / drift;
    }
    double dir[3] = {vx, vy, vz};
    const double mag = std::sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    if (mag == 0) { dir[0] = dir[1] = dir[2] = 0; }
    else { dir[0] /= mag; dir[1] /= mag; dir[2] /= mag; }
    const double gamma = std::sqrt(1.0 + ug * ug + vg * vg + wg * wg);
    const double projn = ug * dir[0] + vg * dir[1] + wg * dir[2];
    ug = ug + (dgamma - 1) * projn * dir[0] + dgamma * dbeta * gamma * dir[0];
    vg = vg + (dgamma - 1) * projn * dir[1] + dgamma * dbeta * gamma * dir[1];
    wg = wg + (dgamma - 1) * projn * dir[2] + dgamma * dbeta * gamma * dir[2];
}

}  // namespace

extern "C" int pictor_setup_weibel(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, int32_t epc, double drift,
                                   double uth, uint64_t seed, int64_t cap, double *q, double *x, double *y, double *z,
                                   double *u, double *v, double *w, int32_t *flv, int64_t *n_out)
{
    if ((ndim != 2 && ndim != 3) || nx < 1 || ny < 1 || (ndim == 3 && nz < 1)) return PICTOR_ERR_ARG;
    if (epc < 2 || (epc % 2) != 0) return PICTOR_ERR_ARG;
    if (!(std::fabs(drift) < 1.0) || uth < 0) return PICTOR_ERR_ARG;
    const int64_t ncell = (int64_t)nx * ny * (ndim == 3 ? nz : 1);
    const int half = epc / 2;
    const int64_t np = ncell * 2 * epc;
    if (n_out) *n_out = np;
    if (cap < np) return PICTOR_ERR_CAPACITY;
    if (!q || !x || !y || !z || !u || !v || !w || !flv) return PICTOR_ERR_ARG;

    auto work = [&](int64_t c0, int64_t c1) {
        for (int64_t cl = c0; cl < c1; cl++) {
            const int i = (int)(cl % nx);
            const int64_t r = cl / nx;
            const int j = (int)(r % ny);
            const int k = (int)(r / ny);
            for (int call = 0; call < 2; call++)
                for (int rr = 0; rr < half; rr++) {
                    const uint64_t p = (uint64_t)((cl * 2 + call) * half + rr);
                    const double px = (double)(3 + i) + upos(hash3(seed, p, 0));
                    const double py = (double)(3 + j) + upos(hash3(seed, p, 1));
                    const double pz = (ndim == 3) ? (double)(3 + k) + upos(hash3(seed, p, 2)) : 1.0 + upos(hash3(seed, p, 2));
                    for (int s = 0; s < 2; s++) {
                        const uint64_t n = 2 * p + s;
                        double g1, g2, g3, g4;
                        gauss2(seed, n, 3, g1, g2);
                        gauss2(seed, n, 5, g3, g4);
                        double ug = uth * g1, vg = uth * g2, wg = uth * g3;
                        // call 0: flavour 1 drifts -x, flavour 2 +x; call 1 the opposite
                        const double sgn = ((call == 0) == (s == 0)) ? -1.0 : 1.0;
                        add_drift(ug, vg, wg, sgn * drift, 0.0, 0.0);
                        const int64_t o = (int64_t)n;
                        q[o] = (s == 0) ? 1.0 : -1.0;
                        x[o] = px; y[o] = py; z[o] = pz;
                        u[o] = ug; v[o] = vg; w[o] = wg;
                        flv[o] = s + 1;
                    }
                }
        }
    };
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    if (ncell < 4096) nt = 1;
    std::vector<std::thread> th;
    const int64_t chunk = (ncell + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const int64_t a = t * chunk, b = std::min<int64_t>(ncell, a + chunk);
        if (a >= b) break;
        th.emplace_back(work, a, b);
    }
    for (auto &t : th) t.join();
    return PICTOR_OK;
}
