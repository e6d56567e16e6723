// noalloc.hip -- energies and momenta without any device allocation.
//
// pictor_energy_noalloc computes out[0..6] of pictor_diagnostics (field
// energies, kinetic energy per flavour, momentum sums) with the same kernels,
// but its partial sums live in `scratch`, a file-scope device array, instead
// of the context's buffer, and the final sums are taken on the device in the
// host loop's order, so the results are bitwise those of pictor_diagnostics.
// It is for callers that may not call hipMalloc (code under stream capture,
// serving loops, harnesses that forbid allocation).
//
// Each GPU holds its own copy of `scratch` until the process ends.  Host code
// never uses the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, and is kept in g_scratch[dev].  Every
// call on one GPU shares that GPU's array, so calls that overlap on one GPU
// (two host threads, or two contexts on different streams) race.
#include <hip/hip_runtime.h>
#include <mutex>

#include "../../include/pictor_hip.h"
#include "ctx.h"
#include "pictor_common.h"

using namespace pictor;

namespace {

constexpr int kFieldBlocks = 1024;   // blocks of the field-energy reduction (as pictor_diagnostics)
constexpr int kOut = 8;              // results at the front of the array
constexpr long long kFieldOff = kOut;
constexpr long long kPartOff = kFieldOff + 2 * kFieldBlocks;

__device__ __align__(256) double scratch[PICTOR_NOALLOC_SCRATCH_DOUBLES];

constexpr int kMaxDev = 64;
double *g_scratch[kMaxDev] = {};
std::mutex g_mu;

// Device address of this GPU's scratch, looked up once per device.
int scratch_ptr(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDev) return fail(PICTOR_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PICTOR_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = (double *)p;
    }
    *out = g_scratch[dev];
    return PICTOR_OK;
}

// One thread adds the partials in the order of pictor_diagnostics' host loops:
// B and E energy over the field blocks, then the five particle sums over tiles.
__global__ void k_final_sums(double *s, int ntiles)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double *fp = s + kFieldOff;
    double be = 0, ee = 0;
    for (int b = 0; b < kFieldBlocks; b++) { be += fp[2 * b]; ee += fp[2 * b + 1]; }
    const double *pp = s + kPartOff;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int t = 0; t < ntiles; t++)
        for (int a = 0; a < 5; a++) acc[a] += pp[(long long)t * 5 + a];
    s[0] = be * 0.5;
    s[1] = ee * 0.5;
    for (int a = 0; a < 5; a++) s[2 + a] = acc[a];
}

}  // namespace

extern "C" int pictor_energy_noalloc(pictor_ctx *c, double out[7])
{
    if (!c) return fail(PICTOR_ERR_ARG, "null context");
    if (!out) return fail(PICTOR_ERR_ARG, "null output");
    HIPTRY(hipSetDevice(c->dev));
    if (c->cyl) return fail(PICTOR_ERR_ARG, "pictor_energy_noalloc: Cartesian geometry only");
    if (c->pending_swap) return fail(PICTOR_ERR_ARG, "diagnostics between move_deposit and exchange_particles");
    const long long need = kPartOff + 5LL * c->tp.ntiles;
    if (need > PICTOR_NOALLOC_SCRATCH_DOUBLES)
        return fail(PICTOR_ERR_ARG, "pictor_energy_noalloc: " + std::to_string(c->tp.ntiles) +
                                        " tiles need " + std::to_string(need) + " scratch doubles, the maximum is " +
                                        std::to_string(PICTOR_NOALLOC_SCRATCH_DOUBLES));
    double *s = nullptr;
    int rc = scratch_ptr(c->dev, &s);
    if (rc) return rc;
    if (c->form == PICTOR_FORM_FP64) {
        double *E[3], *B[3];
        for (int i = 0; i < 3; i++) { E[i] = (double *)c->E[i]; B[i] = (double *)c->B[i]; }
        launch_field_energy<double>(c->g, E, B, s + kFieldOff, kFieldBlocks, c->stream);
        launch_particle_diag_f64(c->g, c->ph, c->tp, c->pf[c->cur], c->d_base, c->d_cnt[c->cur], s + kPartOff,
                                 nullptr, c->stream);
    } else {
        float *E[3], *B[3];
        for (int i = 0; i < 3; i++) { E[i] = (float *)c->E[i]; B[i] = (float *)c->B[i]; }
        launch_field_energy<float>(c->g, E, B, s + kFieldOff, kFieldBlocks, c->stream, c->Elo);
        launch_particle_diag_split(c->g, c->ph, c->tp, c->ps[c->cur], c->d_base, c->d_cnt[c->cur], s + kPartOff,
                                   nullptr, c->stream);
    }
    HIPTRY(hipGetLastError());
    hipLaunchKernelGGL(k_final_sums, dim3(1), dim3(64), 0, c->stream, s, c->tp.ntiles);
    HIPTRY(hipGetLastError());
    HIPTRY(hipMemcpyAsync(out, s, 7 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPTRY(hipStreamSynchronize(c->stream));
    return PICTOR_OK;
}
