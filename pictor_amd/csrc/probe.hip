// probe.hip -- atomic throughput of the current window's accumulation paths on
// the running GPU: LDS atomics (FP32, FP64, 32/64-bit integer) against global
// FP64/FP32 atomics into a field-sized array.  The push and the adjoint
// reverse kernel pick their accumulator types from these rates.
#include <hip/hip_runtime.h>

#include "../../include/pictor_hip.h"

namespace {

constexpr int kLdsWords = 8192;          // 32 KiB of LDS per workgroup
constexpr long long kGlobalN = 1 << 22;  // 4 Mi elements: a 2048^2 field array

enum ProbeKind : int { kLdsF32 = 0, kLdsF64, kLdsU32, kLdsU64, kGlbF64, kGlbF32, kNumKinds };

// each lane adds to a pseudo-random address (LCG) every iteration; all
// addresses are masked into the arrays
template <int KIND>
__global__ __launch_bounds__(256) void k_probe(int iters, unsigned seed, double *gd, float *gf, float *out)
{
    __shared__ float s[kLdsWords];
    __shared__ double sd[kLdsWords / 2];
    for (int i = threadIdx.x; i < kLdsWords; i += 256) s[i] = 0.0f;
    for (int i = threadIdx.x; i < kLdsWords / 2; i += 256) sd[i] = 0.0;
    __syncthreads();
    unsigned h = seed ^ (threadIdx.x * 2654435761u) ^ (blockIdx.x * 40503u);
    for (int it = 0; it < iters; it++) {
        h = h * 1664525u + 1013904223u;
        const unsigned a = h >> 8;
        if (KIND == kLdsF32) atomicAdd(&s[a & (kLdsWords - 1)], 1.0f);
        if (KIND == kLdsF64) atomicAdd(&sd[a & (kLdsWords / 2 - 1)], 1.0);
        if (KIND == kLdsU32) atomicAdd((unsigned *)&s[a & (kLdsWords - 1)], 1u);
        if (KIND == kLdsU64) atomicAdd((unsigned long long *)&sd[a & (kLdsWords / 2 - 1)], 1ull);
        if (KIND == kGlbF64) unsafeAtomicAdd(&gd[(long long)(h & (unsigned)(kGlobalN - 1))], 1.0);
        if (KIND == kGlbF32) unsafeAtomicAdd(&gf[(long long)(h & (unsigned)(kGlobalN - 1))], 1.0f);
    }
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = s[5] + (float)sd[7];
}

template <int KIND>
hipError_t run_kind(int nblk, int iters, double *gd, float *gf, float *out, hipEvent_t e0, hipEvent_t e1,
                    float &ms)
{
    hipLaunchKernelGGL(k_probe<KIND>, dim3(nblk), dim3(256), 0, 0, iters, 1u, gd, gf, out);   // warm-up
    hipError_t e;
    if ((e = hipEventRecord(e0, 0)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_probe<KIND>, dim3(nblk), dim3(256), 0, 0, iters, 7u, gd, gf, out);
    if ((e = hipEventRecord(e1, 0)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
    return hipEventElapsedTime(&ms, e0, e1);
}

}  // namespace

extern "C" int pictor_probe_atomics(int32_t iters, double out[16])
{
    if (iters < 1 || iters > (1 << 20) || !out) return PICTOR_ERR_ARG;
    int dev = 0;
    hipDeviceProp_t p;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&p, dev) != hipSuccess) return PICTOR_ERR_HIP;
    const int ncu = p.multiProcessorCount;
    const double clk = p.clockRate * 1e3;   // Hz
    const int nblk = 4 * ncu;
    double *gd = nullptr;
    float *gf = nullptr, *o = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PICTOR_OK;
    if (hipMalloc(&gd, kGlobalN * sizeof(double)) != hipSuccess || hipMalloc(&gf, kGlobalN * sizeof(float)) != hipSuccess ||
        hipMalloc(&o, nblk * sizeof(float)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = PICTOR_ERR_HIP;
    } else if (hipMemset(gd, 0, kGlobalN * sizeof(double)) != hipSuccess ||
               hipMemset(gf, 0, kGlobalN * sizeof(float)) != hipSuccess) {
        rc = PICTOR_ERR_HIP;
    } else {
        float ms[kNumKinds] = {};
        hipError_t e = run_kind<kLdsF32>(nblk, iters, gd, gf, o, e0, e1, ms[0]);
        if (e == hipSuccess) e = run_kind<kLdsF64>(nblk, iters, gd, gf, o, e0, e1, ms[1]);
        if (e == hipSuccess) e = run_kind<kLdsU32>(nblk, iters, gd, gf, o, e0, e1, ms[2]);
        if (e == hipSuccess) e = run_kind<kLdsU64>(nblk, iters, gd, gf, o, e0, e1, ms[3]);
        if (e == hipSuccess) e = run_kind<kGlbF64>(nblk, iters, gd, gf, o, e0, e1, ms[4]);
        if (e == hipSuccess) e = run_kind<kGlbF32>(nblk, iters, gd, gf, o, e0, e1, ms[5]);
        if (e != hipSuccess) {
            rc = PICTOR_ERR_HIP;
        } else {
            const double ops = (double)nblk * 256.0 * iters;
            for (int k = 0; k < kNumKinds; k++) {
                out[k] = ops / (ms[k] * 1e-3);                    // lane-ops per second, whole GPU
                out[8 + k] = out[k] / ncu / clk;                  // per CU per clock
            }
            out[6] = ncu;
            out[7] = clk;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (gd) (void)hipFree(gd);
    if (gf) (void)hipFree(gf);
    if (o) (void)hipFree(o);
    return rc;
}
