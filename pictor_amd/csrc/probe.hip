// probe.hip -- atomic throughput of the current window's accumulation paths on
// the running GPU: LDS atomics (FP32, FP64, 32/64-bit integer) against global
// FP64/FP32 atomics into a field-sized array.  The push and the adjoint
// reverse kernel pick their accumulator types from these rates.  And the
// stream ceiling of the particle record's access pattern: six FP32 arrays read
// and six written, with one, two or four elements per lane.
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

// six arrays in, six out (the 2-D split record, 24 B each way per element); V
// elements per lane: 1 (4-byte accesses), 2 (two lane-contiguous 4-byte
// accesses 256 elements apart), 20 (float2), 40 (float4)
struct Six {
    float *a[6];
};

template <int V>
__global__ __launch_bounds__(256) void k_stream(Six s, Six d, long long n)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (V == 1) {
        if (t >= n) return;
#pragma unroll
        for (int c = 0; c < 6; c++) d.a[c][t] = s.a[c][t] + 1.0f;
    } else if (V == 2) {
        const long long b = t >> 8, l = t & 255;
        if (b * 512 + 511 >= n) return;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            const float x0 = s.a[c][b * 512 + l], x1 = s.a[c][b * 512 + 256 + l];
            d.a[c][b * 512 + l] = x0 + 1.0f;
            d.a[c][b * 512 + 256 + l] = x1 + 1.0f;
        }
    } else if (V == 20) {
        if (2 * t + 1 >= n) return;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            float2 v = reinterpret_cast<const float2 *>(s.a[c])[t];
            v.x += 1.0f; v.y += 1.0f;
            reinterpret_cast<float2 *>(d.a[c])[t] = v;
        }
    } else {
        if (4 * t + 3 >= n) return;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            float4 v = reinterpret_cast<const float4 *>(s.a[c])[t];
            v.x += 1.0f; v.y += 1.0f; v.z += 1.0f; v.w += 1.0f;
            reinterpret_cast<float4 *>(d.a[c])[t] = v;
        }
    }
}

template <int V>
hipError_t run_stream(const Six &s, const Six &d, long long n, int reps, hipEvent_t e0, hipEvent_t e1, float &ms)
{
    const long long per = V == 1 ? 1 : (V == 2 || V == 20 ? 2 : 4);
    const dim3 grid((unsigned)((n / per + 255) / 256));
    hipError_t e;
    hipLaunchKernelGGL(k_stream<V>, grid, dim3(256), 0, 0, s, d, n);   // warm-up
    if ((e = hipEventRecord(e0, 0)) != hipSuccess) return e;
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_stream<V>, grid, dim3(256), 0, 0, s, d, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(e1, 0)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) return e;
    if ((e = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) return e;
    ms /= reps;
    return hipSuccess;
}

}  // namespace

extern "C" int pictor_probe_stream(int64_t n, double out[8])
{
    if (n < 512 || n > (1LL << 28) || n % 512 || !out) return PICTOR_ERR_ARG;
    Six s{}, d{};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PICTOR_OK;
    bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
    for (int c = 0; c < 6 && ok; c++)
        ok = hipMalloc(&s.a[c], n * sizeof(float)) == hipSuccess && hipMalloc(&d.a[c], n * sizeof(float)) == hipSuccess &&
             hipMemset(s.a[c], 0, n * sizeof(float)) == hipSuccess;
    if (!ok) {
        rc = PICTOR_ERR_HIP;
    } else {
        float ms[4] = {};
        hipError_t e = run_stream<1>(s, d, n, 5, e0, e1, ms[0]);
        if (e == hipSuccess) e = run_stream<2>(s, d, n, 5, e0, e1, ms[1]);
        if (e == hipSuccess) e = run_stream<20>(s, d, n, 5, e0, e1, ms[2]);
        if (e == hipSuccess) e = run_stream<40>(s, d, n, 5, e0, e1, ms[3]);
        if (e != hipSuccess) {
            rc = PICTOR_ERR_HIP;
        } else {
            for (int k = 0; k < 4; k++) {
                out[k] = 48.0 * (double)n / (ms[k] * 1e-3);   // bytes per second
                out[4 + k] = ms[k];
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int c = 0; c < 6; c++) {
        if (s.a[c]) (void)hipFree(s.a[c]);
        if (d.a[c]) (void)hipFree(d.a[c]);
    }
    return rc;
}

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
