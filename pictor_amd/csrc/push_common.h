// push_common.h -- device helpers shared by the 2-D and 3-D push kernels.
#pragma once
#include "pictor_common.h"

namespace pictor {
namespace pushdev {

__device__ inline double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ inline double dmax(double a, double b) { return (b > a) ? b : a; }

__device__ inline unsigned long long lanemask_lt()
{
    unsigned lane = __lane_id();
    return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Per-wave reservation of consecutive slots in the LDS counter of each
// destination; every lane of the wave must call it (d = -1 for idle lanes).
__device__ inline int reserve_slot(int d, int *sCnt)
{
    const bool act = d >= 0;
    unsigned long long pending = __ballot(act);
    int slot = 0;
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int ld = __shfl(d, leader);
        const unsigned long long m = __ballot(act && d == ld);
        int b = 0;
        if ((int)__lane_id() == leader) b = atomicAdd(&sCnt[ld], __popcll(m));
        b = __shfl(b, leader);
        if (act && d == ld) slot = b + __popcll(m & lanemask_lt());
        pending &= ~m;
    }
    return slot;
}

__device__ inline int wrap_tile(int t, int n)
{
    return (t < 0) ? t + n : (t >= n ? t - n : t);
}

__device__ inline int norm_dir(int dd, int n)
{
    if (dd > 1) dd -= n;
    else if (dd < -1) dd += n;
    return dd;
}

// FP32 -> exact integer count of 2^-36 units (the argument is already scaled by 2^36)
__device__ inline double quant(float v) { return (double)__builtin_rintf(v); }

// byte-offset accessors: a uniform base plus a 32-bit lane offset lets the
// compiler use SGPR-base addressing (one VALU op per access)
template <typename T>
__device__ inline T ldo(const T *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}
template <typename T>
__device__ inline void sto(T *base, unsigned off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off) = v;
}


}  // namespace pushdev
}  // namespace pictor
