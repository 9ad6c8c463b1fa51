// Building blocks of the cooperative (one launch, all workgroups resident)
// dense kernels: grid barrier on per-workgroup arrival flags and the
// agent-scope accesses that carry data between workgroups.
//
// Values published for other workgroups are written and read with
// agent-scope atomic accesses (coherent across the XCDs' L2 caches); every
// wave waits for its own stores to complete before the workgroup signals
// arrival, so a workgroup that sees all arrival flags reads complete data.
#pragma once

#include <hip/hip_runtime.h>

namespace rsdft {

constexpr unsigned kCoopSpinLimit = 1u << 24;   // polls before a barrier gives up (seconds)

__device__ __forceinline__ double coop_wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;   // the same value in every lane (addition commutes)
}

// Grid barrier on per-workgroup arrival flags (16 words apart): workgroup g
// stores the step count in flags[16 g] once its stores have completed; the
// first wave polls every flag in parallel.  A wait beyond the poll limit, or
// an error raised by another workgroup, sets *err and makes every workgroup
// leave; returns false then.
__device__ inline bool coop_grid_barrier(unsigned *flags, unsigned target, int *err, int *stop)
{
    __builtin_amdgcn_s_waitcnt(0);   // this wave's publication stores have completed
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x, nwg = gridDim.x;
        if (lane == 0)
            __hip_atomic_store(flags + 16 * blockIdx.x, target, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        unsigned polls = 0;
        while (true) {
            bool ok = true;
            for (int g = lane; g < nwg; g += 64)
                ok = ok && __hip_atomic_load(flags + 16 * g, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT) >= target;
            if (__all(ok)) break;
            const bool bad =
                __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                ++polls > kCoopSpinLimit;
            if (__any(bad)) {
                if (lane == 0) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    *stop = 1;
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
    return *stop == 0;
}

__device__ __forceinline__ void coop_put(double *p, double x)
{
    __hip_atomic_store(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double coop_get(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace rsdft
