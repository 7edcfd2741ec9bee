// device_common.h - wavefront reductions and sample-point lookup shared by
// the kernels of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hase {

struct SampleGeom {
    float px, py;
    int plane;   // level of the sample point (its height in layer units)
    uint32_t p;  // mesh point under the sample
};

__device__ __forceinline__ SampleGeom sample_point(const float2* pts, uint32_t np, uint32_t sample)
{
    uint32_t p = sample % np, lv = sample / np;
    float2 xy = pts[p];
    return SampleGeom{xy.x, xy.y, (int)lv, p};
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// inclusive scan within a 64-lane wavefront
__device__ __forceinline__ double wave_inclusive_scan(double v, int lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

}  // namespace hase
