// device_common.h - wavefront reductions, the wavefront key sort and the
// sample-point lookup shared by the kernels of the library.
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

// Bitonic sort of a wavefront's 256 keys, ascending, key r * 64 + lane in
// v[r] of that lane.  Every lane of the wavefront takes part (the exchanges
// read the partner lane's register).
__device__ __forceinline__ void wave_sort256(uint32_t (&v)[4], int lane)
{
#pragma unroll
    for (uint32_t k = 2u; k <= 256u; k <<= 1) {
#pragma unroll
        for (uint32_t d = k >> 1; d > 0u; d >>= 1) {
            if (d >= 64u) {
                // partners in the same lane: registers r and r + d / 64
                const uint32_t rd = d >> 6;
#pragma unroll
                for (uint32_t r = 0; r < 4u; ++r) {
                    if (r & rd) continue;
                    const bool asc = ((r * 64u + (uint32_t)lane) & k) == 0u;
                    const uint32_t a = v[r], b = v[r | rd];
                    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
                    v[r] = asc ? lo : hi;
                    v[r | rd] = asc ? hi : lo;
                }
            } else {
#pragma unroll
                for (uint32_t r = 0; r < 4u; ++r) {
                    const uint32_t w = (uint32_t)__shfl_xor((int)v[r], (int)d, 64);
                    const bool lower = ((uint32_t)lane & d) == 0u;
                    const bool asc = ((r * 64u + (uint32_t)lane) & k) == 0u;
                    const uint32_t lo = v[r] < w ? v[r] : w, hi = v[r] < w ? w : v[r];
                    v[r] = lower == asc ? lo : hi;
                }
            }
        }
    }
}

}  // namespace hase
