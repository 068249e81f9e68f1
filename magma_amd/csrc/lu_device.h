// Device-side helpers shared by the LU kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace mgi {

constexpr int kWave = 64;

// LAPACK dlamch('S'): the smallest normal double; reciprocal scaling is used
// only when |pivot| >= sfmin (LAPACK dgetf2 / dgetrf2 leaf).
constexpr double kSfminD = 2.2250738585072014e-308;
constexpr float  kSfminF = 1.17549435e-38f;

__device__ __forceinline__ double shfl_xor_d(double v, int mask)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __shfl_xor(lo, mask, kWave);
    hi = __shfl_xor(hi, mask, kWave);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double shfl_d(double v, int src)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __shfl(lo, src, kWave);
    hi = __shfl(hi, src, kWave);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Arg-max of |x| across the wave with LAPACK idamax tie breaking: the largest
// value wins and, among equal values, the smallest index. `v` must be >= 0 for
// real candidates and negative for lanes holding no candidate.
__device__ __forceinline__ void wave_argmax(double& v, int& idx)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = shfl_xor_d(v, off);
        int oi = __shfl_xor(idx, off, kWave);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
}

// Same over float values.
__device__ __forceinline__ void wave_argmax_f(float& v, int& idx)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        float ov = __shfl_xor(v, off, kWave);
        int oi = __shfl_xor(idx, off, kWave);
        if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
}

}  // namespace mgi
