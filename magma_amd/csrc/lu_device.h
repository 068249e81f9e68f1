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

__device__ __forceinline__ float readlane_t(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double readlane_t(double v, int lane) { return readlane_d(v, lane); }
__device__ __forceinline__ float shfl_t(float v, int src) { return __shfl(v, src, kWave); }
__device__ __forceinline__ double shfl_t(double v, int src) { return shfl_d(v, src); }

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

namespace mgi {

// ---------------------------------------------------------------------------
// DPP (data-parallel primitives) cross-lane moves for gfx9-family waves.
// Control words: quad_perm [a,b,c,d] = a | b<<2 | c<<4 | d<<6; row_half_mirror
// 0x141, row_mirror 0x140, row_bcast15 0x142, row_bcast31 0x143.
// ---------------------------------------------------------------------------
constexpr int kDppQuadXor1 = 0xB1;      // [1,0,3,2]
constexpr int kDppQuadXor2 = 0x4E;      // [2,3,0,1]
constexpr int kDppHalfMirror = 0x141;
constexpr int kDppMirror = 0x140;
constexpr int kDppBcast15 = 0x142;
constexpr int kDppBcast31 = 0x143;

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_d(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, CTRL, ROW_MASK, 0xF, false);
}

// Maximum over the 64 lanes, returned as a wave-uniform value. Within each
// 16-lane row: quad xor 1, quad xor 2, half-mirror, mirror; across rows:
// broadcast of lane 15 into rows 1 and 3, then of lane 31 into rows 2 and 3,
// leaving the total in lane 63.
__device__ __forceinline__ double wave_max_dpp(double v)
{
    v = fmax(v, dpp_d<kDppQuadXor1>(v));
    v = fmax(v, dpp_d<kDppQuadXor2>(v));
    v = fmax(v, dpp_d<kDppHalfMirror>(v));
    v = fmax(v, dpp_d<kDppMirror>(v));
    v = fmax(v, dpp_d<kDppBcast15, 0xA>(v));
    v = fmax(v, dpp_d<kDppBcast31, 0xC>(v));
    return readlane_d(v, 63);
}

template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ float dpp_f(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROW_MASK, 0xF,
                                                      false));
}

// Maximum of a float over the 64 lanes (same pattern as wave_max_dpp; each
// stage can issue as a single DPP-sourced v_max_f32).
__device__ __forceinline__ float wave_max_dpp_f(float v)
{
    v = fmaxf(v, dpp_f<kDppQuadXor1>(v));
    v = fmaxf(v, dpp_f<kDppQuadXor2>(v));
    v = fmaxf(v, dpp_f<kDppHalfMirror>(v));
    v = fmaxf(v, dpp_f<kDppMirror>(v));
    v = fmaxf(v, dpp_f<kDppBcast15, 0xA>(v));
    v = fmaxf(v, dpp_f<kDppBcast31, 0xC>(v));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Maximum of an int over the 64 lanes (each stage a DPP-sourced v_max_i32).
__device__ __forceinline__ int wave_max_dpp_i(int v)
{
    v = max(v, dpp_i<kDppQuadXor1>(v));
    v = max(v, dpp_i<kDppQuadXor2>(v));
    v = max(v, dpp_i<kDppHalfMirror>(v));
    v = max(v, dpp_i<kDppMirror>(v));
    v = max(v, dpp_i<kDppBcast15, 0xA>(v));
    v = max(v, dpp_i<kDppBcast31, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ int wave_min_dpp(int v)
{
    v = min(v, dpp_i<kDppQuadXor1>(v));
    v = min(v, dpp_i<kDppQuadXor2>(v));
    v = min(v, dpp_i<kDppHalfMirror>(v));
    v = min(v, dpp_i<kDppMirror>(v));
    v = min(v, dpp_i<kDppBcast15, 0xA>(v));
    v = min(v, dpp_i<kDppBcast31, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}

// Broadcast lane Q of every quad to the whole quad.
template <int Q>
__device__ __forceinline__ double quad_bcast_d(double v)
{
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    return dpp_d<ctrl>(v);
}

}  // namespace mgi
