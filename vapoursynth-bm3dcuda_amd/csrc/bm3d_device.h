// Device helpers shared by the gfx950 kernels (bm3d_kernels.hip,
// bm3d_match_dx.hip, bm3d_filter_mx.hip). Internal.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm3d_kernels.h"

namespace bm3d {
namespace dev {

constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltEpsilon = 1.192092896e-07f;

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ uint32_t pack_pos(int x, int y, int z) {
    return (uint32_t)x | ((uint32_t)y << kPosYShift) | ((uint32_t)z << kPosZShift);
}
__device__ __forceinline__ int pos_x(uint32_t p) { return (int)(p & kPosXMask); }
__device__ __forceinline__ int pos_y(uint32_t p) { return (int)((p >> kPosYShift) & kPosYMask); }
__device__ __forceinline__ int pos_z(uint32_t p) { return (int)(p >> kPosZShift); }

// Sorted (ascending, strict <) list of the 8 best candidates, kept in
// registers. Equivalent to the shuffle-up insertion of source/kernel.cu:473-485
// and the SIMD insertion of cpu_source/source.cpp:179-204.
struct Best8 {
    float e[8];
    uint32_t p[8];

    __device__ __forceinline__ void init(float err, uint32_t pos) {
#pragma unroll
        for (int i = 0; i < 8; ++i) { e[i] = err; p[i] = pos; }
    }

    __device__ __forceinline__ void insert(float err, uint32_t pos) {
        bool c[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) c[i] = err < e[i];
#pragma unroll
        for (int i = 7; i >= 1; --i) {
            float ne = c[i - 1] ? e[i - 1] : (c[i] ? err : e[i]);
            uint32_t np = c[i - 1] ? p[i - 1] : (c[i] ? pos : p[i]);
            e[i] = ne;
            p[i] = np;
        }
        e[0] = c[0] ? err : e[0];
        p[0] = c[0] ? pos : p[0];
    }

    // entry i for a wave-uniform runtime i
    __device__ __forceinline__ float err_at(int i) const {
        float r = e[0];
#pragma unroll
        for (int k = 1; k < 8; ++k) r = (i == k) ? e[k] : r;
        return r;
    }
    __device__ __forceinline__ uint32_t pos_at(int i) const {
        uint32_t r = p[0];
#pragma unroll
        for (int k = 1; k < 8; ++k) r = (i == k) ? p[k] : r;
        return r;
    }
};

// Cross-lane moves with DPP (VALU only, no LDS).
template <int kCtrl>
__device__ __forceinline__ float dpp_mov(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), kCtrl, 0xF, 0xF, false));
}
template <int kCtrl>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_update_dpp(0, v, kCtrl, 0xF, 0xF, false);
}
// Exponent of the fixed-point unit of the accumulation regions. With
// range null, fx_exp (extractor_exp - 24: pre-rounded values are whole
// multiples of the unit). Otherwise fx_exp (-22) plus floor(log2 M) for
// the largest finite |sample| M >= 2 of the call's input: a region pixel's
// sum of weighted estimates (weights <= 1, a few hundred members at most)
// then has the int32 headroom of 2^31 units >= 256 M whatever the scale of
// the samples, and every input within (-2, 2) keeps the unit 2^-22.
__device__ __forceinline__ int fx_exponent(const float* range, int fx_exp) {
    if (!range) return fx_exp;
    const float m = *range;
    return m >= 2.0f ? fx_exp + ilogbf(m) : fx_exp;
}

// xor1 / xor2 are quad permutations; xor4 is a half-row mirror followed by a
// quad reversal (7 - k, then 3 - (k mod 4) within the quad, gives k ^ 4).
template <typename T> __device__ __forceinline__ T lane_xor1(T v) { return dpp_mov<0xB1>(v); }
template <typename T> __device__ __forceinline__ T lane_xor2(T v) { return dpp_mov<0x4E>(v); }
template <typename T> __device__ __forceinline__ T lane_xor4(T v) { return dpp_mov<0x1B>(dpp_mov<0x141>(v)); }
// Value of lane k of the quad (quad_perm [k, k, k, k]).
template <int kLane, typename T> __device__ __forceinline__ T quad_bcast(T v) { return dpp_mov<kLane * 0x55>(v); }

}  // namespace dev
}  // namespace bm3d
