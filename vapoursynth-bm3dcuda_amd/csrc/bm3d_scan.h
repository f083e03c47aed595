// FP32 block-matching scan shared by the FP32 matching kernel
// (bm3d_kernels.hip) and the exact fallback of the mixed-precision kernel
// (bm3d_match_dx.hip). Internal.
#pragma once

#include "bm3d_device.h"

namespace bm3d {
namespace dev {

// LDS tile: image rows [ty0, ty0 + th) x columns [tx0, tx0 + tw), stored as
// `bs` column phases; phase ph holds columns ph, ph + bs, ... at consecutive
// column slots, and each column is `hp` consecutive words (rows). Lanes of a
// work-group row hold blocks bs columns apart, so they read the same phase
// at consecutive column slots.
struct TileGeom {
    int tx0, ty0;  // image coordinates of the tile origin (may be negative)
    int tw, th;    // tile extent
    int bs;        // number of column phases
    int wq;        // column slots per phase
    int hp;        // words per column (>= th)
};

__device__ __forceinline__ int tile_col_base(const TileGeom& g, int tc) {
    int q = tc / g.bs;
    int ph = tc - q * g.bs;
    return (ph * g.wq + q) * g.hp;
}

// Fill the tile from one plane; out-of-image pixels replicate the border
// (they are only ever read for candidates that are masked off).
__device__ __forceinline__ void stage_tile(float* __restrict__ lds, const TileGeom& g,
                           const float* __restrict__ img, int W, int H, int stride) {
    const int n = g.tw * g.th;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        int r = idx / g.tw;
        int c = idx - r * g.tw;
        int gx = clampi(g.tx0 + c, 0, W - 1);
        int gy = clampi(g.ty0 + r, 0, H - 1);
        lds[tile_col_base(g, c) + r] = img[(size_t)gy * stride + gx];
    }
}

// Load one 8-row column of the tile into a ring slot.
__device__ __forceinline__ void load_col(float (&dst)[8], const float* __restrict__ src) {
#pragma unroll
    for (int j = 0; j < 8; ++j) dst[j] = src[j];
}

// Distance in the exact order of cpu_source/source.cpp:123-136 + 98-103:
// per column an even-row and an odd-row FMA chain, then a pairwise tree.
// ring slot of candidate column c is (u + c) & 7.
template <int U>
__device__ __forceinline__ float ring_ssd(const float (&rb)[8][8], const float (&ring)[8][8]) {
    float s[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int sl = (U + c) & 7;
        float d0 = rb[0][c] - ring[sl][0];
        float d1 = rb[1][c] - ring[sl][1];
        float d2 = rb[2][c] - ring[sl][2];
        float d3 = rb[3][c] - ring[sl][3];
        float d4 = rb[4][c] - ring[sl][4];
        float d5 = rb[5][c] - ring[sl][5];
        float d6 = rb[6][c] - ring[sl][6];
        float d7 = rb[7][c] - ring[sl][7];
        float e0 = d0 * d0;
        e0 = fma_(d2, d2, e0);
        e0 = fma_(d4, d4, e0);
        e0 = fma_(d6, d6, e0);
        float e1 = d1 * d1;
        e1 = fma_(d3, d3, e1);
        e1 = fma_(d5, d5, e1);
        e1 = fma_(d7, d7, e1);
        s[c] = e0 + e1;
    }
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

__device__ __forceinline__ float tree8(const float (&s)[8]) {
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// Per-reference-block terms of the distances that need them: the block
// mean (zssd, zsad) and norm (ssd/norm).
struct BlockStat {
    float mean, norm;
};

// Block mean as rtc_source/kernel.hpp:413-421: per column a pairwise sum of
// the 8 rows, the column tree, times 1/64. x(c, j) reads column c, row j.
template <typename F>
__device__ __forceinline__ float block_mean(F x) {
    float s[8];
#pragma unroll
    for (int c = 0; c < 8; ++c)
        s[c] = ((x(c, 0) + x(c, 1)) + (x(c, 2) + x(c, 3))) + ((x(c, 4) + x(c, 5)) + (x(c, 6) + x(c, 7)));
    return tree8(s) * (1.0f / 64.0f);
}

// Per column even-row and odd-row product chains (errors[i % 2] of
// rtc_source/kernel.hpp:372-504, products fused), summed, then the tree.
template <typename FA, typename FB>
__device__ __forceinline__ float dot_chain(FA a, FB b) {
    float s[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float e0 = a(c, 0) * b(c, 0), e1 = a(c, 1) * b(c, 1);
#pragma unroll
        for (int j = 2; j < 8; j += 2) {
            e0 = fma_(a(c, j), b(c, j), e0);
            e1 = fma_(a(c, j + 1), b(c, j + 1), e1);
        }
        s[c] = e0 + e1;
    }
    return tree8(s);
}

__device__ __forceinline__ BlockStat block_stat(const float (&rb)[8][8]) {
    auto x = [&](int c, int j) { return rb[j][c]; };
    return {block_mean(x), __builtin_sqrtf(dot_chain(x, x))};
}

// Distance kCost (BM3D_BM_*) between the reference block rb[row][col] and
// the candidate in the ring (column c in slot (U + c) & 7), in the order of
// the oracle's block_cost: ssd (cpu_source/source.cpp:123-136), sad, zssd,
// zsad, ssd/norm (rtc_source/kernel.hpp:372-504).
template <int kCost, int U>
__device__ __forceinline__ float ring_cost(const float (&rb)[8][8], const float (&ring)[8][8], const BlockStat& st) {
    if constexpr (kCost == 0) {
        return ring_ssd<U>(rb, ring);
    } else if constexpr (kCost == 4) {
        auto r = [&](int c, int j) { return rb[j][c]; };
        auto n = [&](int c, int j) { return ring[(U + c) & 7][j]; };
        const float nn = __builtin_sqrtf(dot_chain(n, n));
        const float dot = dot_chain(r, n);
        return 2.0f - (2.0f * dot) / fma_(st.norm, nn, kFltEpsilon);
    } else {
        float dm = 0.0f;
        if constexpr (kCost == 2 || kCost == 3)
            dm = st.mean - block_mean([&](int c, int j) { return ring[(U + c) & 7][j]; });
        float s[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int sl = (U + c) & 7;
            float e[2];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                float v = rb[j][c] - ring[sl][j];
                if constexpr (kCost == 2 || kCost == 3) v = v - dm;
                if constexpr (kCost == 2) e[j & 1] = j < 2 ? v * v : fma_(v, v, e[j & 1]);
                else e[j & 1] = j < 2 ? fabsf(v) : e[j & 1] + fabsf(v);
            }
            s[c] = e[0] + e[1];
        }
        return tree8(s);
    }
}

template <int kCost, int U, bool kCount>
__device__ __forceinline__ void scan_step(const float* __restrict__ lds, const TileGeom& g,
                                          const float (&rb)[8][8], const BlockStat& st, float (&ring)[8][8],
                                          int& ph, int& colbase, int rt, Best8& best,
                                          int col, int row, bool row_ok, int z, int W,
                                          unsigned& cnt) {
    load_col(ring[(U + 7) & 7], lds + colbase + rt);
    // advance to the next tile column
    ++ph;
    const bool wrap = ph == g.bs;
    ph = wrap ? 0 : ph;
    colbase += wrap ? (g.hp - (g.bs - 1) * g.wq * g.hp) : g.wq * g.hp;
    float e = ring_cost<kCost, U>(rb, ring, st);
    const bool ok = row_ok && col >= 0 && col <= W - 8;
    if constexpr (kCount) cnt += ok ? 1u : 0u;
    if (ok && e < best.e[7]) best.insert(e, pack_pos(col, row, z));
}

// Exhaustive scan of the window of radius `range` around (cx, cy), clamped
// to the plane, row-major (cpu_source/source.cpp:143-215). All lanes walk
// the full (2 range + 1)^2 square so the loop is wave-uniform; candidates
// outside the plane are masked, which leaves the scan order of the valid
// ones -- and hence tie-breaking -- unchanged.
template <int kCost, bool kCount>
__device__ __forceinline__ void scan_window(const float* __restrict__ lds, const TileGeom& g,
                                            const float (&rb)[8][8], const BlockStat& st, Best8& best,
                                            int cx, int cy, int range, int z, int W, int H,
                                            unsigned& cnt) {
    const int n = 2 * range + 1;
    float ring[8][8];
    for (int dyi = 0; dyi < n; ++dyi) {
        const int row = cy - range + dyi;
        const bool row_ok = row >= 0 && row <= H - 8;
        const int rt = row - g.ty0;
        const int tc0 = cx - range - g.tx0;
        int q = tc0 / g.bs;
        int ph = tc0 - q * g.bs;
        int colbase = (ph * g.wq + q) * g.hp;
        // preload the first 7 columns
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            load_col(ring[s], lds + colbase + rt);
            ++ph;
            const bool wrap = ph == g.bs;
            ph = wrap ? 0 : ph;
            colbase += wrap ? (g.hp - (g.bs - 1) * g.wq * g.hp) : g.wq * g.hp;
        }
        const int col0 = cx - range;
        for (int base = 0; base < n; base += 8) {
#define BM3D_SCAN_STEP(U)                                                                    \
    if (base + U < n)                                                                        \
        scan_step<kCost, U, kCount>(lds, g, rb, st, ring, ph, colbase, rt, best, col0 + base + U, row, row_ok, z, \
                                    W, cnt);
            BM3D_SCAN_STEP(0)
            BM3D_SCAN_STEP(1)
            BM3D_SCAN_STEP(2)
            BM3D_SCAN_STEP(3)
            BM3D_SCAN_STEP(4)
            BM3D_SCAN_STEP(5)
            BM3D_SCAN_STEP(6)
            BM3D_SCAN_STEP(7)
#undef BM3D_SCAN_STEP
        }
    }
}

// scan_window reading one plane in global memory instead of an LDS tile,
// for search reaches whose union window does not fit the LDS: the same
// candidates in the same order and the same distances, each candidate
// column loaded as 8 strided words. Pixels outside the plane read the
// clamped border, as the LDS tile holds; they are only ever read for
// candidates that are masked off.
template <int kCost, bool kCount>
__device__ __forceinline__ void scan_window_plane(const float* __restrict__ img, int stride,
                                                  const float (&rb)[8][8], const BlockStat& st, Best8& best,
                                                  int cx, int cy, int range, int z, int W, int H, unsigned& cnt) {
    const int n = 2 * range + 1;
    float ring[8][8];
    for (int dyi = 0; dyi < n; ++dyi) {
        const int row = cy - range + dyi;
        const bool row_ok = row >= 0 && row <= H - 8;
        int ro[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) ro[j] = clampi(row + j, 0, H - 1) * stride;
        auto load = [&](float(&dst)[8], int col) {
            const int c = clampi(col, 0, W - 1);
#pragma unroll
            for (int j = 0; j < 8; ++j) dst[j] = img[ro[j] + c];
        };
        const int col0 = cx - range;
#pragma unroll
        for (int s = 0; s < 7; ++s) load(ring[s], col0 + s);
        for (int base = 0; base < n; base += 8) {
#define BM3D_PLANE_STEP(U)                                                       \
    if (base + U < n) {                                                          \
        load(ring[(U + 7) & 7], col0 + base + U + 7);                            \
        const float e = ring_cost<kCost, U>(rb, ring, st);                       \
        const int col = col0 + base + U;                                         \
        const bool ok = row_ok && col >= 0 && col <= W - 8;                      \
        if constexpr (kCount) cnt += ok ? 1u : 0u;                               \
        if (ok && e < best.e[7]) best.insert(e, pack_pos(col, row, z));          \
    }
            BM3D_PLANE_STEP(0)
            BM3D_PLANE_STEP(1)
            BM3D_PLANE_STEP(2)
            BM3D_PLANE_STEP(3)
            BM3D_PLANE_STEP(4)
            BM3D_PLANE_STEP(5)
            BM3D_PLANE_STEP(6)
            BM3D_PLANE_STEP(7)
#undef BM3D_PLANE_STEP
        }
    }
}

}  // namespace dev
}  // namespace bm3d
