// Mixed-precision block matching (BM3D_PRECISION_MIXED).
//
// The search of source/kernel.cu:432-647 (cpu_source/source.cpp:143-321)
// with the candidate distances pre-scored in FP16 and only the best few
// rescored exactly in the reference's FP32 order (10 for the 8 of the
// spatial search; 4, 8 or 16, at least 2 ps_num, for the ps_num kept of each
// temporal frame: only the first ps_num entries of a frame's list are merged
// into the group and centre the next frame's windows,
// source/kernel.cu:568-611).
//
// Layout as in the FP32 kernel: one lane per reference block, 16 x TBY
// blocks per work-group, the work-group's search windows staged in LDS one
// slot at a time, each lane sliding an 8-column ring across a window. The
// tile is stored as FP16 pairs of vertically adjacent pixels minus an
// offset mu (two copies, pairs starting on even and on odd rows), so a
// candidate column of 8 rows is 4 packed registers. Per candidate:
//
//   N  = |c'|^2              sliding sum of per-column v_dot2c sums
//   A  = (|r'|^2 + N) / 2 - <r', c'>   one chain of 32 v_dot2c_f32_f16 from
//                            the first term (FP32 accumulation): half the
//                            squared distance up to FP16 rounding
//
// A is packed with the candidate's scan index into a 32-bit key (A's
// exponent and 11 mantissa bits above a 12-bit index); the K smallest keys
// are kept by branch-free min/max merges of 8 keys at a time. The kept
// candidates are rescored exactly and ordered by (distance, scan index),
// which is the order the reference's strict sorted insertion in scan order
// produces; the list equals the FP32 path's unless an exact top-k candidate
// ranks below K approximately. A work-group tile whose pixels are not finite or reach
// beyond +-8 around mu (the centre pixel of the slot's frame) is searched
// exhaustively in FP32, exactly as the FP32 kernel.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm3d_inject.h"
#include "bm3d_scan.h"

namespace bm3d {

namespace {

using namespace dev;

typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

constexpr float kInf = __builtin_huge_valf();
constexpr float kDxMaxSpan = 8.0f;
constexpr uint32_t kIdxBits = 12;
constexpr uint32_t kIdxMask = (1u << kIdxBits) - 1;
// keys kept in the spatial search (>= 8), all rescored exactly: 10 keep
// the same groups as 12 on the 1080p test clip; 8 changed 0.13% of them
constexpr int kKSpatial = 10;

__device__ __forceinline__ float dot2(f16x2 a, f16x2 b, float c) { return __builtin_amdgcn_fdot2(a, b, c, false); }
__device__ __forceinline__ f16x2 as_h2(uint32_t u) { return __builtin_bit_cast(f16x2, u); }

__device__ __forceinline__ void cex(uint32_t& a, uint32_t& b) {
    const uint32_t lo = min(a, b);
    b = max(a, b);
    a = lo;
}

// Ascending sort of 8 keys (a 19-comparator network).
__device__ __forceinline__ void sort8(uint32_t (&v)[8]) {
    cex(v[0], v[2]); cex(v[1], v[3]); cex(v[4], v[6]); cex(v[5], v[7]);
    cex(v[0], v[4]); cex(v[1], v[5]); cex(v[2], v[6]); cex(v[3], v[7]);
    cex(v[0], v[1]); cex(v[2], v[3]); cex(v[4], v[5]); cex(v[6], v[7]);
    cex(v[2], v[4]); cex(v[3], v[5]);
    cex(v[1], v[4]); cex(v[3], v[6]);
    cex(v[1], v[2]); cex(v[3], v[4]); cex(v[5], v[6]);
}

// The K smallest 32-bit keys, ascending (branch-free min/max cascade).
template <int K>
struct KeyList {
    uint32_t k[K];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < K; ++i) k[i] = 0xFFFFFFFFu;
    }
    // Sorted insertion: k[0] = min(k[0], key), k[i] = med3(k[i-1], key, k[i])
    // (the median of three is the clamp of key to [k[i-1], k[i]], which the
    // compiler emits as one v_med3_u32): K independent operations per key.
    __device__ __forceinline__ void insert(uint32_t key) {
        uint32_t prev = k[0];
        k[0] = min(k[0], key);
#pragma unroll
        for (int i = 1; i < K; ++i) {
            const uint32_t cur = k[i];
            k[i] = max(min(prev, key), min(max(prev, key), cur));
            prev = cur;
        }
    }
    // (each entry is updated in place, so the list keeps its registers)
    __device__ __forceinline__ void push(uint32_t key) {
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t hi = max(k[i], key);
            k[i] = min(k[i], key);
            key = hi;
        }
    }
    // Adds 8 keys at once. For K >= 8: sort them (b), take
    // c[i] = min(k[i], b[K-1-i]) -- the K smallest of the union, as a
    // sequence that rises then falls -- and sort c with a bitonic merge,
    // padded at the front with zeros up to a power of two (constant
    // comparisons the compiler removes). 11-13 min/max per key instead of
    // the cascade's 2K.
    // K = 4: the 4 smallest of b come from sorting each half and taking
    // min(lo[i], hi[3-i]) (a rise-then-fall sequence, sorted by a 4-entry
    // bitonic merge); the same pairing then merges them into k. 44 min/max
    // for 8 keys instead of the cascade's 64.
    __device__ __forceinline__ void push8(uint32_t (&b)[8]) {
        if constexpr (K == 4) {
            cex(b[0], b[1]); cex(b[2], b[3]); cex(b[0], b[2]); cex(b[1], b[3]); cex(b[1], b[2]);
            cex(b[4], b[5]); cex(b[6], b[7]); cex(b[4], b[6]); cex(b[5], b[7]); cex(b[5], b[6]);
            uint32_t c[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) c[i] = min(b[i], b[7 - i]);
            cex(c[0], c[2]); cex(c[1], c[3]); cex(c[0], c[1]); cex(c[2], c[3]);
#pragma unroll
            for (int i = 0; i < 4; ++i) k[i] = min(k[i], c[3 - i]);
            cex(k[0], k[2]); cex(k[1], k[3]); cex(k[0], k[1]); cex(k[2], k[3]);
        } else if constexpr (K < 8) {
#pragma unroll
            for (int i = 0; i < 8; ++i) push(b[i]);
        } else {
            static_assert(K <= 16, "push8 merges into at most 16 keys");
            constexpr int P = K <= 8 ? 8 : 16;
            sort8(b);
            uint32_t c[P];
#pragma unroll
            for (int i = 0; i < P - K; ++i) c[i] = 0u;
#pragma unroll
            for (int i = 0; i < K; ++i) c[P - K + i] = (K - 1 - i < 8) ? min(k[i], b[K - 1 - i < 8 ? K - 1 - i : 0]) : k[i];
#pragma unroll
            for (int d = P / 2; d >= 1; d /= 2)
#pragma unroll
                for (int i = 0; i < P; ++i)
                    if ((i & d) == 0) cex(c[i], c[i + d]);
#pragma unroll
            for (int i = 0; i < K; ++i) k[i] = c[P - K + i];
        }
    }
    // removes and returns the smallest key
    __device__ __forceinline__ uint32_t pop() {
        const uint32_t r = k[0];
#pragma unroll
        for (int i = 0; i + 1 < K; ++i) k[i] = k[i + 1];
        k[K - 1] = 0xFFFFFFFFu;
        return r;
    }
};

// Exact top 8 by (distance, scan index) for candidates arriving out of scan
// order (the reference's list restricted to them).
struct Best8Id {
    Best8 b;
    uint32_t id[8];
    __device__ __forceinline__ void init(float err, uint32_t pos) {
        b.init(err, pos);
#pragma unroll
        for (int i = 0; i < 8; ++i) id[i] = 0xFFFFFFFFu;
    }
    __device__ __forceinline__ void insert(float err, uint32_t pos, uint32_t idn) {
        bool c[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) c[i] = err < b.e[i] || (err == b.e[i] && idn < id[i]);
#pragma unroll
        for (int i = 7; i >= 1; --i) {
            b.e[i] = c[i - 1] ? b.e[i - 1] : (c[i] ? err : b.e[i]);
            b.p[i] = c[i - 1] ? b.p[i - 1] : (c[i] ? pos : b.p[i]);
            id[i] = c[i - 1] ? id[i - 1] : (c[i] ? idn : id[i]);
        }
        b.e[0] = c[0] ? err : b.e[0];
        b.p[0] = c[0] ? pos : b.p[0];
        id[0] = c[0] ? idn : id[0];
    }
};

// FP16 pair tile: two planes of `plane` dwords; in plane par, column tc's
// word i (at the FP32 kernel's column-phase position, hp2 words per column)
// holds rows (2 i + par, 2 i + par + 1) minus mu.
struct PairGeom {
    int tx0, ty0, tw, th, bs, wq, hp2, plane;
};

__device__ __forceinline__ int pair_col_base(const PairGeom& g, int tc) {
    const int q = tc / g.bs;
    const int ph = tc - q * g.bs;
    return (ph * g.wq + q) * g.hp2;
}

__device__ __forceinline__ void next_pair_col(const PairGeom& g, int& ph, int& colbase) {
    ++ph;
    const bool wrap = ph == g.bs;
    ph = wrap ? 0 : ph;
    colbase += wrap ? (g.hp2 - (g.bs - 1) * g.wq * g.hp2) : g.wq * g.hp2;
}

// Stages the FP16 pairs; returns true when every staged pixel is within
// kDxMaxSpan of mu (finite).
__device__ bool stage_pairs(uint32_t* __restrict__ lds, const PairGeom& g, const float* __restrict__ img,
                            int W, int H, int stride, float mu) {
    const int n = g.tw * g.hp2;
    bool ok = true;
    for (int idx = threadIdx.x; idx < n; idx += blockDim.x) {
        const int c = (int)((unsigned)idx / (unsigned)g.hp2);
        const int i = idx - c * g.hp2;
        const int gx = clampi(g.tx0 + c, 0, W - 1);
        const float* col = img + gx;
        auto px = [&](int r) -> _Float16 {
            const int gy = clampi(g.ty0 + min(r, g.th - 1), 0, H - 1);
            return (_Float16)(col[(size_t)gy * stride] - mu);
        };
        const _Float16 p0 = px(2 * i), p1 = px(2 * i + 1), p2 = px(2 * i + 2);
        ok &= __builtin_fabsf((float)p0) <= kDxMaxSpan && __builtin_fabsf((float)p1) <= kDxMaxSpan &&
              __builtin_fabsf((float)p2) <= kDxMaxSpan;
        const int base = pair_col_base(g, c) + i;
        lds[base] = __builtin_bit_cast(uint32_t, f16x2{p0, p1});
        lds[g.plane + base] = __builtin_bit_cast(uint32_t, f16x2{p1, p2});
    }
    return __syncthreads_and(ok);
}

// Sum of squares of a packed column (exact FP16 products, FP32 sums).
__device__ __forceinline__ float col_sq(const f16x2 (&c)[4]) {
    return dot2(c[3], c[3], dot2(c[2], c[2], dot2(c[1], c[1], dot2(c[0], c[0], 0.0f))));
}

__device__ __forceinline__ void load_pair_col(f16x2 (&dst)[4], const uint32_t* __restrict__ src) {
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[i] = as_h2(src[i]);
}

// One candidate step: loads the window's new column (at `addr`),
// updates the column sums and returns the candidate's key (all ones when it
// is outside the plane). Step u of the row is an in-plane candidate iff
// u - vlo <= vspan (unsigned); kCheck false: the caller knows every step
// is.
template <int U, bool kCount, bool kCheck>
__device__ __forceinline__ uint32_t dx_step(const f16x2 (&rp)[8][4], float nr, f16x2 (&ring)[8][4], float (&csq)[8],
                                            float& N, const uint32_t* __restrict__ addr, int u,
                                            unsigned vlo, unsigned vspan, uint32_t idx, uint32_t hi_mask,
                                            unsigned& cnt) {
    load_pair_col(ring[(U + 7) & 7], addr);
    csq[(U + 7) & 7] = col_sq(ring[(U + 7) & 7]);
    N += csq[(U + 7) & 7];
    // half the distance: (|r'|^2 + |c'|^2) / 2 - <r', c'>, rp holding -r',
    // in one chain
    float d = fma_(N, 0.5f, nr);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
#pragma unroll
        for (int i = 0; i < 4; ++i) d = dot2(rp[c][i], ring[(U + c) & 7][i], d);
    }
    // clamped at zero: a rounded negative would order after every key
    const float approx = fmaxf(d, 0.0f);
    N -= csq[U & 7];  // column leaving the window of the next candidate
    // idx < 2^kIdxBits (the plan allows at most 4096 candidates per list);
    // hi_mask, ~kIdxMask in a register, makes this one and-or with the
    // wave-uniform index as its scalar operand
    const uint32_t key = (__builtin_bit_cast(uint32_t, approx) & hi_mask) | idx;
    if constexpr (!kCheck) {
        if constexpr (kCount) cnt += 1u;
        return key;
    } else {
        const bool ok = (unsigned)u - vlo <= vspan;
        if constexpr (kCount) cnt += ok ? 1u : 0u;
        return ok ? key : 0xFFFFFFFFu;
    }
}

// Approximate scan of the window of radius `range` around (cx, cy) into kl;
// scan index of (dyi, dxi) is idx_base + dyi * n + dxi. kCheck false: every
// candidate of the window is in-plane.
template <int K, bool kCount, bool kCheck, bool kDiv8, bool kIns>
__device__ __forceinline__ void dx_rows(const uint32_t* __restrict__ lds, const PairGeom& g,
                                          const f16x2 (&rp)[8][4], float nr, KeyList<K>& kl, int cx, int cy,
                                          int range, int W, int H, uint32_t idx_base, unsigned& cnt) {
    const int n = 2 * range + 1;
    // kDiv8 (block_step divides 8): 8 columns on are the same phase 8 / bs
    // slots further, so the step addresses of a row advance by a constant
    // (a compile-time choice: a branch here would split the loop body and
    // cost register copies of every column loaded)
    const int inc8 = kDiv8 ? (8 / g.bs) * g.hp2 : 0;
    // ~kIdxMask through an opaque register move: the compiler keeps it in a
    // vector register instead of copying each scan index into one
    uint32_t hi_mask;
    asm volatile("v_mov_b32 %0, %1" : "=v"(hi_mask) : "i"(~kIdxMask));
    f16x2 ring[8][4];
    float csq[8];
    for (int dyi = 0; dyi < n; ++dyi) {
        const int row = cy - range + dyi;
        const int rt = row - g.ty0;
        const int off = (rt & 1) * g.plane + (rt >> 1);
        const int tc0 = cx - range - g.tx0;
        const int q = tc0 / g.bs;
        int ph = tc0 - q * g.bs;
        int colbase = (ph * g.wq + q) * g.hp2;
        float N = 0.0f;
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            load_pair_col(ring[s], lds + off + colbase);
            next_pair_col(g, ph, colbase);
            csq[s] = col_sq(ring[s]);
            N += csq[s];
        }
        const uint32_t* ad[8];  // the new column of steps base .. base + 7
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ad[k] = lds + off + colbase;
            next_pair_col(g, ph, colbase);
        }
        // in-plane steps: row inside, column col0 + u in [0, W - 8]
        const int col0 = cx - range;
        const int lo = max(0, -col0), hi = min(n - 1, W - 8 - col0);
        const bool any = row >= 0 && row <= H - 8 && lo <= hi;
        const unsigned vlo = any ? (unsigned)lo : 0x40000000u;
        const unsigned vspan = any ? (unsigned)(hi - lo) : 0u;
        const uint32_t rowidx = idx_base + (uint32_t)(dyi * n);
        // whole groups of 8 steps unconditionally (the ring returns to the
        // same registers after 8), then the 0..7 remaining steps
        int base = 0;
#define BM3D_DX_KEY(U)                                                                               \
    dx_step<U, kCount, kCheck>(rp, nr, ring, csq, N, ad[U], base + U, vlo, vspan, rowidx + (uint32_t)(base + U), \
                               hi_mask, cnt)
        // kIns: groups of 8 keys merged at once above 8 kept keys (fewer
        // operations per key than sorted insertion), inserted one by one up
        // to 8 (as cheap, with fewer live registers), single steps inserted;
        // else every group merged and single steps pushed through the cascade
        // (the form the one-launch kernel allocates its registers best for)
#define BM3D_DX_STEP(U)                     \
    if constexpr (kIns) kl.insert(BM3D_DX_KEY(U)); \
    else kl.push(BM3D_DX_KEY(U));
#define BM3D_DX_GSTEP(U)                                     \
    if constexpr (K > 8 || !kIns) keys[U] = BM3D_DX_KEY(U);  \
    else kl.insert(BM3D_DX_KEY(U));
        for (; base + 8 <= n; base += 8) {
            uint32_t keys[8];
            BM3D_DX_GSTEP(0)
            BM3D_DX_GSTEP(1)
            BM3D_DX_GSTEP(2)
            BM3D_DX_GSTEP(3)
            BM3D_DX_GSTEP(4)
            BM3D_DX_GSTEP(5)
            BM3D_DX_GSTEP(6)
            BM3D_DX_GSTEP(7)
            if constexpr (K > 8 || !kIns) kl.push8(keys);
            if constexpr (kDiv8) {
#pragma unroll
                for (int k = 0; k < 8; ++k) ad[k] += inc8;
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    ad[k] = lds + off + colbase;
                    next_pair_col(g, ph, colbase);
                }
            }
        }
        // the 0..7 remaining steps, each behind a uniform branch (one copy of
        // each step's code instead of one per remainder length)
        if (base < n) BM3D_DX_STEP(0)
        if (base + 1 < n) BM3D_DX_STEP(1)
        if (base + 2 < n) BM3D_DX_STEP(2)
        if (base + 3 < n) BM3D_DX_STEP(3)
        if (base + 4 < n) BM3D_DX_STEP(4)
        if (base + 5 < n) BM3D_DX_STEP(5)
        if (base + 6 < n) BM3D_DX_STEP(6)
#undef BM3D_DX_GSTEP
#undef BM3D_DX_STEP
#undef BM3D_DX_KEY
    }
}

// dx_rows, without the in-plane checks when every lane's window lies inside
// the plane.
template <int K, bool kCount, bool kDiv8, bool kIns>
__device__ __forceinline__ void dx_window(const uint32_t* __restrict__ lds, const PairGeom& g,
                                          const f16x2 (&rp)[8][4], float nr, KeyList<K>& kl, int cx, int cy,
                                          int range, int W, int H, uint32_t idx_base, unsigned& cnt) {
    const bool inside = cx - range >= 0 && cx + range <= W - 8 && cy - range >= 0 && cy + range <= H - 8;
    if (__all(inside))
        dx_rows<K, kCount, false, kDiv8, kIns>(lds, g, rp, nr, kl, cx, cy, range, W, H, idx_base, cnt);
    else
        dx_rows<K, kCount, true, kDiv8, kIns>(lds, g, rp, nr, kl, cx, cy, range, W, H, idx_base, cnt);
}

// ---------------------------------------------------------------------------
// Spatial search with strips shared between lanes (block steps 2 and 4).
//
// The reference blocks of a work-group row are bs apart, so the 8 columns of
// a block are the strips (bs columns x 8 rows) of 8 / bs consecutive lanes.
// Every lane of a row scans the same displacements (dx, dy) of the window,
// in the per-lane scan's order, but for its own strip only: half the strip's
// distance, |c'_s|^2 / 2 + |r'_s|^2 / 2 - <r'_s, c'_s>, is one chain of 4 bs
// v_dot2 from a sliding sum of per-column squares. The block's half distance
// is the sum over its strips, taken from the next lanes of the row by DPP
// (row_shl within rows of 16 lanes; wave_shl for rows of 32 at bs 4). The
// last 8 / bs - 1 lanes of a row own no block: their strips complete the
// row's last blocks, so a work-group row owns tbw - 8 / bs + 1 blocks.
// Keys, scan indices and rescoring are the per-lane scan's; the pre-scores
// differ from it in FP32 summation order only. Requires every owned block of
// the tile at x = x0 + lx bs (no block clamped to W - 8).

constexpr int kDppRowShl1 = 0x101, kDppRowShl2 = 0x102, kDppWaveShl1 = 0x130;

// v + (s of the lane kCtrl selects; 0 where it selects none)
template <int kCtrl>
__device__ __forceinline__ float add_dpp(float v, float s) {
    return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, s), kCtrl, 0xF, 0xF,
                                                                     true));
}

// Sum of the NS strips starting at this lane.
template <int NS, bool kWave>
__device__ __forceinline__ float strip_sum(float v) {
    if constexpr (NS == 2) {
        return add_dpp<kWave ? kDppWaveShl1 : kDppRowShl1>(v, v);
    } else {
        static_assert(NS == 4 && !kWave, "strips of 2 columns: rows of 16 lanes only");
        const float t = add_dpp<kDppRowShl1>(v, v);
        return add_dpp<kDppRowShl2>(t, t);
    }
}

// The lane's reference strip (SW columns from xs, read clamped to the plane:
// a lane without a block may lie beyond it) as negated packed vertical
// pairs, as ref_pairs; returns |r'_s|^2 / 2.
template <int SW>
__device__ __forceinline__ float ref_strip(f16x2 (&rs)[SW][4], const float* __restrict__ plane, int stride, int W,
                                           int xs, int y, float mu) {
    float nr = 0.0f;
#pragma unroll
    for (int c = 0; c < SW; ++c) {
        const int col = min(xs + c, W - 1);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const _Float16 a = (_Float16)(plane[(size_t)(y + 2 * i) * stride + col] - mu);
            const _Float16 b = (_Float16)(plane[(size_t)(y + 2 * i + 1) * stride + col] - mu);
            rs[c][i] = f16x2{-a, -b};
            nr = dot2(rs[c][i], rs[c][i], nr);
        }
    }
    return 0.5f * nr;
}

// One displacement of the strip scan (step U of a group of 8): the strip's
// new column (at `addr`), its half distance, the block's sum over strips and
// the key, as dx_step.
template <int U, int SW, bool kWave, bool kCount, bool kCheck>
__device__ __forceinline__ uint32_t strip_step(const f16x2 (&rs)[SW][4], float nrs, f16x2 (&ring)[SW][4],
                                               float (&csq)[SW], float& N, const uint32_t* __restrict__ addr, int u,
                                               unsigned vlo, unsigned vspan, uint32_t idx, uint32_t hi_mask,
                                               unsigned& cnt) {
    constexpr int kNew = (U + SW - 1) % SW;
    load_pair_col(ring[kNew], addr);
    csq[kNew] = col_sq(ring[kNew]);
    N += csq[kNew];
    float d = fma_(N, 0.5f, nrs);
#pragma unroll
    for (int c = 0; c < SW; ++c) {
#pragma unroll
        for (int i = 0; i < 4; ++i) d = dot2(rs[c][i], ring[(U + c) % SW][i], d);
    }
    N -= csq[U % SW];
    const float approx = fmaxf(strip_sum<8 / SW, kWave>(d), 0.0f);
    const uint32_t key = (__builtin_bit_cast(uint32_t, approx) & hi_mask) | idx;
    if constexpr (!kCheck) {
        if constexpr (kCount) cnt += 1u;
        return key;
    } else {
        const bool ok = (unsigned)u - vlo <= vspan;
        if constexpr (kCount) cnt += ok ? 1u : 0u;
        return ok ? key : 0xFFFFFFFFu;
    }
}

// Strip scan of the window of radius `range` around the block at (xs, y)
// (xs: the lane's strip, = its block's x when it owns one).
template <int K, int SW, bool kWave, bool kCount, bool kCheck, bool kIns>
__device__ __forceinline__ void strip_rows(const uint32_t* __restrict__ lds, const PairGeom& g,
                                           const f16x2 (&rs)[SW][4], float nrs, KeyList<K>& kl, int xs, int y,
                                           int range, int W, int H, unsigned& cnt) {
    const int n = 2 * range + 1;
    const int inc8 = (8 / SW) * g.hp2;  // 8 columns on: the same phase (bs = SW divides 8)
    uint32_t hi_mask;
    asm volatile("v_mov_b32 %0, %1" : "=v"(hi_mask) : "i"(~kIdxMask));
    f16x2 ring[SW][4];
    float csq[SW];
    for (int dyi = 0; dyi < n; ++dyi) {
        const int row = y - range + dyi;
        const int rt = row - g.ty0;
        const int off = (rt & 1) * g.plane + (rt >> 1);
        const int tc0 = xs - range - g.tx0;
        const int q = tc0 / g.bs;
        int ph = tc0 - q * g.bs;
        int colbase = (ph * g.wq + q) * g.hp2;
        float N = 0.0f;
#pragma unroll
        for (int s = 0; s + 1 < SW; ++s) {
            load_pair_col(ring[s], lds + off + colbase);
            next_pair_col(g, ph, colbase);
            csq[s] = col_sq(ring[s]);
            N += csq[s];
        }
        const uint32_t* ad[8];  // the new column of steps base .. base + 7
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            ad[k] = lds + off + colbase;
            next_pair_col(g, ph, colbase);
        }
        const int col0 = xs - range;
        const int lo = max(0, -col0), hi = min(n - 1, W - 8 - col0);
        const bool any = row >= 0 && row <= H - 8 && lo <= hi;
        const unsigned vlo = any ? (unsigned)lo : 0x40000000u;
        const unsigned vspan = any ? (unsigned)(hi - lo) : 0u;
        const uint32_t rowidx = (uint32_t)(dyi * n);
        int base = 0;
#define BM3D_ST_KEY(U)                                                                                         \
    strip_step<U, SW, kWave, kCount, kCheck>(rs, nrs, ring, csq, N, ad[U], base + U, vlo, vspan,                 \
                                             rowidx + (uint32_t)(base + U), hi_mask, cnt)
#define BM3D_ST_STEP(U)                            \
    if constexpr (kIns) kl.insert(BM3D_ST_KEY(U));  \
    else kl.push(BM3D_ST_KEY(U));
#define BM3D_ST_GSTEP(U)                                     \
    if constexpr (K > 8 || !kIns) keys[U] = BM3D_ST_KEY(U);  \
    else kl.insert(BM3D_ST_KEY(U));
        for (; base + 8 <= n; base += 8) {
            uint32_t keys[8];
            BM3D_ST_GSTEP(0)
            BM3D_ST_GSTEP(1)
            BM3D_ST_GSTEP(2)
            BM3D_ST_GSTEP(3)
            BM3D_ST_GSTEP(4)
            BM3D_ST_GSTEP(5)
            BM3D_ST_GSTEP(6)
            BM3D_ST_GSTEP(7)
            if constexpr (K > 8 || !kIns) kl.push8(keys);
#pragma unroll
            for (int k = 0; k < 8; ++k) ad[k] += inc8;
        }
        if (base < n) BM3D_ST_STEP(0)
        if (base + 1 < n) BM3D_ST_STEP(1)
        if (base + 2 < n) BM3D_ST_STEP(2)
        if (base + 3 < n) BM3D_ST_STEP(3)
        if (base + 4 < n) BM3D_ST_STEP(4)
        if (base + 5 < n) BM3D_ST_STEP(5)
        if (base + 6 < n) BM3D_ST_STEP(6)
#undef BM3D_ST_GSTEP
#undef BM3D_ST_STEP
#undef BM3D_ST_KEY
    }
}

// The spatial search of the lane's block by strips: reference strip, then
// the scan, without the in-plane checks when every owned block's window
// lies inside the plane.
template <int K, int SW, bool kWave, bool kCount, bool kIns>
__device__ __forceinline__ void strip_window(const uint32_t* __restrict__ lds, const PairGeom& g,
                                             const float* __restrict__ plane, int stride, float mu, KeyList<K>& kl,
                                             int xs, int y, bool own, int range, int W, int H, unsigned& cnt) {
    f16x2 rs[SW][4];
    const float nrs = ref_strip<SW>(rs, plane, stride, W, xs, y, mu);
    const bool inside = !own || (xs - range >= 0 && xs + range <= W - 8 && y - range >= 0 && y + range <= H - 8);
    if (__all(inside))
        strip_rows<K, SW, kWave, kCount, false, kIns>(lds, g, rs, nrs, kl, xs, y, range, W, H, cnt);
    else
        strip_rows<K, SW, kWave, kCount, true, kIns>(lds, g, rs, nrs, kl, xs, y, range, W, H, cnt);
}

// ---------------------------------------------------------------------------
// Spatial search with the cross term on the matrix cores.
//
// A wave holds 4 rows (groups) of 16 reference blocks, bs apart; lane
// (q, c) = (lane / 16, lane % 16) owns block c of group q. For one candidate
// row dy of group g, every block's candidates lie in one union row of
// S = 15 bs + 2 R + 1 positions, cut into T <= 5 tiles of 16. Per tile, two
// v_mfma_f32_16x16x32_f16 products give D = -<r'_m, c'_s> for the group's 16
// blocks m and the tile's 16 positions s, with K = 64 pixels ordered
// (column, row): A (lane (q, c): block c of the group, columns q and q + 4,
// rows 0..7) and B (lane (q, c): position c of the tile, the same columns)
// are the FP16 vertical pairs of the staged tile, so B is four words of a
// tile column, as the VALU scan loads them. Lane (q, c) receives D for
// blocks 4 q + i and position c; the values of a block's window (position
// bs m + dxi, dxi in [0, 2R]) go through LDS to lanes (dxi % 4, m), which
// form the keys (N = |c'|^2 from a per-row table of column sums) and keep
// one partial top-K per group. At the end a 4 x 4 lane-group transpose
// (v_permlane32_swap, v_permlane16_swap) brings a block's four partial lists
// to its owner, which merges them.
//
// Keys equal the VALU scan's up to FP32 summation order: the same FP16
// residuals, exact FP16 products, FP32 sums, the same packing and scan
// indices. Requires 16 blocks per row, all bs apart inside the plane,
// bs | 16, S <= 80 and 2 R + 1 <= 28. Opt-in (MatchArgs::mx_spatial): on
// MI355X it issues 13% fewer VALU instructions than the VALU scan but about
// three times the LDS traffic and waits (the exchange and the |c'|^2 table),
// and runs a few percent slower; 76% of each product is outside the blocks'
// windows.

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kMxTiles = 5;
constexpr int kMxMaxP = 28;       // 2 R + 1
constexpr int kMxExWords = 512;   // exchange words per wave: 16 P + 64 spill slots
constexpr int kMxCnWords = 176;   // column sums (S + 7) and N (S) per wave

__device__ __forceinline__ bool mx_eligible(const MatchArgs& a, int bx0) {
    const int R = a.bm_range, P = 2 * R + 1, S = 15 * a.bs + P;
    return a.mx_spatial && a.tbw == 16 && (16 % a.bs) == 0 && S <= 16 * kMxTiles && P <= kMxMaxP &&
           bx0 + 15 < a.nbx && (bx0 + 15) * a.bs <= a.W - 8;
}

// Merges sorted list b into the K smallest of a and b (a sorted ascending).
template <int K>
__device__ __forceinline__ void merge_sorted(KeyList<K>& a, const uint32_t (&b)[K]) {
    constexpr int P = K <= 8 ? 8 : 16;
    uint32_t c[P];
#pragma unroll
    for (int i = 0; i < P - K; ++i) c[i] = 0u;
#pragma unroll
    for (int i = 0; i < K; ++i) c[P - K + i] = min(a.k[i], b[K - 1 - i]);
#pragma unroll
    for (int d = P / 2; d >= 1; d /= 2)
#pragma unroll
        for (int i = 0; i < P; ++i)
            if ((i & d) == 0) cex(c[i], c[i + d]);
#pragma unroll
    for (int i = 0; i < K; ++i) a.k[i] = c[P - K + i];
}

// The spatial search of the centre slot for the 64 blocks of this wave
// into kl (the lane's own block), from the staged tile at `lds`. ex: the
// wave's 512 exchange words; cn: the wave's column-sum table. Groups are
// the outer loop, so only one group's A operands are live at a time.
template <bool kCount>
__device__ __forceinline__ void mx_window(const uint32_t* __restrict__ lds, const PairGeom& pg, const MatchArgs& a,
                                          const float* __restrict__ plane, float mu, int x0, const int (&yg)[4],
                                          const bool (&row_on)[4], uint32_t* __restrict__ ex,
                                          float* __restrict__ cn, KeyList<kKSpatial>& kl, unsigned& cnt) {
    const int lane = threadIdx.x & 63, q = lane >> 4, c = lane & 15;
    const int bs = a.bs, R = a.bm_range, P = 2 * R + 1, S = 15 * bs + P;
    const int T = (S + 15) >> 4, J = (P + 3) >> 2;
    const int cbase = a.margin - R;  // tile column of union position 0, column 0

    // exchange slots: writer (tile t, register i) -> block 4 q + i, window
    // column dxi = 16 t + c - bs (4 q + i), or the lane's spill slot
    int wr[kMxTiles][4];
#pragma unroll
    for (int t = 0; t < kMxTiles; ++t)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = 4 * q + i, dxi = 16 * t + c - bs * m;
            wr[t][i] = (dxi >= 0 && dxi < P) ? m * P + dxi : 16 * P + lane;
        }
    // reader: lane (q, c) takes window columns dxi = q + 4 j of block c
    const int rd = c * P + q;
    unsigned col_ok = 0;  // bit j: dxi in the window and the candidate column in the plane
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const int dxi = q + 4 * j, cx = x0 + bs * c - R + dxi;
        col_ok |= (j < J && dxi < P && cx >= 0 && cx <= a.W - 8) ? 1u << j : 0u;
    }
    // B operand columns: union position 16 t + c, columns q and q + 4; the
    // last tile's positions are clamped to the union (beyond it they feed
    // only spill slots)
    const int tstep = (16 / bs) * pg.hp2;  // 16 columns on: the same phase
    const int cb0 = pair_col_base(pg, cbase + c + q), cb1 = pair_col_base(pg, cbase + c + q + 4);
    const int sl = min(16 * (T - 1) + c, S - 1);
    const int cl0 = pair_col_base(pg, cbase + sl + q), cl1 = pair_col_base(pg, cbase + sl + q + 4);
    // columns of the column-sum table this lane fills (lane, lane + 64)
    const int cs0 = pair_col_base(pg, cbase + min(lane, S + 6));
    const int cs1 = pair_col_base(pg, cbase + min(lane + 64, S + 6));
    uint32_t hi_mask;
    asm volatile("v_mov_b32 %0, %1" : "=v"(hi_mask) : "i"(~kIdxMask));

    KeyList<kKSpatial> L[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        L[g].init();
        // A operands and |r'|^2 / 2 of block c of the group (FP16 residuals
        // as the staged tile holds them, negated: D accumulates -<r', c'>)
        h8 A0, A1;
        float nr;
        {
            const float* blk = plane + (size_t)yg[g] * a.stride + x0 + bs * c + q;
            float part = 0.0f;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                A0[j] = -(_Float16)(blk[(size_t)j * a.stride] - mu);
                A1[j] = -(_Float16)(blk[(size_t)j * a.stride + 4] - mu);
            }
#pragma unroll
            for (int j = 0; j < 8; j += 2) {
                const f16x2 p0{A0[j], A0[j + 1]}, p1{A1[j], A1[j + 1]};
                part = dot2(p1, p1, dot2(p0, p0, part));
            }
            part += __shfl_xor(part, 16);
            part += __shfl_xor(part, 32);
            nr = 0.5f * part;
        }
        for (int dyi = 0; dyi < P; ++dyi) {
            const int cy = yg[g] - R + dyi;
            const int rt = cy - pg.ty0;
            const uint32_t* tile = lds + (rt & 1) * pg.plane + (rt >> 1);
            // column sums of squares over the 8 candidate rows, then N per position
            if (lane < S + 7) {
                f16x2 w[4];
                load_pair_col(w, tile + cs0);
                cn[lane] = col_sq(w);
                if (lane + 64 < S + 7) {
                    load_pair_col(w, tile + cs1);
                    cn[lane + 64] = col_sq(w);
                }
            }
            for (int s = lane; s < S; s += 64) {
                float nsum = cn[s];
#pragma unroll
                for (int u = 1; u < 8; ++u) nsum += cn[s + u];
                cn[kMxCnWords - 80 + s] = nsum;
            }
            // cross terms, tile by tile, into the exchange slots
#pragma unroll
            for (int t = 0; t < kMxTiles; ++t) {
                if (t >= T) break;  // uniform
                const bool last = t == T - 1;
                const uint32_t* p0 = tile + (last ? cl0 : cb0 + t * tstep);
                const uint32_t* p1 = tile + (last ? cl1 : cb1 + t * tstep);
                h8 B0, B1;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const f16x2 v0 = as_h2(p0[w]), v1 = as_h2(p1[w]);
                    B0[2 * w] = v0[0]; B0[2 * w + 1] = v0[1];
                    B1[2 * w] = v1[0]; B1[2 * w + 1] = v1[1];
                }
                f4 D = {0.0f, 0.0f, 0.0f, 0.0f};
                D = __builtin_amdgcn_mfma_f32_16x16x32_f16(A0, B0, D, 0, 0, 0);
                D = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1, B1, D, 0, 0, 0);
                // through scalars (a bit cast of the vector element D[i]
                // compiled to four stores of element 0)
                const float d0 = D[0], d1 = D[1], d2 = D[2], d3 = D[3];
                ex[wr[t][0]] = __float_as_uint(d0);
                ex[wr[t][1]] = __float_as_uint(d1);
                ex[wr[t][2]] = __float_as_uint(d2);
                ex[wr[t][3]] = __float_as_uint(d3);
            }
            // keys of block c of the group, window columns q + 4 j
            const bool row_in = row_on[g] && cy >= 0 && cy <= a.H - 8;
            const uint32_t rowidx = (uint32_t)(dyi * P);
            uint32_t keys[8];
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int dxi = q + 4 * j;
                const float d = __uint_as_float(ex[min(rd + 4 * j, 16 * P - 1)]);
                const float nsum = cn[kMxCnWords - 80 + min(bs * c + dxi, S - 1)];
                const float approx = fmaxf(fma_(nsum, 0.5f, nr) + d, 0.0f);
                const bool ok = row_in && ((col_ok >> j) & 1u);
                if constexpr (kCount) cnt += ok ? 1u : 0u;
                keys[j] = ok ? ((__builtin_bit_cast(uint32_t, approx) & hi_mask) | (rowidx + (uint32_t)dxi))
                             : 0xFFFFFFFFu;
            }
            keys[7] = 0xFFFFFFFFu;
            L[g].push8(keys);
        }
    }

    // lane group g collects the four partial lists of its own blocks
#pragma unroll
    for (int k = 0; k < kKSpatial; ++k) {
        auto s02 = __builtin_amdgcn_permlane32_swap(L[0].k[k], L[2].k[k], false, false);
        auto s13 = __builtin_amdgcn_permlane32_swap(L[1].k[k], L[3].k[k], false, false);
        L[0].k[k] = s02[0]; L[2].k[k] = s02[1]; L[1].k[k] = s13[0]; L[3].k[k] = s13[1];
        auto s01 = __builtin_amdgcn_permlane16_swap(L[0].k[k], L[1].k[k], false, false);
        auto s23 = __builtin_amdgcn_permlane16_swap(L[2].k[k], L[3].k[k], false, false);
        L[0].k[k] = s01[0]; L[1].k[k] = s01[1]; L[2].k[k] = s23[0]; L[3].k[k] = s23[1];
    }
    kl = L[0];
    merge_sorted(kl, L[1].k);
    merge_sorted(kl, L[2].k);
    merge_sorted(kl, L[3].k);
}

// Exact distance between the block at (x, y) of `rplane` and the candidate
// at (cx, cy) of `plane`, in the order of cpu_source/source.cpp:123-136 +
// 98-103 (as ring_ssd); rows read from memory as two 16-byte loads each.
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
__device__ __forceinline__ float exact_ssd(const float* __restrict__ rplane, int x, int y,
                                           const float* __restrict__ plane, int stride, int cx, int cy) {
    const float* p = plane + (size_t)cy * stride + cx;
    const float* q = rplane + (size_t)y * stride + x;
    float d[8][8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const f4u p0 = *reinterpret_cast<const f4u*>(p + j * stride);
        const f4u p1 = *reinterpret_cast<const f4u*>(p + j * stride + 4);
        const f4u q0 = *reinterpret_cast<const f4u*>(q + j * stride);
        const f4u q1 = *reinterpret_cast<const f4u*>(q + j * stride + 4);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            d[j][c] = q0[c] - p0[c];
            d[j][c + 4] = q1[c] - p1[c];
        }
    }
    float s[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        float e0 = d[0][c] * d[0][c];
        e0 = fma_(d[2][c], d[2][c], e0);
        e0 = fma_(d[4][c], d[4][c], e0);
        e0 = fma_(d[6][c], d[6][c], e0);
        float e1 = d[1][c] * d[1][c];
        e1 = fma_(d[3][c], d[3][c], e1);
        e1 = fma_(d[5][c], d[5][c], e1);
        e1 = fma_(d[7][c], d[7][c], e1);
        s[c] = e0 + e1;
    }
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// Exhaustive exact scan of the window of radius `range` around (cx, cy) in
// scan order (the FP32 kernel's scan_window, reading memory directly): the
// path of slots whose range is too wide for FP16.
template <bool kCount>
__device__ void exact_window(const float* __restrict__ rplane, int x, int y, const float* __restrict__ plane,
                             int stride, int cx, int cy, int range, int z, int W, int H, Best8& best,
                             unsigned& cnt) {
    for (int dy = -range; dy <= range; ++dy) {
        const int Y = cy + dy;
        if (Y < 0 || Y > H - 8) continue;
        for (int dx = -range; dx <= range; ++dx) {
            const int X = cx + dx;
            if (X < 0 || X > W - 8) continue;
            if constexpr (kCount) ++cnt;
            const float e = exact_ssd(rplane, x, y, plane, stride, X, Y);
            if (e < best.e[7]) best.insert(e, pack_pos(X, Y, z));
        }
    }
}

// r' = r - mu as negated packed vertical pairs: rp[c][i] = -(r'[2i][c],
// r'[2i+1][c]) (the scan accumulates -<r', c'>); returns |r'|^2 / 2.
__device__ __forceinline__ float ref_pairs(f16x2 (&rp)[8][4], const float* __restrict__ plane, int stride, int x,
                                           int y, float mu) {
    float nr = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const _Float16 a = (_Float16)(plane[(size_t)(y + 2 * i) * stride + x + c] - mu);
            const _Float16 b = (_Float16)(plane[(size_t)(y + 2 * i + 1) * stride + x + c] - mu);
            rp[c][i] = f16x2{-a, -b};
            nr = dot2(rp[c][i], rp[c][i], nr);
        }
    return 0.5f * nr;
}

}  // namespace

// One lane per reference block; a work-group covers TBW x TBY lanes (TBW =
// a.tbw, 16 or 32; TBY = blockDim.x / TBW) and owns a.obw x TBY blocks (obw
// = TBW, or fewer with the strip search: the last lanes of each row then
// own no block); grid (ceil(nbx / obw), ceil(nby / TBY), frames). kKt: keys
// kept per temporal frame (>= 2 ps_num).
// kPh: the whole search (kPhBoth), or at radius > 0 its spatial part alone
// (kPhSpatial: the best 8 go to a.slist) or its temporal part alone
// (kPhTemporal: starts from a.slist), as two launches of one evaluation.
template <bool kCount, int kKt, int kSp, int kPh>
__device__ __forceinline__ void match_dx_body(const MatchArgs& a, Wg w) {
    constexpr bool kDiv8 = kSp != kSpOther;
    // key selection by sorted insertion where it pays (dx_rows): in the
    // two-launch kernels; the one-launch kernel keeps batch merges, which
    // its register allocation handles better (measured)
    constexpr bool kIns = kPh != kPhBoth;
    extern __shared__ __align__(16) float lds_f[];
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds_f);
    __shared__ uint32_t centre_s[256][8];  // the spatial list, per lane (before it: mx_window's exchange)
    __shared__ float mx_cn[4][kMxCnWords];  // mx_window's column sums, per wave
    const int wshift = a.tbw == 32 ? 5 : 4;
    const int tby = blockDim.x >> wshift;
    const int lx = threadIdx.x & (a.tbw - 1), ly = threadIdx.x >> wshift;
    const int f = w.z;
    const int n = a.first_frame + f;
    const int r = a.radius;
    const int bx0 = w.x * a.obw, by0 = w.y * tby;
    const int bxi = bx0 + lx, byi = by0 + ly;
    const bool own = lx < a.obw;
    const bool active = own && bxi < a.nbx && byi < a.nby;
    const int x = min(min(bxi, a.nbx - 1) * a.bs, a.W - 8);
    const int y = min(min(byi, a.nby - 1) * a.bs, a.H - 8);

    TileGeom g;
    g.tx0 = min(bx0 * a.bs, a.W - 8) - a.margin;
    g.ty0 = min(by0 * a.bs, a.H - 8) - a.margin;
    g.tw = a.tile_w;
    g.th = a.tile_h;
    g.bs = a.bs;
    g.wq = a.tile_wq;
    g.hp = a.tile_hp;
    PairGeom pg;
    pg.tx0 = g.tx0;
    pg.ty0 = g.ty0;
    pg.tw = g.tw;
    pg.th = g.th;
    pg.bs = g.bs;
    pg.wq = g.wq;
    pg.hp2 = a.tile_hp2;
    pg.plane = a.bs * a.tile_wq * a.tile_hp2;

    auto slot_plane = [&](int z) -> const float* {
        const int fi = clampi(n - r + z, 0, a.num_frames - 1);
        return a.clip + (size_t)fi * a.frame_stride;
    };
    const float* plane_r = slot_plane(r);

    // Stages slot z as FP16 pairs offset by the centre pixel mu of the
    // slot's frame (the same for every work-group, so the pre-scores do not
    // depend on the launch shape); returns false (the slot is then scanned
    // exactly from memory) when a pixel of the tile is not finite or further
    // than kDxMaxSpan from mu.
    auto stage = [&](const float* img, float& mu) -> bool {
        mu = img[(size_t)(a.H / 2) * a.stride + a.W / 2];
        if (!(__builtin_fabsf(mu) <= 1e30f)) mu = 0.0f;
        return stage_pairs(lds, pg, img, a.W, a.H, a.stride, mu);
    };

    unsigned cnt = 0, cnt_wave = 0;  // cnt_wave: counted for other lanes' blocks too
    Best8 best;
    // the block's spatial list between the two launches (active lanes only)
    auto slist_at = [&]() -> uint32_t* {
        return a.slist + ((size_t)f * a.nblocks + (size_t)byi * a.nbx + bxi) * 16;
    };
    if constexpr (kPh == kPhTemporal) {
        // the spatial best 8 of the first launch: positions, then distances
        best.init(kFltMax, pack_pos(x, y, r));
        if (active) {
            const uint4* in = reinterpret_cast<const uint4*>(slist_at());
            const uint4 p0 = in[0], p1 = in[1], e0 = in[2], e1 = in[3];
            const uint32_t pv[8] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
            const uint32_t ev[8] = {e0.x, e0.y, e0.z, e0.w, e1.x, e1.y, e1.z, e1.w};
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                best.p[i] = pv[i];
                best.e[i] = __builtin_bit_cast(float, ev[i]);
            }
        }
    } else {
        float mu;
        const bool fp16 = stage(plane_r, mu);
        if (fp16) {
            KeyList<kKSpatial> kl;
            kl.init();
            bool mx = false, strip = false;
            if constexpr (kSp == kSpMx) mx = mx_eligible(a, bx0) && (tby & 3) == 0;
            // strips: every owned block of the tile on the grid x0 + lx bs
            if constexpr (kSp >= kSpStrip4) strip = (min(bx0 + a.obw, a.nbx) - 1) * a.bs <= a.W - 8;
            if (strip) {
                const int xs = (bx0 + lx) * a.bs;
                if constexpr (kSp == kSpStrip4)
                    strip_window<kKSpatial, 4, false, kCount, kIns>(lds, pg, plane_r, a.stride, mu, kl, xs, y, own,
                                                              a.bm_range, a.W, a.H, cnt);
                else if constexpr (kSp == kSpStrip4W)
                    strip_window<kKSpatial, 4, true, kCount, kIns>(lds, pg, plane_r, a.stride, mu, kl, xs, y, own,
                                                             a.bm_range, a.W, a.H, cnt);
                else if constexpr (kSp == kSpStrip2)
                    strip_window<kKSpatial, 2, false, kCount, kIns>(lds, pg, plane_r, a.stride, mu, kl, xs, y, own,
                                                              a.bm_range, a.W, a.H, cnt);
            } else if (mx) {
                if constexpr (kSp == kSpMx) {
                const int wv = threadIdx.x >> 6;
                int yg[4];
                bool row_on[4];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int row = by0 + 4 * wv + g;
                    yg[g] = min(min(row, a.nby - 1) * a.bs, a.H - 8);
                    row_on[g] = row < a.nby;
                }
                mx_window<kCount>(lds, pg, a, plane_r, mu, bx0 * a.bs, yg, row_on,
                                  &centre_s[64 * wv][0], mx_cn[wv], kl, cnt_wave);
                }
            } else {
                f16x2 rp[8][4];
                const float nr = ref_pairs(rp, plane_r, a.stride, x, y, mu);
                dx_window<kKSpatial, kCount, kDiv8, kIns>(lds, pg, rp, nr, kl, x, y, a.bm_range, a.W, a.H, 0u, cnt);
            }
            Best8Id bi;
            bi.init(kFltMax, pack_pos(x, y, r));
            const int nW = 2 * a.bm_range + 1;
            const float inv_nW = 1.0f / (float)nW;
#pragma unroll 1
            for (int i = 0; i < kKSpatial; ++i) {
                const uint32_t key = kl.pop();
                if (key != 0xFFFFFFFFu) {
                    const uint32_t idx = key & kIdxMask;
                    const int dyi = (int)(((float)idx + 0.5f) * inv_nW), dxi = (int)idx - dyi * nW;
                    const int cx = x - a.bm_range + dxi, cy = y - a.bm_range + dyi;
                    const float e = exact_ssd(plane_r, x, y, plane_r, a.stride, cx, cy);
                    if (e < bi.b.e[7] || (e == bi.b.e[7] && idx < bi.id[7])) bi.insert(e, pack_pos(cx, cy, r), idx);
                }
            }
            best = bi.b;
        } else {
            best.init(kFltMax, pack_pos(x, y, r));
            exact_window<kCount>(plane_r, x, y, plane_r, a.stride, x, y, a.bm_range, r, a.W, a.H, best, cnt);
        }
    }
    if constexpr (kPh == kPhSpatial) {
        if (r > 0) {  // uniform: the temporal launch continues from slist
            if constexpr (kCount) {
                const unsigned total = (active ? cnt : 0u) + cnt_wave;
                if (total) atomicAdd(a.ncand, (unsigned long long)total);
            }
            if (active) {
                uint4* out = reinterpret_cast<uint4*>(slist_at());
                out[0] = make_uint4(best.p[0], best.p[1], best.p[2], best.p[3]);
                out[1] = make_uint4(best.p[4], best.p[5], best.p[6], best.p[7]);
                out[2] = make_uint4(__builtin_bit_cast(uint32_t, best.e[0]), __builtin_bit_cast(uint32_t, best.e[1]),
                                    __builtin_bit_cast(uint32_t, best.e[2]), __builtin_bit_cast(uint32_t, best.e[3]));
                out[3] = make_uint4(__builtin_bit_cast(uint32_t, best.e[4]), __builtin_bit_cast(uint32_t, best.e[5]),
                                    __builtin_bit_cast(uint32_t, best.e[6]), __builtin_bit_cast(uint32_t, best.e[7]));
            }
            return;
        }
    }

    if (r > 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) centre_s[threadIdx.x][i] = best.p[i];
        const int nWt = 2 * a.ps_range + 1;
        for (int dir = -1; dir <= 1; dir += 2) {
            uint32_t last[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) last[i] = centre_s[threadIdx.x][i];
            for (int t = 1; t <= r; ++t) {
                const int z = r + dir * t;
                const float* plane_z = slot_plane(z);
                __syncthreads();
                float mu;
                const bool fp16 = stage(plane_z, mu);
                Best8 fr;
                fr.init(kFltMax, pack_pos(0, 0, z));
                if (fp16) {
                    f16x2 rp[8][4];
                    const float nr = ref_pairs(rp, plane_r, a.stride, x, y, mu);
                    KeyList<kKt> kl;
                    kl.init();
                    for (int i = 0; i < a.ps_num; ++i) {
                        uint32_t p = last[0];
#pragma unroll
                        for (int k = 1; k < 8; ++k) p = (i == k) ? last[k] : p;
                        dx_window<kKt, kCount, kDiv8, kIns>(lds, pg, rp, nr, kl, pos_x(p), pos_y(p), a.ps_range, a.W, a.H,
                                               (uint32_t)(i * nWt * nWt), cnt);
                    }
                    Best8Id fi;
                    fi.init(kFltMax, pack_pos(0, 0, z));
                    const float inv_nWt = 1.0f / (float)nWt, inv_nWt2 = 1.0f / (float)(nWt * nWt);
#pragma unroll 1
                    for (int i = 0; i < kKt; ++i) {
                        const uint32_t key = kl.pop();
                        if (key != 0xFFFFFFFFu) {
                            const int idx = (int)(key & kIdxMask);
                            const int w = (int)(((float)idx + 0.5f) * inv_nWt2);
                            const int e2 = idx - w * nWt * nWt;
                            const int dyi = (int)(((float)e2 + 0.5f) * inv_nWt), dxi = e2 - dyi * nWt;
                            uint32_t p = last[0];
#pragma unroll
                            for (int k = 1; k < 8; ++k) p = (w == k) ? last[k] : p;
                            const int cx = pos_x(p) - a.ps_range + dxi, cy = pos_y(p) - a.ps_range + dyi;
                            const float e = exact_ssd(plane_r, x, y, plane_z, a.stride, cx, cy);
                            if (e < fi.b.e[7] || (e == fi.b.e[7] && (uint32_t)idx < fi.id[7]))
                                fi.insert(e, pack_pos(cx, cy, z), (uint32_t)idx);
                        }
                    }
                    fr = fi.b;
                } else {
                    for (int i = 0; i < a.ps_num; ++i) {
                        uint32_t p = last[0];
#pragma unroll
                        for (int k = 1; k < 8; ++k) p = (i == k) ? last[k] : p;
                        exact_window<kCount>(plane_r, x, y, plane_z, a.stride, pos_x(p), pos_y(p), a.ps_range, z, a.W,
                                             a.H, fr, cnt);
                    }
                }
                for (int i = 0; i < a.ps_num; ++i) {
                    const float e = fr.err_at(i);
                    if (e < best.e[7]) best.insert(e, fr.pos_at(i));
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) last[i] = fr.p[i];
            }
        }
    }

    // centre insert (source/kernel.cu:616-647)
    const uint32_t cpos = pack_pos(x, y, r);
    bool found = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) found |= best.p[i] == cpos;
    if (!found) {
#pragma unroll
        for (int i = 7; i >= 1; --i) best.p[i] = best.p[i - 1];
        best.p[0] = cpos;
    }
    if constexpr (kCount) {
        const unsigned total = (active ? cnt : 0u) + cnt_wave;
        if (total) atomicAdd(a.ncand, (unsigned long long)total);
    }
    if (active && a.groups) {
        uint4* out = reinterpret_cast<uint4*>(a.groups + ((size_t)f * a.nblocks + (size_t)byi * a.nbx + bxi) * 8);
        out[0] = make_uint4(best.p[0], best.p[1], best.p[2], best.p[3]);
        out[1] = make_uint4(best.p[4], best.p[5], best.p[6], best.p[7]);
    }
}

// kSp: the spatial search variant (kSp* above; dx_spatial_variant); kPh:
// the part of the search (match_dx_body).
template <bool kCount, int kKt, int kSp, int kPh>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void bm3d_match_dx(MatchArgs a) {
    match_dx_body<kCount, kKt, kSp, kPh>(a, launch_wg_xcd());
}

// The spatial launch of the two-launch strip search: its tile and the
// spatial search alone fit 4 waves per SIMD (128 VGPRs).
template <bool kCount, int kSp>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void bm3d_match_dx_spatial(MatchArgs a) {
    match_dx_body<kCount, 8, kSp, kPhSpatial>(a, launch_wg_xcd());
}

template <int kKt, int kSp>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void bm3d_match_dx_multi(
    const MatchArgs* __restrict__ jobs, const JobGrid* __restrict__ grids, int njobs) {
    const int wg = multi_wg_xcd();
    const int j = find_job(grids, njobs, wg);
    const MatchArgs& a = jobs[j];
    match_dx_body<false, kKt, kSp, kPhBoth>(a, job_wg(grids[j], wg));
}

// Multi-job forms of the two-launch strip search's launches.
template <int kSp>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void bm3d_match_dx_spatial_multi(
    const MatchArgs* __restrict__ jobs, const JobGrid* __restrict__ grids, int njobs) {
    const int wg = multi_wg_xcd();
    const int j = find_job(grids, njobs, wg);
    const MatchArgs& a = jobs[j];
    match_dx_body<false, 8, kSp, kPhSpatial>(a, job_wg(grids[j], wg));
}

template <int kKt>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void bm3d_match_dx_temporal_multi(
    const MatchArgs* __restrict__ jobs, const JobGrid* __restrict__ grids, int njobs) {
    const int wg = multi_wg_xcd();
    const int j = find_job(grids, njobs, wg);
    const MatchArgs& a = jobs[j];
    match_dx_body<false, kKt, kSpLane, kPhTemporal>(a, job_wg(grids[j], wg));
}

namespace {
template <int kKt, int kSp, int kPh = kPhBoth>
hipError_t launch_dx(const MatchArgs& a, int tby, size_t lds_bytes, int frames, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.nbx + a.obw - 1) / a.obw, (a.nby + tby - 1) / tby, frames);
    dim3 block(a.tbw * tby);
    if constexpr (kPh == kPhSpatial) {
        if (a.ncand) hipLaunchKernelGGL((bm3d_match_dx_spatial<true, kSp>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_match_dx_spatial<false, kSp>), grid, block, lds_bytes, s, a);
    } else {
        if (a.ncand) hipLaunchKernelGGL((bm3d_match_dx<true, kKt, kSp, kPh>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_match_dx<false, kKt, kSp, kPh>), grid, block, lds_bytes, s, a);
    }
    return hipGetLastError();
}
template <int kKt, int kSp>
hipError_t launch_dx_multi(const MatchArgs* jobs, const JobGrid* grids, int njobs, int wgs, int threads,
                           size_t lds_bytes, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    hipLaunchKernelGGL((bm3d_match_dx_multi<kKt, kSp>), dim3(wgs), dim3(threads), lds_bytes, s, jobs, grids,
                       njobs);
    return hipGetLastError();
}
template <int kKt, int kSp>
hipError_t set_dx_limit(size_t bytes) {
    const void* fns[] = {reinterpret_cast<const void*>(&bm3d_match_dx<false, kKt, kSp, kPhBoth>),
                         reinterpret_cast<const void*>(&bm3d_match_dx<true, kKt, kSp, kPhBoth>),
                         reinterpret_cast<const void*>(&bm3d_match_dx_multi<kKt, kSp>),
                         reinterpret_cast<const void*>(&bm3d_match_dx<false, kKt, kSpLane, kPhTemporal>),
                         reinterpret_cast<const void*>(&bm3d_match_dx<true, kKt, kSpLane, kPhTemporal>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    if constexpr (kSp >= kSpStrip4) {  // the spatial launches exist for the strip searches
        const void* sp[] = {reinterpret_cast<const void*>(&bm3d_match_dx_spatial<false, kSp>),
                            reinterpret_cast<const void*>(&bm3d_match_dx_spatial<true, kSp>),
                            reinterpret_cast<const void*>(&bm3d_match_dx_spatial_multi<kSp>),
                            reinterpret_cast<const void*>(&bm3d_match_dx_temporal_multi<kKt>)};
        for (const void* f : sp) {
            hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e != hipSuccess) return e;
        }
    }
    return hipSuccess;
}
template <int kSp>
hipError_t launch_dx_keys(const MatchArgs& a, int tby, size_t lds_bytes, int frames, hipStream_t s) {
    switch (a.kt) {
        case 4: return launch_dx<4, kSp>(a, tby, lds_bytes, frames, s);
        case 8: return launch_dx<8, kSp>(a, tby, lds_bytes, frames, s);
        default: return launch_dx<16, kSp>(a, tby, lds_bytes, frames, s);
    }
}
template <int kSp>
hipError_t launch_dx_multi_keys(const MatchArgs* jobs, const JobGrid* grids, int njobs, int wgs, int threads,
                                size_t lds_bytes, int kt, hipStream_t s) {
    switch (kt) {
        case 4: return launch_dx_multi<4, kSp>(jobs, grids, njobs, wgs, threads, lds_bytes, s);
        case 8: return launch_dx_multi<8, kSp>(jobs, grids, njobs, wgs, threads, lds_bytes, s);
        default: return launch_dx_multi<16, kSp>(jobs, grids, njobs, wgs, threads, lds_bytes, s);
    }
}
template <int kSp>
hipError_t set_dx_limits(size_t bytes) {
    hipError_t e = set_dx_limit<4, kSp>(bytes);
    if (e == hipSuccess) e = set_dx_limit<8, kSp>(bytes);
    if (e == hipSuccess) e = set_dx_limit<16, kSp>(bytes);
    return e;
}
}  // namespace

int dx_temporal_keys(int ps_num) { return ps_num <= 4 ? 8 : 16; }

bool dx_block_step_divides_8(int block_step) { return 8 % block_step == 0; }

int dx_spatial_variant(const MatchArgs& a) {
    if (!dx_block_step_divides_8(a.bs)) return kSpOther;
    if (a.strip) return a.bs == 2 ? kSpStrip2 : a.tbw == 32 ? kSpStrip4W : kSpStrip4;
    return a.mx_spatial ? kSpMx : kSpLane;
}

hipError_t launch_match_dx(const MatchArgs& a, int tby, size_t lds_bytes, int frames, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    if (a.phase == kPhTemporal) {
        switch (a.kt) {
            case 4: return launch_dx<4, kSpLane, kPhTemporal>(a, tby, lds_bytes, frames, s);
            case 8: return launch_dx<8, kSpLane, kPhTemporal>(a, tby, lds_bytes, frames, s);
            default: return launch_dx<16, kSpLane, kPhTemporal>(a, tby, lds_bytes, frames, s);
        }
    }
    if (a.phase == kPhSpatial) {  // the strip searches (the temporal keys do not matter)
        switch (dx_spatial_variant(a)) {
            case kSpStrip4W: return launch_dx<8, kSpStrip4W, kPhSpatial>(a, tby, lds_bytes, frames, s);
            case kSpStrip2: return launch_dx<8, kSpStrip2, kPhSpatial>(a, tby, lds_bytes, frames, s);
            default: return launch_dx<8, kSpStrip4, kPhSpatial>(a, tby, lds_bytes, frames, s);
        }
    }
    switch (dx_spatial_variant(a)) {
        case kSpOther: return launch_dx_keys<kSpOther>(a, tby, lds_bytes, frames, s);
        case kSpMx: return launch_dx_keys<kSpMx>(a, tby, lds_bytes, frames, s);
        case kSpStrip4: return launch_dx_keys<kSpStrip4>(a, tby, lds_bytes, frames, s);
        case kSpStrip4W: return launch_dx_keys<kSpStrip4W>(a, tby, lds_bytes, frames, s);
        case kSpStrip2: return launch_dx_keys<kSpStrip2>(a, tby, lds_bytes, frames, s);
        default: return launch_dx_keys<kSpLane>(a, tby, lds_bytes, frames, s);
    }
}

hipError_t launch_match_dx_multi(const MatchArgs* jobs, const JobGrid* grids, int njobs, int wgs, int threads,
                                 size_t lds_bytes, int kt, int variant, hipStream_t s, int phase) {
    BM3D_INJECT_LAUNCH();
    const dim3 g(wgs), blk(threads);
    if (phase == kPhSpatial) {
        switch (variant) {
            case kSpStrip4W:
                hipLaunchKernelGGL((bm3d_match_dx_spatial_multi<kSpStrip4W>), g, blk, lds_bytes, s, jobs, grids, njobs);
                break;
            case kSpStrip2:
                hipLaunchKernelGGL((bm3d_match_dx_spatial_multi<kSpStrip2>), g, blk, lds_bytes, s, jobs, grids, njobs);
                break;
            default:
                hipLaunchKernelGGL((bm3d_match_dx_spatial_multi<kSpStrip4>), g, blk, lds_bytes, s, jobs, grids, njobs);
        }
        return hipGetLastError();
    }
    if (phase == kPhTemporal) {
        switch (kt) {
            case 4: hipLaunchKernelGGL((bm3d_match_dx_temporal_multi<4>), g, blk, lds_bytes, s, jobs, grids, njobs); break;
            case 8: hipLaunchKernelGGL((bm3d_match_dx_temporal_multi<8>), g, blk, lds_bytes, s, jobs, grids, njobs); break;
            default:
                hipLaunchKernelGGL((bm3d_match_dx_temporal_multi<16>), g, blk, lds_bytes, s, jobs, grids, njobs);
        }
        return hipGetLastError();
    }
    switch (variant) {
        case kSpOther: return launch_dx_multi_keys<kSpOther>(jobs, grids, njobs, wgs, threads, lds_bytes, kt, s);
        case kSpMx: return launch_dx_multi_keys<kSpMx>(jobs, grids, njobs, wgs, threads, lds_bytes, kt, s);
        case kSpStrip4: return launch_dx_multi_keys<kSpStrip4>(jobs, grids, njobs, wgs, threads, lds_bytes, kt, s);
        case kSpStrip4W: return launch_dx_multi_keys<kSpStrip4W>(jobs, grids, njobs, wgs, threads, lds_bytes, kt, s);
        case kSpStrip2: return launch_dx_multi_keys<kSpStrip2>(jobs, grids, njobs, wgs, threads, lds_bytes, kt, s);
        default: return launch_dx_multi_keys<kSpLane>(jobs, grids, njobs, wgs, threads, lds_bytes, kt, s);
    }
}

namespace {
// Writes one launch's arguments and its one-job grid where a multi-job kernel
// reads them, in stream order (a kernel node, so graph captures hold it).
__global__ void bm3d_store_match_table(MatchArgs a, JobGrid g, MatchArgs* __restrict__ da, JobGrid* __restrict__ dg) {
    if (threadIdx.x == 0) {
        *da = a;
        *dg = g;
    }
}
}  // namespace

hipError_t launch_match_dx_table(const MatchArgs& a, int tby, size_t lds_bytes, int frames, void* table,
                                 hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    MatchArgs* da = static_cast<MatchArgs*>(table);
    JobGrid* dg = reinterpret_cast<JobGrid*>(static_cast<char*>(table) + kMatchTableGridOffset);
    const int gx = (a.nbx + a.obw - 1) / a.obw, gy = (a.nby + tby - 1) / tby;
    hipLaunchKernelGGL(bm3d_store_match_table, dim3(1), dim3(64), 0, s, a, JobGrid{0, gx, gy}, da, dg);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_match_dx_multi(da, dg, 1, gx * gy * frames, a.tbw * tby, lds_bytes, a.kt, dx_spatial_variant(a), s,
                                 a.phase);
}

hipError_t set_match_dx_lds_limit(size_t bytes) {
    hipError_t e = hipSuccess;
    for (hipError_t (*f)(size_t) : {set_dx_limits<kSpLane>, set_dx_limits<kSpOther>, set_dx_limits<kSpMx>,
                                    set_dx_limits<kSpStrip4>, set_dx_limits<kSpStrip4W>, set_dx_limits<kSpStrip2>})
        if (e == hipSuccess) e = f(bytes);
    return e;
}

}  // namespace bm3d
