// Block matching with FP16 MFMA pre-scoring and certified exact rescoring
// (BM3D_PRECISION_MIXED). Produces the same group lists as the FP32 kernel
// bm3d_match (the reference's search: source/kernel.cu:432-647,
// cpu_source/source.cpp:143-321) for every input.
//
// Work-group: 4 waves, 8 x 8 reference blocks, 16 per wave. Per searched
// slot the work-group stages the union of its search windows once:
//
//   T16  the slot's pixels minus a tile offset mu, rounded to FP16
//   NC   for every candidate position, the sum of squares of its 8 x 8 FP16
//        block (sum of 8 row sums; every term is exact in FP32)
//
// MFMA phase (a whole wave per reference block). The distance of the block
// r to the candidate at (Y, X) is ||r~ - c~||^2 = |r~|^2 + NC(Y, X) -
// 2 <r~, c~>. The dot products of one block with a 16 x 16 square of
// candidates are one 16 x 16 x 192 product: A (16 x 192) holds, for
// candidate row m, image rows Yb + m + j, columns Xb .. Xb + 23 (j = 0..7);
// B (192 x 16) holds, for column shift s, the block's row j placed at
// columns s .. s + 7 of the 24 (a Toeplitz band). So A is read straight out
// of T16 (16-byte aligned, Xb a multiple of 8) and B is built once per
// block and slot. Six v_mfma_f32_16x16x32_f16 per 256 candidates; the
// approximate distances of the window entries go to an LDS map.
//
// Scan phase (a quad of lanes per reference block, the block's lists
// replicated in its 4 lanes):
//   1. the 8th smallest of 16 interleaved group minima of the map, u, bounds
//      the approximate 8th-best distance from above;
//   2. a rigorous bound |A - E| <= alpha sqrt(E) + beta + rho E between the
//      stored approximation A and the exact FP32 distance E (FP16 rounding of
//      pixels and map, FP32 rounding of the sums) turns u into a threshold
//      theta such that every entry whose exact distance can reach the top 8
//      has A <= theta;
//   3. the entries with A <= theta are compacted, in scan order, into a
//      queue (in place, over the map) and rescored exactly in the
//      reference's FP32 order, the 4 lanes taking 2 columns each and adding
//      the column sums with DPP in the reference's tree order; the sorted
//      insertion (strict <) in scan order then reproduces the reference's
//      list.
// A slot whose pixels are not finite or span more than +-8 around mu is
// rescored exhaustively (every entry queued); the result is the same.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm3d_device.h"

namespace bm3d {

namespace {

using namespace dev;

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr float kInf = __builtin_huge_valf();
constexpr int kMxThreads = 64 * kMxWaves;
// Largest |pixel - mu| for which the FP16 path is used; keeps every map
// value (|A| <= 4 * 64 * M^2) finite in FP16.
constexpr float kMxMaxSpan = 8.0f;

// Per-slot statistics of the staged tile, in LDS.
struct SlotStats {
    float mu, M, lo, hi;
    int all;  // rescore every entry
};

// Writes of one wave to LDS are visible to its later reads.
__device__ __forceinline__ void wave_lds_sync() {
    asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off));
    return v;
}

// Bounds between the stored approximate distance and the exact one for a
// search of slot z: pixel rounding error h <= (2^-11 + 2^-23) M + 2^-25 for
// |v - mu| <= M (FP32 subtraction, then FP16 round-to-nearest), so each
// difference d~ = r~ - c~ is within eta = h_r + h_c of the true difference,
// and sum |d~^2 - d^2| <= 16 eta sqrt(S) + 64 eta^2 (Cauchy-Schwarz).
// Accumulation error of |r~|^2 + NC - 2 <r~, c~> in FP32: at most
// 6.1e-6 (|r~|^2 + |c~|^2) <= 3.9e-4 (Mz^2 + Mr^2); taken 4x. FP16 storage
// of A and the FP32 rounding of the exact sum add at most 5e-4 E.
struct Cert {
    float alpha, beta, rho;
    __device__ __forceinline__ void init(float Mz, float Mr) {
        const float hz = 5.5e-4f * Mz + 1e-7f;
        const float hr = 5.5e-4f * Mr + 1e-7f;
        const float eta = hz + hr;
        alpha = 16.02f * eta;
        beta = 64.1f * eta * eta + 1.6e-3f * (Mz * Mz + Mr * Mr) + 1e-6f;
        rho = 5e-4f;
    }
    // Every entry with exact distance <= the 8th smallest exact distance
    // has A <= theta(u), when at least 8 entries have A <= u.
    __device__ __forceinline__ float theta(float u) const {
        const float v = fmaxf(u + beta, 0.0f);
        const float om = 1.0f - rho;
        const float sq = (alpha + __builtin_sqrtf(alpha * alpha + 4.0f * om * v)) / (2.0f * om);
        const float U = sq * sq;
        return (U + alpha * sq + beta + rho * U) * 1.0001f + 1e-6f;
    }
};

// Compare-exchange network: ascending bitonic sort of 16 values.
__device__ __forceinline__ void sort16(float (&v)[16]) {
#pragma unroll
    for (int k = 2; k <= 16; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const float a = v[i], b = v[l];
                    v[i] = up ? fminf(a, b) : fmaxf(a, b);
                    v[l] = up ? fmaxf(a, b) : fminf(a, b);
                }
            }
}

// Stages slot image `img` for the tile at (tx0, ty0): T16 and NC, plus the
// slot statistics. Whole work-group; ends with a barrier. `hs` (row sums,
// column-major, hs_pitch floats per column) is scratch that aliases the
// approximate-distance maps, which are dead while a slot is staged.
__device__ void stage_slot(const MxArgs& a, const float* __restrict__ img, int tx0, int ty0,
                           _Float16* __restrict__ t16, float* __restrict__ nc, float* __restrict__ hs,
                           float* red, SlotStats* st) {
    const int tid = threadIdx.x;
    const int area = a.t16_rows * a.t16_pitch;
    constexpr int kBatch = 8;  // loads in flight per thread
    float lo = kInf, hi = -kInf;
    bool bad = false;
    for (int base = 0; base < area; base += kBatch * kMxThreads) {
        float v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = min(base + u * kMxThreads + tid, area - 1);
            const int rr = idx / a.t16_pitch;
            const int cc = idx - rr * a.t16_pitch;
            v[u] = img[(size_t)clampi(ty0 + rr, 0, a.H - 1) * a.stride + clampi(tx0 + cc, 0, a.W - 1)];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            bad |= !(__builtin_fabsf(v[u]) <= 1e30f);
            lo = fminf(lo, v[u]);
            hi = fmaxf(hi, v[u]);
        }
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off));
        hi = fmaxf(hi, __shfl_xor(hi, off));
    }
    const int anybad = __any(bad);
    const int wave = tid >> 6;
    __syncthreads();  // previous users of red / st / the maps are done
    if ((tid & 63) == 0) {
        red[wave] = lo;
        red[kMxWaves + wave] = hi;
        red[2 * kMxWaves + wave] = anybad ? 1.0f : 0.0f;
    }
    __syncthreads();
    lo = red[0];
    hi = red[kMxWaves];
    float badf = red[2 * kMxWaves];
#pragma unroll
    for (int w = 1; w < kMxWaves; ++w) {
        lo = fminf(lo, red[w]);
        hi = fmaxf(hi, red[kMxWaves + w]);
        badf += red[2 * kMxWaves + w];
    }
    const float mu = 0.5f * lo + 0.5f * hi;
    const float M = fmaxf(hi - mu, mu - lo) * (1.0f + 1.0f / (1 << 20));
    const bool all = badf != 0.0f || !(M <= kMxMaxSpan);
    if (tid == 0) {
        st->mu = mu;
        st->M = M;
        st->lo = lo;
        st->hi = hi;
        st->all = all ? 1 : 0;
    }
    if (all) {
        __syncthreads();
        return;
    }
    for (int base = 0; base < area; base += kBatch * kMxThreads) {
        float v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = min(base + u * kMxThreads + tid, area - 1);
            const int rr = idx / a.t16_pitch;
            const int cc = idx - rr * a.t16_pitch;
            v[u] = img[(size_t)clampi(ty0 + rr, 0, a.H - 1) * a.stride + clampi(tx0 + cc, 0, a.W - 1)];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const int idx = base + u * kMxThreads + tid;
            if (idx < area) t16[idx] = (_Float16)(v[u] - mu);
        }
    }
    __syncthreads();
    // row sums hs[col][row] = sum_k t16[row][col + k]^2: a task is one row
    // and 8 columns (16 halves, two aligned 16-byte reads)
    const int ngrp = (a.nc_cols + 7) >> 3;
    for (int task = tid; task < a.t16_rows * ngrp; task += kMxThreads) {
        const int row = task / ngrp;
        const int c0 = (task - row * ngrp) << 3;
        const h8 x0 = *reinterpret_cast<const h8*>(t16 + row * a.t16_pitch + c0);
        const h8 x1 = *reinterpret_cast<const h8*>(t16 + row * a.t16_pitch + c0 + 8);
        float q[16];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float f0 = (float)x0[e], f1 = (float)x1[e];
            q[e] = f0 * f0;
            q[8 + e] = f1 * f1;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float sum = ((q[c] + q[c + 1]) + (q[c + 2] + q[c + 3])) +
                              ((q[c + 4] + q[c + 5]) + (q[c + 6] + q[c + 7]));
            if (c0 + c < a.nc_cols) hs[(c0 + c) * a.hs_pitch + row] = sum;
        }
    }
    __syncthreads();
    // NC[col][y] = sum_j hs[col][y + j]: a task is one column, 16 rows
    const int chunks = (a.nc_rows + 15) >> 4;
    for (int task = tid; task < a.nc_cols * chunks; task += kMxThreads) {
        const int cc = task % a.nc_cols;
        const int y0 = (task / a.nc_cols) << 4;
        float h[24];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const f4 v4 = *reinterpret_cast<const f4*>(hs + cc * a.hs_pitch + y0 + 4 * u);
            h[4 * u] = v4[0];
            h[4 * u + 1] = v4[1];
            h[4 * u + 2] = v4[2];
            h[4 * u + 3] = v4[3];
        }
#pragma unroll
        for (int yy = 0; yy < 16; ++yy) {
            const float sum = ((h[yy] + h[yy + 1]) + (h[yy + 2] + h[yy + 3])) +
                              ((h[yy + 4] + h[yy + 5]) + (h[yy + 6] + h[yy + 7]));
            if (y0 + yy < a.nc_rows) nc[cc * a.nc_pitch + y0 + yy] = sum;
        }
    }
    __syncthreads();
}

// One search of a reference block: `nwin` windows of half-size `hw` around
// the centres `cen` (packed positions), in slot `z`.
struct Search {
    int nwin, hw, nW, win_stride;
    int z;
};

// Cursor over the MFMA blocks (16 x 16 candidates) that cover the windows of
// one search, window by window, rows then columns. Wave-uniform.
struct BlockCursor {
    int w, bxi, byi, nbx, nby, wx, wy, xb0, yb0;
    __device__ __forceinline__ void window(const Search& s, const uint32_t* cen, int tx0, int ty0) {
        const uint32_t c = cen[w];
        wx = pos_x(c) - s.hw - tx0;  // window origin, tile coordinates (>= 0)
        wy = pos_y(c) - s.hw - ty0;
        xb0 = wx & ~7;
        yb0 = wy & ~3;
        nbx = (wx - xb0 + s.nW + 15) >> 4;
        nby = (wy - yb0 + s.nW + 15) >> 4;
        bxi = 0;
        byi = 0;
    }
    __device__ __forceinline__ void next(const Search& s, const uint32_t* cen, int tx0, int ty0) {
        if (++bxi < nbx) return;
        bxi = 0;
        if (++byi < nby) return;
        if (++w < s.nwin) window(s, cen, tx0, ty0);
    }
};

// MFMA phase for one reference block: approximate distances of every
// in-window entry into map (halves, window w at w * win_stride, row-major).
// rv: this lane's pixel (row lane / 8, column lane % 8) of the block; nr:
// |r~|^2 of the block for this slot.
__device__ void mfma_block(const MxArgs& a, float rv, float nr, const Search& s, const uint32_t* cen,
                           float mu, int tx0, int ty0, const _Float16* __restrict__ t16,
                           const float* __restrict__ nc, _Float16* __restrict__ map,
                           uint32_t* __restrict__ rowbuf) {
    const int lane = threadIdx.x & 63;
    // The block minus mu in FP16 as Toeplitz rows: padded row P_j[t] =
    // r~[j][t - 16] (t in [16, 24), else 0) is stored twice, as dwords
    // (P[2q], P[2q+1]) and (P[2q+1], P[2q+2]), so that any 8 consecutive
    // halves are 4 dword-aligned dwords of one of the copies.
    {
        const int j = lane >> 3, k = lane & 7;
        const _Float16 hv = (_Float16)(rv - mu);
        _Float16* rb16 = reinterpret_cast<_Float16*>(rowbuf);
        rb16[(0 * 8 + j) * 42 + 16 + k] = hv;
        rb16[(1 * 8 + j) * 42 + 15 + k] = hv;
    }
    wave_lds_sync();
    // B fragments: lane (s = lane & 15, g = lane >> 4); K-block b covers
    // K = 32 b + 8 g + e = 24 j + k', value r~[j][k' - s]
    const int sft = lane & 15, g = lane >> 4, m = lane & 15;
    h8 bf[6];
    int offA[6];
#pragma unroll
    for (int b = 0; b < 6; ++b) {
        const int K0 = 32 * b + 8 * g;
        const int jj = K0 / 24, k0 = K0 - 24 * jj;
        const int t0 = k0 - sft + 16;
        const uint32_t* q = rowbuf + ((t0 & 1) * 8 + jj) * 21 + (t0 >> 1);
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const u4 u = {q[0], q[1], q[2], q[3]};
        bf[b] = __builtin_bit_cast(h8, u);
        // A operand: candidate row m, image row + jj, column + k0
        offA[b] = jj * a.t16_pitch + k0;
    }
    BlockCursor cur;
    cur.w = 0;
    cur.window(s, cen, tx0, ty0);
    auto epilogue = [&](const f4& acc, int w, int wx, int wy, int Xb, int Yb) {
        const f4 nv = *reinterpret_cast<const f4*>(nc + (Xb + sft) * a.nc_pitch + Yb + 4 * g);
        const int dx = Xb + sft - wx;
        const bool colok = dx >= 0 && dx < s.nW;
        _Float16* mw = map + w * s.win_stride;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int dy = Yb + 4 * g + i - wy;
            const float A = fma_(-2.0f, acc[i], nv[i] + nr);
            if (colok && dy >= 0 && dy < s.nW) mw[dy * s.nW + dx] = (_Float16)A;
        }
    };
    // two blocks at a time: independent accumulator chains
    while (cur.w < s.nwin) {
        const int w0 = cur.w, wx0 = cur.wx, wy0 = cur.wy;
        const int X0 = cur.xb0 + 16 * cur.bxi, Y0 = cur.yb0 + 16 * cur.byi;
        cur.next(s, cen, tx0, ty0);
        const bool two = cur.w < s.nwin;
        const int w1 = cur.w, wx1 = cur.wx, wy1 = cur.wy;
        const int X1 = two ? cur.xb0 + 16 * cur.bxi : X0, Y1 = two ? cur.yb0 + 16 * cur.byi : Y0;
        if (two) cur.next(s, cen, tx0, ty0);
        const _Float16* base0 = t16 + (Y0 + m) * a.t16_pitch + X0;
        const _Float16* base1 = t16 + (Y1 + m) * a.t16_pitch + X1;
        f4 acc0 = {0.0f, 0.0f, 0.0f, 0.0f}, acc1 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int b = 0; b < 6; ++b) {
            const h8 af0 = *reinterpret_cast<const h8*>(base0 + offA[b]);
            const h8 af1 = *reinterpret_cast<const h8*>(base1 + offA[b]);
            acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(af0, bf[b], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(af1, bf[b], acc1, 0, 0, 0);
        }
        epilogue(acc0, w0, wx0, wy0, X0, Y0);
        if (two) epilogue(acc1, w1, wx1, wy1, X1, Y1);
    }
}

// Scan phase for the reference block of this quad: fills `list` (sorted,
// strict-< insertion in scan order) with the exact top 8 of the search.
template <bool kCount>
__device__ void scan_block(const MxArgs& a, const float* __restrict__ plane_z, const Search& s,
                           const uint32_t* cen, bool active, bool all, const Cert& cert,
                           const float (&rc)[8][2], _Float16* __restrict__ map, Best8& list,
                           unsigned& cnt) {
    const int g = threadIdx.x & 3;
    const int nEnt = s.nW * s.nW;
    const int chunks = (nEnt + 3) >> 2;
    const float inv_nW = 1.0f / (float)s.nW;
    // pass 1: 16 group minima (entry chunk c goes to group (c & 3, lane))
    float gm[4] = {kInf, kInf, kInf, kInf};
    int nvalid = 0;
    for (int w = 0; w < s.nwin; ++w) {
        const uint32_t c = cen[w];
        const int wx0 = pos_x(c) - s.hw, wy0 = pos_y(c) - s.hw;
        const _Float16* mw = map + w * s.win_stride;
        for (int c0 = 0; c0 < chunks; c0 += 4) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = 4 * (c0 + u) + g;
                const int dy = (int)(((float)e + 0.5f) * inv_nW);
                const int dx = e - dy * s.nW;
                const int X = wx0 + dx, Y = wy0 + dy;
                const bool valid = active && e < nEnt && X >= 0 && X <= a.W - 8 && Y >= 0 && Y <= a.H - 8;
                const float A = (float)mw[min(e, s.win_stride - 1)];
                gm[u] = valid ? fminf(gm[u], A) : gm[u];
                nvalid += valid ? 1 : 0;
            }
        }
    }
    if constexpr (kCount) cnt += (unsigned)nvalid;
    const int tot_valid = quad_bcast<0>(nvalid) + quad_bcast<1>(nvalid) + quad_bcast<2>(nvalid) +
                          quad_bcast<3>(nvalid);
    float v16[16];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v16[u] = quad_bcast<0>(gm[u]);
        v16[4 + u] = quad_bcast<1>(gm[u]);
        v16[8 + u] = quad_bcast<2>(gm[u]);
        v16[12 + u] = quad_bcast<3>(gm[u]);
    }
    sort16(v16);
    const bool take_all = all || tot_valid <= 8 || !(v16[7] < kInf);
    const float theta = take_all ? kInf : cert.theta(v16[7]);

    // pass 2: queue (packed w << 12 | dy << 6 | dx, scan order) in place;
    // four chunks are read before any of them is queued
    uint16_t* q16 = reinterpret_cast<uint16_t*>(map);
    int qn = 0;
    for (int w = 0; w < s.nwin; ++w) {
        const uint32_t c = cen[w];
        const int wx0 = pos_x(c) - s.hw, wy0 = pos_y(c) - s.hw;
        const _Float16* mw = map + w * s.win_stride;
        for (int c0 = 0; c0 < chunks; c0 += 4) {
            float A[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) A[u] = (float)mw[min(4 * (c0 + u) + g, s.win_stride - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int e = 4 * (c0 + u) + g;
                const int dy = (int)(((float)e + 0.5f) * inv_nW);
                const int dx = e - dy * s.nW;
                const int X = wx0 + dx, Y = wy0 + dy;
                const bool valid = active && e < nEnt && X >= 0 && X <= a.W - 8 && Y >= 0 && Y <= a.H - 8;
                const bool hit = valid && (take_all || A[u] <= theta);
                const int hv = hit ? 1 : 0;
                const int h0 = quad_bcast<0>(hv), h1 = quad_bcast<1>(hv), h2 = quad_bcast<2>(hv),
                          h3 = quad_bcast<3>(hv);
                const int pre = (g > 0 ? h0 : 0) + (g > 1 ? h1 : 0) + (g > 2 ? h2 : 0);
                if (hit) q16[qn + pre] = (uint16_t)((w << 12) | (dy << 6) | dx);
                qn += h0 + h1 + h2 + h3;
            }
        }
    }
    wave_lds_sync();

    // exact rescoring of the queue in scan order, kResc entries per round
    // (their pixel loads in flight together)
    constexpr int kResc = 4;
    const int maxq = wave_max(qn);
    const int cx = 2 * g;
    for (int q0 = 0; q0 < maxq; q0 += kResc) {
        int X[kResc], Y[kResc];
        bool live[kResc];
        float cv[kResc][8][2];
#pragma unroll
        for (int b = 0; b < kResc; ++b) {
            live[b] = q0 + b < qn;
            const uint32_t code = q16[live[b] ? q0 + b : 0];
            const int w = (int)(code >> 12);
            const int dy = (int)((code >> 6) & 63u), dx = (int)(code & 63u);
            const uint32_t c = cen[w];
            X[b] = live[b] ? pos_x(c) - s.hw + dx : 0;
            Y[b] = live[b] ? pos_y(c) - s.hw + dy : 0;
        }
#pragma unroll
        for (int b = 0; b < kResc; ++b) {
            const float* cp = plane_z + (size_t)Y[b] * a.stride + X[b] + cx;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                cv[b][j][0] = cp[(size_t)j * a.stride];
                cv[b][j][1] = cp[(size_t)j * a.stride + 1];
            }
        }
#pragma unroll
        for (int b = 0; b < kResc; ++b) {
            float sc[2];
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                float d[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) d[j] = rc[j][k] - cv[b][j][k];
                float e0 = d[0] * d[0];
                e0 = fma_(d[2], d[2], e0);
                e0 = fma_(d[4], d[4], e0);
                e0 = fma_(d[6], d[6], e0);
                float e1 = d[1] * d[1];
                e1 = fma_(d[3], d[3], e1);
                e1 = fma_(d[5], d[5], e1);
                e1 = fma_(d[7], d[7], e1);
                sc[k] = e0 + e1;
            }
            // ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)), lane g
            // holding columns 2g, 2g + 1 (cpu_source/source.cpp:98-103)
            const float p = sc[0] + sc[1];
            const float pq = p + lane_xor1(p);
            const float E = pq + lane_xor2(pq);
            if (live[b] && E < list.e[7]) list.insert(E, pack_pos(X[b], Y[b], s.z));
        }
    }
    wave_lds_sync();
}

}  // namespace

template <bool kCount>
__global__ __launch_bounds__(kMxThreads) void bm3d_match_mx(MxArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    _Float16* t16 = reinterpret_cast<_Float16*>(smem);
    float* nc = reinterpret_cast<float*>(smem + a.off_nc);
    _Float16* maps = reinterpret_cast<_Float16*>(smem + a.off_map);
    uint32_t* rowbufs = reinterpret_cast<uint32_t*>(smem + a.off_row);
    float* red = reinterpret_cast<float*>(smem + a.off_misc);
    SlotStats* st = reinterpret_cast<SlotStats*>(red + 4 * kMxWaves);
    uint32_t* cens = reinterpret_cast<uint32_t*>(red + 4 * kMxWaves + 8);  // [64][8] window centres
    float* nrs = reinterpret_cast<float*>(cens + kMxRefsX * kMxRefsY * 8);  // [64] |r~|^2
    float* hs = reinterpret_cast<float*>(maps);  // row sums while staging

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int f = blockIdx.z;
    const int n = a.first_frame + f;
    const int r = a.radius;
    const int bx0 = blockIdx.x * kMxRefsX, by0 = blockIdx.y * kMxRefsY;
    const int tx0 = min(bx0 * a.bs, a.W - 8) - a.rho;
    const int ty0 = min(by0 * a.bs, a.H - 8) - a.rho;

    auto slot_plane = [&](int z) -> const float* {
        const int fi = clampi(n - r + z, 0, a.num_frames - 1);
        return a.clip + (size_t)fi * a.frame_stride;
    };
    auto ref_of = [&](int t, int& x, int& y) -> bool {
        const int bxi = bx0 + (t % kMxRefsX), byi = by0 + (t / kMxRefsX);
        x = min(min(bxi, a.nbx - 1) * a.bs, a.W - 8);
        y = min(min(byi, a.nby - 1) * a.bs, a.H - 8);
        return bxi < a.nbx && byi < a.nby;
    };

    // this quad's reference block
    const int qd = lane >> 2, g = lane & 3;
    const int tq = wave * 16 + qd;
    int x, y;
    const bool active = ref_of(tq, x, y);
    _Float16* myrow_map = maps + (size_t)tq * a.map_stride;
    uint32_t* rowbuf = rowbufs + wave * (2 * 8 * 21);
    for (int i = lane; i < 2 * 8 * 21; i += 64) rowbuf[i] = 0u;

    const float* plane_r = slot_plane(r);
    float rc[8][2];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        rc[j][0] = plane_r[(size_t)(y + j) * a.stride + x + 2 * g];
        rc[j][1] = plane_r[(size_t)(y + j) * a.stride + x + 2 * g + 1];
    }
    unsigned cnt = 0;

    // |r~|^2 of this quad's block for slot offset mu (exact squares of the
    // FP16 values; summation order is covered by the error bound)
    auto block_norm = [&](float mu) {
        float sum = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const float h = (float)(_Float16)(rc[j][k] - mu);
                sum = fma_(h, h, sum);
            }
        sum += lane_xor1(sum);
        sum += lane_xor2(sum);
        if (g == 0) nrs[tq] = sum;
    };
    // MFMA phase of this wave's 16 blocks; the next block's pixel is loaded
    // while the current one is processed
    auto mfma_phase = [&](const Search& sr, float mu) {
        const int pj = lane >> 3, pk = lane & 7;
        int xi, yi;
        bool act = ref_of(wave * 16, xi, yi);
        float rv = plane_r[(size_t)(yi + pj) * a.stride + xi + pk];
        for (int i = 0; i < 16; ++i) {
            const int t = wave * 16 + i;
            int xn = xi, yn = yi;
            const bool actn = i + 1 < 16 ? ref_of(t + 1, xn, yn) : false;
            const float rvn = plane_r[(size_t)(yn + pj) * a.stride + xn + pk];
            if (act)
                mfma_block(a, rv, nrs[t], sr, cens + t * 8, mu, tx0, ty0, t16, nc,
                           maps + (size_t)t * a.map_stride, rowbuf);
            act = actn;
            rv = rvn;
        }
        wave_lds_sync();
    };

    // ---- spatial search in slot r
    stage_slot(a, plane_r, tx0, ty0, t16, nc, hs, red, st);
    const float mu_r = st->mu, M_r = st->M, lo_r = st->lo, hi_r = st->hi;
    const bool all_r = st->all != 0;
    Search ss;
    ss.nwin = 1;
    ss.hw = a.bm_range;
    ss.nW = 2 * a.bm_range + 1;
    ss.win_stride = a.map_stride;
    ss.z = r;
    if (g == 0) cens[tq * 8] = pack_pos(x, y, r);
    block_norm(mu_r);
    wave_lds_sync();
    if (!all_r) mfma_phase(ss, mu_r);
    Cert cert;
    cert.init(M_r, M_r);
    Best8 best;
    best.init(kFltMax, pack_pos(x, y, r));
    scan_block<kCount>(a, plane_r, ss, cens + tq * 8, active, all_r, cert, rc, myrow_map, best, cnt);

    // ---- temporal search (source/kernel.cu:495-614)
    if (r > 0) {
        uint32_t centre[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) centre[i] = best.p[i];
        Search ts;
        ts.nwin = a.ps_num;
        ts.hw = a.ps_range;
        ts.nW = 2 * a.ps_range + 1;
        ts.win_stride = a.win_stride;
        for (int dir = -1; dir <= 1; dir += 2) {
            uint32_t last[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) last[i] = centre[i];
            for (int t = 1; t <= r; ++t) {
                const int z = r + dir * t;
                ts.z = z;
                const float* plane_z = slot_plane(z);
                stage_slot(a, plane_z, tx0, ty0, t16, nc, hs, red, st);
                const float mu_z = st->mu, M_z = st->M;
                const bool all_z = st->all != 0 || all_r;
                // window centres of every block of this wave, for the MFMA phase
                if (g == 0) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) cens[tq * 8 + k] = last[k];
                }
                block_norm(mu_z);
                wave_lds_sync();
                if (!all_z) mfma_phase(ts, mu_z);
                const float M_rz = fmaxf(hi_r - mu_z, mu_z - lo_r) * (1.0f + 1.0f / (1 << 20));
                Cert cz;
                cz.init(M_z, M_rz);
                Best8 fr;
                fr.init(kFltMax, pack_pos(0, 0, z));
                scan_block<kCount>(a, plane_z, ts, cens + tq * 8, active, all_z, cz, rc, myrow_map, fr, cnt);
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
        if (active) atomicAdd(a.ncand, (unsigned long long)cnt);
    }
    if (active && g == 0 && a.groups) {
        const int bxi = bx0 + (tq % kMxRefsX), byi = by0 + (tq / kMxRefsX);
        uint4* out = reinterpret_cast<uint4*>(a.groups + ((size_t)f * a.nblocks + (size_t)byi * a.nbx + bxi) * 8);
        out[0] = make_uint4(best.p[0], best.p[1], best.p[2], best.p[3]);
        out[1] = make_uint4(best.p[4], best.p[5], best.p[6], best.p[7]);
    }
}

hipError_t launch_match_mx(const MxArgs& a, size_t lds_bytes, int frames, hipStream_t s) {
    dim3 grid((a.nbx + kMxRefsX - 1) / kMxRefsX, (a.nby + kMxRefsY - 1) / kMxRefsY, frames);
    if (a.ncand) hipLaunchKernelGGL((bm3d_match_mx<true>), grid, dim3(kMxThreads), lds_bytes, s, a);
    else hipLaunchKernelGGL((bm3d_match_mx<false>), grid, dim3(kMxThreads), lds_bytes, s, a);
    return hipGetLastError();
}

hipError_t set_match_mx_lds_limit(size_t bytes) {
    const void* fns[] = {reinterpret_cast<const void*>(&bm3d_match_mx<false>),
                         reinterpret_cast<const void*>(&bm3d_match_mx<true>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace bm3d
