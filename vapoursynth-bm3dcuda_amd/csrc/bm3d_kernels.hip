// gfx950 (MI355X) kernels for BM3D / V-BM3D.
//
// The reference runs the whole per-block pipeline in one CUDA kernel with a
// 32-thread block per 4 reference blocks (source/kernel.cu:366-757). Here the
// pipeline is split by resource profile:
//
//   bm3d_match     one lane per reference block. The work-group stages the
//                  union search window of its 16 x TBY blocks in LDS, split
//                  into block_step column phases so that the 16 lanes of a
//                  row read consecutive words. Each lane slides an 8-column
//                  register ring across the window (8 LDS reads per
//                  candidate) and keeps its top-8 in registers.
//   bm3d_filter    8 lanes per group, 8 groups per wave: gather, 3-D DCT,
//                  hard threshold or Wiener shrinkage, inverse, and
//                  aggregation (pre-rounded when extractor_exp is set) into
//                  int32 fixed-point LDS regions, per slot of an evaluation
//                  or per output frame of a run of evaluations. Per-member
//                  weights are added once at the member's corner.
//   bm3d_finalize / bm3d_den_box   stacks from the slot regions and corner
//                  maps (8x8 box sum of the corners).
//   bm3d_vagg_targets   VAggregate straight from the output-frame regions.
//   bm3d_zero / bm3d_divide / bm3d_vaggregate   elementwise passes.
//
// Every float operation is written out (the build uses -ffp-contract=off),
// so results match the CPU oracle bit for bit up to the order of the
// aggregation atomics.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bm3d_inject.h"
#include "bm3d_scan.h"
#include "bm3d_xform.h"

namespace bm3d {

namespace {

using namespace dev;

// Cosine constants of the 8-point DCT, written as double literals rounded
// once to float (the oracle rounds the same literals the same way).
constexpr float kR2 = static_cast<float>(1.41421356237309504880168872420969808);
constexpr float kK1 = static_cast<float>(1.84775906502257351225636637879357657);  // 2 cos(pi/8)
constexpr float kK3 = static_cast<float>(0.76536686473017954345691996806916186);  // 2 cos(3pi/8)
constexpr float kC1 = static_cast<float>(1.96157056080646086901156744484086917);  // 2 cos(pi/16)
constexpr float kC3 = static_cast<float>(1.66293922460509004760575233432468773);  // 2 cos(3pi/16)
constexpr float kC5 = static_cast<float>(1.11114046603920391150735136234616736);  // 2 cos(5pi/16)
constexpr float kC7 = static_cast<float>(0.39018064403225654488047858302562963);  // 2 cos(7pi/16)


// 8-point DCT-II scaled to 4 x orthonormal (the reference's normalisation,
// cpu_source/source.cpp:444-536), by an even/odd split: a = x + rev(x) feeds
// the even outputs through a 4-point DCT, b = x - rev(x) the odd outputs
// through a symmetric 4x4 cosine matrix.
__device__ __forceinline__ void dct8_fwd(float& v0, float& v1, float& v2, float& v3,
                                         float& v4, float& v5, float& v6, float& v7) {
    float a0 = v0 + v7, a1 = v1 + v6, a2 = v2 + v5, a3 = v3 + v4;
    float b0 = v0 - v7, b1 = v1 - v6, b2 = v2 - v5, b3 = v3 - v4;
    float c0 = a0 + a3, c1 = a1 + a2, d0 = a0 - a3, d1 = a1 - a2;
    v0 = kR2 * (c0 + c1);
    v4 = kR2 * (c0 - c1);
    v2 = fma_(kK1, d0, kK3 * d1);
    v6 = fma_(kK3, d0, -(kK1 * d1));
    v1 = fma_(kC1, b0, fma_(kC3, b1, fma_(kC5, b2, kC7 * b3)));
    v3 = fma_(kC3, b0, fma_(-kC7, b1, fma_(-kC1, b2, -(kC5 * b3))));
    v5 = fma_(kC5, b0, fma_(-kC1, b1, fma_(kC7, b2, kC3 * b3)));
    v7 = fma_(kC7, b0, fma_(-kC5, b1, fma_(kC3, b2, -(kC1 * b3))));
}

// Transpose of dct8_fwd (DCT-III, 4 x orthonormal): dct8_inv(dct8_fwd(x)) = 16 x.
__device__ __forceinline__ void dct8_inv(float& v0, float& v1, float& v2, float& v3,
                                         float& v4, float& v5, float& v6, float& v7) {
    float gc0 = kR2 * (v0 + v4), gc1 = kR2 * (v0 - v4);
    float gd0 = fma_(kK1, v2, kK3 * v6);
    float gd1 = fma_(kK3, v2, -(kK1 * v6));
    float ga0 = gc0 + gd0, ga3 = gc0 - gd0, ga1 = gc1 + gd1, ga2 = gc1 - gd1;
    float gb0 = fma_(kC1, v1, fma_(kC3, v3, fma_(kC5, v5, kC7 * v7)));
    float gb1 = fma_(kC3, v1, fma_(-kC7, v3, fma_(-kC1, v5, -(kC5 * v7))));
    float gb2 = fma_(kC5, v1, fma_(-kC1, v3, fma_(kC7, v5, kC3 * v7)));
    float gb3 = fma_(kC7, v1, fma_(-kC5, v3, fma_(kC3, v5, -(kC1 * v7))));
    v0 = ga0 + gb0; v7 = ga0 - gb0;
    v1 = ga1 + gb1; v6 = ga1 - gb1;
    v2 = ga2 + gb2; v5 = ga2 - gb2;
    v3 = ga3 + gb3; v4 = ga3 - gb3;
}

template <bool kInverse>
__device__ __forceinline__ void dct8(float* v) {
    if constexpr (kInverse) dct8_inv(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
    else dct8_fwd(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
}

// Along the first index of v[8][8] (v[i][c] for fixed c).
template <bool kInverse>
__device__ __forceinline__ void dct8_strided(float (&v)[8][8], int c) {
    if constexpr (kInverse)
        dct8_inv(v[0][c], v[1][c], v[2][c], v[3][c], v[4][c], v[5][c], v[6][c], v[7][c]);
    else
        dct8_fwd(v[0][c], v[1][c], v[2][c], v[3][c], v[4][c], v[5][c], v[6][c], v[7][c]);
}

// Transforms other than the DCT (bm3d_xform.h), rounded once to float, in
// constant memory: a wave-uniform kind reads its matrix through the scalar
// cache.
struct XformF32 {
    float m[3][2][64];
};
constexpr XformF32 xform_f32(const XformMatrices& d) {
    XformF32 f{};
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 64; ++j) f.m[k][i][j] = static_cast<float>(d.m[k][i][j]);
    return f;
}
__constant__ XformF32 kXformF32 = xform_f32(kXformMatrices);

// v <- M v: each output a product with v[0], then an FMA chain over
// v[1] .. v[7] (zero entries included), as the oracle's xform8.
__device__ __forceinline__ void xform_mat(float& v0, float& v1, float& v2, float& v3, float& v4, float& v5,
                                          float& v6, float& v7, const float* __restrict__ M) {
    const float x[8] = {v0, v1, v2, v3, v4, v5, v6, v7};
    float o[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        float acc = M[r * 8] * x[0];
#pragma unroll
        for (int k = 1; k < 8; ++k) acc = fma_(M[r * 8 + k], x[k], acc);
        o[r] = acc;
    }
    v0 = o[0]; v1 = o[1]; v2 = o[2]; v3 = o[3]; v4 = o[4]; v5 = o[5]; v6 = o[6]; v7 = o[7];
}

// Transform `kind` (0: the DCT butterflies, else the matrix) of v[0..7] or,
// strided, of v[0..7][c]. kind is wave-uniform.
template <bool kInverse>
__device__ __forceinline__ void xform8(float* v, int kind) {
    if (kind == 0) dct8<kInverse>(v);
    else xform_mat(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], kXformF32.m[kind - 1][kInverse]);
}
template <bool kInverse>
__device__ __forceinline__ void xform8_strided(float (&v)[8][8], int c, int kind) {
    if (kind == 0) dct8_strided<kInverse>(v, c);
    else xform_mat(v[0][c], v[1][c], v[2][c], v[3][c], v[4][c], v[5][c], v[6][c], v[7][c],
                   kXformF32.m[kind - 1][kInverse]);
}

// ---------------------------------------------------------------- matching

// Block matching: one lane per reference block; a work-group covers
// TBW x TBY reference blocks (TBW = a.tbw, 16 or 32; TBY = blockDim.x /
// TBW). Grid: (ceil(nbx/TBW), ceil(nby/TBY), frames). Spatial search on the centre slot (source/kernel.cu
// :432-492), then for r > 0 the predictive temporal search
// (source/kernel.cu:495-614, cpu_source/source.cpp:220-321), then the centre
// insert (source/kernel.cu:616-647). Writes 8 packed positions per block;
// with kCount, also adds the number of in-plane candidates to *a.ncand.
// kCost selects the block distance (MatchArgs::cost, the bm_error_s option
// of rtc_source/kernel.hpp:372-504). kPlane: the candidates are read from
// the planes in global memory instead of an LDS tile (search reaches whose
// union window does not fit the LDS); same order, same distances.
template <int kCost, bool kCount, bool kPlane>
__device__ __forceinline__ void match_body(const MatchArgs& a, Wg w) {
    extern __shared__ float lds[];
    const int wshift = a.tbw == 32 ? 5 : 4;
    const int tby = blockDim.x >> wshift;
    const int lx = threadIdx.x & (a.tbw - 1);
    const int ly = threadIdx.x >> wshift;
    const int f = w.z;
    const int n = a.first_frame + f;
    const int r = a.radius;

    const int bx0 = w.x * a.tbw;
    const int by0 = w.y * tby;
    const int bxi = bx0 + lx;
    const int byi = by0 + ly;
    const bool active = bxi < a.nbx && byi < a.nby;
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

    auto slot_plane = [&](int z) -> const float* {
        int fi = clampi(n - r + z, 0, a.num_frames - 1);
        return a.clip + (size_t)fi * a.frame_stride;
    };

    float rb[8][8];
    if constexpr (kPlane) {
        const float* blk = slot_plane(r) + (size_t)y * a.stride + x;
#pragma unroll
        for (int j = 0; j < 8; ++j)
#pragma unroll
            for (int c = 0; c < 8; ++c) rb[j][c] = blk[(size_t)j * a.stride + c];
    } else {
        stage_tile(lds, g, slot_plane(r), a.W, a.H, a.stride);
        __syncthreads();
        const int tc = x - g.tx0;
        const int rt = y - g.ty0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float* colp = lds + tile_col_base(g, tc + c) + rt;
#pragma unroll
            for (int j = 0; j < 8; ++j) rb[j][c] = colp[j];
        }
    }

    BlockStat st{0.0f, 0.0f};
    if constexpr (kCost >= 2) st = block_stat(rb);

    unsigned cnt = 0;
    Best8 best;
    best.init(kFltMax, pack_pos(x, y, r));
    if constexpr (kPlane)
        scan_window_plane<kCost, kCount>(slot_plane(r), a.stride, rb, st, best, x, y, a.bm_range, r, a.W, a.H, cnt);
    else
        scan_window<kCost, kCount>(lds, g, rb, st, best, x, y, a.bm_range, r, a.W, a.H, cnt);

    if (r > 0) {
        uint32_t centre[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) centre[i] = best.p[i];
        for (int dir = -1; dir <= 1; dir += 2) {
            uint32_t last[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) last[i] = centre[i];
            for (int t = 1; t <= r; ++t) {
                const int z = r + dir * t;
                if constexpr (!kPlane) {
                    __syncthreads();
                    stage_tile(lds, g, slot_plane(z), a.W, a.H, a.stride);
                    __syncthreads();
                }
                Best8 fr;
                fr.init(kFltMax, pack_pos(0, 0, z));
                for (int i = 0; i < a.ps_num; ++i) {
                    uint32_t p = last[0];
#pragma unroll
                    for (int k = 1; k < 8; ++k) p = (i == k) ? last[k] : p;
                    if constexpr (kPlane)
                        scan_window_plane<kCost, kCount>(slot_plane(z), a.stride, rb, st, fr, pos_x(p), pos_y(p),
                                                         a.ps_range, z, a.W, a.H, cnt);
                    else
                        scan_window<kCost, kCount>(lds, g, rb, st, fr, pos_x(p), pos_y(p), a.ps_range, z, a.W, a.H,
                                                   cnt);
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

    // centre insert
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
    if (active && a.groups) {
        uint4* out = reinterpret_cast<uint4*>(
            a.groups + ((size_t)f * a.nblocks + (size_t)byi * a.nbx + bxi) * 8);
        out[0] = make_uint4(best.p[0], best.p[1], best.p[2], best.p[3]);
        out[1] = make_uint4(best.p[4], best.p[5], best.p[6], best.p[7]);
    }
}

template <int kCost, bool kCount, bool kPlane>
__global__ __launch_bounds__(256) void bm3d_match(MatchArgs a) {
    match_body<kCost, kCount, kPlane>(a, launch_wg_xcd());
}

template <int kCost>
__global__ __launch_bounds__(256) void bm3d_match_multi(const MatchArgs* __restrict__ jobs,
                                                        const JobGrid* __restrict__ grids, int njobs) {
    const int wg = multi_wg_xcd();
    const int j = find_job(grids, njobs, wg);
    const MatchArgs& a = jobs[j];
    match_body<kCost, false, false>(a, job_wg(grids[j], wg));
}

// ---------------------------------------------------------------- filtering

// 8x8 transpose across the 8 lanes of a group: lane k holds x[j] = M[j][k]
// before and x[j] = M[k][j] after. Three exchange stages (d = 4, 2, 1): the
// element at (register j, lane k) trades places with (j ^ d, k ^ d) when bit
// d of j and k differ.
__device__ __forceinline__ void transpose_member(float (&x)[8], int k) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j & 4) continue;
        const float a = lane_xor4(x[j + 4]), b = lane_xor4(x[j]);
        x[j] = (k & 4) ? a : x[j];
        x[j + 4] = (k & 4) ? x[j + 4] : b;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j & 2) continue;
        const float a = lane_xor2(x[j + 2]), b = lane_xor2(x[j]);
        x[j] = (k & 2) ? a : x[j];
        x[j + 2] = (k & 2) ? x[j + 2] : b;
    }
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
        const float a = lane_xor1(x[j + 1]), b = lane_xor1(x[j]);
        x[j] = (k & 1) ? a : x[j];
        x[j + 1] = (k & 1) ? x[j + 1] : b;
    }
}

// Forward 3-D transform: per member along j (column layout), transpose,
// along k (row layout); then along i. Coefficients end in row layout.
// kGeneric: transform kinds t2 (rows, columns) and t1 (members) as chosen by
// the plan (rtc_source/kernel.hpp:585-605); otherwise the DCT throughout.
template <bool kGeneric>
__device__ __forceinline__ void forward3d(float (&v)[8][8], int k, int t2, int t1) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if constexpr (kGeneric) xform8<false>(v[i], t2);
        else dct8<false>(v[i]);
        transpose_member(v[i], k);
        if constexpr (kGeneric) xform8<false>(v[i], t2);
        else dct8<false>(v[i]);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if constexpr (kGeneric) xform8_strided<false>(v, c, t1);
        else dct8_strided<false>(v, c);
    }
}

// Inverse 3-D transform from row layout: along i; then per member along k,
// transpose, along j. The estimate ends in column layout.
template <bool kGeneric>
__device__ __forceinline__ void inverse3d(float (&v)[8][8], int k, int t2, int t1) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        if constexpr (kGeneric) xform8_strided<true>(v, c, t1);
        else dct8_strided<true>(v, c);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        if constexpr (kGeneric) xform8<true>(v[i], t2);
        else dct8<true>(v[i]);
        transpose_member(v[i], k);
        if constexpr (kGeneric) xform8<true>(v[i], t2);
        else dct8<true>(v[i]);
    }
}

// Collaborative filtering (source/kernel.cu:649-756). A work-group of
// blockDim.x / 64 waves owns a tile of tbx x tby reference blocks and filters
// their groups 8 per wave (lane k of a group holds column k; member
// transposes go through DPP lane exchanges, not LDS). Weighted estimates are
// accumulated as int32 fixed point (units of 2^fx_exponent) with integer LDS
// atomics -- LDS float atomics are several times slower -- into regions
// covering the tile plus `rm` pixels (every member of a tile's groups lies
// inside them when rm >= bm_range + radius * ps_range), per temporal slot of
// one evaluation or per output frame over a run of evaluations
// (FilterArgs::mode, filter_walk). With extractor_exp the pre-rounded values
// are whole multiples of the unit, so the integer sums are exact. Members
// outside the regions (mode 0 only) go to global float atomics on the
// stack. Member weights go to a corner map (one atomic per member).
template <bool kFinal, bool kChroma, bool kGeneric>
__device__ __forceinline__ void filter_body(const FilterArgs& a, Wg wg) {
    extern __shared__ float lds[];
    __shared__ SlotTable st;
    int* acc = reinterpret_cast<int*>(lds);
    const int nwaves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 3;
    const int k = lane & 7;
    const int tx = wg.x, ty = wg.y;
    const int rx0 = min(tx * a.tbx * a.bs, a.W - 8) - a.rm;
    const int ry0 = min(ty * a.tby * a.bs, a.H - 8) - a.rm;
    const int ngroups = a.tbx * a.tby;
    const float E = a.extractor;
    const float fx_scale = ldexpf(1.0f, -fx_exponent(a.fx_range, a.fx_exp));
    const WrapSink ws{a.wrap, a.wrap_flag};
    // one weighted estimate can reach 2^30 units (tiny extractor_exp on
    // large samples): every value then takes the split 64-bit add
    const bool wide = (a.vbound * *a.range + E) * fx_scale >= 0x1p30f;

    filter_walk(a, wg, acc, st, kChroma ? 3 : 1, kChroma, [&](int p, int n, int gf) {
        const float sigma = a.sigma[p];
        const float* src = a.src + (size_t)p * a.plane_stride;
        float* res = a.res ? a.res + (size_t)gf * a.res_frame_stride + (size_t)p * a.res_plane_stride : nullptr;

        for (int base = 0; base < ngroups; base += nwaves * 8) {
            const int li = base + wave * 8 + g;
            const int bx = tx * a.tbx + li % a.tbx;
            const int by = ty * a.tby + li / a.tbx;
            const bool active = li < ngroups && bx < a.nbx && by < a.nby;
            const int blk = active ? by * a.nbx + bx : 0;

            uint32_t pk[8];
            {
                const uint4* gp = reinterpret_cast<const uint4*>(a.groups + ((size_t)gf * a.nblocks + blk) * 8);
                const uint4 q0 = gp[0], q1 = gp[1];
                pk[0] = q0.x; pk[1] = q0.y; pk[2] = q0.z; pk[3] = q0.w;
                pk[4] = q1.x; pk[5] = q1.y; pk[6] = q1.z; pk[7] = q1.w;
            }
            const float* mp[8];  // member block, column k, in the noisy clip
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                // positions come from bm3d_match and are in range; the clamps
                // only keep a corrupted list from addressing outside the planes
                const int z = min(pos_z(pk[i]), 2 * a.radius);
                const int px = min(pos_x(pk[i]), a.W - 8);
                const int py = min(pos_y(pk[i]), a.H - 8);
                pk[i] = pack_pos(px, py, z);
                const int fi = clampi(n - a.radius + z, 0, a.num_frames - 1);
                mp[i] = src + (size_t)fi * a.frame_stride + (size_t)py * a.stride + px + k;
            }
            float v[8][8];
            float w;
            if constexpr (kFinal) {
                // Wiener shrinkage (source/kernel.cu:284-326). The basic
                // estimate's group is transformed and turned into coefficients
                // before the noisy group is loaded, so at most two groups are
                // live. Lane k holds row k of every member (row layout).
                const ptrdiff_t roff = a.ref - a.src;  // ref shares the clip layout
                float cf[8][8];
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) cf[i][j] = mp[i][roff + (ptrdiff_t)j * a.stride];
                forward3d<kGeneric>(cf, k, a.t2d, a.t1d);
                const float s2 = sigma * sigma;
                float sum = 0.0f;
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const float sq = cf[i][c] * cf[i][c];
                        float coef = sq / (sq + s2);
                        if (i == 0 && c == 0) coef = (k == 0) ? 1.0f : coef;
                        sum = fma_(coef, coef, sum);
                        cf[i][c] = coef;
                    }
                sum = sum + lane_xor1(sum);
                sum = sum + lane_xor2(sum);
                sum = sum + lane_xor4(sum);
                w = 1.0f / sum;
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[i][j] = mp[i][(size_t)j * a.stride];
                forward3d<kGeneric>(v, k, a.t2d, a.t1d);
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int c = 0; c < 8; ++c) v[i][c] = (v[i][c] * cf[i][c]) * (1.0f / 4096.0f);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int j = 0; j < 8; ++j) v[i][j] = mp[i][(size_t)j * a.stride];
                forward3d<kGeneric>(v, k, a.t2d, a.t1d);
                // hard thresholding (source/kernel.cu:208-252)
                int kept = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const float val = v[i][c];
                        const float thr = (i == 0 && c == 0 && k == 0) ? 0.0f : sigma;
                        const bool keep = fabsf(val) >= thr;
                        kept += keep ? 1 : 0;
                        v[i][c] = keep ? val * (1.0f / 4096.0f) : 0.0f;
                    }
                kept += lane_xor1(kept);
                kept += lane_xor2(kept);
                kept += lane_xor4(kept);
                w = 1.0f / (float)kept;
            }
            inverse3d<kGeneric>(v, k, a.t2d, a.t1d);

            if (active) {
                // aggregation with pre-rounding (source/kernel.cu:717-756)
                const float wr = (w + E) - E;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int z = pos_z(pk[i]), px = pos_x(pk[i]), py = pos_y(pk[i]);
                    const int mult = st.mult[z];
                    if (mult == 0) continue;
                    const int lx = px - rx0, ly = py - ry0;
                    if (lx >= 0 && ly >= 0 && lx + 8 <= a.rw && ly + 8 <= a.rh) {
                        int* q = acc + st.region[z] + ly * a.rw + lx + k;
                        const size_t widx = st.corner[z] + (size_t)py * a.stride + px + k;
                        // a slot read `mult` times by VAggregate is added as often
                        if (!wide) {
                            int iv[8][1];
#pragma unroll
                            for (int j = 0; j < 8; ++j) {
                                float val = w * v[i][j];
                                val = (val + E) - E;
                                iv[j][0] = (int)__builtin_rintf(val * fx_scale);
                            }
                            for (int c = 0; c < mult; ++c) add_units<8, 1>(q, a.rw, iv, ws, widx, a.stride);
                        } else {
                            for (int c = 0; c < mult; ++c)
#pragma unroll
                                for (int j = 0; j < 8; ++j) {
                                    float val = w * v[i][j];
                                    val = (val + E) - E;
                                    add_units_wide(q + j * a.rw, __builtin_rintf(val * fx_scale), ws,
                                                   widx + (size_t)j * a.stride);
                                }
                        }
                    } else if (res) {
                        float* q = res + (size_t)(2 * z) * a.plane_size + (size_t)py * a.stride + px + k;
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            float val = w * v[i][j];
                            val = (val + E) - E;
                            unsafeAtomicAdd(q + (size_t)j * a.stride, val);
                        }
                    }
                    if (k == 0) {
                        float* cq = a.corner + st.corner[z] + (size_t)py * a.stride + px;
                        for (int c = 0; c < mult; ++c) unsafeAtomicAdd(cq, wr);
                    }
                }
            }
        }
    });
}

template <bool kFinal, bool kChroma, bool kGeneric>
__global__ __launch_bounds__(256) void bm3d_filter(FilterArgs a) {
    filter_body<kFinal, kChroma, kGeneric>(a, launch_wg());
}

template <bool kFinal, bool kChroma, bool kGeneric>
__global__ __launch_bounds__(256) void bm3d_filter_multi(const FilterArgs* __restrict__ jobs,
                                                         const JobGrid* __restrict__ grids, int njobs) {
    const int j = find_job(grids, njobs, blockIdx.x);
    const FilterArgs& a = jobs[j];
    filter_body<kFinal, kChroma, kGeneric>(a, job_wg(grids[j], blockIdx.x));
}

}  // namespace

// ---------------------------------------------------------------- elementwise

// Zeroes grid.y runs of n floats, `step` floats apart (a fill kernel rather
// than hipMemsetAsync, so the sequence can be captured in a graph).
// Zero fill; with `flag`, only when *flag is set (uniform early exit).
__global__ void bm3d_zero(float* __restrict__ p, size_t n, size_t step, const int* __restrict__ flag) {
    if (flag && *flag == 0) return;
    float* q = p + (size_t)blockIdx.y * step;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t inc = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += inc) q[i] = 0.0f;
}

// Largest finite |sample| over the planes of `frames` frames: grid (rows,
// frames * nplanes), 256 threads over the columns; one atomic per
// work-group on the float's bit pattern (non-negative floats order like
// their bits).
__device__ __forceinline__ void absmax_body(const float* __restrict__ clip, long long frame_stride,
                                            long long plane_stride, int nplanes, int W, int H, int stride,
                                            float* out, Wg w, int rows_step) {
    const int f = w.y / nplanes, p = w.y % nplanes;
    const float* base = clip + (size_t)f * frame_stride + (size_t)p * plane_stride;
    float m = 0.0f;
    const int bd = blockDim.x;
    for (int y = w.x; y < H; y += rows_step) {
        const float* row = base + (size_t)y * stride;
        // four independent loads in flight per thread
        for (int x = threadIdx.x; x < W; x += 4 * bd) {
            float v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) v[k] = x + k * bd < W ? fabsf(row[x + k * bd]) : 0.0f;
#pragma unroll
            for (int k = 0; k < 4; ++k) m = v[k] <= kFltMax ? fmaxf(m, v[k]) : m;  // skips NaN and inf
        }
    }
    m = fmaxf(m, lane_xor1(m));
    m = fmaxf(m, lane_xor2(m));
    m = fmaxf(m, dpp_mov<0x141>(m));
    m = fmaxf(m, dpp_mov<0x140>(m));
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    // one atomic per work-group: thousands of waves maxing into one word
    // serialise at its L2 channel
    __shared__ float wmax[16];
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i) m = fmaxf(m, wmax[i]);
        atomicMax(reinterpret_cast<unsigned int*>(out), __float_as_uint(m));
    }
}

__global__ __launch_bounds__(256) void bm3d_absmax(const float* __restrict__ clip, long long frame_stride,
                                                   long long plane_stride, int nplanes, int W, int H, int stride,
                                                   float* out) {
    absmax_body(clip, frame_stride, plane_stride, nplanes, W, H, stride, out, launch_wg(), gridDim.x);
}

__global__ __launch_bounds__(256) void bm3d_absmax_multi(const AbsmaxArgs* __restrict__ jobs,
                                                         const JobGrid* __restrict__ grids, int njobs) {
    const int j = find_job(grids, njobs, blockIdx.x);
    const AbsmaxArgs& a = jobs[j];
    absmax_body(a.clip, a.frame_stride, a.plane_stride, a.nplanes, a.W, a.H, a.stride, a.out,
                job_wg(grids[j], blockIdx.x), grids[j].gx);
}

// Weighted-sum rows of the stacks: the integer sum of the tile regions
// bm3d_filter wrote to scratch that cover the pixel (at most 2 x 2 tiles),
// converted once to float, plus the stack's own value when members could
// fall outside the regions. Grid: (ceil(W/256), H, frames * planes * slots).
__global__ void bm3d_finalize(FinalizeArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int fps = blockIdx.z;
    if (x >= a.W) return;
    const int z = fps % a.slots;
    const int p = (fps / a.slots) % a.nplanes;
    const int f = fps / (a.slots * a.nplanes);
    float* num = a.res + (size_t)f * a.res_frame_stride + (size_t)p * a.res_plane_stride +
                 (size_t)(2 * z) * a.plane_size + (size_t)y * a.stride + x;
    float s = a.accumulate ? num[0] : 0.0f;
    long long isum = 0;
    if (!(a.chroma && a.sigma[p] < kFltEpsilon)) {
        const int region = a.slots * a.rh * a.rw;
        const int th = a.tby * a.bs, tw_ = a.tbx * a.bs;
        const int ylo_n = y - a.rm - 7 - (a.tby - 1) * a.bs;
        const int xlo_n = x - a.rm - 7 - (a.tbx - 1) * a.bs;
        const int ty_lo = max((ylo_n < 0 ? 0 : ylo_n / th) - 1, 0);
        const int ty_hi = min((y + a.rm) / th + 1, a.nty - 1);
        const int tx_lo = max((xlo_n < 0 ? 0 : xlo_n / tw_) - 1, 0);
        const int tx_hi = min((x + a.rm) / tw_ + 1, a.ntx - 1);
        for (int ty = ty_lo; ty <= ty_hi; ++ty) {
            const int ly = y - (min(ty * th, a.H - 8) - a.rm);
            if (ly < 0 || ly >= a.rh) continue;
            for (int tx = tx_lo; tx <= tx_hi; ++tx) {
                const int lx = x - (min(tx * tw_, a.W - 8) - a.rm);
                if (lx < 0 || lx >= a.rw) continue;
                const int* reg = a.scratch + ((((size_t)f * a.nty + ty) * a.ntx + tx) * a.nplanes + p) * region;
                isum += reg[(z * a.rh + ly) * a.rw + lx];
            }
        }
        if (*a.wrap_flag)  // uniform: some region word wrapped
            isum += (long long)a.wrap[(size_t)f * a.corner_frame_stride + (size_t)(p * a.slots + z) * a.plane_size +
                                      (size_t)y * a.stride + x]
                    << 32;
    }
    num[0] = s + (float)isum * ldexpf(1.0f, fx_exponent(a.fx_range, a.fx_exp));
}

// den[y][x] = sum over a, b in [0, 8) of corner[y - a][x - b]: each member
// block's weight added once at its corner covers its 8x8 footprint. Rows of
// a 64 x 8 output tile; the sum runs rows-then-columns in a fixed order.
// Grid: (ceil(W/64), ceil(H/8), frames * planes * slots).
__global__ __launch_bounds__(256) void bm3d_den_box(DenArgs a) {
    __shared__ float rowsum[15][64];
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // 0..3
    const int x = blockIdx.x * 64 + tx;
    const int y0 = blockIdx.y * 8;
    const int fps = blockIdx.z;
    const int slot = fps % a.slots;
    const int p = (fps / a.slots) % a.nplanes;
    const int f = fps / (a.slots * a.nplanes);
    const float* cmap = a.corner + (size_t)f * a.corner_frame_stride +
                        (size_t)(p * a.slots + slot) * a.plane_size;
    float* den = a.res + (size_t)f * a.res_frame_stride + (size_t)p * a.res_plane_stride +
                 (size_t)(slot * 2 + 1) * a.plane_size;
    // horizontal 8-sums for rows y0-7 .. y0+7
    for (int rr = ty; rr < 15; rr += 4) {
        const int yy = y0 - 7 + rr;
        float s = 0.0f;
        if (yy >= 0 && yy < a.H) {
            const float* row = cmap + (size_t)yy * a.stride;
#pragma unroll
            for (int b = 0; b < 8; ++b) {
                const int xx = x - b;
                s = s + ((xx >= 0 && xx < a.W) ? row[xx] : 0.0f);
            }
        }
        rowsum[rr][tx] = s;
    }
    __syncthreads();
    for (int jj = ty; jj < 8; jj += 4) {
        const int y = y0 + jj;
        if (y < a.H && x < a.W) {
            float s = 0.0f;
#pragma unroll
            for (int aa = 0; aa < 8; ++aa) s = s + rowsum[jj + 7 - aa][tx];
            den[(size_t)y * a.stride + x] = s;
        }
    }
}

// out = num / den per pixel (r = 0; source/source.cpp:170-188).
__global__ void bm3d_divide(DivideArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int fp = blockIdx.z;  // frame * planes + plane
    if (x >= a.W) return;
    const int f = fp / a.nplanes, p = fp % a.nplanes;
    const float* num = a.res + (size_t)f * a.res_frame_stride + (size_t)p * a.res_plane_stride +
                       (size_t)y * a.stride + x;
    const float nv = num[0];
    const float dv = num[a.plane_size];
    a.dst[(size_t)f * a.dst_frame_stride + (size_t)p * a.dst_plane_stride + (size_t)y * a.dst_stride + x] = nv / dv;
}

// VAggregate (source/source.cpp:816-846): output frame n sums slot
// s(i) = clamp(2r - i, n - N + 1 + r, n + r) of the stack of frame
// clamp(n - r + i, 0, N - 1), i ascending from 0, then divides.
__global__ void bm3d_vaggregate(VAggArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int fp = blockIdx.z;  // out-frame-in-batch * planes + plane
    if (x >= a.W) return;
    const int fo = fp / a.nplanes, p = fp % a.nplanes;
    const int n = a.first_frame + fo;
    const int r = a.radius;
    const int tw = 2 * r + 1;
    float num = 0.0f, den = 0.0f;
    for (int i = 0; i < tw; ++i) {
        const int fi = clampi(n - r + i, 0, a.num_frames - 1);
        const int s = clampi(2 * r - i, n - a.num_frames + 1 + r, n + r);
        const float* st = a.stacks + (size_t)(fi - a.stack_first) * a.stack_frame_stride +
                          (size_t)p * a.stack_plane_stride + (size_t)(2 * s) * a.plane_size +
                          (size_t)y * a.stride + x;
        num = num + st[0];
        den = den + st[a.plane_size];
    }
    a.dst[(size_t)fo * a.dst_frame_stride + (size_t)p * a.dst_plane_stride + (size_t)y * a.dst_stride + x] = num / den;
}

// Tiles whose accumulation regions cover pixel coordinate v along one axis:
// the contiguous range [lo, hi] (origins min(t * span, limit) - rm are
// non-decreasing in t), found with a float reciprocal and exact integer
// corrections instead of integer divisions.
__device__ __forceinline__ void covering_tiles(int v, int span, float inv_span, int limit, int rm, int rlen,
                                               int ntiles, int& lo, int& hi) {
    // hi: last t with min(t * span, limit) - rm <= v
    const int u = v + rm;
    int t = (int)((float)u * inv_span);
    t += (t + 1) * span <= u ? 1 : 0;
    t -= t * span > u ? 1 : 0;
    hi = min(t, ntiles - 1);
    if (limit - rm <= v) hi = ntiles - 1;  // the clamped last tile
    // lo: first t with min(t * span, limit) - rm + rlen > v
    const int w = v + rm - rlen + 1;  // need t * span >= w (or the clamped origin)
    if (w <= 0) {
        lo = 0;
    } else {
        int c = (int)((float)w * inv_span);
        c += c * span < w ? 1 : 0;
        c -= (c - 1) * span >= w ? 1 : 0;
        lo = c;
    }
    while (lo > 0 && min((lo - 1) * span, limit) - rm + rlen > v) --lo;
    while (lo <= hi && min(lo * span, limit) - rm + rlen <= v) ++lo;
}

// VAggregate of output frames first_frame .. first_frame + count - 1 from
// target-mode accumulation (FilterArgs::mode 1): the filter has already
// summed, per output frame n, every slot VAggregate reads for n
// (source/source.cpp:816-846) into integer regions and one corner weight
// map. num = the covering tiles' region values summed as integers, times
// the unit; den = the corner map's 8 x 8 box sum (rows, then columns, as
// bm3d_den_box); out = num / den. Work-group: 64 x 8 pixels, grid
// (ceil(W/64), ceil(H/8), count * nplanes).
__device__ __forceinline__ void vagg_targets_body(const VTargetArgs& a, Wg w) {
    // A work-group produces 64 x kVaggRows output pixels; thread (tx, ty)
    // the column x0 + tx of rows ty, ty + 4, ...
    constexpr int kR = kVaggRows, kRH = kVaggRows + 7;
    __shared__ float cm[kRH][72];      // corner-map window: rows y0-7.., columns x0-7..
    __shared__ float rowsum[kRH][64];
    __shared__ int2 ycover[kR];  // covering tile rows [lo, hi] of each output row
    const int tx = threadIdx.x & 63;
    const int ty = threadIdx.x >> 6;  // 0..3
    const int x0 = w.x * 64;
    const int x = x0 + tx;
    const int y0 = w.y * kR;
    const int fo = w.z / a.nplanes, p = w.z % a.nplanes;
    if (!((a.plane_mask >> p) & 1u)) return;  // uniform per work-group
    const int n = a.first_frame + fo;
    const int slot = n % a.ring;
    const int th = a.tby * a.bs, tw_ = a.tbx * a.bs;
    const float fx_unit = ldexpf(1.0f, fx_exponent(a.fx_range, a.fx_exp));
    const float* cmap = a.corner + (size_t)slot * a.corner_frame + (size_t)p * a.plane_size;
    for (int k = threadIdx.x; k < kRH * 71; k += 256) {
        const int rr = k / 71, cc = k - rr * 71;
        const int yy = y0 - 7 + rr, xx = x0 - 7 + cc;
        cm[rr][cc] = (yy >= 0 && yy < a.H && xx >= 0 && xx < a.W) ? cmap[(size_t)yy * a.stride + xx] : 0.0f;
    }
    // covering tiles of this thread's column: xlo and (usually) xlo + 1
    int xlo, xhi;
    const int xc = min(x, a.W - 1);
    covering_tiles(xc, tw_, 1.0f / (float)tw_, a.W - 8, a.rm, a.rw, a.ntx, xlo, xhi);
    const int lxa = xc - (min(xlo * tw_, a.W - 8) - a.rm);
    const int lxb = xc - (min((xlo + 1) * tw_, a.W - 8) - a.rm);
    const bool xa = (unsigned)lxa < (unsigned)a.rw;
    const bool xb = xlo + 1 <= xhi && (unsigned)lxb < (unsigned)a.rw;
    const int region = a.rh * a.rw;
    const size_t tstride = (size_t)a.nplanes * region;  // ints between consecutive tiles
    const int* scr = a.scratch + (size_t)slot * a.target_stride + (size_t)p * region;
    const int* wmap = *a.wrap_flag ? a.wrap + (size_t)slot * a.corner_frame + (size_t)p * a.plane_size : nullptr;
    const long long* pmap = ((a.partial[fo >> 5] >> (fo & 31)) & 1u)
                                ? a.part + (size_t)slot * a.corner_frame + (size_t)p * a.plane_size
                                : nullptr;
    if (threadIdx.x < kR) {  // once per row, not once per lane
        int ylo, yhi;
        covering_tiles(min(y0 + (int)threadIdx.x, a.H - 1), th, 1.0f / (float)th, a.H - 8, a.rm, a.rh, a.nty, ylo,
                       yhi);
        ycover[threadIdx.x] = make_int2(ylo, yhi);
    }
    __syncthreads();
    for (int rr = ty; rr < kRH; rr += 4) {
        float s = 0.0f;
#pragma unroll
        for (int b = 0; b < 8; ++b) s = s + cm[rr][tx + 7 - b];
        rowsum[rr][tx] = s;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < kR / 4; ++h) {
        const int jj = ty + 4 * h;
        const int y = y0 + jj;
        const int yc = min(y, a.H - 1);
        const int2 yr = ycover[jj];
        const int ylo = yr.x, yhi = yr.y;
        float dv = 0.0f;
#pragma unroll
        for (int aa = 0; aa < 8; ++aa) dv = dv + rowsum[jj + 7 - aa][tx];
        long long isum = 0;
        if (pmap) {
            isum = pmap[(size_t)yc * a.stride + xc];
        } else if (xhi - xlo <= 1 && yhi - ylo <= 1) {
            // at most 2 x 2 covering regions: four independent, predicated loads
            const int lya = yc - (min(ylo * th, a.H - 8) - a.rm);
            const int lyb = yc - (min((ylo + 1) * th, a.H - 8) - a.rm);
            const bool ya = (unsigned)lya < (unsigned)a.rh;
            const bool yb = ylo + 1 <= yhi && (unsigned)lyb < (unsigned)a.rh;
            const int* r0 = scr + ((size_t)ylo * a.ntx + xlo) * tstride;
            const int* r1 = r0 + (size_t)a.ntx * tstride;
            const int v00 = (ya && xa) ? r0[lya * a.rw + lxa] : 0;
            const int v01 = (ya && xb) ? r0[tstride + lya * a.rw + lxb] : 0;
            const int v10 = (yb && xa) ? r1[lyb * a.rw + lxa] : 0;
            const int v11 = (yb && xb) ? r1[tstride + lyb * a.rw + lxb] : 0;
            isum = ((long long)v00 + v01) + ((long long)v10 + v11);
        } else {
            // clamped edge tiles can cover a pixel three times
            int ly = yc - (min(ylo * th, a.H - 8) - a.rm);
            for (int tyy = ylo; tyy <= yhi; ++tyy) {
                int lx = lxa;
                for (int txx = xlo; txx <= xhi; ++txx) {
                    const int* reg = scr + ((size_t)tyy * a.ntx + txx) * tstride;
                    if ((unsigned)lx < (unsigned)a.rw && (unsigned)ly < (unsigned)a.rh) isum += reg[ly * a.rw + lx];
                    lx -= min((txx + 1) * tw_, a.W - 8) - min(txx * tw_, a.W - 8);
                }
                ly -= min((tyy + 1) * th, a.H - 8) - min(tyy * th, a.H - 8);
            }
        }
        if (wmap) isum += (long long)wmap[(size_t)yc * a.stride + xc] << 32;  // counted wraps
        const float num = 0.0f + (float)isum * fx_unit;
        if (y < a.H && x < a.W)
            a.dst[(size_t)fo * a.dst_frame_stride + (size_t)p * a.dst_plane_stride + (size_t)y * a.dst_stride + x] =
                num / dv;
    }
}

__global__ __launch_bounds__(256) void bm3d_vagg_targets(VTargetArgs a) { vagg_targets_body(a, launch_wg()); }

__global__ __launch_bounds__(256) void bm3d_vagg_targets_multi(const VTargetArgs* __restrict__ jobs,
                                                               const JobGrid* __restrict__ grids, int njobs) {
    const int j = find_job(grids, njobs, blockIdx.x);
    const VTargetArgs& a = jobs[j];
    vagg_targets_body(a, job_wg(grids[j], blockIdx.x));
}

// ---------------------------------------------------------------- launchers

template <int kCost>
void launch_match_cost(const MatchArgs& a, dim3 grid, dim3 block, size_t lds_bytes, hipStream_t s) {
    if (lds_bytes == 0) {
        if (a.ncand) hipLaunchKernelGGL((bm3d_match<kCost, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((bm3d_match<kCost, false, true>), grid, block, 0, s, a);
    } else {
        if (a.ncand) hipLaunchKernelGGL((bm3d_match<kCost, true, false>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_match<kCost, false, false>), grid, block, lds_bytes, s, a);
    }
}

// lds_bytes 0: the plane-reading variant.
hipError_t launch_match(const MatchArgs& a, int tby, size_t lds_bytes, int frames, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.nbx + a.tbw - 1) / a.tbw, (a.nby + tby - 1) / tby, frames);
    dim3 block(a.tbw * tby);
    switch (a.cost) {
        case 0: launch_match_cost<0>(a, grid, block, lds_bytes, s); break;
        case 1: launch_match_cost<1>(a, grid, block, lds_bytes, s); break;
        case 2: launch_match_cost<2>(a, grid, block, lds_bytes, s); break;
        case 3: launch_match_cost<3>(a, grid, block, lds_bytes, s); break;
        case 4: launch_match_cost<4>(a, grid, block, lds_bytes, s); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_match_multi(const MatchArgs* jobs, const JobGrid* grids, int njobs, int wgs, int threads,
                              size_t lds_bytes, int cost, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid(wgs), block(threads);
    switch (cost) {
        case 0: hipLaunchKernelGGL((bm3d_match_multi<0>), grid, block, lds_bytes, s, jobs, grids, njobs); break;
        case 1: hipLaunchKernelGGL((bm3d_match_multi<1>), grid, block, lds_bytes, s, jobs, grids, njobs); break;
        case 2: hipLaunchKernelGGL((bm3d_match_multi<2>), grid, block, lds_bytes, s, jobs, grids, njobs); break;
        case 3: hipLaunchKernelGGL((bm3d_match_multi<3>), grid, block, lds_bytes, s, jobs, grids, njobs); break;
        case 4: hipLaunchKernelGGL((bm3d_match_multi<4>), grid, block, lds_bytes, s, jobs, grids, njobs); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int kCost>
hipError_t set_match_cost_lds_limit(size_t bytes) {
    const void* fns[] = {reinterpret_cast<const void*>(&bm3d_match<kCost, false, false>),
                         reinterpret_cast<const void*>(&bm3d_match<kCost, true, false>),
                         reinterpret_cast<const void*>(&bm3d_match_multi<kCost>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t set_match_lds_limit(size_t bytes) {
    hipError_t e = set_match_cost_lds_limit<0>(bytes);
    if (e == hipSuccess) e = set_match_cost_lds_limit<1>(bytes);
    if (e == hipSuccess) e = set_match_cost_lds_limit<2>(bytes);
    if (e == hipSuccess) e = set_match_cost_lds_limit<3>(bytes);
    if (e == hipSuccess) e = set_match_cost_lds_limit<4>(bytes);
    return e;
}

template <bool kGeneric>
void launch_filter_kind(const FilterArgs& a, bool final_, bool chroma, dim3 grid, dim3 block, size_t lds_bytes,
                        hipStream_t s) {
    if (final_) {
        if (chroma) hipLaunchKernelGGL((bm3d_filter<true, true, kGeneric>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_filter<true, false, kGeneric>), grid, block, lds_bytes, s, a);
    } else {
        if (chroma) hipLaunchKernelGGL((bm3d_filter<false, true, kGeneric>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_filter<false, false, kGeneric>), grid, block, lds_bytes, s, a);
    }
}

hipError_t launch_filter(const FilterArgs& a, bool final_, bool chroma, int grid_z, int waves, size_t lds_bytes,
                         hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid(a.ntx, a.nty, grid_z);
    dim3 block(64 * waves);
    if (a.t2d < 0 || a.t2d >= kNumTransforms || a.t1d < 0 || a.t1d >= kNumTransforms) return hipErrorInvalidValue;
    if (a.t2d || a.t1d) launch_filter_kind<true>(a, final_, chroma, grid, block, lds_bytes, s);
    else launch_filter_kind<false>(a, final_, chroma, grid, block, lds_bytes, s);
    return hipGetLastError();
}

template <bool kGeneric>
void launch_filter_multi_kind(const FilterArgs* jobs, const JobGrid* grids, int njobs, bool final_, bool chroma,
                              dim3 grid, dim3 block, size_t lds_bytes, hipStream_t s) {
    if (final_) {
        if (chroma)
            hipLaunchKernelGGL((bm3d_filter_multi<true, true, kGeneric>), grid, block, lds_bytes, s, jobs, grids, njobs);
        else
            hipLaunchKernelGGL((bm3d_filter_multi<true, false, kGeneric>), grid, block, lds_bytes, s, jobs, grids,
                               njobs);
    } else {
        if (chroma)
            hipLaunchKernelGGL((bm3d_filter_multi<false, true, kGeneric>), grid, block, lds_bytes, s, jobs, grids,
                               njobs);
        else
            hipLaunchKernelGGL((bm3d_filter_multi<false, false, kGeneric>), grid, block, lds_bytes, s, jobs, grids,
                               njobs);
    }
}

hipError_t launch_filter_multi(const FilterArgs* jobs, const JobGrid* grids, int njobs, int wgs, bool final_,
                               bool chroma, bool generic, int waves, size_t lds_bytes, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid(wgs), block(64 * waves);
    if (generic) launch_filter_multi_kind<true>(jobs, grids, njobs, final_, chroma, grid, block, lds_bytes, s);
    else launch_filter_multi_kind<false>(jobs, grids, njobs, final_, chroma, grid, block, lds_bytes, s);
    return hipGetLastError();
}

template <bool kGeneric>
hipError_t set_filter_kind_lds_limit(size_t bytes) {
    const void* fns[] = {reinterpret_cast<const void*>(&bm3d_filter<true, true, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter<true, false, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter<false, true, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter<false, false, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter_multi<true, true, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter_multi<true, false, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter_multi<false, true, kGeneric>),
                         reinterpret_cast<const void*>(&bm3d_filter_multi<false, false, kGeneric>)};
    for (const void* f : fns) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t set_filter_lds_limit(size_t bytes) {
    hipError_t e = set_filter_kind_lds_limit<false>(bytes);
    if (e == hipSuccess) e = set_filter_kind_lds_limit<true>(bytes);
    return e;
}

hipError_t launch_finalize(const FinalizeArgs& a, int count, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.W + 255) / 256, a.H, count);
    hipLaunchKernelGGL(bm3d_finalize, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zero(float* p, size_t n, size_t step, int count, hipStream_t s, const int* flag) {
    BM3D_INJECT_LAUNCH();
    if (n == 0 || count == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(bm3d_zero, dim3((unsigned)blocks, count), dim3(256), 0, s, p, n, step, flag);
    return hipGetLastError();
}

hipError_t launch_absmax(const float* clip, long long frame_stride, long long plane_stride, int nplanes, int W,
                         int H, int stride, int frames, float* out, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid(H < 64 ? H : 64, frames * nplanes);
    hipLaunchKernelGGL(bm3d_absmax, grid, dim3(256), 0, s, clip, frame_stride, plane_stride, nplanes, W, H, stride,
                       out);
    return hipGetLastError();
}

hipError_t launch_den_box(const DenArgs& a, int count, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.W + 63) / 64, (a.H + 7) / 8, count);
    hipLaunchKernelGGL(bm3d_den_box, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_divide(const DivideArgs& a, int count, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.W + 255) / 256, a.H, count);
    hipLaunchKernelGGL(bm3d_divide, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vagg_targets_multi(const VTargetArgs* jobs, const JobGrid* grids, int njobs, int wgs,
                                     hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    hipLaunchKernelGGL(bm3d_vagg_targets_multi, dim3(wgs), dim3(256), 0, s, jobs, grids, njobs);
    return hipGetLastError();
}

hipError_t launch_absmax_multi(const AbsmaxArgs* jobs, const JobGrid* grids, int njobs, int wgs, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    hipLaunchKernelGGL(bm3d_absmax_multi, dim3(wgs), dim3(256), 0, s, jobs, grids, njobs);
    return hipGetLastError();
}

hipError_t launch_vagg_targets(const VTargetArgs& a, int count, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.W + 63) / 64, (a.H + kVaggRows - 1) / kVaggRows, count * a.nplanes);
    hipLaunchKernelGGL(bm3d_vagg_targets, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_vaggregate(const VAggArgs& a, int count, hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid((a.W + 255) / 256, a.H, count);
    hipLaunchKernelGGL(bm3d_vaggregate, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace bm3d
