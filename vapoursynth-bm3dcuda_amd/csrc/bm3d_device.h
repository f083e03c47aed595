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
// the largest finite |sample| M >= 2 of the call's input, so the unit keeps
// 22 bits below the samples' scale whatever that scale is, and every input
// within (-2, 2) keeps the unit 2^-22. A pixel's sum may still pass 2^31
// units (many members of weight near 1 on flat content); the region words
// then wrap, and the wraps are counted exactly (add_wrap below).
__device__ __forceinline__ int fx_exponent(const float* range, int fx_exp) {
    if (!range) return fx_exp;
    const float m = *range;
    return m >= 2.0f ? fx_exp + ilogbf(m) : fx_exp;
}

// ---- int32 region words with exact wrap accounting ----------------------
// A region word holds the low 32 bits (two's complement) of a pixel's sum of
// fixed-point units. Every add that carries the word past INT_MAX or below
// INT_MIN is counted, +1 or -1, in a 32-bit wrap map with the layout of the
// corner weight map, and readers add count * 2^32: the sum is exact for any
// magnitude. Adds only check the returned old value against 2^30 (two VALU
// operations per add) while every added value is below 2^30 in magnitude;
// the exact wrap is computed only when that check fires.

// Wrap of old + v: +1 (carried past INT_MAX), -1 (below INT_MIN), else 0.
__device__ __forceinline__ int add_wrap(int old, int v) {
    const int s = (int)((unsigned)old + (unsigned)v);
    return ((old ^ s) & (v ^ s)) < 0 ? (v < 0 ? -1 : 1) : 0;
}
// Bit 31 set iff |old| >= 2^30, when an add of |v| < 2^30 can wrap.
__device__ __forceinline__ unsigned wrap_risk(int old) { return (unsigned)old + 0x40000000u; }

// Where wraps go: the count map (corner-map layout) and a flag that readers
// test before reading it (set once any wrap was counted).
struct WrapSink {
    int* map;
    int* flag;
    __device__ __forceinline__ void count(size_t idx, int c) const {
        atomicAdd(map + idx, c);
        atomicOr(flag, 1);
    }
};

// Adds whole-number units x of any magnitude below 2^62 to region word q:
// its low 32 bits go to the word, the rest and the add's own wrap to the map.
__device__ __forceinline__ void add_units_wide(int* q, float x, const WrapSink& ws, size_t idx) {
    const long long xi = (long long)x;
    const int lo = (int)(unsigned)(unsigned long long)xi;
    const int hi = (int)((xi - (long long)lo) >> 32);
    const int old = atomicAdd(q, lo);
    const int c = hi + add_wrap(old, lo);
    if (c) ws.count(idx, c);
}

// Adds values iv[j][i] (|iv| < 2^30) to words q + j * qrow + i, whose
// pixels are map indices idx0 + j * irow + i, counting wraps exactly.
template <int kR, int kC>
__device__ __forceinline__ void add_units(int* q, int qrow, const int (&iv)[kR][kC], const WrapSink& ws,
                                          size_t idx0, size_t irow) {
    int old[kR][kC];
    unsigned risk = 0;
#pragma unroll
    for (int j = 0; j < kR; ++j)
#pragma unroll
        for (int i = 0; i < kC; ++i) {
            old[j][i] = atomicAdd(q + j * qrow + i, iv[j][i]);
            risk |= wrap_risk(old[j][i]);
        }
    if (__builtin_expect((int)risk < 0, 0)) {
#pragma unroll
        for (int j = 0; j < kR; ++j)
#pragma unroll
            for (int i = 0; i < kC; ++i) {
                const int c = add_wrap(old[j][i], iv[j][i]);
                if (c) ws.count(idx0 + j * irow + i, c);
            }
    }
}

// xor1 / xor2 are quad permutations; xor4 is a half-row mirror followed by a
// quad reversal (7 - k, then 3 - (k mod 4) within the quad, gives k ^ 4).
template <typename T> __device__ __forceinline__ T lane_xor1(T v) { return dpp_mov<0xB1>(v); }
template <typename T> __device__ __forceinline__ T lane_xor2(T v) { return dpp_mov<0x4E>(v); }
template <typename T> __device__ __forceinline__ T lane_xor4(T v) { return dpp_mov<0x1B>(dpp_mov<0x141>(v)); }
// Value of lane k of the quad (quad_perm [k, k, k, k]).
template <int kLane, typename T> __device__ __forceinline__ T quad_bcast(T v) { return dpp_mov<kLane * 0x55>(v); }

// Work-group coordinates of a kernel body: blockIdx in a single-job launch,
// the coordinates within its job's grid in a multi-job launch.
struct Wg {
    int x, y, z;
};
__device__ __forceinline__ Wg launch_wg() { return {(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z}; }

// Work-groups are dispatched to the 8 XCDs of an MI355X in turn (flat id L
// runs on XCD L % 8), and each XCD has its own L2. xcd_order renumbers them
// so that each XCD takes runs of kXcdRun consecutive ids, the runs dealt to
// the XCDs in turn: tiles that neighbour in the grid, which read
// overlapping rows and frames, then meet in one L2, while every XCD still
// gets an even share of each part of the grid (multi-job launches mix
// light and heavy jobs). The ids past the last whole round of runs are
// split into 8 contiguous shares. A permutation of [0, T), so every
// work-group still runs once. The block matchers use it (their search
// windows overlap from tile to tile); the filters and VAggregate keep the
// plain order, which measured faster for them.
constexpr int kXcdRun = 64;
__device__ __forceinline__ int xcd_order(int L, int T) {
    constexpr int kRound = 8 * kXcdRun;
    const int whole = T - T % kRound;
    if (L < whole) {
        const int o = L % kRound;
        return L - o + (o & 7) * kXcdRun + (o >> 3);
    }
    const int t = T - whole, l = L - whole;
    const int q = t >> 3, r = t & 7, k = l & 7;
    return whole + k * q + min(k, r) + (l >> 3);
}
// launch_wg in XCD order
__device__ __forceinline__ Wg launch_wg_xcd() {
    const int gx = gridDim.x, gxy = gx * gridDim.y;
    const int n = xcd_order(blockIdx.x + gx * blockIdx.y + gxy * blockIdx.z, gxy * gridDim.z);
    const int z = n / gxy, rem = n - z * gxy, y = rem / gx;
    return {rem - y * gx, y, z};
}
// flat work-group index of a multi-job launch in XCD order
__device__ __forceinline__ int multi_wg_xcd() { return xcd_order(blockIdx.x, gridDim.x); }

// Job of flat work-group index wg (binary search over the sorted begins;
// every value is uniform over the work-group, so this runs on the scalar
// unit).
__device__ __forceinline__ int find_job(const JobGrid* __restrict__ grids, int njobs, int wg) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (grids[mid].begin <= wg) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
__device__ __forceinline__ Wg job_wg(const JobGrid& g, int wg) {
    const int local = wg - g.begin;
    const int plane = g.gx * g.gy;
    const int z = local / plane;
    const int rem = local - z * plane;
    const int y = rem / g.gx;
    return {rem - y * g.gx, y, z};
}

// Number of i in [0, 2r] for which VAggregate's output frame
// n = clamp(m - r + z, 0, N - 1) reads slot z of the evaluation of frame m:
// it reads slot clamp(2r - i, n - N + 1 + r, n + r) of the evaluation of
// frame clamp(n - r + i, 0, N - 1) (source/source.cpp:834-836). 1 inside
// the clip; 0, 2 or more for evaluations within r of its ends.
__device__ __forceinline__ int vagg_multiplicity(int m, int z, int r, int N) {
    const int n = clampi(m - r + z, 0, N - 1);
    int c = 0;
    for (int i = 0; i <= 2 * r; ++i) {
        const int mi = clampi(n - r + i, 0, N - 1);
        const int si = clampi(2 * r - i, n - N + 1 + r, n + r);
        c += (mi == m && si == z) ? 1 : 0;
    }
    return c;
}

// Where the members of one evaluation go, per temporal slot z: the offset
// of the LDS accumulation region, the number of times each member is added,
// and the offset of the weight map (floats from FilterArgs::corner).
struct SlotTable {
    int region[kMaxSlots];
    int mult[kMaxSlots];
    long long corner[kMaxSlots];
};

// Walks the evaluations of a filter work-group's tile (grid (ntx, nty, z);
// FilterArgs::mode) plane by plane: zeroes the LDS regions `acc`, fills the
// slot table of each evaluation, runs body(p, m, gf) -- every group of the
// tile for evaluation frame m, group lists at index gf -- and writes the
// regions to scratch: per evaluation (mode 0), or per output frame once the
// evaluations that add to it are done (mode 1). LDS: tw regions of rh x rw.
template <typename Body>
__device__ __forceinline__ void filter_walk(const FilterArgs& a, Wg w, int* acc, SlotTable& st, int nplanes,
                                            bool chroma, Body&& body) {
    const int r = a.radius, tw = 2 * r + 1, N = a.num_frames;
    const int regsize = a.rh * a.rw;
    const int tx = w.x, ty = w.y;
    const int rx0 = min(tx * a.tbx * a.bs, a.W - 8) - a.rm;  // region origin in the plane
    const int ry0 = min(ty * a.tby * a.bs, a.H - 8) - a.rm;
    int e0, e1;
    if (a.mode == 0) {
        e0 = a.first_frame + w.z;
        e1 = e0 + 1;
    } else {
        e0 = a.eval_lo + w.z * a.chunk;
        e1 = min(e0 + a.chunk, a.eval_hi);
        if (e0 >= e1) return;  // uniform over the work-group
    }
    for (int p = 0; p < nplanes; ++p) {
        if (chroma && a.sigma[p] < kFltEpsilon) continue;  // source/kernel.cu:657-668
        for (int i = threadIdx.x; i < tw * regsize; i += blockDim.x) acc[i] = 0;
        int flushed = max(e0 - r, 0) - 1;  // mode 1: last output frame written
        for (int m = e0; m < e1; ++m) {
            __syncthreads();  // the previous evaluation's table and regions are done with
            if ((int)threadIdx.x < tw) {
                const int z = threadIdx.x;
                if (a.mode == 0) {
                    st.region[z] = z * regsize;
                    st.mult[z] = 1;
                    st.corner[z] = (long long)w.z * a.corner_frame_stride + (long long)(p * tw + z) * a.plane_size;
                } else {
                    const int t = clampi(m - r + z, 0, N - 1);
                    st.region[z] = (t % tw) * regsize;
                    st.mult[z] = vagg_multiplicity(m, z, r, N);
                    st.corner[z] = (long long)(t % a.ring) * a.corner_frame_stride + (long long)p * a.plane_size;
                }
            }
            __syncthreads();
            body(p, m, a.mode == 0 ? w.z : m - a.eval_lo);
            __syncthreads();
            if (a.mode == 1) {
                // frames whose every contributing evaluation (within r) is done
                const int hi = (m == N - 1 || m == e1 - 1) ? min(m + r, N - 1) : m - r;
                for (int t = flushed + 1; t <= hi; ++t) {
                    const bool whole = max(t - r, 0) >= e0 && min(t + r, N - 1) <= e1 - 1;
                    int* src = acc + (t % tw) * regsize;
                    int* dst = a.scratch + (size_t)(t % a.ring) * a.target_stride +
                               (((size_t)ty * a.ntx + tx) * nplanes + p) * regsize;
                    if (whole) {
                        for (int i = threadIdx.x; i < regsize; i += blockDim.x) {
                            dst[i] = src[i];
                            src[i] = 0;
                        }
                    } else {
                        // a frame whose contributors span several runs: its
                        // region adds to a 64-bit per-pixel sum (zeroed by the
                        // caller), which cannot wrap
                        unsigned long long* q = reinterpret_cast<unsigned long long*>(a.part) +
                                                (size_t)(t % a.ring) * a.corner_frame_stride +
                                                (size_t)p * a.plane_size;
                        for (int i = threadIdx.x; i < regsize; i += blockDim.x) {
                            const int v = src[i];
                            if (v != 0) {
                                // region pixel (lx, ly) is plane pixel (rx0 + lx, ry0 + ly)
                                const int ly = i / a.rw, lx = i - ly * a.rw;
                                atomicAdd(q + (size_t)(ry0 + ly) * a.stride + rx0 + lx,
                                          (unsigned long long)(long long)v);
                            }
                            src[i] = 0;
                        }
                    }
                }
                flushed = max(flushed, hi);
            }
        }
        if (a.mode == 0) {
            __syncthreads();
            int* out = a.scratch + ((((size_t)w.z * a.nty + ty) * a.ntx + tx) * nplanes + p) * tw * regsize;
            for (int i = threadIdx.x; i < tw * regsize; i += blockDim.x) out[i] = acc[i];
        }
        __syncthreads();
    }
}

}  // namespace dev
}  // namespace bm3d
