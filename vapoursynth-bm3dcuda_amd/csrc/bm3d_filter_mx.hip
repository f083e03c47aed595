// Mixed-precision collaborative filtering (BM3D_PRECISION_MIXED).
//
// The filter of source/kernel.cu:649-756 (hard threshold 208-281, Wiener
// 284-363) with the separable 3-D DCT evaluated as FP16 matrix products on
// the MFMA pipe (FP32 accumulation); shrinkage, weights and aggregation stay
// in FP32 and use the FP32 kernel's accumulation scheme (int32 fixed-point
// LDS regions per work-group tile, corner weight map), so the consumers of
// the regions (bm3d_finalize, bm3d_vagg_targets) are shared.
//
// A wave filters two groups (16 member blocks b) per step. With D the 8-point
// DCT-II (4 x orthonormal) and K = D (x) D acting on a block's 64 pixels:
//
//   Y^T = X^T K^T      16 blocks x 64 coefficients  (v_mfma_f32_16x16x32_f16)
//   Z^T = Dbd Y^T      along the members            (v_mfma_f32_16x16x16_f16)
//   shrink Z in FP32
//   W   = Z Dbd        members back                 (v_mfma_f32_16x16x16_f16)
//   X'  = K^T W        pixels back                  (v_mfma_f32_16x16x32_f16)
//
// where Dbd = diag(D, D) keeps the two groups apart. Each product's result
// registers are the next product's operand registers as they stand (the
// output layout of one MFMA is the B layout of the next, or the A layout of
// the transposed matrix), so no lane exchange or LDS staging is needed
// between the four stages. The operand registers of K, K^T and Dbd are
// precomputed per lane on the host (build_filter_mx_ops). The other
// transforms of the transform_2d_s / transform_1d_s options
// (rtc_source/kernel.hpp:200-361) replace D and D^T by their forward and
// inverse matrices in those operands; the kernel is the same.
//
// Lane l = 16 q + c. Loads: lane (q, c) reads rows q and q + 4 of block c
// (member c % 8 of group c / 8). Coefficients: lane (q, c) holds Z[m][u]
// for member frequency m = 4 q + i (group q / 2) and 2-D coefficient
// u = 16 t + c, i, t = 0..3. Estimates: lane (q, c) holds pixels
// (row 2 T + q / 2, column 4 (q % 2) + i) of block c, T, i = 0..3.

#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>
#include <stdint.h>

#include "bm3d_inject.h"
#include "bm3d_device.h"
#include "bm3d_xform.h"

namespace bm3d {

namespace {

using namespace dev;

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

__device__ __forceinline__ h8 cvt8(f4u lo, f4u hi) {
    return h8{(_Float16)lo[0], (_Float16)lo[1], (_Float16)lo[2], (_Float16)lo[3],
              (_Float16)hi[0], (_Float16)hi[1], (_Float16)hi[2], (_Float16)hi[3]};
}
__device__ __forceinline__ h4 cvt4(f4 v) {
    return h4{(_Float16)v[0], (_Float16)v[1], (_Float16)v[2], (_Float16)v[3]};
}
__device__ __forceinline__ f4 mfma32(h8 a, h8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4 mfma16(h4 a, h4 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
}

// Per-lane transform operands (kMxOpVectors uint4 per lane, lane-interleaved),
// read from a work-group copy in LDS where each product needs them: held in
// registers for the whole kernel they took 68 VGPRs, and the filter, bound
// by memory latency, needs the occupancy more than the LDS bandwidth.
struct MxOps {
    const uint4* s;  // LDS copy (this lane's entries: s[v * 64])
    __device__ __forceinline__ uint4 at(int i) const { return s[i]; }
    // The copy as seen from one wave step: an opaque zero offset keeps the
    // compiler from hoisting the reads out of the group loop into registers.
    __device__ __forceinline__ MxOps step() const {
        int z;
        asm volatile("v_mov_b32 %0, 0" : "=v"(z));
        return MxOps{s + z};
    }
    // K^T as the B operand of Y^T = X^T K^T: [tile t][k-step k]
    __device__ __forceinline__ h8 kf(int t, int k) const { return __builtin_bit_cast(h8, at((2 * t + k) * 64)); }
    // K^T as the A operand of X' = K^T W: [tile T][k-step k]
    __device__ __forceinline__ h8 ki(int t, int k) const {
        return __builtin_bit_cast(h8, at((8 + 2 * t + k) * 64));
    }
    // Dbd as the A operand (forward along members) and the B operand (inverse)
    __device__ __forceinline__ h4 dA() const {
        const uint4 d = at(16 * 64);
        return __builtin_bit_cast(h4, uint2{d.x, d.y});
    }
    __device__ __forceinline__ h4 dB() const {
        const uint4 d = at(16 * 64);
        return __builtin_bit_cast(h4, uint2{d.z, d.w});
    }
};

// Largest |v| over the 64 lanes (every lane gets it).
__device__ __forceinline__ float wave_max_abs(float m) {
    m = fmaxf(m, lane_xor1(m));
    m = fmaxf(m, lane_xor2(m));
    m = fmaxf(m, dpp_mov<0x141>(m));
    m = fmaxf(m, dpp_mov<0x140>(m));
    const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, m), 0));
    const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, m), 16));
    const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, m), 32));
    const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, m), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}

// Forward 3-D transform of the two groups whose member blocks start at mp
// (lane (q, c): block c): z[t][i] = Z[4 q + i][16 t + c] of the residual
// (X - mu) / s. mu is the group's offset (top-left pixel of its first
// member, the reference block; returned for the lane's block's group) and s
// a power of two (returned, shared by both groups) that keeps every
// residual within +-4, so no FP16 operand of the forward or inverse
// products can overflow: |Y| <= 128 x 4, |Z| <= 8 sqrt(2) |Y|, and W, the
// member inverse of shrunk Z, has 2-norm at most 16 x the 2-norm of Y's
// 8 members (|W| < 23200). A constant
// offset only moves the DC coefficient, which both shrinkage rules pass
// unchanged, so mu is added back exactly after the inverse.
// The lane's two rows of its block (rows q and q + 4, 16 bytes each half).
__device__ __forceinline__ void load_rows(f4u (&v)[4], const float* __restrict__ mp, int stride, int q) {
    const float* r0 = mp + (size_t)q * stride;
    const float* r1 = r0 + (size_t)4 * stride;
    v[0] = *reinterpret_cast<const f4u*>(r0);
    v[1] = *reinterpret_cast<const f4u*>(r0 + 4);
    v[2] = *reinterpret_cast<const f4u*>(r1);
    v[3] = *reinterpret_cast<const f4u*>(r1 + 4);
}

// forward_mx on rows already loaded by load_rows (v is overwritten).
__device__ __forceinline__ void forward_mx_rows(f4 (&z)[4], float& mu, float& s, const MxOps& o, f4u (&v)[4],
                                                int c) {
    const float mu0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v[0][0]), 0));
    const float mu1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v[0][0]), 8));
    mu = (c >> 3) ? mu1 : mu0;
    float m = 0.0f;
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v[h][e] = v[h][e] - mu;
            m = fmaxf(m, fabsf(v[h][e]));
        }
    const float g = wave_max_abs(m);
    s = 1.0f;
    if (g > 4.0f && g <= kFltMax) {  // rare: residuals beyond +-4 (not NaN / inf)
        const int k = ilogbf(g) - 1;  // g / 2^k < 4
        s = ldexpf(1.0f, k);
        const float inv = ldexpf(1.0f, -k);
#pragma unroll
        for (int h = 0; h < 4; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) v[h][e] = v[h][e] * inv;
    }
    const h8 a0 = cvt8(v[0], v[1]);
    const h8 a1 = cvt8(v[2], v[3]);
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    const h4 dA = o.dA();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f4 y = mfma32(a0, o.kf(t, 0), zero);
        y = mfma32(a1, o.kf(t, 1), y);
        z[t] = mfma16(dA, cvt4(y), zero);
    }
}

// Inverse 3-D transform of shrunk coefficients z (layout of forward_mx):
// x[T][i] = estimate at pixel (2 T + q / 2, 4 (q % 2) + i) of block c,
// scaled by 4096 (the transform pair's gain).
__device__ __forceinline__ void inverse_mx(f4 (&x)[4], const MxOps& o, const f4 (&z)[4]) {
    const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
    h4 w[4];
    const h4 dB = o.dB();
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] = cvt4(mfma16(cvt4(z[t]), dB, zero));
    const h8 b0 = h8{w[0][0], w[0][1], w[0][2], w[0][3], w[1][0], w[1][1], w[1][2], w[1][3]};
    const h8 b1 = h8{w[2][0], w[2][1], w[2][2], w[2][3], w[3][0], w[3][1], w[3][2], w[3][3]};
#pragma unroll
    for (int T = 0; T < 4; ++T) {
        f4 v = mfma32(o.ki(T, 0), b0, zero);
        x[T] = mfma32(o.ki(T, 1), b1, v);
    }
}

// Sum over the 16 lanes of each row of the wave (every lane gets its row's sum).
__device__ __forceinline__ float row_sum16(float s) {
    s = s + lane_xor1(s);
    s = s + lane_xor2(s);
    s = s + dpp_mov<0x141>(s);  // half-row mirror: quad 0 <-> quad 1 of each 8 lanes
    s = s + dpp_mov<0x140>(s);  // row mirror: lanes 0-7 <-> 8-15
    return s;
}

// Grid (ntx, nty, grid_z), blockDim.x = 64 x waves. Same tile, region and
// scratch layout, and the same accumulation modes, as bm3d_filter
// (bm3d_kernels.hip, filter_walk).
template <bool kFinal, bool kChroma, bool kWide>
__device__ __forceinline__ void filter_mx_body(const FilterArgs& a, Wg wg) {
    extern __shared__ float lds[];
    __shared__ SlotTable st;
    __shared__ uint4 ops_s[kMxOpVectors * 64];
    int* acc = reinterpret_cast<int*>(lds);
    const int nwaves = blockDim.x >> 6;
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int q = lane >> 4, c = lane & 15;
    const int tx = wg.x, ty = wg.y;
    const int rx0 = min(tx * a.tbx * a.bs, a.W - 8) - a.rm;
    const int ry0 = min(ty * a.tby * a.bs, a.H - 8) - a.rm;
    const int ngroups = a.tbx * a.tby;
    const float E = a.extractor;
    const float fx_scale = ldexpf(1.0f, -fx_exponent(a.fx_range, a.fx_exp));
    const WrapSink wsink{a.wrap, a.wrap_flag};
    // When one weighted estimate can reach 2^30 units (a small extractor_exp
    // on large samples; 5/4 covers the FP16 transforms' rounding), every
    // value takes the split 64-bit add: the kWide variant of the kernel,
    // launched beside the other, does the work then and the other returns.
    const bool wide = (1.25f * a.vbound * *a.range + E) * fx_scale >= 0x1p30f;
    if (wide != kWide) return;  // uniform over the work-group

    // (read after filter_walk's first barrier)
    for (int i = threadIdx.x; i < kMxOpVectors * 64; i += blockDim.x)
        ops_s[i] = reinterpret_cast<const uint4*>(a.mx_ops)[i];
    const MxOps ops{ops_s + lane};

    filter_walk(a, wg, acc, st, kChroma ? 3 : 1, kChroma, [&](int p, int n, int gf) {
        const float sigma = a.sigma[p];
        const float* src = a.src + (size_t)p * a.plane_stride;
        float* res = a.res ? a.res + (size_t)gf * a.res_frame_stride + (size_t)p * a.res_plane_stride : nullptr;

        // The group entry of a wave step is loaded one step ahead, so the
        // member rows it addresses are requested as soon as the step starts
        // (the compiler cannot move the load above the previous step's
        // global atomics itself). Prefetching the rows too measured slower.
        // (li / tbx through a float reciprocal: exact for li < 2^20)
        const float inv_tbx = 1.0f / (float)a.tbx;
        const uint32_t* gbase = a.groups + (size_t)gf * a.nblocks * 8 + (c & 7);
        auto entry = [&](int base, bool& act) -> uint32_t {
            const int li = base + wave * 2 + (c >> 3);
            const int qy = (int)(((float)li + 0.5f) * inv_tbx);
            const int bx = tx * a.tbx + (li - qy * a.tbx);
            const int by = ty * a.tby + qy;
            act = li < ngroups && bx < a.nbx && by < a.nby;
            const int blk = act ? by * a.nbx + bx : 0;
            return gbase[(size_t)blk * 8];
        };
        bool next_active = false;
        uint32_t next_pk = entry(0, next_active);
        for (int base = 0; base < ngroups; base += nwaves * 2) {
            const bool active = next_active;
            const uint32_t pk = next_pk;
            if (base + nwaves * 2 < ngroups) next_pk = entry(base + nwaves * 2, next_active);
            // positions come from block matching and are in range; the clamps
            // only keep a corrupted list from addressing outside the planes
            const int z = min(pos_z(pk), 2 * a.radius);
            const int px = min(pos_x(pk), a.W - 8);
            const int py = min(pos_y(pk), a.H - 8);
            const int fi = clampi(n - a.radius + z, 0, a.num_frames - 1);
            const float* mp = src + (size_t)fi * a.frame_stride + (size_t)py * a.stride + px;
            const MxOps o = ops.step();

            f4 zc[4];
            float w0, w1, mu, sc;
            if constexpr (kFinal) {
                // Wiener shrinkage (source/kernel.cu:284-326): coefficients
                // from the basic estimate's group, applied to the noisy group.
                // The DC coefficient is 1, so the offsets pass unchanged.
                // both groups' rows are requested before either transform
                f4u vb[4], vn[4];
                load_rows(vb, mp + (a.ref - a.src), a.stride, q);
                load_rows(vn, mp, a.stride, q);
                f4 cf[4];
                float mub, sb;
                forward_mx_rows(cf, mub, sb, o, vb, c);
                const float s2 = (sigma * sigma) / (sb * sb);  // sigma^2 in units of the scaled residual
                float sum = 0.0f;
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float sq = cf[t][i] * cf[t][i];
                        float coef = sq * __builtin_amdgcn_rcpf(sq + s2);
                        if (t == 0 && i == 0) coef = (c == 0 && (q & 1) == 0) ? 1.0f : coef;
                        sum = fma_(coef, coef, sum);
                        cf[t][i] = coef;
                    }
                sum = row_sum16(sum);
                // rows 0-1 hold group 0's coefficients, rows 2-3 group 1's
                const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sum), 0));
                const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sum), 16));
                const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sum), 32));
                const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, sum), 48));
                w0 = 1.0f / (r0 + r1);
                w1 = 1.0f / (r2 + r3);
                forward_mx_rows(zc, mu, sc, o, vn, c);
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i) zc[t][i] = zc[t][i] * cf[t][i];
            } else {
                // hard thresholding (source/kernel.cu:208-252); the DC
                // coefficient is compared with 0 instead of sigma
                f4u v[4];
                load_rows(v, mp, a.stride, q);
                forward_mx_rows(zc, mu, sc, o, v, c);
                const float thr_ac = sigma / sc;  // sigma in units of the scaled residual
                int kept0 = 0, kept1 = 0;
#pragma unroll
                for (int t = 0; t < 4; ++t)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float v = zc[t][i];
                        const float thr = (t == 0 && i == 0 && c == 0 && (q & 1) == 0) ? 0.0f : thr_ac;
                        const bool keep = fabsf(v) >= thr;
                        const unsigned long long m = __ballot(keep);
                        kept0 += __builtin_popcount((uint32_t)m);
                        kept1 += __builtin_popcount((uint32_t)(m >> 32));
                        zc[t][i] = keep ? v : 0.0f;
                    }
                w0 = 1.0f / (float)kept0;
                w1 = 1.0f / (float)kept1;
            }
            f4 x[4];
            inverse_mx(x, o, zc);

            const int mult = st.mult[z];
            if (active && mult > 0) {
                // aggregation with pre-rounding (source/kernel.cu:717-756);
                // the transform gain 4096 is divided out with the weight
                const float w = (c >> 3) ? w1 : w0;
                const float ws = w * (sc * (1.0f / 4096.0f));  // estimate = x s / 4096 + mu
                const float wm = w * mu;
                const int lx = px - rx0, ly = py - ry0;
                const int r = q >> 1, col = 4 * (q & 1);
                if (kWide && lx >= 0 && ly >= 0 && lx + 8 <= a.rw && ly + 8 <= a.rh) {
                    // split 64-bit adds (one estimate can reach 2^30 units)
                    int* qa = acc + st.region[z] + (ly + r) * a.rw + lx + col;
                    const size_t widx = st.corner[z] + (size_t)(py + r) * a.stride + px + col;
                    for (int k = 0; k < mult; ++k)
#pragma unroll
                        for (int T = 0; T < 4; ++T)
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                float val = fma_(ws, x[T][i], wm);
                                val = (val + E) - E;
                                add_units_wide(qa + 2 * T * a.rw + i, __builtin_rintf(val * fx_scale), wsink,
                                               widx + (size_t)(2 * T) * a.stride + i);
                            }
                } else if (lx >= 0 && ly >= 0 && lx + 8 <= a.rw && ly + 8 <= a.rh) {
                    int* qa = acc + st.region[z] + (ly + r) * a.rw + lx + col;
                    const size_t widx = st.corner[z] + (size_t)(py + r) * a.stride + px + col;
                    int iv[4][4];
                    if (E == 0.0f) {
                        // no pre-rounding: the unit's power-of-two scale folds
                        // into the weight exactly
                        const float wsf = ws * fx_scale, wmf = wm * fx_scale;
#pragma unroll
                        for (int T = 0; T < 4; ++T)
#pragma unroll
                            for (int i = 0; i < 4; ++i) iv[T][i] = (int)__builtin_rintf(fma_(wsf, x[T][i], wmf));
                    } else {
#pragma unroll
                        for (int T = 0; T < 4; ++T)
#pragma unroll
                            for (int i = 0; i < 4; ++i) {
                                float val = fma_(ws, x[T][i], wm);
                                val = (val + E) - E;
                                iv[T][i] = (int)__builtin_rintf(val * fx_scale);
                            }
                    }
                    // a slot read `mult` times by VAggregate is added as often
                    for (int k = 0; k < mult; ++k) add_units<4, 4>(qa, 2 * a.rw, iv, wsink, widx, 2 * (size_t)a.stride);
                } else if (res) {
                    float* qr = res + (size_t)(2 * z) * a.plane_size + (size_t)(py + r) * a.stride + px + col;
#pragma unroll
                    for (int T = 0; T < 4; ++T)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            float val = fma_(ws, x[T][i], wm);
                            val = (val + E) - E;
                            unsafeAtomicAdd(qr + (size_t)(2 * T) * a.stride + i, val);
                        }
                }
                if (q == 0) {
                    const float wr = (w + E) - E;
                    float* cq = a.corner + st.corner[z] + (size_t)py * a.stride + px;
                    for (int k = 0; k < mult; ++k) unsafeAtomicAdd(cq, wr);
                }
            }
        }
    });
}

// 4 waves per SIMD: 128 VGPRs (the transform operands are read from LDS)
template <bool kFinal, bool kChroma, bool kWide>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void bm3d_filter_mx(FilterArgs a) {
    filter_mx_body<kFinal, kChroma, kWide>(a, launch_wg());
}

template <bool kFinal, bool kChroma, bool kWide>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void bm3d_filter_mx_multi(
    const FilterArgs* __restrict__ jobs,
                                                            const JobGrid* __restrict__ grids, int njobs) {
    const int j = find_job(grids, njobs, blockIdx.x);
    const FilterArgs& a = jobs[j];
    filter_mx_body<kFinal, kChroma, kWide>(a, job_wg(grids[j], blockIdx.x));
}

}  // namespace

// 8-point DCT-II, 4 x orthonormal: D[0][x] = sqrt(2), D[m][x] = 2 cos((2x + 1) m pi / 16).
static double dct_entry(int m, int x) {
    const double pi = 3.14159265358979323846264338327950288;
    return m == 0 ? ::sqrt(2.0) : 2.0 * ::cos((2 * x + 1) * m * pi / 16.0);
}

// Entry (r, k) of the forward (inverse = 0) or inverse matrix of transform
// `kind`: the DCT pair (inverse = transpose), else bm3d_xform.h.
static double xform_entry(int kind, int inverse, int r, int k) {
    if (kind == 0) return inverse ? dct_entry(k, r) : dct_entry(r, k);
    return kXformMatrices.m[kind - 1][inverse][r * 8 + k];
}

// With F2, I2 the forward and inverse matrices of transform t2d (rows and
// columns) and F1, I1 those of t1d (members): K[u][p] = F2[u/8][p/8]
// F2[u%8][p%8] takes pixels p to coefficients u, Kinv[p][u] = I2[p/8][u/8]
// I2[p%8][u%8] takes them back, and Fbd, Ibd are F1, I1 per group. The four
// products are Y^T = X^T K^T, Z^T = Fbd Y^T, W = Z Ibd^T, X' = Kinv W; for
// the DCT (I = F^T) they are the products of the file comment.
void build_filter_mx_ops(uint32_t* out, int t2d, int t1d) {
    auto h = [](double v) -> uint16_t { return __builtin_bit_cast(uint16_t, (_Float16)v); };
    // K[u][p] for coefficient u and pixel p
    auto K = [&](int u, int p) { return xform_entry(t2d, 0, u >> 3, p >> 3) * xform_entry(t2d, 0, u & 7, p & 7); };
    // Kinv[p][u]: pixel p from coefficient u
    auto Kinv = [&](int p, int u) {
        return xform_entry(t2d, 1, p >> 3, u >> 3) * xform_entry(t2d, 1, p & 7, u & 7);
    };
    // member frequency m from member b (forward), member b from m (inverse),
    // within each of the two groups
    auto Fbd = [&](int m, int b) { return (m >> 3) == (b >> 3) ? xform_entry(t1d, 0, m & 7, b & 7) : 0.0; };
    auto Ibd = [&](int b, int m) { return (m >> 3) == (b >> 3) ? xform_entry(t1d, 1, b & 7, m & 7) : 0.0; };
    for (int lane = 0; lane < 64; ++lane) {
        const int q = lane >> 4, c = lane & 15;
        uint16_t e[kMxOpVectors][8];
        for (int v = 0; v < 8; ++v) {
            const int t = v >> 1, s = v & 1;
            // B[k = 8 q + kk][n = c] of Y^T = X^T K^T: pixel (row q + 4 s, column kk), coefficient 16 t + c
            for (int kk = 0; kk < 8; ++kk) e[v][kk] = h(K(16 * t + c, (q + 4 * s) * 8 + kk));
            // A[r = c][k = 8 q + j] of X' = Kinv W: pixel 16 T + c, coefficient 32 s + 16 (j / 4) + 4 q + j % 4
            for (int j = 0; j < 8; ++j) e[8 + v][j] = h(Kinv(16 * t + c, 32 * s + 16 * (j >> 2) + 4 * q + (j & 3)));
        }
        for (int i = 0; i < 4; ++i) {
            e[16][i] = h(Fbd(c, 4 * q + i));      // A[m = c][k = 4 q + i]
            e[16][4 + i] = h(Ibd(c, 4 * q + i));  // B[k = m = 4 q + i][n = b = c]
        }
        for (int v = 0; v < kMxOpVectors; ++v)
            for (int w = 0; w < 4; ++w)
                out[(v * 64 + lane) * 4 + w] = (uint32_t)e[v][2 * w] | ((uint32_t)e[v][2 * w + 1] << 16);
    }
}

// Per axis with forward rows F[k][.] and inverse columns I[.][k]: a
// coefficient k of samples bounded by M is at most ||F_k||_1 M. Hard
// thresholding: |w est(q)| = |sum over kept k of c_k I(q, k)| / (4096 N_kept)
// <= max_k |c_k I(q, k)| / 4096, a product over the three axes of
// max_k ||F_k||_1 max_q |I(q, k)|. Wiener: w = 1 / sum coef^2 with the DC
// coefficient 1, so by Cauchy-Schwarz |w est(q)| <= sqrt(sum_k c_k^2
// I(q, k)^2) / 4096, a product over the axes of max_q sqrt(sum_k
// (||F_k||_1 I(q, k))^2).
float contribution_bound(int t2d, int t1d, bool final_) {
    auto axis = [&](int kind) {
        double hard = 0.0, wien = 0.0;
        double l1[8];
        for (int k = 0; k < 8; ++k) {
            l1[k] = 0.0;
            for (int x = 0; x < 8; ++x) l1[k] += ::fabs(xform_entry(kind, 0, k, x));
        }
        for (int k = 0; k < 8; ++k)
            for (int q = 0; q < 8; ++q) hard = std::max(hard, l1[k] * ::fabs(xform_entry(kind, 1, q, k)));
        for (int q = 0; q < 8; ++q) {
            double s = 0.0;
            for (int k = 0; k < 8; ++k) {
                const double t = l1[k] * xform_entry(kind, 1, q, k);
                s += t * t;
            }
            wien = std::max(wien, ::sqrt(s));
        }
        return final_ ? wien : hard;
    };
    const double a2 = axis(t2d), a1 = axis(t1d);
    return (float)(a2 * a2 * a1 / 4096.0 * 1.0001);
}

template <bool kWide>
void launch_filter_mx_kind(const FilterArgs& a, bool final_, bool chroma, dim3 grid, dim3 block, size_t lds_bytes,
                           hipStream_t s) {
    if (final_) {
        if (chroma) hipLaunchKernelGGL((bm3d_filter_mx<true, true, kWide>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_filter_mx<true, false, kWide>), grid, block, lds_bytes, s, a);
    } else {
        if (chroma) hipLaunchKernelGGL((bm3d_filter_mx<false, true, kWide>), grid, block, lds_bytes, s, a);
        else hipLaunchKernelGGL((bm3d_filter_mx<false, false, kWide>), grid, block, lds_bytes, s, a);
    }
}

// Without extractor_exp one weighted estimate stays below 2^28 units (the
// unit keeps 22 bits below the samples' scale and vbound <= 23), so only
// plans with it can need the kWide variant; both are launched then and one
// of them returns at once.
hipError_t launch_filter_mx(const FilterArgs& a, bool final_, bool chroma, int grid_z, int waves, size_t lds_bytes,
                            hipStream_t s) {
    BM3D_INJECT_LAUNCH();
    dim3 grid(a.ntx, a.nty, grid_z), block(64 * waves);
    launch_filter_mx_kind<false>(a, final_, chroma, grid, block, lds_bytes, s);
    if (a.extractor != 0.0f) launch_filter_mx_kind<true>(a, final_, chroma, grid, block, lds_bytes, s);
    return hipGetLastError();
}

template <bool kWide>
void launch_filter_mx_multi_kind(const FilterArgs* jobs, const JobGrid* grids, int njobs, dim3 grid, dim3 block,
                                 bool final_, bool chroma, size_t lds_bytes, hipStream_t s) {
    if (final_) {
        if (chroma)
            hipLaunchKernelGGL((bm3d_filter_mx_multi<true, true, kWide>), grid, block, lds_bytes, s, jobs, grids, njobs);
        else
            hipLaunchKernelGGL((bm3d_filter_mx_multi<true, false, kWide>), grid, block, lds_bytes, s, jobs, grids,
                               njobs);
    } else {
        if (chroma)
            hipLaunchKernelGGL((bm3d_filter_mx_multi<false, true, kWide>), grid, block, lds_bytes, s, jobs, grids,
                               njobs);
        else
            hipLaunchKernelGGL((bm3d_filter_mx_multi<false, false, kWide>), grid, block, lds_bytes, s, jobs, grids,
                               njobs);
    }
}

hipError_t launch_filter_mx_multi(const FilterArgs* jobs, const JobGrid* grids, int njobs, int wgs, bool final_,
                                  bool chroma, int waves, size_t lds_bytes, hipStream_t s, bool any_extractor) {
    BM3D_INJECT_LAUNCH();
    dim3 grid(wgs), block(64 * waves);
    launch_filter_mx_multi_kind<false>(jobs, grids, njobs, grid, block, final_, chroma, lds_bytes, s);
    if (any_extractor) launch_filter_mx_multi_kind<true>(jobs, grids, njobs, grid, block, final_, chroma, lds_bytes, s);
    return hipGetLastError();
}

template <bool kWide>
hipError_t set_filter_mx_kind_lds_limit(size_t bytes) {
    const void* ks[] = {
        reinterpret_cast<const void*>(&bm3d_filter_mx<false, false, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx<false, true, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx<true, false, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx<true, true, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx_multi<false, false, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx_multi<false, true, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx_multi<true, false, kWide>),
        reinterpret_cast<const void*>(&bm3d_filter_mx_multi<true, true, kWide>),
    };
    for (const void* k : ks) {
        hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t set_filter_mx_lds_limit(size_t bytes) {
    hipError_t e = set_filter_mx_kind_lds_limit<false>(bytes);
    if (e == hipSuccess) e = set_filter_mx_kind_lds_limit<true>(bytes);
    return e;
}

}  // namespace bm3d
