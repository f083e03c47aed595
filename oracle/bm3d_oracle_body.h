/*
 * Body of the CPU oracle, instantiated twice by bm3d_oracle.c:
 *   T = float  (suffix _f32): the FP32 mirror of the reference op order
 *   T = double (suffix _f64): the same pipeline in double precision
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links or calls this.
 *
 * Every function cites the reference lines it restates. Paths are relative
 * to the reference repository root (WolframRhodium/VapourSynth-BM3DCUDA).
 * Where the reference CPU plugin (cpu_source/source.cpp) and the reference
 * GPU kernel (source/kernel.cu) disagree, the choice is stated in place.
 */

/* ---- 8-point scaled DCT-II / DCT-III (4 x orthonormal) -------------------
 * cpu_source/source.cpp:444-536 and source/kernel.cu:84-176 use a scaled
 * DCT pair whose forward is 4 x orthonormal DCT-II and whose inverse is
 * 4 x orthonormal DCT-III, so inverse(forward(x)) = 16 x per dimension.
 * This restatement uses its own even/odd factorization of the same matrix
 * (not the reference's FFTW-derived one); the result agrees with the
 * reference to float rounding. The op sequence below is the one the HIP
 * kernels use, so FP32 results are bit-identical to them.                  */
static inline void SFX(dct8_fwd)(T v[8]) {
    const T R2 = (T)DCT_R2, K1 = (T)DCT_K1, K3 = (T)DCT_K3;
    const T C1 = (T)DCT_C1, C3 = (T)DCT_C3, C5 = (T)DCT_C5, C7 = (T)DCT_C7;
    T a0 = v[0] + v[7], a1 = v[1] + v[6], a2 = v[2] + v[5], a3 = v[3] + v[4];
    T b0 = v[0] - v[7], b1 = v[1] - v[6], b2 = v[2] - v[5], b3 = v[3] - v[4];
    T c0 = a0 + a3, c1 = a1 + a2, d0 = a0 - a3, d1 = a1 - a2;
    v[0] = R2 * (c0 + c1);
    v[4] = R2 * (c0 - c1);
    v[2] = FMA(K1, d0, K3 * d1);
    v[6] = FMA(K3, d0, -(K1 * d1));
    v[1] = FMA(C1, b0, FMA(C3, b1, FMA(C5, b2, C7 * b3)));
    v[3] = FMA(C3, b0, FMA(-C7, b1, FMA(-C1, b2, -(C5 * b3))));
    v[5] = FMA(C5, b0, FMA(-C1, b1, FMA(C7, b2, C3 * b3)));
    v[7] = FMA(C7, b0, FMA(-C5, b1, FMA(C3, b2, -(C1 * b3))));
}

static inline void SFX(dct8_inv)(T v[8]) {
    const T R2 = (T)DCT_R2, K1 = (T)DCT_K1, K3 = (T)DCT_K3;
    const T C1 = (T)DCT_C1, C3 = (T)DCT_C3, C5 = (T)DCT_C5, C7 = (T)DCT_C7;
    T gc0 = R2 * (v[0] + v[4]), gc1 = R2 * (v[0] - v[4]);
    T gd0 = FMA(K1, v[2], K3 * v[6]);
    T gd1 = FMA(K3, v[2], -(K1 * v[6]));
    T ga0 = gc0 + gd0, ga3 = gc0 - gd0, ga1 = gc1 + gd1, ga2 = gc1 - gd1;
    T gb0 = FMA(C1, v[1], FMA(C3, v[3], FMA(C5, v[5], C7 * v[7])));
    T gb1 = FMA(C3, v[1], FMA(-C7, v[3], FMA(-C1, v[5], -(C5 * v[7]))));
    T gb2 = FMA(C5, v[1], FMA(-C1, v[3], FMA(C7, v[5], C3 * v[7])));
    T gb3 = FMA(C7, v[1], FMA(-C5, v[3], FMA(C3, v[5], -(C1 * v[7]))));
    v[0] = ga0 + gb0; v[7] = ga0 - gb0;
    v[1] = ga1 + gb1; v[6] = ga1 - gb1;
    v[2] = ga2 + gb2; v[5] = ga2 - gb2;
    v[3] = ga3 + gb3; v[4] = ga3 - gb3;
}

void SFX(oracle_dct8)(T *v, int n, int inverse) {
    for (int i = 0; i < n; ++i) {
        if (inverse) SFX(dct8_inv)(&v[8 * i]);
        else SFX(dct8_fwd)(&v[8 * i]);
    }
}

/* Transform `kind` (0 dct, 1 haar, 2 wht, 3 bior1.5) of one 8-vector. The
 * DCT takes its butterfly; the others multiply by their matrix, each output
 * as a product with v[0] followed by an FMA chain over v[1] .. v[7] (zero
 * entries included), the order the HIP kernels use. The reference's NVRTC
 * kernels evaluate the same linear maps by butterflies
 * (rtc_source/kernel.hpp:200-361) under -use_fast_math
 * (rtc_source/source.cpp:262-267); results agree to float rounding. */
static inline void SFX(xform8)(T v[8], int kind, int inverse) {
    if (kind == 0) {
        if (inverse) SFX(dct8_inv)(v);
        else SFX(dct8_fwd)(v);
        return;
    }
    const double *M = xform_matrix(kind, inverse);
    T o[8];
    for (int r = 0; r < 8; ++r) {
        T acc = (T)M[r * 8] * v[0];
        for (int k = 1; k < 8; ++k) acc = FMA((T)M[r * 8 + k], v[k], acc);
        o[r] = acc;
    }
    for (int r = 0; r < 8; ++r) v[r] = o[r];
}

void SFX(oracle_xform8)(T *v, int n, int kind, int inverse) {
    for (int i = 0; i < n; ++i) SFX(xform8)(&v[8 * i], kind, inverse);
}

/* 3-D transform of a group g[i][j][k] (member i, row j, column k).
 * Forward: along j for every (i,k), then along k for every (i,j), then along
 * i -- the order of cpu_source/source.cpp:596-603 (kernel.cu:264-270).
 * Inverse: along i, then k, then j, i.e. the forward order reversed. The
 * reference inverts in forward order (cpu_source/source.cpp:607-612); the
 * two differ only in float rounding, and the reversed order is the one the
 * HIP kernels use (fewer cross-lane transposes). */
/* t2 is the transform along j and k (transform_2d), t1 along i
 * (transform_1d), as rtc_source/kernel.hpp:585-605; the bm3dcuda kernels
 * are the case t2 = t1 = dct. */
static void SFX(transform3d_kinds)(T g[8][8][8], int inverse, int t2, int t1) {
    T v[8];
    if (!inverse) {
        for (int i = 0; i < 8; ++i)
            for (int k = 0; k < 8; ++k) {
                for (int j = 0; j < 8; ++j) v[j] = g[i][j][k];
                SFX(xform8)(v, t2, 0);
                for (int j = 0; j < 8; ++j) g[i][j][k] = v[j];
            }
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) SFX(xform8)(g[i][j], t2, 0);
        for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 8; ++k) {
                for (int i = 0; i < 8; ++i) v[i] = g[i][j][k];
                SFX(xform8)(v, t1, 0);
                for (int i = 0; i < 8; ++i) g[i][j][k] = v[i];
            }
    } else {
        for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 8; ++k) {
                for (int i = 0; i < 8; ++i) v[i] = g[i][j][k];
                SFX(xform8)(v, t1, 1);
                for (int i = 0; i < 8; ++i) g[i][j][k] = v[i];
            }
        for (int i = 0; i < 8; ++i)
            for (int j = 0; j < 8; ++j) SFX(xform8)(g[i][j], t2, 1);
        for (int i = 0; i < 8; ++i)
            for (int k = 0; k < 8; ++k) {
                for (int j = 0; j < 8; ++j) v[j] = g[i][j][k];
                SFX(xform8)(v, t2, 1);
                for (int j = 0; j < 8; ++j) g[i][j][k] = v[j];
            }
    }
}

static void SFX(transform3d)(T g[8][8][8], int inverse) {
    SFX(transform3d_kinds)(g, inverse, 0, 0);
}

void SFX(oracle_transform3d)(T *g, int inverse) {
    SFX(transform3d)((T (*)[8][8])g, inverse);
}

void SFX(oracle_transform3d_kinds)(T *g, int inverse, int t2, int t1) {
    SFX(transform3d_kinds)((T (*)[8][8])g, inverse, t2, t1);
}

/* ---- block distance -------------------------------------------------------
 * Sum of squared differences in the exact FP32 order of
 * cpu_source/source.cpp:123-136 (compute_distance) + 98-103 (reduce_add),
 * equivalently source/kernel.cu:455-471: per column c an even-row and an
 * odd-row FMA chain, s_c = e0 + e1, then a pairwise tree over columns.      */
static inline T SFX(ssd)(const T *ref /* 8x8 contiguous */, const T *p, int stride) {
    T s[8];
    for (int c = 0; c < 8; ++c) {
        T d0 = ref[0 * 8 + c] - p[0 * stride + c];
        T d1 = ref[1 * 8 + c] - p[1 * stride + c];
        T d2 = ref[2 * 8 + c] - p[2 * stride + c];
        T d3 = ref[3 * 8 + c] - p[3 * stride + c];
        T d4 = ref[4 * 8 + c] - p[4 * stride + c];
        T d5 = ref[5 * 8 + c] - p[5 * stride + c];
        T d6 = ref[6 * 8 + c] - p[6 * stride + c];
        T d7 = ref[7 * 8 + c] - p[7 * stride + c];
        T e0 = d0 * d0;
        e0 = FMA(d2, d2, e0);
        e0 = FMA(d4, d4, e0);
        e0 = FMA(d6, d6, e0);
        T e1 = d1 * d1;
        e1 = FMA(d3, d3, e1);
        e1 = FMA(d5, d5, e1);
        e1 = FMA(d7, d7, e1);
        s[c] = e0 + e1;
    }
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

T SFX(oracle_ssd)(const T *ref, const T *p, int stride) {
    return SFX(ssd)(ref, p, stride);
}

/* Pairwise tree over the 8 per-column partial results, the order of
 * reduce_subwarp (rtc_source/kernel.hpp:363-370) and reduce_add
 * (cpu_source/source.cpp:98-103). */
static inline T SFX(tree8)(const T s[8]) {
    return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

/* Block mean, rtc_source/kernel.hpp:413-421: per column a pairwise sum of
 * its 8 rows, the column tree, times 1/64. */
static inline T SFX(block_mean)(const T *b, int stride) {
    T s[8];
    for (int c = 0; c < 8; ++c)
        s[c] = ((b[0 * stride + c] + b[1 * stride + c]) + (b[2 * stride + c] + b[3 * stride + c])) +
               ((b[4 * stride + c] + b[5 * stride + c]) + (b[6 * stride + c] + b[7 * stride + c]));
    return SFX(tree8)(s) * ((T)1 / (T)64);
}

/* The block distances of the NVRTC variant (rtc_source/kernel.hpp:372-504):
 * 0 ssd, 1 sad, 2 zssd, 3 zsad, 4 ssd/norm. Per column an even-row and an
 * odd-row accumulation (errors[i % 2] over rows i, starting from 0, with the
 * products fused as nvcc contracts them), the column sum e0 + e1, then the
 * pairwise tree over columns. Means subtract as (c - n) - (mean_c - mean_n).
 * ssd/norm: 2 - (2 <c, n>) / fma(|c|, |n|, FLT_EPSILON) with correctly
 * rounded sqrt and division (the reference's -use_fast_math versions are
 * approximations of these). */
static inline T SFX(block_cost)(const T *ref /* 8x8 contiguous */, const T *p, int stride, int cost) {
    if (cost == 0) return SFX(ssd)(ref, p, stride);
    T s[8];
    if (cost == 4) {
        T sc[8], sn[8];
        for (int c = 0; c < 8; ++c) {
            T a[2] = {0, 0}, b[2] = {0, 0}, d[2] = {0, 0};
            for (int j = 0; j < 8; ++j) {
                T cv = ref[j * 8 + c], nv = p[j * stride + c];
                if (j < 2) {
                    a[j] = cv * cv;
                    b[j] = nv * nv;
                    d[j] = cv * nv;
                } else {
                    a[j & 1] = FMA(cv, cv, a[j & 1]);
                    b[j & 1] = FMA(nv, nv, b[j & 1]);
                    d[j & 1] = FMA(cv, nv, d[j & 1]);
                }
            }
            sc[c] = a[0] + a[1];
            sn[c] = b[0] + b[1];
            s[c] = d[0] + d[1];
        }
        T cn = SQRT(SFX(tree8)(sc)), nn = SQRT(SFX(tree8)(sn));
        T dot = SFX(tree8)(s);
        return (T)2 - ((T)2 * dot) / FMA(cn, nn, (T)FLT_EPSILON);
    }
    T dm = 0;
    if (cost == 2 || cost == 3) dm = SFX(block_mean)(ref, 8) - SFX(block_mean)(p, stride);
    for (int c = 0; c < 8; ++c) {
        T e[2] = {0, 0};
        for (int j = 0; j < 8; ++j) {
            T v = ref[j * 8 + c] - p[j * stride + c];
            if (cost >= 2) v = v - dm;
            if (cost == 2) e[j & 1] = j < 2 ? v * v : FMA(v, v, e[j & 1]);
            else e[j & 1] = j < 2 ? FABS(v) : e[j & 1] + FABS(v);
        }
        s[c] = e[0] + e[1];
    }
    return SFX(tree8)(s);
}

T SFX(oracle_block_cost)(const T *ref, const T *p, int stride, int cost) {
    return SFX(block_cost)(ref, p, stride, cost);
}

/* ---- sorted list of the 8 best matches ---------------------------------- */
typedef struct {
    T err[8];
    int x[8], y[8], z[8];
} SFX(list8);

/* Insert at the first slot whose error is strictly greater; the last entry
 * drops. cpu_source/source.cpp:179-204 (SIMD form) and kernel.cu:473-485. */
static inline void SFX(list_insert)(SFX(list8) *L, T e, int x, int y, int z) {
    int i;
    for (i = 0; i < 8; ++i)
        if (e < L->err[i]) break;
    if (i == 8) return;
    for (int j = 7; j > i; --j) {
        L->err[j] = L->err[j - 1];
        L->x[j] = L->x[j - 1];
        L->y[j] = L->y[j - 1];
        L->z[j] = L->z[j - 1];
    }
    L->err[i] = e;
    L->x[i] = x;
    L->y[i] = y;
    L->z[i] = z;
}

/* Exhaustive search over the clamped window around (cx, cy), row-major.
 * cpu_source/source.cpp:143-215 (block_matching). */
static void SFX(search_window)(SFX(list8) *L, const T *refblk, const T *img, int stride,
                               int W, int H, int range, int cx, int cy, int z,
                               long long *ncand, int cost) {
    int left = cx - range > 0 ? cx - range : 0;
    int right = cx + range < W - 8 ? cx + range : W - 8;
    int top = cy - range > 0 ? cy - range : 0;
    int bottom = cy + range < H - 8 ? cy + range : H - 8;
    for (int row = top; row <= bottom; ++row)
        for (int col = left; col <= right; ++col) {
            T e = SFX(block_cost)(refblk, &img[(size_t)row * stride + col], stride, cost);
            SFX(list_insert)(L, e, col, row, z);
        }
    if (ncand) *ncand += (long long)(bottom - top + 1) * (right - left + 1);
}

/* Spatial + predictive temporal search for the reference block at (x, y).
 * cpu_source/source.cpp:220-321 (block_matching_temporal): spatial search on
 * the centre slot, then per direction (-1 then +1) and t = 1..radius a fresh
 * per-frame list F filled from ps_num windows (duplicates kept), of which the
 * first ps_num entries are merged into L; the next frame searches around F.
 * Initial positions of unfilled entries are (x, y, centre) as on the CPU
 * (cpu_source/source.cpp:807-812); the GPU kernel starts them at (0, 0),
 * which only matters when a window holds fewer than 8 candidates. */
static void SFX(match_block)(SFX(list8) *L, const T *refblk, const T *const *slots, int stride,
                             int W, int H, int R, int radius, int ps_num, int ps_range,
                             int x, int y, long long *ncand, int cost) {
    const float big = FLT_MAX;
    for (int i = 0; i < 8; ++i) {
        L->err[i] = (T)big;
        L->x[i] = x;
        L->y[i] = y;
        L->z[i] = radius;
    }
    SFX(search_window)(L, refblk, slots[radius], stride, W, H, R, x, y, radius, ncand, cost);
    if (radius == 0) return;
    int cxs[8], cys[8];
    for (int i = 0; i < 8; ++i) { cxs[i] = L->x[i]; cys[i] = L->y[i]; }
    for (int dir = -1; dir <= 1; dir += 2) {
        int lx[8], ly[8];
        memcpy(lx, cxs, sizeof lx);
        memcpy(ly, cys, sizeof ly);
        for (int t = 1; t <= radius; ++t) {
            int z = radius + dir * t;
            SFX(list8) F;
            for (int i = 0; i < 8; ++i) {
                F.err[i] = (T)big;
                F.x[i] = 0; F.y[i] = 0; F.z[i] = z;
            }
            for (int i = 0; i < ps_num; ++i)
                SFX(search_window)(&F, refblk, slots[z], stride, W, H, ps_range, lx[i], ly[i], z, ncand, cost);
            for (int i = 0; i < ps_num; ++i)
                SFX(list_insert)(L, F.err[i], F.x[i], F.y[i], z);
            memcpy(lx, F.x, sizeof lx);
            memcpy(ly, F.y, sizeof ly);
        }
    }
}

/* If (x, y, centre) is absent, prepend it and drop the last entry.
 * cpu_source/source.cpp:325-391 (insert_if_not_in[_temporal]),
 * source/kernel.cu:616-647. */
static void SFX(insert_centre)(SFX(list8) *L, int x, int y, int z) {
    for (int i = 0; i < 8; ++i)
        if (L->x[i] == x && L->y[i] == y && L->z[i] == z) return;
    for (int j = 7; j > 0; --j) {
        L->err[j] = L->err[j - 1];
        L->x[j] = L->x[j - 1];
        L->y[j] = L->y[j - 1];
        L->z[j] = L->z[j - 1];
    }
    L->err[0] = 0;
    L->x[0] = x;
    L->y[0] = y;
    L->z[0] = z;
}

/* Hard thresholding, source/kernel.cu:208-252 (cpu_source/source.cpp:562-594):
 * keep |c| >= thr (DC always kept), kept values scaled by 1/4096, weight =
 * 1 / number kept. Exact division as on the GPU (the CPU plugin uses the
 * approximate _mm256_rcp_ps). */
static T SFX(hard_threshold)(T g[8][8][8], T thr) {
    int kept = 0;
    for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j)
            for (int k = 0; k < 8; ++k) {
                T v = g[i][j][k];
                T t = (i == 0 && j == 0 && k == 0) ? (T)0 : thr;
                int keep = FABS(v) >= t;
                kept += keep;
                g[i][j][k] = keep ? v * ((T)1 / (T)4096) : (T)0;
            }
    return (T)1 / (T)kept;
}

/* Empirical Wiener shrinkage, source/kernel.cu:284-326
 * (cpu_source/source.cpp:617-643): coef = B^2 / (B^2 + sigma^2), DC coef = 1,
 * N <- N * coef / 4096, weight = 1 / sum coef^2. B^2 + sigma^2 is a plain
 * (unfused) add as on the CPU. The sum of coef^2 runs per row j over the 64
 * (i, k) in index order with FMA, then a pairwise tree over rows; the
 * reference sums per column k then over columns -- same terms, different
 * float rounding. */
static T SFX(wiener)(T n[8][8][8], T b[8][8][8], T sigma) {
    T s2 = sigma * sigma;
    T acc[8];
    for (int j = 0; j < 8; ++j) {
        T a = 0;
        for (int i = 0; i < 8; ++i)
            for (int k = 0; k < 8; ++k) {
                T bv = b[i][j][k];
                T sq = bv * bv;
                T coef = sq / (sq + s2);
                if (i == 0 && j == 0 && k == 0) coef = 1;
                a = FMA(coef, coef, a);
                n[i][j][k] = (n[i][j][k] * coef) * ((T)1 / (T)4096);
            }
        acc[j] = a;
    }
    T norm = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    return (T)1 / norm;
}

/* Group-level collaborative filtering for unit tests: forward 3-D
 * transform, shrinkage, inverse; returns the aggregation weight. */
T SFX(oracle_collab_hard)(T *g, T sigma) {
    T (*G)[8][8] = (T (*)[8][8])g;
    SFX(transform3d)(G, 0);
    T w = SFX(hard_threshold)(G, sigma);
    SFX(transform3d)(G, 1);
    return w;
}

T SFX(oracle_collab_wiener)(T *n, T *b, T sigma) {
    T (*N)[8][8] = (T (*)[8][8])n;
    T (*B)[8][8] = (T (*)[8][8])b;
    SFX(transform3d)(N, 0);
    SFX(transform3d)(B, 0);
    T w = SFX(wiener)(N, B, sigma);
    SFX(transform3d)(N, 1);
    return w;
}

/* One BM3D(n) evaluation for one plane group, the per-frame work of
 * source/kernel.cu:366-757 (grid of reference blocks) with the
 * accumulation of cpu_source/source.cpp:674-735 in block-scan order.
 *
 * src/ref: pointers [plane * (2r+1) + slot]; ref only for final.
 * res: [planes][2r+1][2][H][stride]; slot z holds num rows then den rows.
 * sigma: already scaled as source/source.cpp:520-523.
 * groups (nullable): per reference block 8 x (x, y, z) after centre insert.
 * Returns the number of block-matching candidates evaluated. */
long long SFX(oracle_bm3d_ex)(T *res, const T *const *src, const T *const *ref,
                              int W, int H, int stride, const float *sigma,
                              int bs, int R, int radius, int ps_num, int ps_range,
                              int chroma, int final_, float extractor, int *groups,
                              int cost, int t2, int t1) {
    const int tw = 2 * radius + 1;
    const int nplanes = chroma ? 3 : 1;
    const size_t plane_sz = (size_t)H * stride;
    const T E = (T)extractor;
    long long ncand = 0;
    memset(res, 0, sizeof(T) * (size_t)nplanes * tw * 2 * plane_sz);

    const T *const *bm_slots = final_ ? ref : src;
    T (*grp)[8][8] = malloc(sizeof(T) * 512);
    T (*grb)[8][8] = malloc(sizeof(T) * 512);
    size_t blk = 0;
    for (int _y = 0; _y < H - 8 + bs; _y += bs) {
        int y = _y < H - 8 ? _y : H - 8;
        for (int _x = 0; _x < W - 8 + bs; _x += bs, ++blk) {
            int x = _x < W - 8 ? _x : W - 8;
            T refblk[64];
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 8; ++k)
                    refblk[j * 8 + k] = bm_slots[radius][(size_t)(y + j) * stride + x + k];
            SFX(list8) L;
            SFX(match_block)(&L, refblk, bm_slots, stride, W, H, R, radius, ps_num, ps_range,
                             x, y, &ncand, cost);
            SFX(insert_centre)(&L, x, y, radius);
            if (groups) {
                for (int i = 0; i < 8; ++i) {
                    groups[blk * 24 + i * 3 + 0] = L.x[i];
                    groups[blk * 24 + i * 3 + 1] = L.y[i];
                    groups[blk * 24 + i * 3 + 2] = L.z[i];
                }
            }
            for (int p = 0; p < nplanes; ++p) {
                /* chroma planes with (scaled) sigma < FLT_EPSILON are skipped:
                 * kernel.cu:657-668, cpu_source/source.cpp:846-849 */
                if (chroma && sigma[p] < FLT_EPSILON) continue;
                const T *const *ps = &src[p * tw];
                for (int i = 0; i < 8; ++i)
                    for (int j = 0; j < 8; ++j)
                        for (int k = 0; k < 8; ++k)
                            grp[i][j][k] = ps[L.z[i]][(size_t)(L.y[i] + j) * stride + L.x[i] + k];
                T w;
                SFX(transform3d_kinds)(grp, 0, t2, t1);
                if (final_) {
                    const T *const *pr = &ref[p * tw];
                    for (int i = 0; i < 8; ++i)
                        for (int j = 0; j < 8; ++j)
                            for (int k = 0; k < 8; ++k)
                                grb[i][j][k] = pr[L.z[i]][(size_t)(L.y[i] + j) * stride + L.x[i] + k];
                    SFX(transform3d_kinds)(grb, 0, t2, t1);
                    w = SFX(wiener)(grp, grb, (T)sigma[p]);
                } else {
                    w = SFX(hard_threshold)(grp, (T)sigma[p]);
                }
                SFX(transform3d_kinds)(grp, 1, t2, t1);
                /* aggregation with pre-rounding, kernel.cu:717-756 */
                T *rp = res + (size_t)p * tw * 2 * plane_sz;
                T wr = (w + E) - E;
                for (int i = 0; i < 8; ++i) {
                    T *num = rp + (size_t)(2 * L.z[i]) * plane_sz;
                    T *den = num + plane_sz;
                    for (int j = 0; j < 8; ++j)
                        for (int k = 0; k < 8; ++k) {
                            size_t o = (size_t)(L.y[i] + j) * stride + L.x[i] + k;
                            T v = w * grp[i][j][k];
                            v = (v + E) - E;
                            num[o] += v;
                            den[o] += wr;
                        }
                }
            }
        }
    }
    free(grp);
    free(grb);
    return ncand;
}

long long SFX(oracle_bm3d)(T *res, const T *const *src, const T *const *ref,
                           int W, int H, int stride, const float *sigma,
                           int bs, int R, int radius, int ps_num, int ps_range,
                           int chroma, int final_, float extractor, int *groups) {
    return SFX(oracle_bm3d_ex)(res, src, ref, W, H, stride, sigma, bs, R, radius, ps_num, ps_range,
                               chroma, final_, extractor, groups, 0, 0, 0);
}

/* Block matching only: the group lists of every reference block. */
long long SFX(oracle_match_ex)(const T *const *slots, int W, int H, int stride,
                               int bs, int R, int radius, int ps_num, int ps_range,
                               int *groups, T *errors, int cost) {
    long long ncand = 0;
    size_t blk = 0;
    for (int _y = 0; _y < H - 8 + bs; _y += bs) {
        int y = _y < H - 8 ? _y : H - 8;
        for (int _x = 0; _x < W - 8 + bs; _x += bs, ++blk) {
            int x = _x < W - 8 ? _x : W - 8;
            T refblk[64];
            for (int j = 0; j < 8; ++j)
                for (int k = 0; k < 8; ++k)
                    refblk[j * 8 + k] = slots[radius][(size_t)(y + j) * stride + x + k];
            SFX(list8) L;
            SFX(match_block)(&L, refblk, slots, stride, W, H, R, radius, ps_num, ps_range,
                             x, y, &ncand, cost);
            SFX(insert_centre)(&L, x, y, radius);
            for (int i = 0; i < 8; ++i) {
                groups[blk * 24 + i * 3 + 0] = L.x[i];
                groups[blk * 24 + i * 3 + 1] = L.y[i];
                groups[blk * 24 + i * 3 + 2] = L.z[i];
                if (errors) errors[blk * 8 + i] = L.err[i];
            }
        }
    }
    return ncand;
}

long long SFX(oracle_match)(const T *const *slots, int W, int H, int stride,
                            int bs, int R, int radius, int ps_num, int ps_range,
                            int *groups, T *errors) {
    return SFX(oracle_match_ex)(slots, W, H, stride, bs, R, radius, ps_num, ps_range, groups, errors, 0);
}
