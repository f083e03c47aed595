// Plugin-level filters over device-resident clips: BM3D, VAggregate and
// BM3Dv2 with the argument handling and frame semantics of the bm3dcuda
// VapourSynth plugin (source/source.cpp:457-989). A VapourSynth adapter
// maps its VSMap and frames onto these calls one to one; tests and the
// Python mirror (bm3d_amd.vs) call them directly.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cfloat>
#include <cmath>
#include <memory>
#include <string>

#include "bm3d_inject.h"
#include "bm3d_internal.h"
#include "bm3d_mi355x.h"

using bm3d::fail;
using bm3d::Kind;

namespace {

bool same_format(const bm3d_video_info& a, const bm3d_video_info& b) {
    return a.color_family == b.color_family && a.sample_type == b.sample_type &&
           a.bits_per_sample == b.bits_per_sample && a.subsampling_w == b.subsampling_w &&
           a.subsampling_h == b.subsampling_h && a.num_planes == b.num_planes;
}

bool is_yuv444ps(const bm3d_video_info& v) {
    return v.color_family == BM3D_YUV && v.sample_type == 1 && v.bits_per_sample == 32 &&
           v.subsampling_w == 0 && v.subsampling_h == 0 && v.num_planes == 3;
}

#define BM3D_HIP_F(expr)                                                                                   \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            return fail(BM3D_EHIP, std::string("BM3D: '") + #expr + "' failed: " + hipGetErrorString(_e)); \
    } while (0)

}  // namespace

namespace {

int plane_w(const bm3d_video_info& v, int p) { return p == 0 ? v.width : v.width >> v.subsampling_w; }
int plane_h(const bm3d_video_info& v, int p) { return p == 0 ? v.height : v.height >> v.subsampling_h; }

int ensure_scratch(bm3d_filter* f, size_t floats) {
    if (floats <= f->scratch_cap) return BM3D_OK;
    if (f->scratch) (void)hipFree(f->scratch);
    f->scratch = nullptr;
    f->scratch_cap = 0;
    BM3D_HIP_F(bm3d::malloc_device(&f->scratch, floats * sizeof(float)));
    f->scratch_cap = floats;
    return BM3D_OK;
}

// Plan for plane p (or the chroma group when p == 0 and chroma), created on
// first use with the clip's row stride.
int get_plan(bm3d_filter* f, int p, int stride, bm3d_plan** out) {
    if (f->plans[p] && f->plan_stride[p] == stride) {
        *out = f->plans[p];
        return BM3D_OK;
    }
    bm3d_plan_destroy(f->plans[p]);
    f->plans[p] = nullptr;
    bm3d_desc d{};
    d.width = plane_w(f->vi, p);
    d.height = plane_h(f->vi, p);
    d.stride = stride;
    for (int i = 0; i < 3; ++i) d.sigma[i] = f->chroma ? f->sigma[i] : f->sigma[p];
    d.block_step = f->block_step[p];
    d.bm_range = f->bm_range[p];
    d.radius = f->radius;
    d.ps_num = f->ps_num[p];
    d.ps_range = f->ps_range[p];
    d.chroma = f->chroma ? 1 : 0;
    d.final_ = f->final_ ? 1 : 0;
    d.extractor_exp = f->extractor_exp;
    d.precision = f->precision;
    // with chroma, plane 0's options serve all planes (rtc_source/source.cpp:987-1000)
    d.bm_error = f->bm_error[p];
    d.transform_2d = f->transform_2d[p];
    d.transform_1d = f->transform_1d[p];
    if (int rc = bm3d_plan_create(&d, &f->plans[p])) return rc;
    f->plan_stride[p] = stride;
    *out = f->plans[p];
    return BM3D_OK;
}

// Copies (or zeroes, when src is null) plane rows of frames first..first+count-1;
// frame n of dst is at dst + (n - dst_base) * dfs.
int copy_plane(float* dst, long long dfs, int ds, const float* src, long long sfs, int ss, int w, int h,
               int first, int count, hipStream_t st, int dst_base = 0) {
    for (int n = first; n < first + count; ++n) {
        float* d = dst + (size_t)(n - dst_base) * dfs;
        if (src) {
            BM3D_HIP_F(hipMemcpy2DAsync(d, (size_t)ds * 4, src + (size_t)n * sfs, (size_t)ss * 4, (size_t)w * 4, h,
                                        hipMemcpyDeviceToDevice, st));
        } else {
            // a fill kernel, not hipMemset2DAsync, so that a caller can
            // capture the call in a graph; at most 65535 rows per launch
            for (int r0 = 0; r0 < h; r0 += 65535)
                BM3D_HIP_F(bm3d::launch_zero(d + (size_t)r0 * ds, (size_t)w, (size_t)ds, std::min(65535, h - r0), st));
        }
    }
    return BM3D_OK;
}

// Reads the per-plane arguments with the reference's tail-fill rule
// (a missing entry repeats the previous one, the first defaults).
template <typename T, typename V>
void fill3(T (&dst)[3], int n, const V* vals, T dflt) {
    for (int i = 0; i < 3; ++i) dst[i] = i < n ? (T)vals[i] : (i == 0 ? dflt : dst[i - 1]);
}

// A string option of the NVRTC variant (rtc_source/source.cpp:854-912):
// per plane, lower-cased, one of `names` (code = index); a missing entry
// repeats the previous one and the first defaults to names[0].
int parse_option(const char* opt, int n, const char* const* vals, const char* const* names, int count,
                 int (&dst)[3]) {
    for (int i = 0; i < 3; ++i) {
        if (i >= n) {
            dst[i] = i == 0 ? 0 : dst[i - 1];
            continue;
        }
        std::string v = vals[i] ? vals[i] : "";
        for (char& ch : v) ch = (char)std::tolower((unsigned char)ch);
        int code = -1;
        for (int k = 0; k < count; ++k)
            if (v == names[k]) code = k;
        if (code < 0) return fail(BM3D_EINVAL, std::string("BM3D: invalid '") + opt + "': " + v);
        dst[i] = code;
    }
    return BM3D_OK;
}

int create_bm3d(const bm3d_args* a, const bm3d_video_info* vi, const bm3d_video_info* ref, bm3d_filter** out) {
    *out = nullptr;
    if (!a || !vi) return fail(BM3D_EINVAL, "BM3D: null argument");
    // source/source.cpp:476-482
    if (vi->width <= 0 || vi->height <= 0 || vi->sample_type != 1 || vi->bits_per_sample != 32)
        return fail(BM3D_EINVAL, "BM3D: only constant format 32bit float input supported");
    auto f = std::make_unique<bm3d_filter>();
    f->kind = Kind::kBM3D;
    f->vi = *vi;
    f->num_frames = vi->num_frames;
    if (ref) {  // 486-501
        if (!same_format(*ref, *vi)) return fail(BM3D_EINVAL, "BM3D: \"ref\" must be of the same format as \"clip\"");
        if (ref->width != vi->width || ref->height != vi->height)
            return fail(BM3D_EINVAL, "BM3D: \"ref\" must be of the same dimensions as \"clip\"");
        if (ref->num_frames != vi->num_frames)
            return fail(BM3D_EINVAL, "BM3D: \"ref\" must be of the same number of frames as \"clip\"");
        f->final_ = true;
    }
    // sigma, 504-523 (scaling happens in the plan)
    for (int i = 0; i < 3; ++i) {
        if (i < a->n_sigma) {
            f->sigma[i] = (float)a->sigma[i];
            if (f->sigma[i] < 0.0f) return fail(BM3D_EINVAL, "BM3D: \"sigma\" must be non-negative");
        } else {
            f->sigma[i] = i == 0 ? 3.0f : f->sigma[i - 1];
        }
        f->process[i] = !(f->sigma[i] < FLT_EPSILON);
    }
    fill3(f->block_step, a->n_block_step, a->block_step, 8);  // 525-535
    for (int i = 0; i < std::min(a->n_block_step, 3); ++i)
        if (a->block_step[i] <= 0 || a->block_step[i] > 8)
            return fail(BM3D_EINVAL, "BM3D: \"block_step\" must be in range [1, 8]");
    fill3(f->bm_range, a->n_bm_range, a->bm_range, 9);  // 537-547
    for (int i = 0; i < std::min(a->n_bm_range, 3); ++i)
        if (a->bm_range[i] <= 0) return fail(BM3D_EINVAL, "BM3D: \"bm_range\" must be positive");
    f->radius = a->n_radius > 0 ? (int)a->radius : 0;  // 549-559
    if (f->radius < 0) return fail(BM3D_EINVAL, "BM3D: \"radius\" must be non-negative");
    fill3(f->ps_num, a->n_ps_num, a->ps_num, 2);  // 561-571
    for (int i = 0; i < std::min(a->n_ps_num, 3); ++i)
        if (a->ps_num[i] <= 0 || a->ps_num[i] > 8) return fail(BM3D_EINVAL, "BM3D: \"ps_num\" must be in range [1, 8]");
    fill3(f->ps_range, a->n_ps_range, a->ps_range, 4);  // 573-583
    for (int i = 0; i < std::min(a->n_ps_range, 3); ++i)
        if (a->ps_range[i] <= 0) return fail(BM3D_EINVAL, "BM3D: \"ps_range\" must be positive");
    f->chroma = a->n_chroma > 0 && a->chroma != 0;  // 585-595
    if (f->chroma && !is_yuv444ps(*vi))
        return fail(BM3D_EINVAL, "BM3D: clip format must be YUV444 when \"chroma\" is true");
    {  // the NVRTC variant's options, checked before the device (rtc_source/source.cpp:854-912)
        static const char* const kErrors[] = {"ssd", "sad", "zssd", "zsad", "ssd/norm"};
        static const char* const kTransforms[] = {"dct", "haar", "wht", "bior1.5"};
        if (int rc = parse_option("bm_error_s", std::min(a->n_bm_error_s, 3), a->bm_error_s, kErrors, 5,
                                  f->bm_error))
            return rc;
        if (int rc = parse_option("transform_2d_s", std::min(a->n_transform_2d_s, 3), a->transform_2d_s,
                                  kTransforms, 4, f->transform_2d))
            return rc;
        if (int rc = parse_option("transform_1d_s", std::min(a->n_transform_1d_s, 3), a->transform_1d_s,
                                  kTransforms, 4, f->transform_1d))
            return rc;
    }
    const int dev = a->n_device_id > 0 ? (int)a->device_id : 0;  // 597-611
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    if (dev < 0 || dev >= ndev) return fail(BM3D_EINVAL, "BM3D: invalid device ID (" + std::to_string(dev) + ")");
    f->device = dev;
    f->extractor_exp = a->n_extractor_exp > 0 ? (int)a->extractor_exp : 0;  // 623-629
    f->zero_init = a->n_zero_init > 0 ? a->zero_init != 0 : true;          // 631-634
    f->fast = a->n_fast > 0 ? a->fast != 0 : true;                        // 636-639
    f->precision = a->n_precision > 0 ? (int)a->precision : BM3D_PRECISION_FP32;
    // the reference leaves tiny planes unchecked; the kernels need 8x8
    for (int p = 0; p < vi->num_planes && p < 3; ++p)
        if ((f->chroma ? p == 0 : f->process[p]) && (plane_w(*vi, p) < 8 || plane_h(*vi, p) < 8))
            return fail(BM3D_EINVAL, "BM3D: plane must be at least 8x8");
    f->out_vi = *vi;
    if (f->radius) f->out_vi.height *= 2 * (2 * f->radius + 1);  // BM3DInit, 190-204
    *out = f.release();
    return BM3D_OK;
}

// BM3DGetFrame (source/source.cpp:206-441) for a range of frames; output
// frame n goes to o->data[p] + (n - ob) * o->frame_stride[p].
int frames_bm3d(bm3d_filter* f, const bm3d_frames* c, const bm3d_frames* ref, const bm3d_frames* o, int first,
                int count, hipStream_t st, int ob = 0) {
    const int of = first - ob;  // first output frame, relative to o
    BM3D_HIP_F(hipSetDevice(f->device));
    const int N = f->num_frames;
    const int r = f->radius;
    const int tw = 2 * r + 1;
    const int np = std::min(f->vi.num_planes, 3);
    if (f->final_ && !ref) return fail(BM3D_EINVAL, "BM3D: \"ref\" frames missing");
    if (f->chroma) {
        // one plan for the three planes: they must share one layout
        const long long ps = c->data[1] - c->data[0];
        if (c->data[2] - c->data[1] != ps || c->frame_stride[1] != c->frame_stride[0] ||
            c->frame_stride[2] != c->frame_stride[0] || c->stride[1] != c->stride[0] || c->stride[2] != c->stride[0])
            return fail(BM3D_EINVAL, "BM3D: chroma planes must share one layout");
        bm3d_plan* plan;
        if (int rc = get_plan(f, 0, c->stride[0], &plan)) return rc;
        bm3d_clip src{c->data[0], c->frame_stride[0], ps, N};
        bm3d_clip rc_{ref ? ref->data[0] : nullptr, ref ? ref->frame_stride[0] : 0, ref ? ref->data[1] - ref->data[0] : 0, N};
        const int H = f->vi.height, W = f->vi.width, S = c->stride[0];
        if (r == 0) {
            const long long stack = bm3d_stack_size(plan);
            if (int rc = ensure_scratch(f, (size_t)count * stack)) return rc;
            if (int rc = bm3d_execute(plan, f->scratch, 0, 0, &src, f->final_ ? &rc_ : nullptr, first, count, st))
                return rc;
            for (int p = 0; p < 3; ++p) {
                if (f->process[p]) {
                    if (int rc2 = bm3d_aggregate(o->data[p] + (size_t)of * o->frame_stride[p], o->frame_stride[p], 0,
                                                 o->stride[p], f->scratch + (size_t)p * 2 * H * S, stack, 0, W, H, S, 1,
                                                 count, st))
                        return rc2;
                } else if (int rc2 = copy_plane(o->data[p], o->frame_stride[p], o->stride[p], c->data[p],
                                                c->frame_stride[p], c->stride[p], W, H, first, count, st, ob)) {
                    return rc2;
                }
            }
        } else {
            const long long ops = o->data[1] - o->data[0];
            if (o->data[2] - o->data[1] != ops || o->frame_stride[1] != o->frame_stride[0] ||
                o->frame_stride[2] != o->frame_stride[0] || o->stride[0] != S || o->stride[1] != S || o->stride[2] != S)
                return fail(BM3D_EINVAL, "BM3D: output planes must share the input layout");
            if (int rc = bm3d_execute(plan, o->data[0] + (size_t)of * o->frame_stride[0], o->frame_stride[0], ops,
                                      &src, f->final_ ? &rc_ : nullptr, first, count, st))
                return rc;
            for (int p = 0; p < 3; ++p)
                if (!f->process[p] && f->zero_init)
                    if (int rc = copy_plane(o->data[p], o->frame_stride[p], o->stride[p], nullptr, 0, 0, W,
                                            H * 2 * tw, first, count, st, ob))
                        return rc;
        }
        return BM3D_OK;
    }
    for (int p = 0; p < np; ++p) {
        const int W = plane_w(f->vi, p), H = plane_h(f->vi, p);
        if (!f->process[p]) {
            // unprocessed planes: copied (r = 0) or zeroed with zero_init (r > 0)
            if (r == 0) {
                if (int rc = copy_plane(o->data[p], o->frame_stride[p], o->stride[p], c->data[p], c->frame_stride[p],
                                        c->stride[p], W, H, first, count, st, ob))
                    return rc;
            } else if (f->zero_init) {
                if (int rc = copy_plane(o->data[p], o->frame_stride[p], o->stride[p], nullptr, 0, 0, W, H * 2 * tw,
                                        first, count, st, ob))
                    return rc;
            }
            continue;
        }
        bm3d_plan* plan;
        if (int rc = get_plan(f, p, c->stride[p], &plan)) return rc;
        bm3d_clip src{c->data[p], c->frame_stride[p], 0, N};
        bm3d_clip rc_{ref ? ref->data[p] : nullptr, ref ? ref->frame_stride[p] : 0, 0, N};
        if (r == 0) {
            const long long stack = bm3d_stack_size(plan);
            if (int rc = ensure_scratch(f, (size_t)count * stack)) return rc;
            if (int rc = bm3d_execute(plan, f->scratch, 0, 0, &src, f->final_ ? &rc_ : nullptr, first, count, st))
                return rc;
            if (int rc = bm3d_aggregate(o->data[p] + (size_t)of * o->frame_stride[p], o->frame_stride[p], 0,
                                        o->stride[p], f->scratch, 0, 0, W, H, c->stride[p], 1, count, st))
                return rc;
        } else {
            if (o->stride[p] != c->stride[p])
                return fail(BM3D_EINVAL, "BM3D: output planes must share the input row stride");
            if (int rc = bm3d_execute(plan, o->data[p] + (size_t)of * o->frame_stride[p], o->frame_stride[p], 0,
                                      &src, f->final_ ? &rc_ : nullptr, first, count, st))
                return rc;
        }
    }
    return BM3D_OK;
}

// BM3Dv2 with radius > 0 as one fused pass per plane group
// (bm3d_execute_vagg: the BM3D evaluations summed straight into the output
// frames VAggregate reads, bit for bit the stacks route's values with
// extractor_exp). Unprocessed planes are the source's (VAggregate copies
// them, source/source.cpp:838-846). BM3D_EUNSUPPORTED when a plan's tiles
// cannot hold every group member or the chroma planes are not equally
// spaced; the caller then takes the stacks route.
int frames_v2_fused(bm3d_filter* f, const bm3d_frames* c, const bm3d_frames* ref, const bm3d_frames* o, int first,
                    int count, hipStream_t st) {
    bm3d_filter* b = f->bm3d.get();
    BM3D_HIP_F(hipSetDevice(b->device));
    const int N = f->num_frames;
    const int np = std::min(f->vi.num_planes, 3);
    if (b->final_ && !ref) return fail(BM3D_EINVAL, "BM3D: \"ref\" frames missing");
    auto copy_unprocessed = [&]() -> int {
        for (int p = 0; p < np; ++p)
            if (!f->process[p])
                if (int rc = copy_plane(o->data[p], o->frame_stride[p], o->stride[p], c->data[p], c->frame_stride[p],
                                        c->stride[p], plane_w(f->vi, p), plane_h(f->vi, p), first, count, st))
                    return rc;
        return BM3D_OK;
    };
    if (b->chroma) {
        const long long ps = c->data[1] - c->data[0], ops = o->data[1] - o->data[0];
        const bool even = c->data[2] - c->data[1] == ps && o->data[2] - o->data[1] == ops &&
                          c->frame_stride[1] == c->frame_stride[0] && c->frame_stride[2] == c->frame_stride[0] &&
                          o->frame_stride[1] == o->frame_stride[0] && o->frame_stride[2] == o->frame_stride[0] &&
                          c->stride[1] == c->stride[0] && c->stride[2] == c->stride[0] && o->stride[1] == o->stride[0] &&
                          o->stride[2] == o->stride[0] &&
                          (!ref || (ref->data[1] - ref->data[0] == ps && ref->data[2] - ref->data[1] == ps &&
                                    ref->frame_stride[0] == c->frame_stride[0]));
        if (!even) return BM3D_EUNSUPPORTED;
        bm3d_plan* plan;
        if (int rc = get_plan(b, 0, c->stride[0], &plan)) return rc;
        bm3d_clip src{c->data[0], c->frame_stride[0], ps, N};
        bm3d_clip rc_{ref ? ref->data[0] : nullptr, ref ? ref->frame_stride[0] : 0, ps, N};
        if (int rc = bm3d_execute_vagg(plan, o->data[0] + (size_t)first * o->frame_stride[0], o->frame_stride[0], ops,
                                       o->stride[0], &src, b->final_ ? &rc_ : nullptr, first, count, st))
            return rc;
        return copy_unprocessed();
    }
    for (int p = 0; p < np; ++p) {
        if (!f->process[p]) continue;
        bm3d_plan* plan;
        if (int rc = get_plan(b, p, c->stride[p], &plan)) return rc;
        bm3d_clip src{c->data[p], c->frame_stride[p], 0, N};
        bm3d_clip rc_{ref ? ref->data[p] : nullptr, ref ? ref->frame_stride[p] : 0, 0, N};
        if (ref && ref->frame_stride[p] != c->frame_stride[p]) return BM3D_EUNSUPPORTED;
        if (int rc = bm3d_execute_vagg(plan, o->data[p] + (size_t)first * o->frame_stride[p], o->frame_stride[p], 0,
                                       o->stride[p], &src, b->final_ ? &rc_ : nullptr, first, count, st))
            return rc;
    }
    return copy_unprocessed();
}

// VAggregateGetFrame (source/source.cpp:746-862) for a range of frames.
// stacks: the BM3D output clip; stack_first: frame index of its data[p].
int frames_vagg(bm3d_filter* f, const bm3d_frames* stacks, int stack_first, const bm3d_frames* src,
                const bm3d_frames* o, int first, int count, hipStream_t st) {
    const int N = f->num_frames;
    for (int p = 0; p < std::min(f->vi.num_planes, 3); ++p) {
        const int W = plane_w(f->vi, p), H = plane_h(f->vi, p);
        if (!f->vprocess[p]) {
            if (int rc = copy_plane(o->data[p], o->frame_stride[p], o->stride[p], src->data[p], src->frame_stride[p],
                                    src->stride[p], W, H, first, count, st))
                return rc;
            continue;
        }
        if (int rc = bm3d_vaggregate(o->data[p] + (size_t)first * o->frame_stride[p], o->frame_stride[p], 0, o->stride[p],
                                     stacks->data[p], stacks->frame_stride[p], 0, stack_first, W, H, stacks->stride[p], 1,
                                     f->radius, N, first, count, st))
            return rc;
    }
    return BM3D_OK;
}

// Keeps the creation arguments of a filter (strings copied).
void keep_recipe(bm3d_filter* f, Kind kind, const bm3d_args* a, const bm3d_video_info* clip,
                 const bm3d_video_info* ref, const long long* planes, int n_planes) {
    bm3d::FilterRecipe& r = f->recipe;
    r.kind = kind;
    if (a) r.args = *a;
    if (clip) r.clip = *clip;
    r.has_ref = ref != nullptr;
    if (ref) r.ref = *ref;
    if (planes) r.planes.assign(planes, planes + n_planes);
    if (!a) return;
    const char* const* src[3] = {a->bm_error_s, a->transform_2d_s, a->transform_1d_s};
    const int n[3] = {a->n_bm_error_s, a->n_transform_2d_s, a->n_transform_1d_s};
    r.strings.clear();
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < std::min(n[k], 3); ++i) r.strings.emplace_back(src[k][i] ? src[k][i] : "");
    const char** dst[3] = {r.args.bm_error_s, r.args.transform_2d_s, r.args.transform_1d_s};
    size_t at = 0;
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < std::min(n[k], 3); ++i) dst[k][i] = src[k][i] ? r.strings[at++].c_str() : (at++, nullptr);
}

}  // namespace

namespace bm3d {

int clone_filter(const bm3d_filter* f, bm3d_filter** out) {
    const FilterRecipe& r = f->recipe;
    switch (r.kind) {
        case Kind::kBM3D:
            return bm3d_filter_create(&r.args, &r.clip, r.has_ref ? &r.ref : nullptr, out);
        case Kind::kVAggregate:
            return bm3d_vaggregate_create(r.planes.data(), (int)r.planes.size(), &r.clip, &r.ref, out);
        default:
            return bm3dv2_filter_create(&r.args, &r.clip, r.has_ref ? &r.ref : nullptr, out);
    }
}

}  // namespace bm3d

extern "C" {

int bm3d_filter_create(const bm3d_args* a, const bm3d_video_info* clip, const bm3d_video_info* ref,
                       bm3d_filter** out) {
    if (!out) return fail(BM3D_EINVAL, "BM3D: null argument");
    const int rc = create_bm3d(a, clip, ref, out);
    if (rc == BM3D_OK) keep_recipe(*out, Kind::kBM3D, a, clip, ref, nullptr, 0);
    return rc;
}

int bm3d_vaggregate_create(const long long* planes, int n_planes, const bm3d_video_info* clip,
                           const bm3d_video_info* src, bm3d_filter** out) {
    if (!out || !clip || !src) return fail(BM3D_EINVAL, "VAggregate: null argument");
    *out = nullptr;
    auto f = std::make_unique<bm3d_filter>();
    f->kind = Kind::kVAggregate;
    f->vi = *src;
    f->out_vi = *src;
    f->num_frames = src->num_frames;
    if (src->height <= 0 || clip->height <= 0) return fail(BM3D_EINVAL, "VAggregate: variable-size clips are not supported");
    f->radius = (clip->height / src->height - 2) / 4;  // source/source.cpp:892
    if (f->radius < 1 || clip->height != src->height * 2 * (2 * f->radius + 1))
        return fail(BM3D_EINVAL, "VAggregate: \"clip\" is not a BM3D output with radius > 0 for \"src\"");
    // indices beyond the clip's planes are accepted and ignored, as in the
    // reference (BM3Dv2 passes all processed indices of 0..2,
    // source/source.cpp:979-983; VAggregateGetFrame loops over numPlanes)
    for (int i = 0; i < n_planes; ++i) {
        if (!planes || planes[i] < 0 || planes[i] > 2) return fail(BM3D_EINVAL, "VAggregate: invalid plane index");
        f->vprocess[planes[i]] = true;
    }
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) f->device = dev;
    keep_recipe(f.get(), Kind::kVAggregate, nullptr, clip, src, planes, n_planes);
    *out = f.release();
    return BM3D_OK;
}

int bm3dv2_filter_create(const bm3d_args* a, const bm3d_video_info* clip, const bm3d_video_info* ref,
                         bm3d_filter** out) {
    if (!out || !a || !clip) return fail(BM3D_EINVAL, "BM3D: null argument");
    *out = nullptr;
    // process flags from sigma (source/source.cpp:916-930)
    bool process[3] = {true, true, true};
    for (int i = 0; i < std::min(3, a->n_sigma); ++i)
        if ((float)a->sigma[i] < FLT_EPSILON) process[i] = false;
    if (a->n_sigma > 0)
        for (int i = a->n_sigma; i < 3; ++i) process[i] = process[i - 1];
    auto f = std::make_unique<bm3d_filter>();
    f->kind = Kind::kBM3Dv2;
    f->vi = *clip;
    f->out_vi = *clip;
    f->num_frames = clip->num_frames;
    bool skip = true;
    for (int i = 0; i < std::min(clip->num_planes, 3); ++i) skip &= !process[i];
    keep_recipe(f.get(), Kind::kBM3Dv2, a, clip, ref, nullptr, 0);
    if (skip) {  // 932-942: the source clip itself
        f->passthrough = true;
        *out = f.release();
        return BM3D_OK;
    }
    bm3d_filter* b = nullptr;
    if (int rc = create_bm3d(a, clip, ref, &b)) return rc;
    f->bm3d.reset(b);
    f->radius = b->radius;
    f->fast = b->fast;
    f->device = b->device;
    for (int i = 0; i < 3; ++i) f->process[i] = process[i];
    if (b->radius > 0) {  // 956-977
        long long pl[3];
        int n = 0;
        for (int i = 0; i < 3; ++i)
            if (process[i]) pl[n++] = i;
        bm3d_filter* v = nullptr;
        if (int rc = bm3d_vaggregate_create(pl, n, &b->out_vi, clip, &v)) return rc;
        f->vagg.reset(v);
    }
    *out = f.release();
    return BM3D_OK;
}

int bm3d_filter_info(const bm3d_filter* f, bm3d_video_info* out, int* radius, int process[3]) {
    if (!f) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (out) *out = f->out_vi;
    if (radius) *radius = f->radius;
    if (process)
        for (int i = 0; i < 3; ++i) process[i] = f->kind == Kind::kVAggregate ? f->vprocess[i] : f->process[i];
    return BM3D_OK;
}

int bm3d_filter_get_frames(bm3d_filter* f, const bm3d_frames* clip, const bm3d_frames* ref, const bm3d_frames* out,
                           int first, int count, void* stream) {
    if (!f || !clip || !out) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (first < 0 || count < 0 || first + count > f->num_frames) return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    if (count == 0) return BM3D_OK;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    switch (f->kind) {
        case Kind::kBM3D:
            return frames_bm3d(f, clip, ref, out, first, count, st);
        case Kind::kVAggregate:
            if (!ref) return fail(BM3D_EINVAL, "VAggregate: \"src\" frames missing");
            return frames_vagg(f, clip, 0, ref, out, first, count, st);
        case Kind::kBM3Dv2: {
            const int np = std::min(f->vi.num_planes, 3);
            if (f->passthrough) {
                for (int p = 0; p < np; ++p)
                    if (int rc = copy_plane(out->data[p], out->frame_stride[p], out->stride[p], clip->data[p],
                                            clip->frame_stride[p], clip->stride[p], plane_w(f->vi, p), plane_h(f->vi, p),
                                            first, count, st))
                        return rc;
                return BM3D_OK;
            }
            if (!f->vagg) return frames_bm3d(f->bm3d.get(), clip, ref, out, first, count, st);
            {
                const int rc = frames_v2_fused(f, clip, ref, out, first, count, st);
                if (rc != BM3D_EUNSUPPORTED) return rc;
                bm3d::clear_error();
            }
            // BM3D stacks of the frames the temporal window touches, then VAggregate
            const int r = f->radius;
            const int lo = std::max(first - r, 0), hi = std::min(first + count + r, f->num_frames);
            bm3d_frames st_frames{};
            size_t total = 0, offs[3] = {0, 0, 0};
            for (int p = 0; p < np; ++p) {
                const int H = plane_h(f->vi, p);
                st_frames.stride[p] = clip->stride[p];
                st_frames.frame_stride[p] = (long long)2 * (2 * r + 1) * H * clip->stride[p];
                offs[p] = total;
                total += (size_t)(hi - lo) * st_frames.frame_stride[p];
            }
            if (int rc = ensure_scratch(f, total)) return rc;
            for (int p = 0; p < np; ++p) st_frames.data[p] = f->scratch + offs[p];
            if (f->bm3d->chroma) {
                // the chroma plan writes three equally spaced planes
                const long long one = st_frames.frame_stride[0];
                for (int p = 0; p < 3; ++p) {
                    st_frames.frame_stride[p] = 3 * one;
                    st_frames.data[p] = f->scratch + (size_t)p * one;
                }
            }
            // stacks of frame lo .. hi-1 start at data[p]
            if (int rc = frames_bm3d(f->bm3d.get(), clip, ref, &st_frames, lo, hi - lo, st, lo)) return rc;
            return frames_vagg(f->vagg.get(), &st_frames, lo, clip, out, first, count, st);
        }
    }
    return fail(BM3D_EINVAL, "BM3D: unknown filter");
}

void bm3d_filter_destroy(bm3d_filter* f) { delete f; }

}  // extern "C"
