// Launch layer: the extern "C" entry points of include/bm3d_mi355x.h.
//
// Replaces get_graphexec (source/kernel.cu:759-864): instead of a graph with
// baked HtoD / memset / kernel / DtoH nodes per frame, a plan holds the
// parameters and the device scratch (group lists, corner-weight maps), and
// bm3d_execute runs zero -> match -> filter -> den-box on device-resident
// clips for a whole batch of output frames per launch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bm3d_internal.h"
#include "bm3d_kernels.h"
#include "bm3d_mi355x.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

int bm3d::fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

namespace {

using bm3d::fail;

int hip_fail(hipError_t e, const char* what) {
    return fail(BM3D_EHIP, std::string("BM3D: '") + what + "' failed: " + hipGetErrorString(e));
}

#define BM3D_HIP(expr)                                         \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return hip_fail(_e, #expr);      \
    } while (0)

// Frames per launch: bounds the scratch (group lists and corner maps).
constexpr int kMaxBatchFrames = 16;

// Optional per-kernel timing with HIP events (bm3d_profile_*).
struct Profiler {
    std::mutex mu;
    bool on = false;
    struct Rec { int kernel; hipEvent_t a, b; };
    std::vector<Rec> recs;
};
Profiler& profiler() {
    static Profiler p;
    return p;
}

// Brackets one launch with events when profiling is on.
struct ProfScope {
    hipEvent_t a = nullptr, b = nullptr;
    int kernel;
    hipStream_t s;
    ProfScope(int k, hipStream_t st) : kernel(k), s(st) {
        Profiler& P = profiler();
        std::lock_guard<std::mutex> g(P.mu);
        if (!P.on) return;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { a = b = nullptr; return; }
        (void)hipEventRecord(a, s);
    }
    ~ProfScope() {
        if (!a) return;
        (void)hipEventRecord(b, s);
        Profiler& P = profiler();
        std::lock_guard<std::mutex> g(P.mu);
        P.recs.push_back({kernel, a, b});
    }
};

}  // namespace

struct bm3d_plan {
    bm3d_desc d{};
    float sigma[3]{};   // scaled thresholds
    float extractor = 0.0f;
    int nplanes = 1, tw = 1;
    int nbx = 0, nby = 0, nblocks = 0;
    int margin = 0, tby = 16, tile_w = 0, tile_h = 0, tile_wq = 0, tile_hp = 0;
    size_t lds_bytes = 0;
    long long plane_size = 0, stack_size = 0, corner_frame = 0;
    int device = 0;
    // filter work-group tiling and LDS accumulation region
    int ftbx = 16, ftby = 8, fntx = 0, fnty = 0, frm = 0, frw = 0, frh = 0;
    size_t filter_lds_bytes = 0;
    long long scratch_frame = 0;  // floats of region scratch per frame
    uint32_t* groups = nullptr;
    float* corner = nullptr;
    int* scratch = nullptr;
    // fixed-point unit of the accumulation regions: 2^fx_exp, scaled by the
    // input's range (measured into *range per call) unless extractor_exp is set
    int fx_exp = 0;
    float* range = nullptr;
    int cap_frames = 0;
    // mixed-precision matching kernel (precision MIXED)
    bool use_dx = false, dx_ok = false;
    int tile_hp2 = 0;
    size_t dx_lds_bytes = 0;
    // mixed-precision filter kernel (precision MIXED): per-lane transform operands
    bool use_mx_filter = false;
    void* mx_ops = nullptr;
};

namespace {

int count_positions(int n, int bs) {
    int c = 0;
    for (int v = 0; v < n - 8 + bs; v += bs) ++c;
    return c;
}

// Tile geometry for a work-group of 16 x tby reference blocks.
void tile_geometry(bm3d_plan* p, int tby) {
    const int bs = p->d.block_step;
    p->tby = tby;
    p->tile_w = 15 * bs + 8 + 2 * p->margin;
    p->tile_h = (tby - 1) * bs + 8 + 2 * p->margin;
    p->tile_wq = (p->tile_w + bs - 1) / bs;
    p->tile_hp = p->tile_h | 1;  // odd column pitch: lanes of a row hit distinct banks
    p->lds_bytes = (size_t)bs * p->tile_wq * p->tile_hp * sizeof(float);
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// Mixed-precision matching kernel (bm3d_match_dx.hip): the FP32 kernel's
// tile, stored as two planes of FP16 pairs (hp2 words per column, odd so
// that the 16 lanes of a row hit distinct banks). Scan indices are packed
// in 12 bits.
void dx_geometry(bm3d_plan* p, int max_lds) {
    const bm3d_desc& d = p->d;
    p->use_dx = false;
    if (d.precision != BM3D_PRECISION_MIXED) return;
    const int nWs = 2 * d.bm_range + 1, nWt = 2 * d.ps_range + 1;
    if (nWs * nWs > 4096 || (d.radius > 0 && d.ps_num * nWt * nWt > 4096)) return;
    p->tile_hp2 = ((p->tile_h + 1) / 2 + 1) | 1;
    const size_t bytes = (size_t)2 * d.block_step * p->tile_wq * p->tile_hp2 * 4;
    if (bytes > (size_t)max_lds) return;
    p->dx_lds_bytes = bytes;
    p->use_dx = p->dx_ok = true;
}

// Work-group size of the mixed filter kernel.
constexpr int kMxFilterWaves = 8;

// LDS bytes of the filter kernel for a tile of tbx x tby blocks and margin rm.
size_t filter_lds(const bm3d_plan* p, int tbx, int tby, int rm, int* rw, int* rh) {
    const int bs = p->d.block_step;
    *rw = (tbx - 1) * bs + 8 + 2 * rm;
    *rh = (tby - 1) * bs + 8 + 2 * rm;
    return (size_t)p->tw * *rw * *rh * sizeof(float);
}

// Filter tiling: the largest accumulation margin up to the full search reach
// (bm_range + radius * ps_range) whose region fits the LDS budget; members
// beyond a reduced margin fall back to global atomics.
void filter_geometry(bm3d_plan* p, size_t budget) {
    const int full = p->margin;
    const int tiles[][2] = {{8, 8}, {8, 4}};
    for (auto& t : tiles) {
        for (int rm = full; rm >= 0; --rm) {
            int rw, rh;
            size_t b = filter_lds(p, t[0], t[1], rm, &rw, &rh);
            if (b <= budget && (rm == full || (t[0] == 8 && t[1] == 4))) {
                p->ftbx = t[0]; p->ftby = t[1]; p->frm = rm; p->frw = rw; p->frh = rh;
                p->filter_lds_bytes = b;
                p->fntx = (p->nbx + t[0] - 1) / t[0];
                p->fnty = (p->nby + t[1] - 1) / t[1];
                p->scratch_frame = (long long)p->fntx * p->fnty * p->nplanes * p->tw * rw * rh;
                return;
            }
        }
    }
}

int ensure_scratch(bm3d_plan* p, int frames) {
    if (frames <= p->cap_frames) return BM3D_OK;
    if (p->groups) (void)hipFree(p->groups);
    if (p->corner) (void)hipFree(p->corner);
    if (p->scratch) (void)hipFree(p->scratch);
    p->groups = nullptr;
    p->corner = nullptr;
    p->scratch = nullptr;
    p->cap_frames = 0;
    BM3D_HIP(hipMalloc(&p->groups, (size_t)frames * p->nblocks * 8 * sizeof(uint32_t)));
    BM3D_HIP(hipMalloc(&p->corner, (size_t)frames * p->corner_frame * sizeof(float)));
    BM3D_HIP(hipMalloc(&p->scratch, (size_t)frames * p->scratch_frame * sizeof(int)));
    p->cap_frames = frames;
    return BM3D_OK;
}

// Block matching for frames first .. first + count - 1 into `groups` (or
// counting candidates into `ncand`): the mixed-precision kernel when the
// plan selected it, else the FP32 kernel.
hipError_t launch_matching(const bm3d_plan* p, uint32_t* groups, unsigned long long* ncand,
                           const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count,
                           hipStream_t stream) {
    const bm3d_desc& d = p->d;
    const float* clip = d.final_ ? ref->data : src->data;  // matching runs on ref in the final pass
    bm3d::MatchArgs m{};
    m.clip = clip;
    m.frame_stride = src->frame_stride;
    m.num_frames = src->num_frames;
    m.first_frame = first_frame;
    m.W = d.width; m.H = d.height; m.stride = d.stride;
    m.bs = d.block_step; m.bm_range = d.bm_range; m.radius = d.radius;
    m.ps_num = d.ps_num; m.ps_range = d.ps_range;
    m.nbx = p->nbx; m.nby = p->nby; m.nblocks = p->nblocks;
    m.margin = p->margin;
    m.tile_w = p->tile_w; m.tile_h = p->tile_h; m.tile_wq = p->tile_wq; m.tile_hp = p->tile_hp;
    m.groups = groups;
    m.ncand = ncand;
    if (p->use_dx) {
        m.tile_hp2 = p->tile_hp2;
        return bm3d::launch_match_dx(m, p->tby, p->dx_lds_bytes, count, stream);
    }
    return bm3d::launch_match(m, p->tby, p->lds_bytes, count, stream);
}

// Range pointer of the fixed-point unit (null: the unit is fixed).
const float* fx_range(const bm3d_plan* p) { return p->d.extractor_exp ? nullptr : p->range; }

// Measures the largest finite |sample| of the processed planes of frames
// lo .. hi of src into p->range (the fixed-point unit of every evaluation
// and aggregation of the call follows from it).
hipError_t measure_range(bm3d_plan* p, const bm3d_clip* src, int lo, int hi, hipStream_t stream) {
    if (!fx_range(p)) return hipSuccess;
    hipError_t e = bm3d::launch_zero(p->range, 1, 0, 1, stream);
    if (e != hipSuccess) return e;
    return bm3d::launch_absmax(src->data + (size_t)lo * src->frame_stride, src->frame_stride, src->plane_stride,
                               p->nplanes, p->d.width, p->d.height, p->d.stride, hi - lo + 1, p->range, stream);
}

// Block matching and filtering of frames n0 .. n0 + nf - 1: group lists,
// the tile accumulation regions (scratch) and the corner weight maps. With
// ring > 0, frame n's regions and corner maps go to index n % ring of the
// plan's buffers (which hold `ring` frames), else to index n - n0.
hipError_t evaluate(bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref, int n0, int nf, float* res,
                    long long res_frame_stride, long long res_plane_stride, int ring, hipStream_t stream) {
    const bm3d_desc& d = p->d;
    hipError_t e = hipSuccess;
    if (ring > 0) {
        // corner maps of the ring slots of n0 .. n0 + nf - 1 (may wrap)
        int f = 0;
        while (f < nf && e == hipSuccess) {
            const int slot = (n0 + f) % ring;
            const int run = std::min(nf - f, ring - slot);
            e = bm3d::launch_zero(p->corner + (size_t)slot * p->corner_frame, (size_t)run * p->corner_frame, 0, 1,
                                  stream);
            f += run;
        }
    } else {
        e = bm3d::launch_zero(p->corner, (size_t)nf * p->corner_frame, 0, 1, stream);
    }
    if (e != hipSuccess) return e;
    {
        ProfScope ps(0, stream);
        e = launch_matching(p, p->groups, nullptr, src, ref, n0, nf, stream);
    }
    if (e != hipSuccess) return e;
    bm3d::FilterArgs fa{};
    fa.src = src->data;
    fa.ref = d.final_ ? ref->data : nullptr;
    fa.frame_stride = src->frame_stride;
    fa.plane_stride = src->plane_stride;
    fa.num_frames = src->num_frames;
    fa.first_frame = n0;
    fa.W = d.width; fa.H = d.height; fa.stride = d.stride;
    fa.plane_size = p->plane_size;
    fa.radius = d.radius;
    fa.nblocks = p->nblocks;
    for (int i = 0; i < 3; ++i) fa.sigma[i] = p->sigma[i];
    fa.extractor = p->extractor;
    fa.groups = p->groups;
    fa.res = res;
    fa.res_frame_stride = res_frame_stride;
    fa.res_plane_stride = res_plane_stride;
    fa.corner = p->corner;
    fa.corner_frame_stride = p->corner_frame;
    fa.bs = d.block_step; fa.nbx = p->nbx; fa.nby = p->nby;
    fa.tbx = p->ftbx; fa.tby = p->ftby; fa.ntx = p->fntx; fa.nty = p->fnty;
    fa.rm = p->frm; fa.rw = p->frw; fa.rh = p->frh;
    fa.scratch = p->scratch;
    fa.fx_range = fx_range(p);
    fa.fx_exp = p->fx_exp;
    fa.ring = ring;
    {
        ProfScope ps(1, stream);
        if (p->use_mx_filter) {
            fa.mx_ops = p->mx_ops;
            e = bm3d::launch_filter_mx(fa, d.final_ != 0, d.chroma != 0, nf, kMxFilterWaves, p->filter_lds_bytes,
                                       stream);
        } else {
            e = bm3d::launch_filter(fa, d.final_ != 0, d.chroma != 0, nf, p->filter_lds_bytes, stream);
        }
    }
    return e;
}


}  // namespace

extern "C" {

const char* bm3d_last_error(void) { return g_last_error.c_str(); }

const char* bm3d_version(void) { return "bm3d_mi355x 0.1 (gfx950)"; }

int bm3d_plan_create(const bm3d_desc* desc, bm3d_plan** out) {
    g_last_error.clear();
    if (!desc || !out) return fail(BM3D_EINVAL, "BM3D: null argument");
    *out = nullptr;
    const bm3d_desc& d = *desc;
    // argument checks, same order and wording as source/source.cpp:503-611
    for (int i = 0; i < 3; ++i)
        if (d.sigma[i] < 0.0f) return fail(BM3D_EINVAL, "BM3D: \"sigma\" must be non-negative");
    if (d.block_step <= 0 || d.block_step > 8)
        return fail(BM3D_EINVAL, "BM3D: \"block_step\" must be in range [1, 8]");
    if (d.bm_range <= 0) return fail(BM3D_EINVAL, "BM3D: \"bm_range\" must be positive");
    if (d.radius < 0) return fail(BM3D_EINVAL, "BM3D: \"radius\" must be non-negative");
    if (d.ps_num <= 0 || d.ps_num > 8) return fail(BM3D_EINVAL, "BM3D: \"ps_num\" must be in range [1, 8]");
    if (d.ps_range <= 0) return fail(BM3D_EINVAL, "BM3D: \"ps_range\" must be positive");
    // limits the reference leaves unchecked
    if (d.width < 8 || d.height < 8) return fail(BM3D_EINVAL, "BM3D: plane must be at least 8x8");
    if (d.stride < d.width) return fail(BM3D_EINVAL, "BM3D: stride must be >= width");
    if (d.width > bm3d::kMaxWidth || d.height > bm3d::kMaxHeight)
        return fail(BM3D_EUNSUPPORTED, "BM3D: plane larger than 16384x8192 is not supported");
    if (d.radius > bm3d::kMaxRadius) return fail(BM3D_EUNSUPPORTED, "BM3D: \"radius\" above 15 is not supported");
    if (d.extractor_exp < 0 || d.extractor_exp > 100)
        return fail(BM3D_EINVAL, "BM3D: \"extractor_exp\" out of range");
    if (d.precision != BM3D_PRECISION_FP32 && d.precision != BM3D_PRECISION_MIXED)
        return fail(BM3D_EINVAL, "BM3D: unknown precision");

    auto* p = new bm3d_plan;
    p->d = d;
    p->d.chroma = d.chroma ? 1 : 0;
    p->d.final_ = d.final_ ? 1 : 0;
    // sigma scaling, source/source.cpp:520-523 (float, left to right)
    const float f = (3.0f / 4.0f) / 255.0f * 64.0f * (p->d.final_ ? 1.0f : 2.7f);
    for (int i = 0; i < 3; ++i) p->sigma[i] = d.sigma[i] * f;
    // extractor = 2^e, source/source.cpp:623-629
    p->extractor = d.extractor_exp ? std::ldexp(1.0f, d.extractor_exp) : 0.0f;
    // Fixed-point unit of the LDS accumulation: with extractor 2^e, every
    // pre-rounded value in (-E/2, 2E) is a multiple of 2^(e-24), so sums are
    // exact; otherwise 2^-22 (1.2e-7 rounding, sums up to +-512).
    p->fx_exp = d.extractor_exp ? d.extractor_exp - 24 : -22;
    p->nplanes = p->d.chroma ? 3 : 1;
    p->tw = 2 * d.radius + 1;
    p->nbx = count_positions(d.width, d.block_step);
    p->nby = count_positions(d.height, d.block_step);
    p->nblocks = p->nbx * p->nby;
    p->margin = d.bm_range + d.radius * d.ps_range;
    p->plane_size = (long long)d.height * d.stride;
    p->stack_size = (long long)p->nplanes * p->tw * 2 * p->plane_size;
    p->corner_frame = (long long)p->nplanes * p->tw * p->plane_size;

    hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) {
        delete p;
        return hip_fail(e, "hipGetDevice");
    }
    int max_lds = 0;
    e = hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device);
    if (e != hipSuccess) {
        delete p;
        return hip_fail(e, "hipDeviceGetAttribute");
    }
    tile_geometry(p, 16);
    if (p->lds_bytes > 65536) tile_geometry(p, 4);
    if (p->lds_bytes > (size_t)max_lds) {
        delete p;
        return fail(BM3D_EUNSUPPORTED, "BM3D: search window (bm_range + radius * ps_range) too large for LDS");
    }
    if (p->lds_bytes > 65536) {
        e = bm3d::set_match_lds_limit(p->lds_bytes);
        if (e != hipSuccess) {
            delete p;
            return hip_fail(e, "hipFuncSetAttribute");
        }
    }
    filter_geometry(p, std::min<size_t>((size_t)max_lds, 52 * 1024));
    if (p->filter_lds_bytes == 0 || p->filter_lds_bytes > (size_t)max_lds) {
        delete p;
        return fail(BM3D_EUNSUPPORTED, "BM3D: no filter tiling fits the device's LDS");
    }
    if (p->filter_lds_bytes > 65536) {
        e = bm3d::set_filter_lds_limit(p->filter_lds_bytes);
        if (e != hipSuccess) {
            delete p;
            return hip_fail(e, "hipFuncSetAttribute");
        }
    }
    dx_geometry(p, max_lds);
    if (p->use_dx && p->dx_lds_bytes > 65536) {
        e = bm3d::set_match_dx_lds_limit(p->dx_lds_bytes);
        if (e != hipSuccess) {
            delete p;
            return hip_fail(e, "hipFuncSetAttribute");
        }
    }
    e = hipMalloc(&p->range, sizeof(float));
    if (e != hipSuccess) {
        bm3d_plan_destroy(p);
        return hip_fail(e, "hipMalloc");
    }
    if (d.precision == BM3D_PRECISION_MIXED) {
        std::vector<uint32_t> ops((size_t)bm3d::kMxOpVectors * 64 * 4);
        bm3d::build_filter_mx_ops(ops.data());
        const size_t bytes = ops.size() * sizeof(uint32_t);
        e = hipMalloc(&p->mx_ops, bytes);
        if (e == hipSuccess) e = hipMemcpy(p->mx_ops, ops.data(), bytes, hipMemcpyHostToDevice);
        if (e == hipSuccess && p->filter_lds_bytes > 65536) e = bm3d::set_filter_mx_lds_limit(p->filter_lds_bytes);
        if (e != hipSuccess) {
            bm3d_plan_destroy(p);
            return hip_fail(e, "mixed filter set-up");
        }
        p->use_mx_filter = true;
    }
    *out = p;
    return BM3D_OK;
}

void bm3d_plan_destroy(bm3d_plan* p) {
    if (!p) return;
    if (p->groups) (void)hipFree(p->groups);
    if (p->corner) (void)hipFree(p->corner);
    if (p->scratch) (void)hipFree(p->scratch);
    if (p->mx_ops) (void)hipFree(p->mx_ops);
    if (p->range) (void)hipFree(p->range);
    delete p;
}

long long bm3d_stack_size(const bm3d_plan* p) { return p ? p->stack_size : 0; }

int bm3d_execute(bm3d_plan* p, float* d_res, long long res_frame_stride, long long res_plane_stride,
                 const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count, void* stream_) {
    g_last_error.clear();
    if (!p || !d_res || !src || !src->data) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (p->d.final_ && (!ref || !ref->data)) return fail(BM3D_EINVAL, "BM3D: final pass needs \"ref\"");
    if (p->d.final_ && (ref->frame_stride != src->frame_stride || ref->plane_stride != src->plane_stride ||
                        ref->num_frames != src->num_frames))
        return fail(BM3D_EINVAL, "BM3D: \"ref\" must have the layout and number of frames of \"clip\"");
    if (src->num_frames <= 0 || count < 0 || first_frame < 0 || first_frame + count > src->num_frames)
        return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    const long long plane_stack = (long long)p->tw * 2 * p->plane_size;
    if (res_plane_stride == 0) res_plane_stride = plane_stack;
    if (res_frame_stride == 0) res_frame_stride = p->nplanes * res_plane_stride;
    if (res_plane_stride < plane_stack || res_frame_stride < (p->nplanes - 1) * res_plane_stride + plane_stack)
        return fail(BM3D_EINVAL, "BM3D: result strides overlap");
    if (count == 0) return BM3D_OK;
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    int dev = 0;
    BM3D_HIP(hipGetDevice(&dev));
    if (dev != p->device) return fail(BM3D_EINVAL, "BM3D: plan used on a different device than it was created on");

    const int batch = std::min(count, kMaxBatchFrames);
    if (int rc = ensure_scratch(p, batch)) return rc;

    const bm3d_desc& d = p->d;
    BM3D_HIP(measure_range(p, src, std::max(first_frame - d.radius, 0),
                           std::min(first_frame + count - 1 + d.radius, src->num_frames - 1), stream));
    for (int f0 = 0; f0 < count; f0 += batch) {
        const int nf = std::min(batch, count - f0);
        const int n0 = first_frame + f0;
        float* res = d_res + (size_t)f0 * res_frame_stride;
        // Stacks are fully written by finalize (sums) and den-box (weights);
        // they start from zero only when members can land outside the LDS
        // regions and go to global atomics.
        const bool outliers = p->frm < p->margin;
        if (outliers) {
            if (res_plane_stride == plane_stack) {
                BM3D_HIP(bm3d::launch_zero(res, (size_t)p->nplanes * plane_stack, (size_t)res_frame_stride, nf, stream));
            } else {
                for (int f = 0; f < nf; ++f)
                    BM3D_HIP(bm3d::launch_zero(res + (size_t)f * res_frame_stride, (size_t)plane_stack,
                                               (size_t)res_plane_stride, p->nplanes, stream));
            }
        }
        BM3D_HIP(evaluate(p, src, ref, n0, nf, res, res_frame_stride, res_plane_stride, 0, stream));

        bm3d::FinalizeArgs fz{};
        fz.scratch = p->scratch;
        fz.fx_range = fx_range(p);
        fz.fx_exp = p->fx_exp;
        fz.res = res;
        fz.res_frame_stride = res_frame_stride;
        fz.res_plane_stride = res_plane_stride;
        fz.W = d.width; fz.H = d.height; fz.stride = d.stride;
        fz.plane_size = p->plane_size;
        fz.nplanes = p->nplanes;
        fz.slots = p->tw;
        fz.bs = d.block_step; fz.tbx = p->ftbx; fz.tby = p->ftby; fz.ntx = p->fntx; fz.nty = p->fnty;
        fz.rm = p->frm; fz.rw = p->frw; fz.rh = p->frh;
        fz.accumulate = outliers ? 1 : 0;
        fz.chroma = d.chroma;
        for (int i = 0; i < 3; ++i) fz.sigma[i] = p->sigma[i];
        BM3D_HIP(bm3d::launch_finalize(fz, nf * p->nplanes * p->tw, stream));

        bm3d::DenArgs da{};
        da.corner = p->corner;
        da.corner_frame_stride = p->corner_frame;
        da.res = res;
        da.res_frame_stride = res_frame_stride;
        da.res_plane_stride = res_plane_stride;
        da.W = d.width; da.H = d.height; da.stride = d.stride;
        da.plane_size = p->plane_size;
        da.nplanes = p->nplanes;
        da.slots = p->tw;
        BM3D_HIP(bm3d::launch_den_box(da, nf * p->nplanes * p->tw, stream));
    }
    return BM3D_OK;
}

long long bm3d_num_blocks(const bm3d_plan* p) { return p ? p->nblocks : 0; }

int bm3d_profile_enable(int on) {
    Profiler& P = profiler();
    std::lock_guard<std::mutex> g(P.mu);
    P.on = on != 0;
    return BM3D_OK;
}

int bm3d_profile_reset(void) {
    Profiler& P = profiler();
    std::lock_guard<std::mutex> g(P.mu);
    for (auto& r : P.recs) {
        (void)hipEventSynchronize(r.b);
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    P.recs.clear();
    return BM3D_OK;
}

int bm3d_profile_read(int kernel, double* total_ms, long long* launches) {
    g_last_error.clear();
    if (!total_ms || !launches) return fail(BM3D_EINVAL, "BM3D: null argument");
    Profiler& P = profiler();
    std::lock_guard<std::mutex> g(P.mu);
    double t = 0.0;
    long long n = 0;
    for (auto& r : P.recs) {
        if (r.kernel != kernel) continue;
        BM3D_HIP(hipEventSynchronize(r.b));
        float ms = 0.0f;
        BM3D_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        ++n;
    }
    *total_ms = t;
    *launches = n;
    return BM3D_OK;
}

int bm3d_plan_matching_kernel(const bm3d_plan* p) { return p && p->use_dx ? 1 : 0; }

int bm3d_plan_filter_kernel(const bm3d_plan* p) { return p && p->use_mx_filter ? 1 : 0; }

int bm3d_plan_select_kernels(bm3d_plan* p, int matching, int filter) {
    g_last_error.clear();
    if (!p) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (matching < 0 || matching > 1 || filter < 0 || filter > 1)
        return fail(BM3D_EINVAL, "BM3D: kernel selector must be 0 (FP32) or 1 (mixed precision)");
    if ((matching || filter) && p->d.precision != BM3D_PRECISION_MIXED)
        return fail(BM3D_EINVAL, "BM3D: mixed-precision kernels need a plan created with precision MIXED");
    if (matching && !p->dx_ok)
        return fail(BM3D_EUNSUPPORTED, "BM3D: search windows too large for the mixed-precision matching kernel");
    p->use_dx = matching != 0;
    p->use_mx_filter = filter != 0;
    return BM3D_OK;
}

int bm3d_count_candidates(bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count,
                          unsigned long long* ncand, void* stream) {
    g_last_error.clear();
    if (!p || !src || !src->data || !ncand) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (p->d.final_ && (!ref || !ref->data)) return fail(BM3D_EINVAL, "BM3D: final pass needs \"ref\"");
    if (src->num_frames <= 0 || count < 0 || first_frame < 0 || first_frame + count > src->num_frames)
        return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    unsigned long long* d_cnt = nullptr;
    BM3D_HIP(hipMalloc(&d_cnt, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), st);
    if (e == hipSuccess && count > 0) e = launch_matching(p, nullptr, d_cnt, src, ref, first_frame, count, st);
    if (e == hipSuccess) e = hipMemcpyAsync(ncand, d_cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_cnt);
    if (e != hipSuccess) return hip_fail(e, "bm3d_count_candidates");
    return BM3D_OK;
}

int bm3d_match(bm3d_plan* p, unsigned int* d_groups, const bm3d_clip* src, const bm3d_clip* ref,
               int first_frame, int count, void* stream) {
    g_last_error.clear();
    if (!p || !d_groups || !src || !src->data) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (p->d.final_ && (!ref || !ref->data)) return fail(BM3D_EINVAL, "BM3D: final pass needs \"ref\"");
    if (src->num_frames <= 0 || count < 0 || first_frame < 0 || first_frame + count > src->num_frames)
        return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    if (count == 0) return BM3D_OK;
    BM3D_HIP(launch_matching(p, d_groups, nullptr, src, ref, first_frame, count,
                             reinterpret_cast<hipStream_t>(stream)));
    return BM3D_OK;
}

int bm3d_execute_vagg(bm3d_plan* p, float* dst, long long dst_frame_stride, long long dst_plane_stride,
                      int dst_stride, const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count,
                      void* stream_) {
    g_last_error.clear();
    if (!p || !dst || !src || !src->data) return fail(BM3D_EINVAL, "BM3D: null argument");
    const bm3d_desc& d = p->d;
    if (d.final_ && (!ref || !ref->data)) return fail(BM3D_EINVAL, "BM3D: final pass needs \"ref\"");
    if (d.final_ && (ref->frame_stride != src->frame_stride || ref->plane_stride != src->plane_stride ||
                     ref->num_frames != src->num_frames))
        return fail(BM3D_EINVAL, "BM3D: \"ref\" must have the layout and number of frames of \"clip\"");
    if (src->num_frames <= 0 || count < 0 || first_frame < 0 || first_frame + count > src->num_frames)
        return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    if (d.radius <= 0) return fail(BM3D_EINVAL, "BM3D: bm3d_execute_vagg needs radius > 0");
    if (p->frm < p->margin)
        return fail(BM3D_EUNSUPPORTED, "BM3D: search reach too large for region aggregation "
                                       "(use bm3d_execute and bm3d_vaggregate)");
    if (dst_stride == 0) dst_stride = d.width;
    if (dst_stride < d.width) return fail(BM3D_EINVAL, "BM3D: destination stride < width");
    if (dst_plane_stride == 0) dst_plane_stride = (long long)d.height * dst_stride;
    if (dst_frame_stride == 0) dst_frame_stride = p->nplanes * dst_plane_stride;
    if (count == 0) return BM3D_OK;
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    int dev = 0;
    BM3D_HIP(hipGetDevice(&dev));
    if (dev != p->device) return fail(BM3D_EINVAL, "BM3D: plan used on a different device than it was created on");

    const int N = src->num_frames, r = d.radius;
    // evaluations needed: frames clamp(n - r .. n + r) of every output n
    const int mlo = std::max(first_frame - r, 0);
    const int mhi = std::min(first_frame + count - 1 + r, N - 1);
    const int batch = std::min(kMaxBatchFrames, mhi - mlo + 1);
    // resident evaluations: an output's 2r + 1 plus the batch in flight
    const int ring = std::min(batch + 2 * r, mhi - mlo + 1);
    if (int rc = ensure_scratch(p, ring)) return rc;

    bm3d::VRegArgs va{};
    va.scratch = p->scratch;
    va.scratch_frame = p->scratch_frame;
    va.corner = p->corner;
    va.corner_frame = p->corner_frame;
    va.ring = ring;
    va.fx_range = fx_range(p);
    va.fx_exp = p->fx_exp;
    va.dst_frame_stride = dst_frame_stride;
    va.dst_plane_stride = dst_plane_stride;
    va.dst_stride = dst_stride;
    va.W = d.width; va.H = d.height; va.stride = d.stride;
    va.plane_size = p->plane_size;
    va.nplanes = p->nplanes;
    va.plane_mask = 0;
    for (int i = 0; i < p->nplanes; ++i)
        if (!(d.chroma && p->sigma[i] < std::numeric_limits<float>::epsilon())) va.plane_mask |= 1u << i;
    va.radius = r;
    va.num_frames = N;
    va.bs = d.block_step; va.tbx = p->ftbx; va.tby = p->ftby; va.ntx = p->fntx; va.nty = p->fnty;
    va.rm = p->frm; va.rw = p->frw; va.rh = p->frh;

    BM3D_HIP(measure_range(p, src, std::max(mlo - r, 0), std::min(mhi + r, N - 1), stream));
    int next_out = first_frame;
    for (int b0 = mlo; b0 <= mhi;) {
        const int b1 = std::min(b0 + batch, mhi + 1);
        BM3D_HIP(evaluate(p, src, ref, b0, b1 - b0, nullptr, 0, 0, ring, stream));
        // outputs whose evaluations are all resident
        const int out_hi = (b1 - 1 == N - 1) ? first_frame + count - 1
                                              : std::min(first_frame + count - 1, b1 - 1 - r);
        if (out_hi >= next_out) {
            va.first_frame = next_out;
            va.dst = dst + (size_t)(next_out - first_frame) * dst_frame_stride;
            BM3D_HIP(bm3d::launch_vagg_regions(va, out_hi - next_out + 1, stream));
            next_out = out_hi + 1;
        }
        b0 = b1;
    }
    if (next_out != first_frame + count) return fail(BM3D_EHIP, "BM3D: internal frame schedule error");
    return BM3D_OK;
}

int bm3d_execute_window(bm3d_plan* p, float* d_res, const float* d_src, void* stream) {
    if (!p || !d_res || !d_src) return fail(BM3D_EINVAL, "BM3D: null argument");
    const long long fs = p->plane_size;           // one slot
    const long long ps = (long long)p->tw * fs;   // one plane of slots
    const long long clip_stride = p->nplanes * ps;
    bm3d_clip ref{d_src, fs, ps, p->tw};
    bm3d_clip src{d_src + (p->d.final_ ? clip_stride : 0), fs, ps, p->tw};
    return bm3d_execute(p, d_res, 0, 0, &src, p->d.final_ ? &ref : nullptr, p->d.radius, 1, stream);
}

int bm3d_aggregate(float* dst, long long dst_frame_stride, long long dst_plane_stride, int dst_stride,
                   const float* res, long long res_frame_stride, long long res_plane_stride, int width,
                   int height, int stride, int nplanes, int count, void* stream) {
    g_last_error.clear();
    if (!dst || !res) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (width <= 0 || height <= 0 || stride < width || dst_stride < width || nplanes <= 0 || count < 0)
        return fail(BM3D_EINVAL, "BM3D: invalid aggregation shape");
    if (count == 0) return BM3D_OK;
    bm3d::DivideArgs a{};
    a.res = res;
    a.plane_size = (long long)height * stride;
    a.res_plane_stride = res_plane_stride ? res_plane_stride : 2 * a.plane_size;
    a.res_frame_stride = res_frame_stride ? res_frame_stride : nplanes * a.res_plane_stride;
    a.dst = dst;
    a.dst_frame_stride = dst_frame_stride;
    a.dst_plane_stride = dst_plane_stride;
    a.dst_stride = dst_stride;
    a.W = width; a.H = height; a.stride = stride;
    a.nplanes = nplanes;
    BM3D_HIP(bm3d::launch_divide(a, count * nplanes, reinterpret_cast<hipStream_t>(stream)));
    return BM3D_OK;
}

int bm3d_vaggregate(float* dst, long long dst_frame_stride, long long dst_plane_stride, int dst_stride,
                    const float* stacks, long long stack_frame_stride, long long stack_plane_stride,
                    int stack_first, int width, int height, int stride, int nplanes, int radius,
                    int num_frames, int first_frame, int count, void* stream) {
    g_last_error.clear();
    if (!dst || !stacks) return fail(BM3D_EINVAL, "VAggregate: null argument");
    if (width <= 0 || height <= 0 || stride < width || dst_stride < width || nplanes <= 0 || radius < 0 ||
        count < 0 || first_frame < 0 || first_frame + count > num_frames)
        return fail(BM3D_EINVAL, "VAggregate: invalid shape or frame range");
    if (count == 0) return BM3D_OK;
    // the stacks must cover every frame the window touches
    if (stack_first > std::max(first_frame - radius, 0))
        return fail(BM3D_EINVAL, "VAggregate: stacks do not cover the temporal window");
    bm3d::VAggArgs a{};
    a.stacks = stacks;
    a.plane_size = (long long)height * stride;
    a.stack_plane_stride = stack_plane_stride ? stack_plane_stride : (long long)(2 * radius + 1) * 2 * a.plane_size;
    a.stack_frame_stride = stack_frame_stride ? stack_frame_stride : nplanes * a.stack_plane_stride;
    a.stack_first = stack_first;
    a.dst = dst;
    a.dst_frame_stride = dst_frame_stride;
    a.dst_plane_stride = dst_plane_stride;
    a.dst_stride = dst_stride;
    a.W = width; a.H = height; a.stride = stride;
    a.nplanes = nplanes;
    a.radius = radius;
    a.num_frames = num_frames;
    a.first_frame = first_frame;
    BM3D_HIP(bm3d::launch_vaggregate(a, count * nplanes, reinterpret_cast<hipStream_t>(stream)));
    return BM3D_OK;
}

}  // extern "C"
