// VapourSynth (API 3) plugin entry of the bm3dcuda filters on MI355X:
// VapourSynthPluginInit registers BM3D, VAggregate and BM3Dv2 with the
// reference's identifiers and argument strings (source/source.cpp:991-1027;
// built with BM3D_VS_RTC, those of the NVRTC variant,
// rtc_source/source.cpp:1358-1397; with BM3D_VS_OWN_NAMESPACE, the NVRTC
// variant's arguments under an identifier and namespace of its own), and the
// filters run on libbm3d_mi355x through its C ABI:
//   * creation parses the argument map into bm3d_args and calls
//     bm3d_filter_create / bm3dv2_filter_create / bm3d_vaggregate_create,
//     which apply the reference's defaults, checks and messages;
//   * GetFrame requests the input frames bm3d_filter_host_range names and
//     hands the host frames to bm3d_filter_get_frames_host, which uploads
//     each once and returns only the output frame (the reference stages the
//     whole window through pinned memory per frame, source/source.cpp:
//     296-426).
// BM3Dv2 is one filter here (BM3D and VAggregate fused on the device), not
// an invoke of BM3D then VAggregate as in source/source.cpp:911-989: its
// per-frame stacks never leave the GPU. One optional argument has no
// counterpart in the reference: precision ("fp32", the default, or
// "mixed").
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "bm3d_mi355x.h"
#include "vs_api3.h"

namespace {

#if defined(BM3D_VS_OWN_NAMESPACE)
// its own identifier and namespace, to install beside a bm3dcuda plugin
constexpr const char* kIdentifier = "com.mi355x.bm3d";
constexpr const char* kNamespace = "bm3dmi355x";
constexpr const char* kDescription = "BM3D algorithm implemented in HIP for MI355X";
constexpr const char* kBm3dArgs =
    "clip:clip;ref:clip:opt;sigma:float[]:opt;block_step:int[]:opt;bm_range:int[]:opt;radius:int:opt;"
    "ps_num:int[]:opt;ps_range:int[]:opt;chroma:int:opt;device_id:int:opt;fast:int:opt;extractor_exp:int:opt;"
    "bm_error_s:data[]:opt;transform_2d_s:data[]:opt;transform_1d_s:data[]:opt;zero_init:int:opt;"
    "precision:data:opt;";
#elif defined(BM3D_VS_RTC)
constexpr const char* kIdentifier = "com.wolframrhodium.bm3dcuda_rtc";
constexpr const char* kNamespace = "bm3dcuda_rtc";
constexpr const char* kDescription = "BM3D algorithm implemented in HIP for MI355X (bm3dcuda_rtc interface)";
constexpr const char* kBm3dArgs =
    "clip:clip;ref:clip:opt;sigma:float[]:opt;block_step:int[]:opt;bm_range:int[]:opt;radius:int:opt;"
    "ps_num:int[]:opt;ps_range:int[]:opt;chroma:int:opt;device_id:int:opt;fast:int:opt;extractor_exp:int:opt;"
    "bm_error_s:data[]:opt;transform_2d_s:data[]:opt;transform_1d_s:data[]:opt;zero_init:int:opt;"
    "precision:data:opt;";
#else
constexpr const char* kIdentifier = "com.wolframrhodium.bm3dcuda";
constexpr const char* kNamespace = "bm3dcuda";
constexpr const char* kDescription = "BM3D algorithm implemented in HIP for MI355X (bm3dcuda interface)";
constexpr const char* kBm3dArgs =
    "clip:clip;ref:clip:opt;sigma:float[]:opt;block_step:int[]:opt;bm_range:int[]:opt;radius:int:opt;"
    "ps_num:int[]:opt;ps_range:int[]:opt;chroma:int:opt;device_id:int:opt;fast:int:opt;extractor_exp:int:opt;"
    "zero_init:int:opt;precision:data:opt;";
#endif
constexpr const char* kVAggregateArgs = "clip:clip;src:clip;planes:int[];";

enum class Kind { kBM3D, kVAggregate, kBM3Dv2 };

struct Instance {
    Kind kind;
    bm3d_filter* filter = nullptr;
    VSNodeRef* clip = nullptr;  // BM3D, BM3Dv2: "clip"; VAggregate: the BM3D output
    VSNodeRef* ref = nullptr;   // BM3D, BM3Dv2: "ref" (may be null); VAggregate: "src"
    VSVideoInfo vi{};           // output
    int num_frames = 0;
    int radius = 0;
    int process[3] = {0, 0, 0};
};

int num_elements(const VSMap* in, const char* key, const VSAPI* vsapi) {
    const int n = vsapi->propNumElements(in, key);
    return n < 0 ? 0 : n;
}

bm3d_video_info video_info(const VSVideoInfo* v) {
    bm3d_video_info o{};
    if (v->format) {
        const VSFormat* f = v->format;
        o.color_family = f->colorFamily == cmGray  ? BM3D_GRAY
                         : f->colorFamily == cmRGB ? BM3D_RGB
                         : f->colorFamily == cmYUV ? BM3D_YUV
                                                   : 0;
        o.sample_type = f->sampleType;
        o.bits_per_sample = f->bitsPerSample;
        o.subsampling_w = f->subSamplingW;
        o.subsampling_h = f->subSamplingH;
        o.num_planes = f->numPlanes;
        o.width = v->width;
        o.height = v->height;
    } else {
        o.sample_type = -1;  // variable format: rejected like the reference (isConstantFormat)
    }
    o.num_frames = v->numFrames;
    return o;
}

// The BM3D / BM3Dv2 argument map as bm3d_args. String values point into
// `in`, which outlives the create call.
int parse_args(const VSMap* in, const VSAPI* vsapi, bm3d_args* a, std::string* err) {
    std::memset(a, 0, sizeof(*a));
    int e = 0;
    a->n_sigma = num_elements(in, "sigma", vsapi);
    for (int i = 0; i < a->n_sigma && i < 3; ++i) a->sigma[i] = vsapi->propGetFloat(in, "sigma", i, &e);
    auto ints = [&](const char* key, int* n, long long* v, int cap) {
        *n = num_elements(in, key, vsapi);
        for (int i = 0; i < *n && i < cap; ++i) v[i] = vsapi->propGetInt(in, key, i, &e);
    };
    ints("block_step", &a->n_block_step, a->block_step, 3);
    ints("bm_range", &a->n_bm_range, a->bm_range, 3);
    ints("radius", &a->n_radius, &a->radius, 1);
    ints("ps_num", &a->n_ps_num, a->ps_num, 3);
    ints("ps_range", &a->n_ps_range, a->ps_range, 3);
    ints("chroma", &a->n_chroma, &a->chroma, 1);
    ints("device_id", &a->n_device_id, &a->device_id, 1);
    ints("fast", &a->n_fast, &a->fast, 1);
    ints("extractor_exp", &a->n_extractor_exp, &a->extractor_exp, 1);
    ints("zero_init", &a->n_zero_init, &a->zero_init, 1);
    auto strings = [&](const char* key, int* n, const char** v) {
        *n = num_elements(in, key, vsapi);
        for (int i = 0; i < *n && i < 3; ++i) v[i] = vsapi->propGetData(in, key, i, &e);
    };
    strings("bm_error_s", &a->n_bm_error_s, a->bm_error_s);
    strings("transform_2d_s", &a->n_transform_2d_s, a->transform_2d_s);
    strings("transform_1d_s", &a->n_transform_1d_s, a->transform_1d_s);
    if (num_elements(in, "precision", vsapi) > 0) {
        const std::string p = vsapi->propGetData(in, "precision", 0, &e);
        a->n_precision = 1;
        if (p == "fp32") {
            a->precision = BM3D_PRECISION_FP32;
        } else if (p == "mixed") {
            a->precision = BM3D_PRECISION_MIXED;
        } else {
            *err = "BM3D: \"precision\" must be \"fp32\" or \"mixed\"";
            return BM3D_EINVAL;
        }
    }
    return BM3D_OK;
}

void free_instance(Instance* d, const VSAPI* vsapi) {
    if (d->filter) bm3d_filter_destroy(d->filter);
    if (d->clip) vsapi->freeNode(d->clip);
    if (d->ref) vsapi->freeNode(d->ref);
    delete d;
}

// Output video info from bm3d_filter_info: the input's format and rate,
// the filter's size (BM3D with radius > 0 stacks 2 (2r + 1) planes' worth
// of rows, source/source.cpp:190-200).
void finish_instance(Instance* d, const VSVideoInfo* in_vi, const VSAPI* vsapi) {
    bm3d_video_info o{};
    bm3d_filter_info(d->filter, &o, &d->radius, d->process);
    d->vi = *in_vi;
    d->vi.width = o.width;
    d->vi.height = o.height;
    d->vi.numFrames = o.num_frames;
    (void)vsapi;
}

void VS_CC filter_init(VSMap*, VSMap*, void** instanceData, VSNode* node, VSCore*, const VSAPI* vsapi) {
    auto* d = static_cast<Instance*>(*instanceData);
    vsapi->setVideoInfo(&d->vi, 1, node);
}

bm3d_host_frame host_frame(const VSFrameRef* f, int planes, const VSAPI* vsapi) {
    bm3d_host_frame h{};
    for (int p = 0; p < planes && p < 3; ++p) {
        h.data[p] = reinterpret_cast<float*>(const_cast<uint8_t*>(vsapi->getReadPtr(f, p)));
        h.stride[p] = vsapi->getStride(f, p) / (int)sizeof(float);
    }
    return h;
}

const VSFrameRef* VS_CC filter_get_frame(int n, int activationReason, void** instanceData, void**,
                                         VSFrameContext* frameCtx, VSCore* core, const VSAPI* vsapi) {
    auto* d = static_cast<Instance*>(*instanceData);
    int lo = 0, hi = 0;
    if (bm3d_filter_host_range(d->filter, n, 1, &lo, &hi) != BM3D_OK) {
        vsapi->setFilterError(bm3d_last_error(), frameCtx);
        return nullptr;
    }
    if (activationReason == arInitial) {
        for (int i = lo; i <= hi; ++i) vsapi->requestFrameFilter(i, d->clip, frameCtx);
        if (d->kind == Kind::kVAggregate) {
            vsapi->requestFrameFilter(n, d->ref, frameCtx);
        } else if (d->ref) {
            for (int i = lo; i <= hi; ++i) vsapi->requestFrameFilter(i, d->ref, frameCtx);
        }
        return nullptr;
    }
    if (activationReason != arAllFramesReady) return nullptr;

    const int planes = d->vi.format->numPlanes;
    const int count = hi - lo + 1;
    std::vector<const VSFrameRef*> held;
    held.reserve(2 * count + 1);
    std::vector<bm3d_host_frame> cin(count), rin;
    for (int i = 0; i < count; ++i) {
        held.push_back(vsapi->getFrameFilter(lo + i, d->clip, frameCtx));
        cin[i] = host_frame(held.back(), planes, vsapi);
    }
    const VSFrameRef* prop_src = nullptr;
    if (d->kind == Kind::kVAggregate) {
        // VAggregate reads only frame n of "src" (source/source.cpp:762)
        rin.assign(count, bm3d_host_frame{});
        prop_src = vsapi->getFrameFilter(n, d->ref, frameCtx);
        held.push_back(prop_src);
        rin[n - lo] = host_frame(prop_src, planes, vsapi);
    } else {
        prop_src = held[n - lo];
        if (d->ref) {
            rin.resize(count);
            for (int i = 0; i < count; ++i) {
                held.push_back(vsapi->getFrameFilter(lo + i, d->ref, frameCtx));
                rin[i] = host_frame(held.back(), planes, vsapi);
            }
        }
    }
    VSFrameRef* dst = vsapi->newVideoFrame(d->vi.format, d->vi.width, d->vi.height, prop_src, core);
    bm3d_host_frame out{};
    for (int p = 0; p < planes && p < 3; ++p) {
        out.data[p] = reinterpret_cast<float*>(vsapi->getWritePtr(dst, p));
        out.stride[p] = vsapi->getStride(dst, p) / (int)sizeof(float);
    }
    const int rc = bm3d_filter_get_frames_host(d->filter, cin.data(), rin.empty() ? nullptr : rin.data(), lo, &out,
                                               n, 1);
    for (const VSFrameRef* f : held) vsapi->freeFrame(f);
    if (rc != BM3D_OK) {
        vsapi->freeFrame(dst);
        vsapi->setFilterError(bm3d_last_error(), frameCtx);
        return nullptr;
    }
    if (d->kind == Kind::kBM3D && d->radius > 0) {
        // source/source.cpp:428-435
        VSMap* props = vsapi->getFramePropsRW(dst);
        vsapi->propSetInt(props, "BM3D_V_radius", d->radius, paReplace);
        const int64_t process[3] = {d->process[0], d->process[1], d->process[2]};
        vsapi->propSetIntArray(props, "BM3D_V_process", process, 3);
    }
    return dst;
}

void VS_CC filter_free(void* instanceData, VSCore*, const VSAPI* vsapi) {
    free_instance(static_cast<Instance*>(instanceData), vsapi);
}

const char* filter_name(Kind k) {
    return k == Kind::kBM3D ? "BM3D" : k == Kind::kVAggregate ? "VAggregate" : "BM3Dv2";
}

// BM3D and BM3Dv2 (userData selects which).
void VS_CC bm3d_create(const VSMap* in, VSMap* out, void* userData, VSCore* core, const VSAPI* vsapi) {
    const Kind kind = userData ? Kind::kBM3Dv2 : Kind::kBM3D;
    auto* d = new Instance{kind};
    d->clip = vsapi->propGetNode(in, "clip", 0, nullptr);
    int e = 0;
    d->ref = vsapi->propGetNode(in, "ref", 0, &e);
    if (e) d->ref = nullptr;
    const VSVideoInfo* vi = vsapi->getVideoInfo(d->clip);
    const bm3d_video_info cvi = video_info(vi);
    bm3d_video_info rvi{};
    if (d->ref) rvi = video_info(vsapi->getVideoInfo(d->ref));
    bm3d_args a;
    std::string err;
    int rc = parse_args(in, vsapi, &a, &err);
    if (rc == BM3D_OK) {
        rc = kind == Kind::kBM3D ? bm3d_filter_create(&a, &cvi, d->ref ? &rvi : nullptr, &d->filter)
                                 : bm3dv2_filter_create(&a, &cvi, d->ref ? &rvi : nullptr, &d->filter);
        if (rc != BM3D_OK) err = bm3d_last_error();
    }
    if (rc != BM3D_OK) {
        vsapi->setError(out, err.c_str());
        free_instance(d, vsapi);
        return;
    }
    finish_instance(d, vi, vsapi);
    if (kind == Kind::kBM3Dv2) {
        bool any = false;
        for (int p = 0; p < vi->format->numPlanes && p < 3; ++p) any |= d->process[p] != 0;
        if (!any) {  // no plane processed: the clip itself (source/source.cpp:932-943)
            vsapi->propSetNode(out, "clip", d->clip, paReplace);
            free_instance(d, vsapi);
            return;
        }
    }
    vsapi->createFilter(in, out, filter_name(kind), filter_init, filter_get_frame, filter_free, fmParallel, 0, d,
                        core);
}

void VS_CC vaggregate_create(const VSMap* in, VSMap* out, void*, VSCore* core, const VSAPI* vsapi) {
    auto* d = new Instance{Kind::kVAggregate};
    d->clip = vsapi->propGetNode(in, "clip", 0, nullptr);
    d->ref = vsapi->propGetNode(in, "src", 0, nullptr);
    const VSVideoInfo* src_vi = vsapi->getVideoInfo(d->ref);
    const bm3d_video_info cvi = video_info(vsapi->getVideoInfo(d->clip));
    const bm3d_video_info svi = video_info(src_vi);
    const int n = num_elements(in, "planes", vsapi);
    std::vector<long long> planes(n);
    for (int i = 0; i < n; ++i) planes[i] = vsapi->propGetInt(in, "planes", i, nullptr);
    if (bm3d_vaggregate_create(planes.data(), n, &cvi, &svi, &d->filter) != BM3D_OK) {
        vsapi->setError(out, bm3d_last_error());
        free_instance(d, vsapi);
        return;
    }
    finish_instance(d, src_vi, vsapi);
    vsapi->createFilter(in, out, "VAggregate", filter_init, filter_get_frame, filter_free, fmParallel, 0, d, core);
}

}  // namespace

VS_EXTERNAL_API(void) VapourSynthPluginInit(VSConfigPlugin configFunc, VSRegisterFunction registerFunc,
                                            VSPlugin* plugin) {
    configFunc(kIdentifier, kNamespace, kDescription, VAPOURSYNTH_API_VERSION, 1, plugin);
    registerFunc("BM3D", kBm3dArgs, bm3d_create, nullptr, plugin);
    registerFunc("VAggregate", kVAggregateArgs, vaggregate_create, nullptr, plugin);
    static int v2_tag;
    registerFunc("BM3Dv2", kBm3dArgs, bm3d_create, &v2_tag, plugin);
}
