// Batched launches over many clips (bm3d_batch_* in include/bm3d_mi355x.h).
//
// Each job is one plane group of one clip with its own plan(s) (parameters,
// tiles). Jobs are packed into waves whose scratch fits a device-memory
// budget; within a wave every stage runs as one multi-job launch (the
// kernels' *_multi entries find their job from the flat work-group index),
// grouped only where jobs need different kernel variants or block sizes.
// A job's frames are all evaluated in one run, so every output frame's
// regions are complete within its work-group (target mode, no atomics), and
// the ring of output frames is the whole clip. The whole sequence is
// captured once into a HIP graph.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "bm3d_inject.h"
#include "bm3d_internal.h"

namespace {

using bm3d::fail;

int hip_err(hipError_t e, const char* what) {
    return fail(BM3D_EHIP, std::string("BM3D: '") + what + "' failed: " + hipGetErrorString(e));
}

#define BATCH_HIP(expr)                                   \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return hip_err(_e, #expr);  \
    } while (0)

// Default device memory one wave's scratch may take.
constexpr long long kDefaultWaveBudget = 16LL << 30;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Job {
    bm3d_batch_job job{};
    bm3d_plan* plan[2] = {nullptr, nullptr};  // basic, final
    unsigned process = 0;                     // planes BM3Dv2 processes (bit p)
    int nplanes = 1;
    int wave = 0;
    // scratch of the job within its wave's arena (bytes)
    size_t corner_off = 0, wrap_off = 0, range_off = 0, groups_off = 0, scratch_off = 0;
    size_t slist_off = 0;  // spatial lists between the two matching launches (two_launch_match)
};

// Whether a job's matching runs as the two-launch strip search (spatial by
// strips, then temporal with one lane per block, bm3d_capi.cpp
// launch_matching), for either pass.
bool two_launch_match(const Job& j) {
    for (int pass = 0; pass < j.job.passes; ++pass)
        if (bm3d::plan_uses_strip(j.plan[pass]) && j.plan[pass]->d.radius > 0) return true;
    return false;
}

// One kernel launch of the graph.
struct Launch {
    int kind;  // 0 absmax, 1 match, 2 mixed match, 3 filter, 4 mixed filter, 5 vagg
    std::vector<int> jobs;
    int pass = 0;
    int block = 0;  // threads (matching) or waves (filtering)
    int keys = 0, final_ = 0, chroma = 0;  // keys: the mixed matcher's temporal keys
    int phase = 0;                         // mixed matching: MatchArgs::phase
    int any_extractor = 0;                 // mixed filtering: some job has extractor_exp
    int variant = 0;  // matching: block distance (FP32) or dx_spatial_variant (mixed);
                      // filtering: the FP32 generic-transform variant
    size_t lds = 0;
    size_t args_off = 0, grids_off = 0;  // in the table buffer
    int wgs = 0;
};

}  // namespace

struct bm3d_batch {
    int device = 0;
    std::vector<Job> jobs;
    bm3d_tuning tune{};
    int nwaves = 0;
    long long wave_budget = kDefaultWaveBudget;
    size_t arena_bytes = 0;
    char* arena = nullptr;
    std::vector<std::pair<size_t, size_t>> wave_zero;  // per wave: bytes of corner and wrap maps + ranges from 0
    char* tables = nullptr;
    size_t tables_bytes = 0;
    std::vector<std::vector<Launch>> wave_launches;  // per wave, in order
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int launches = 0;
};

namespace {

void release_graph(bm3d_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    if (b->tables) (void)hipFree(b->tables);
    b->tables = nullptr;
    b->tables_bytes = 0;
}

bm3d_clip ref_clip(const Job& j) {
    return bm3d_clip{j.job.basic, j.job.src.frame_stride, j.job.src.plane_stride, j.job.src.num_frames};
}

int passes(const Job& j) { return j.job.passes; }

// Lays out the waves and their scratch for the jobs' current geometry.
void layout(bm3d_batch* b) {
    b->nwaves = 0;
    b->arena_bytes = 0;
    b->wave_zero.clear();
    size_t used = 0, zero_bytes = 0;
    std::vector<int> wave_jobs;
    auto close_wave = [&]() {
        if (wave_jobs.empty()) return;
        // corner maps, wrap maps and ranges (with the wrap flag after each
        // range) first, zeroed together before each pass; then groups and
        // regions
        size_t off = 0;
        for (int i : wave_jobs) {
            Job& j = b->jobs[i];
            const size_t map = align_up((size_t)j.job.src.num_frames * j.nplanes * j.plan[0]->plane_size * 4, 256);
            j.corner_off = off;
            j.wrap_off = off + map;
            off += 2 * map;
        }
        for (int i : wave_jobs) {
            b->jobs[i].range_off = off;
            off += 256;
        }
        zero_bytes = off;
        for (int i : wave_jobs) {
            Job& j = b->jobs[i];
            j.groups_off = off;
            off += align_up((size_t)j.job.src.num_frames * j.plan[0]->nblocks * 8 * sizeof(uint32_t), 256);
            j.scratch_off = off;
            off += align_up((size_t)j.job.src.num_frames * j.plan[0]->region_ints * sizeof(int), 256);
            if (two_launch_match(j)) {
                j.slist_off = off;
                off += align_up((size_t)j.job.src.num_frames * j.plan[0]->nblocks * 16 * sizeof(uint32_t), 256);
            }
        }
        b->wave_zero.emplace_back(0, zero_bytes);
        b->arena_bytes = std::max(b->arena_bytes, off);
        ++b->nwaves;
        wave_jobs.clear();
        used = 0;
    };
    for (int i = 0; i < (int)b->jobs.size(); ++i) {
        Job& j = b->jobs[i];
        const bm3d_plan* p = j.plan[0];
        const long long N = j.job.src.num_frames;
        const size_t need = (size_t)N * ((size_t)j.nplanes * p->plane_size * 8 + (size_t)p->nblocks * 32 +
                                         (size_t)p->region_ints * 4 + (two_launch_match(j) ? p->nblocks * 64 : 0)) +
                            4096;
        if (!wave_jobs.empty() && used + need > (size_t)b->wave_budget) close_wave();
        j.wave = b->nwaves;
        wave_jobs.push_back(i);
        used += need;
    }
    close_wave();
}

template <typename T>
size_t table_push(std::vector<char>& tab, const std::vector<T>& v) {
    const size_t off = align_up(tab.size(), 256);
    tab.resize(off + v.size() * sizeof(T));
    std::memcpy(tab.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// Builds every launch's argument and grid tables (host copy in `tab`).
int build_launches(bm3d_batch* b, std::vector<char>& tab) {
    b->wave_launches.assign(b->nwaves, {});
    b->launches = 0;
    for (int w = 0; w < b->nwaves; ++w) {
        char* base = b->arena;
        for (int pass = 0; pass < 2; ++pass) {
            std::vector<int> js;
            for (int i = 0; i < (int)b->jobs.size(); ++i) {
                const Job& j = b->jobs[i];
                if (j.wave == w && passes(j) > pass && j.process) js.push_back(i);
            }
            if (js.empty()) continue;
            auto& L = b->wave_launches[w];
            auto range = [&](const Job& j) -> float* { return reinterpret_cast<float*>(base + j.range_off); };
            auto fx_range = [&](const Job& j) -> float* { return j.job.desc.extractor_exp ? nullptr : range(j); };
            auto wrap_flag = [&](const Job& j) -> int* { return reinterpret_cast<int*>(base + j.range_off) + 1; };
            // fixed-point range
            {
                Launch l{};
                l.kind = 0;
                l.pass = pass;
                std::vector<bm3d::AbsmaxArgs> args;
                std::vector<bm3d::JobGrid> grids;
                for (int i : js) {
                    const Job& j = b->jobs[i];
                    bm3d::AbsmaxArgs a{};
                    a.clip = j.job.src.data;
                    a.frame_stride = j.job.src.frame_stride;
                    a.plane_stride = j.job.src.plane_stride;
                    a.nplanes = j.nplanes;
                    a.W = j.job.desc.width;
                    a.H = j.job.desc.height;
                    a.stride = j.job.desc.stride;
                    a.out = range(j);
                    const int gx = std::min(a.H, 64), gy = j.job.src.num_frames * j.nplanes;
                    grids.push_back({l.wgs, gx, gy});
                    l.wgs += gx * gy;
                    args.push_back(a);
                    l.jobs.push_back(i);
                }
                if (!args.empty()) {
                    l.args_off = table_push(tab, args);
                    l.grids_off = table_push(tab, grids);
                    L.push_back(l);
                }
            }
            // block matching, grouped by kernel variant and block size (the
            // variant: the distance for the FP32 matcher, the spatial search
            // for the mixed one)
            // (the two-launch strip searches as a spatial launch, phase 1, then
            // a temporal one, phase 2, each over all such jobs)
            std::map<std::tuple<int, int, int, int, int>, std::vector<int>> mgroups;
            for (int i : js) {
                const bm3d_plan* p = b->jobs[i].plan[pass];
                const int keys = p->use_dx ? bm3d::plan_temporal_keys(p) : 0;
                const int variant = p->use_dx ? bm3d::plan_dx_variant(p) : p->d.bm_error;
                if (p->use_dx && bm3d::plan_uses_strip(p) && p->d.radius > 0) {
                    mgroups[{1, bm3d::kPhSpatial, 0, p->tbw * p->tby, variant}].push_back(i);
                    mgroups[{1, bm3d::kPhTemporal, keys, p->tbw * p->tby, bm3d::kSpLane}].push_back(i);
                } else {
                    mgroups[{p->use_dx ? 1 : 0, bm3d::kPhBoth, keys, p->tbw * p->tby, variant}].push_back(i);
                }
            }
            for (auto& [key, ids] : mgroups) {
                Launch l{};
                l.kind = std::get<0>(key) ? 2 : 1;
                l.pass = pass;
                l.phase = std::get<1>(key);
                l.block = std::get<3>(key);
                l.variant = std::get<4>(key);
                std::vector<bm3d::MatchArgs> args;
                std::vector<bm3d::JobGrid> grids;
                for (int i : ids) {
                    const Job& j = b->jobs[i];
                    const bm3d_plan* p = j.plan[pass];
                    const bm3d_clip ref = ref_clip(j);
                    bm3d::MatchArgs a = bm3d::plan_match_args(p, &j.job.src, &ref);
                    a.groups = reinterpret_cast<uint32_t*>(base + j.groups_off);
                    a.first_frame = 0;
                    a.phase = l.phase;
                    if (l.phase != bm3d::kPhBoth) a.slist = reinterpret_cast<uint32_t*>(base + j.slist_off);
                    if (l.phase == bm3d::kPhTemporal) {  // one lane per block
                        a.strip = 0;
                        a.obw = a.tbw;
                    }
                    const int obw = a.obw;
                    const int gx = (p->nbx + obw - 1) / obw, gy = (p->nby + p->tby - 1) / p->tby;
                    grids.push_back({l.wgs, gx, gy});
                    l.wgs += gx * gy * j.job.src.num_frames;
                    l.lds = std::max(l.lds, p->use_dx ? p->dx_lds_bytes : p->lds_bytes);
                    l.keys = bm3d::plan_temporal_keys(p);
                    args.push_back(a);
                    l.jobs.push_back(i);
                }
                l.args_off = table_push(tab, args);
                l.grids_off = table_push(tab, grids);
                L.push_back(l);
            }
            // filtering, grouped by kernel variant and waves
            std::map<std::tuple<int, int, int, int>, std::vector<int>> fgroups;
            for (int i : js) {
                const bm3d_plan* p = b->jobs[i].plan[pass];
                const int waves = std::min(p->tune.filter_waves,
                                           p->use_mx_filter ? bm3d::kMaxMxFilterWaves : bm3d::kMaxFilterWaves);
                // the FP32 filter has a variant for transforms other than the DCT
                const int generic = !p->use_mx_filter && (p->d.transform_2d || p->d.transform_1d);
                fgroups[{p->use_mx_filter ? 1 : 0, p->d.chroma, waves, generic}].push_back(i);
            }
            for (auto& [key, ids] : fgroups) {
                Launch l{};
                l.kind = std::get<0>(key) ? 4 : 3;
                l.pass = pass;
                l.chroma = std::get<1>(key);
                l.final_ = pass;
                l.block = std::get<2>(key);
                l.variant = std::get<3>(key);
                std::vector<bm3d::FilterArgs> args;
                std::vector<bm3d::JobGrid> grids;
                for (int i : ids) {
                    const Job& j = b->jobs[i];
                    const bm3d_plan* p = j.plan[pass];
                    const bm3d_clip ref = ref_clip(j);
                    bm3d::FilterArgs a = bm3d::plan_filter_args(p, &j.job.src, &ref);
                    const int N = j.job.src.num_frames;
                    a.groups = reinterpret_cast<const uint32_t*>(base + j.groups_off);
                    a.corner = reinterpret_cast<float*>(base + j.corner_off);
                    a.scratch = reinterpret_cast<int*>(base + j.scratch_off);
                    a.fx_range = fx_range(j);
                    a.range = range(j);
                    a.wrap = reinterpret_cast<int*>(base + j.wrap_off);
                    a.wrap_flag = wrap_flag(j);
                    a.res = nullptr;
                    a.mode = 1;
                    a.eval_lo = 0;
                    a.eval_hi = N;
                    a.chunk = N;
                    a.ring = N;
                    a.target_stride = p->region_ints;
                    a.corner_frame_stride = (long long)p->nplanes * p->plane_size;
                    grids.push_back({l.wgs, p->fntx, p->fnty});
                    l.wgs += p->fntx * p->fnty;
                    l.lds = std::max(l.lds, p->filter_lds_bytes);
                    l.any_extractor |= p->d.extractor_exp != 0;
                    args.push_back(a);
                    l.jobs.push_back(i);
                }
                l.args_off = table_push(tab, args);
                l.grids_off = table_push(tab, grids);
                L.push_back(l);
            }
            // VAggregate from the output-frame regions
            {
                Launch l{};
                l.kind = 5;
                l.pass = pass;
                std::vector<bm3d::VTargetArgs> args;
                std::vector<bm3d::JobGrid> grids;
                for (int i : js) {
                    const Job& j = b->jobs[i];
                    const bm3d_plan* p = j.plan[pass];
                    const int N = j.job.src.num_frames;
                    bm3d::VTargetArgs a = bm3d::plan_vtarget_args(p);
                    a.scratch = reinterpret_cast<const int*>(base + j.scratch_off);
                    a.corner = reinterpret_cast<const float*>(base + j.corner_off);
                    a.ring = N;
                    a.fx_range = fx_range(j);
                    a.wrap = reinterpret_cast<const int*>(base + j.wrap_off);
                    a.wrap_flag = wrap_flag(j);
                    a.plane_mask = j.process;
                    a.dst = (passes(j) == 2 && pass == 0) ? j.job.basic : j.job.out;
                    a.dst_frame_stride = j.job.src.frame_stride;
                    a.dst_plane_stride = j.job.src.plane_stride;
                    a.dst_stride = j.job.desc.stride;
                    a.first_frame = 0;
                    const int gx = (a.W + 63) / 64, gy = (a.H + bm3d::kVaggRows - 1) / bm3d::kVaggRows;
                    grids.push_back({l.wgs, gx, gy});
                    l.wgs += gx * gy * N * j.nplanes;
                    args.push_back(a);
                    l.jobs.push_back(i);
                }
                l.args_off = table_push(tab, args);
                l.grids_off = table_push(tab, grids);
                L.push_back(l);
            }
        }
        b->launches += (int)b->wave_launches[w].size();
    }
    return BM3D_OK;
}

hipError_t enqueue(const bm3d_batch* b, const Launch& l, hipStream_t s) {
    const char* t = b->tables;
    const bm3d::JobGrid* grids = reinterpret_cast<const bm3d::JobGrid*>(t + l.grids_off);
    const int n = (int)l.jobs.size();
    switch (l.kind) {
        case 0:
            return bm3d::launch_absmax_multi(reinterpret_cast<const bm3d::AbsmaxArgs*>(t + l.args_off), grids, n,
                                             l.wgs, s);
        case 1:
            return bm3d::launch_match_multi(reinterpret_cast<const bm3d::MatchArgs*>(t + l.args_off), grids, n,
                                            l.wgs, l.block, l.lds, l.variant, s);
        case 2:
            return bm3d::launch_match_dx_multi(reinterpret_cast<const bm3d::MatchArgs*>(t + l.args_off), grids, n,
                                               l.wgs, l.block, l.lds, l.keys, l.variant, s, l.phase);
        case 3:
            return bm3d::launch_filter_multi(reinterpret_cast<const bm3d::FilterArgs*>(t + l.args_off), grids, n,
                                             l.wgs, l.final_ != 0, l.chroma != 0, l.variant != 0, l.block, l.lds, s);
        case 4:
            return bm3d::launch_filter_mx_multi(reinterpret_cast<const bm3d::FilterArgs*>(t + l.args_off), grids,
                                                n, l.wgs, l.final_ != 0, l.chroma != 0, l.block, l.lds, s,
                                                l.any_extractor != 0);
        default:
            return bm3d::launch_vagg_targets_multi(reinterpret_cast<const bm3d::VTargetArgs*>(t + l.args_off),
                                                   grids, n, l.wgs, s);
    }
}

// Copies src planes that BM3Dv2 does not process into the outputs of a pass.
hipError_t copy_unprocessed(const Job& j, int pass, hipStream_t s) {
    const bm3d_batch_job& J = j.job;
    float* dst = (J.passes == 2 && pass == 0) ? J.basic : J.out;
    const size_t row_bytes = (size_t)J.desc.height * J.desc.stride * sizeof(float);
    const size_t pitch = (size_t)J.src.frame_stride * sizeof(float);
    for (int p = 0; p < j.nplanes; ++p) {
        if ((j.process >> p) & 1u) continue;
        hipError_t e = hipMemcpy2DAsync(dst + (size_t)p * J.src.plane_stride, pitch,
                                        J.src.data + (size_t)p * J.src.plane_stride, pitch, row_bytes,
                                        J.src.num_frames, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Zero kernel over a byte range (a float fill of zeros; no memset node, so
// the captured graph replays it faithfully).
hipError_t zero_bytes(char* p, size_t bytes, hipStream_t s) {
    return bm3d::launch_zero(reinterpret_cast<float*>(p), bytes / sizeof(float), 0, 1, s);
}

// (Re)builds scratch, tables and the graph for the jobs' current plans.
int rebuild(bm3d_batch* b) {
    release_graph(b);
    layout(b);
    if (b->arena) (void)hipFree(b->arena);
    b->arena = nullptr;
    BATCH_HIP(bm3d::malloc_device(&b->arena, std::max<size_t>(b->arena_bytes, 256)));
    std::vector<char> tab;
    if (int rc = build_launches(b, tab)) return rc;
    b->tables_bytes = std::max<size_t>(tab.size(), 256);
    BATCH_HIP(bm3d::malloc_device(&b->tables, b->tables_bytes));
    if (!tab.empty()) BATCH_HIP(hipMemcpy(b->tables, tab.data(), tab.size(), hipMemcpyHostToDevice));

    hipStream_t cs = nullptr;
    BATCH_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    for (int w = 0; w < b->nwaves && e == hipSuccess; ++w) {
        for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
            bool any = false;
            for (const Launch& l : b->wave_launches[w]) any |= l.pass == pass;
            if (any) e = zero_bytes(b->arena + b->wave_zero[w].first, b->wave_zero[w].second, cs);
            for (const Launch& l : b->wave_launches[w])
                if (l.pass == pass && e == hipSuccess) e = enqueue(b, l, cs);
            for (const Job& j : b->jobs)
                if (j.wave == w && passes(j) > pass && e == hipSuccess) e = copy_unprocessed(j, pass, cs);
        }
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &g);
    if (e == hipSuccess) e = ec;
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_err(e, "batch graph capture");
    }
    b->graph = g;
    BATCH_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return BM3D_OK;
}

int apply_to_jobs(bm3d_batch* b, const bm3d_tuning& t) {
    for (Job& j : b->jobs)
        for (bm3d_plan* p : j.plan)
            if (p)
                if (int rc = bm3d::plan_apply_tuning(p, t)) return rc;
    for (Job& j : b->jobs)
        for (bm3d_plan* p : j.plan)
            if (p && (p->frm < p->margin || p->lds_bytes == 0))
                return fail(BM3D_EUNSUPPORTED, "BM3D: search reach too large for a batch job's LDS tiles "
                                               "(use per-clip plans)");
    b->tune = t;
    return BM3D_OK;
}

}  // namespace

extern "C" {

int bm3d_batch_create(const bm3d_batch_job* jobs, int njobs, long long scratch_budget, bm3d_batch** out) {
    if (!jobs || !out || njobs <= 0) return fail(BM3D_EINVAL, "BM3D: batch needs jobs");
    if (scratch_budget < 0) return fail(BM3D_EINVAL, "BM3D: negative scratch budget");
    *out = nullptr;
    auto* b = new bm3d_batch;
    if (scratch_budget > 0) b->wave_budget = scratch_budget;
    int rc = BM3D_OK;
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) rc = hip_err(e, "hipGetDevice");
    for (int i = 0; i < njobs && rc == BM3D_OK; ++i) {
        const bm3d_batch_job& J = jobs[i];
        Job j;
        j.job = J;
        if (J.passes != 1 && J.passes != 2) rc = fail(BM3D_EINVAL, "BM3D: batch job passes must be 1 or 2");
        else if (!J.src.data || !J.out || (J.passes == 2 && !J.basic))
            rc = fail(BM3D_EINVAL, "BM3D: batch job buffer missing");
        else if (J.src.num_frames <= 0) rc = fail(BM3D_EINVAL, "BM3D: batch job has no frames");
        if (rc) break;
        for (int pass = 0; pass < J.passes && rc == BM3D_OK; ++pass) {
            bm3d_desc d = J.desc;
            d.final_ = pass;
            rc = bm3d_plan_create(&d, &j.plan[pass]);
        }
        if (rc == BM3D_OK) {
            const bm3d_plan* p = j.plan[0];
            j.nplanes = p->nplanes;
            if (J.src.plane_stride < p->plane_size && p->nplanes > 1)
                rc = fail(BM3D_EINVAL, "BM3D: batch job planes overlap");
            // planes BM3Dv2 processes: sigma >= FLT_EPSILON (source/source.cpp:916-930)
            for (int q = 0; q < p->nplanes; ++q)
                if (!(J.desc.sigma[q] < std::numeric_limits<float>::epsilon())) j.process |= 1u << q;
        }
        b->jobs.push_back(j);
    }
    if (rc == BM3D_OK) {
        b->tune = b->jobs[0].plan[0]->tune;
        rc = apply_to_jobs(b, b->tune);
    }
    if (rc == BM3D_OK) rc = rebuild(b);
    if (rc != BM3D_OK) {
        const std::string msg = bm3d_last_error();
        bm3d_batch_destroy(b);
        return fail(rc, msg);
    }
    *out = b;
    return BM3D_OK;
}

int bm3d_batch_execute(bm3d_batch* b, void* stream) {
    if (!b || !b->exec) return fail(BM3D_EINVAL, "BM3D: null batch");
    int dev = 0;
    BATCH_HIP(hipGetDevice(&dev));
    if (dev != b->device) return fail(BM3D_EINVAL, "BM3D: batch used on a different device than it was created on");
    BATCH_HIP(hipGraphLaunch(b->exec, reinterpret_cast<hipStream_t>(stream)));
    return BM3D_OK;
}

int bm3d_batch_get_tuning(const bm3d_batch* b, bm3d_tuning* t) {
    if (!b || !t) return fail(BM3D_EINVAL, "BM3D: null argument");
    *t = b->tune;
    return BM3D_OK;
}

int bm3d_batch_set_tuning(bm3d_batch* b, const bm3d_tuning* t) {
    if (!b || !t) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (const char* why = bm3d::plan_tuning_error(*t)) return fail(BM3D_EINVAL, std::string("BM3D: tuning: ") + why);
    const bm3d_tuning old = b->tune;
    int rc = apply_to_jobs(b, *t);
    if (rc != BM3D_OK) {
        const std::string msg = bm3d_last_error();
        (void)apply_to_jobs(b, old);
        (void)rebuild(b);
        return fail(rc, msg);
    }
    return rebuild(b);
}

int bm3d_batch_autotune(bm3d_batch* b, void* stream_) {
    if (!b) return fail(BM3D_EINVAL, "BM3D: null batch");
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BATCH_HIP(hipEventCreate(&e0));
    BATCH_HIP(hipEventCreate(&e1));
    auto time_one = [&](const bm3d_tuning& t) -> double {
        if (bm3d::plan_tuning_error(t) || bm3d_batch_set_tuning(b, &t) != BM3D_OK) return -1.0;
        if (bm3d_batch_execute(b, stream_) != BM3D_OK) return -1.0;
        float best = -1.0f;
        for (int rep = 0; rep < 2; ++rep) {
            if (hipEventRecord(e0, stream) != hipSuccess || bm3d_batch_execute(b, stream_) != BM3D_OK ||
                hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                return -1.0;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
            best = best < 0.0f ? ms : std::min(best, ms);
        }
        return best;
    };
    bm3d_tuning best = b->tune;
    double best_ms = time_one(best);
    int rc = BM3D_OK;
    if (best_ms < 0.0) {
        rc = fail(BM3D_EHIP, std::string("BM3D: batch autotune: the current launch shapes failed: ") +
                                 bm3d_last_error());
    } else {
        auto consider = [&](const bm3d_tuning& t) {
            const double ms = time_one(t);
            if (ms >= 0.0 && ms < best_ms) {
                best_ms = ms;
                best = t;
            }
        };
        const int shapes[][2] = {{16, 16}, {16, 8}, {16, 4}, {32, 8}, {32, 4}};
        for (auto& sh : shapes) {
            bm3d_tuning t = best;
            t.match_width = sh[0];
            t.match_rows = sh[1];
            consider(t);
        }
        const int tiles[][3] = {{8, 8, 4}, {8, 8, 8}, {16, 8, 4}, {16, 8, 8}, {8, 16, 8}, {16, 16, 8}, {8, 4, 4}};
        const bm3d_tuning base = best;
        for (auto& f : tiles) {
            bm3d_tuning t = base;
            t.filter_tile_x = f[0];
            t.filter_tile_y = f[1];
            t.filter_waves = f[2];
            consider(t);
        }
        rc = bm3d_batch_set_tuning(b, &best);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int bm3d_batch_info(const bm3d_batch* b, int* waves, int* launches, long long* scratch_bytes) {
    if (!b) return fail(BM3D_EINVAL, "BM3D: null batch");
    if (waves) *waves = b->nwaves;
    if (launches) *launches = b->launches;
    if (scratch_bytes) *scratch_bytes = (long long)b->arena_bytes;
    return BM3D_OK;
}

void bm3d_batch_destroy(bm3d_batch* b) {
    if (!b) return;
    release_graph(b);
    if (b->arena) (void)hipFree(b->arena);
    for (Job& j : b->jobs)
        for (bm3d_plan* p : j.plan)
            if (p) bm3d_plan_destroy(p);
    delete b;
}

}  // extern "C"
