// Shared definitions of the host-side translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <memory>
#include <string>
#include <vector>

#include "bm3d_kernels.h"
#include "bm3d_mi355x.h"

namespace bm3d {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);
// Clears the calling thread's last error (after a handled failure).
void clear_error();

}  // namespace bm3d

struct bm3d_plan {
    bm3d_desc d{};
    float sigma[3]{};   // scaled thresholds
    float extractor = 0.0f;
    int nplanes = 1, tw = 1;
    int nbx = 0, nby = 0, nblocks = 0;
    int margin = 0;     // search reach: bm_range + radius * ps_range
    long long plane_size = 0, stack_size = 0;
    int device = 0;
    int max_lds = 0;
    bm3d_tuning tune{};
    // matching work-group tile: tby rows of tbw reference blocks plus the margin
    int tbw = 16, tby = 16, tile_w = 0, tile_h = 0, tile_wq = 0, tile_hp = 0;
    size_t lds_bytes = 0;  // 0: the union window exceeds the LDS; the matcher reads the planes
    // filter work-group tiling and LDS accumulation region (margin frm)
    int ftbx = 8, ftby = 8, fntx = 0, fnty = 0, frm = 0, frw = 0, frh = 0;
    size_t filter_lds_bytes = 0;
    long long region_ints = 0;  // one region per tile and plane, all tiles: per slot or output frame
    // device buffers, capacities in elements
    uint32_t* groups = nullptr;
    size_t groups_cap = 0;
    float* corner = nullptr;
    size_t corner_cap = 0;
    int* scratch = nullptr;
    size_t scratch_cap = 0;
    // fixed-point unit of the accumulation regions: 2^fx_exp, scaled by the
    // input's range (measured into *range per call) unless extractor_exp is set
    int fx_exp = 0;
    float* range = nullptr;
    // wrap counts of the int32 region words (corner-map layout, corner_cap
    // entries) and the flag readers test before reading them: the int just
    // after *range, zero at creation and set for good once a wrap is counted
    // (from then on the counts are zeroed with the corner maps)
    int* wrap = nullptr;
    int* wrap_flag = nullptr;
    // bm3d_execute_vagg: 64-bit per-pixel sums of output frames whose
    // contributing evaluations span several runs (corner-map layout)
    long long* part = nullptr;
    size_t part_cap = 0;
    // |one weighted estimate| <= vbound * (largest |sample|)
    float vbound = 0.0f;
    // set by a call made while its stream was capturing: the captured graph
    // holds the plan's buffers and tiles, so they may no longer change
    bool pinned = false;
    // mixed-precision matching kernel (precision MIXED)
    bool use_dx = false, dx_ok = false;
    bool mx_spatial = false;  // the mixed matcher's spatial search on the matrix cores (where eligible; opt-in)
    // the mixed matcher's spatial search by strips shared between lanes where
    // the block step and tile allow it (strip_supported)
    bool strip = false;
    // at radius > 0 the strip search runs as two launches (spatial by strips,
    // then temporal, one lane per block) handing the spatial best 8 over in
    // slist (16 words per block and frame, slist_cap words)
    uint32_t* slist = nullptr;
    size_t slist_cap = 0;
    // kMatchTableBytes of device memory after the range (launch_match_dx_table)
    void* match_table = nullptr;
    int tile_hp2 = 0;
    size_t dx_lds_bytes = 0;
    // mixed-precision filter kernel (precision MIXED): per-lane transform operands
    bool use_mx_filter = false;
    void* mx_ops = nullptr;
    // bm3d_execute_vagg: a second stream for the work matching does not wait
    // for (input range, zero fills), forked from and joined back into the
    // caller's stream, so that it overlaps the first matching launch
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

namespace bm3d {

// Plan internals for the batch launcher (bm3d_batch.cpp): apply a tuning
// (geometry, LDS limits); kernel arguments over the plan's geometry and
// parameters, with the plan's own buffers -- callers replace the buffer
// pointers (groups, corner, scratch, fixed-point range) with their own.
int plan_apply_tuning(bm3d_plan* p, const bm3d_tuning& t);
const char* plan_tuning_error(const bm3d_tuning& t);
MatchArgs plan_match_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref);
// Keys the mixed matcher keeps per temporal frame: desc.mixed_temporal_keys,
// or dx_temporal_keys(ps_num) when that is 0.
int plan_temporal_keys(const bm3d_plan* p);
// Whether the plan's matching uses the strip search (at radius > 0 as two
// launches, spatial then temporal); whether a batch's matching arguments
// describe it (as the plan's, with the same lists); the blocks a batch's
// matching work-group row owns (tbw without strips); the batch's
// dx_spatial_variant of the plan. plan_match_args describes the batch's
// launches (the batch clears the strip fields for a temporal launch).
bool plan_uses_strip(const bm3d_plan* p);
bool plan_batch_strip(const bm3d_plan* p);
int plan_match_owned_width(const bm3d_plan* p);
int plan_dx_variant(const bm3d_plan* p);
FilterArgs plan_filter_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref);
VTargetArgs plan_vtarget_args(const bm3d_plan* p);

}  // namespace bm3d

namespace bm3d {

enum class Kind { kBM3D, kVAggregate, kBM3Dv2 };

// The arguments a filter was created from, kept so that the host-frame path
// can create independent copies of the filter (one per resource set).
struct FilterRecipe {
    Kind kind = Kind::kBM3D;
    bm3d_args args{};              // string pointers re-aimed at `strings`
    std::vector<std::string> strings;
    bm3d_video_info clip{}, ref{};
    bool has_ref = false;
    std::vector<long long> planes;  // VAggregate
};

// Host-frame state of a filter (bm3d_host.cpp).
struct HostPool;
void destroy_host_pool(HostPool* pool);

}  // namespace bm3d

// A plugin-level filter (bm3d_filters.cpp): BM3D, VAggregate or BM3Dv2.
struct bm3d_filter {
    bm3d::Kind kind = bm3d::Kind::kBM3D;
    bm3d_video_info vi{};      // input clip
    bm3d_video_info out_vi{};  // output clip
    int num_frames = 0;
    // BM3D
    bool final_ = false;
    bool chroma = false;
    bool zero_init = true;
    bool fast = true;          // 4 resource sets for host frames (source/source.cpp:636-639)
    int radius = 0;
    int extractor_exp = 0;
    int precision = BM3D_PRECISION_FP32;
    int device = 0;
    bool process[3]{};
    float sigma[3]{};
    int block_step[3]{}, bm_range[3]{}, ps_num[3]{}, ps_range[3]{};
    int bm_error[3]{}, transform_2d[3]{}, transform_1d[3]{};  // BM3D_BM_*, BM3D_TRANSFORM_*
    bm3d_plan* plans[3]{};
    int plan_stride[3]{};
    float* scratch = nullptr;
    size_t scratch_cap = 0;
    // BM3Dv2
    std::unique_ptr<bm3d_filter> bm3d;  // the BM3D stage
    std::unique_ptr<bm3d_filter> vagg;  // the VAggregate stage (radius > 0)
    bool passthrough = false;
    // VAggregate: planes aggregated, radius from the stack height
    bool vprocess[3]{};
    // creation arguments and host-frame resource sets
    bm3d::FilterRecipe recipe;
    bm3d::HostPool* host = nullptr;

    ~bm3d_filter() {
        bm3d::destroy_host_pool(host);
        for (auto* p : plans) bm3d_plan_destroy(p);
        if (scratch) (void)hipFree(scratch);
    }
};

namespace bm3d {

// Creates a filter like `f` from its recipe (same kind, arguments, clips).
int clone_filter(const bm3d_filter* f, bm3d_filter** out);

}  // namespace bm3d
