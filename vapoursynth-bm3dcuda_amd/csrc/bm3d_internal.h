// Shared definitions of the host-side translation units (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <string>

#include "bm3d_kernels.h"
#include "bm3d_mi355x.h"

namespace bm3d {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

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
    size_t lds_bytes = 0;
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
    // mixed-precision matching kernel (precision MIXED)
    bool use_dx = false, dx_ok = false;
    int tile_hp2 = 0;
    size_t dx_lds_bytes = 0;
    // mixed-precision filter kernel (precision MIXED): per-lane transform operands
    bool use_mx_filter = false;
    void* mx_ops = nullptr;
};

namespace bm3d {

// Plan internals for the batch launcher (bm3d_batch.cpp): apply a tuning
// (geometry, LDS limits); kernel arguments over the plan's geometry and
// parameters, with the plan's own buffers -- callers replace the buffer
// pointers (groups, corner, scratch, fixed-point range) with their own.
int plan_apply_tuning(bm3d_plan* p, const bm3d_tuning& t);
const char* plan_tuning_error(const bm3d_tuning& t);
MatchArgs plan_match_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref);
FilterArgs plan_filter_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref);
VTargetArgs plan_vtarget_args(const bm3d_plan* p);

}  // namespace bm3d
