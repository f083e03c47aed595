// Kernel argument blocks and launchers shared by bm3d_kernels.hip and the
// launch layer (bm3d_capi.cpp). Internal: not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace bm3d {

// Packed block position: x in bits 0-13, y in bits 14-26, slot z in 27-31.
constexpr int kPosYShift = 14;
constexpr int kPosZShift = 27;
constexpr uint32_t kPosXMask = (1u << 14) - 1;
constexpr uint32_t kPosYMask = (1u << 13) - 1;
constexpr int kMaxWidth = 1 << 14;
constexpr int kMaxHeight = 1 << 13;
constexpr int kMaxRadius = 15;

// Largest temporal width 2r + 1 (the filters' per-slot tables).
constexpr int kMaxSlots = 2 * kMaxRadius + 1;

// Multi-job launches (one launch over several clips' work): a 1-D grid
// whose work-groups are the concatenated grids of the jobs. Job j's grid is
// gx x gy x gz work-groups starting at flat index `begin`; entries are
// sorted by begin and the launch has jobs[n-1].begin + its size work-groups.
struct JobGrid {
    int begin;
    int gx, gy;
};

// Block distances (BM3D_BM_*) and 8-point transforms (BM3D_TRANSFORM_*).
constexpr int kNumCosts = 5;
constexpr int kNumTransforms = 4;

struct MatchArgs {
    const float* clip;       // plane used for matching, frame 0
    long long frame_stride;  // floats between frames
    int num_frames;          // frames in the clip (slot indices clamp to it)
    int first_frame;         // output frame of grid z = 0
    int W, H, stride;
    int bs, bm_range, radius, ps_num, ps_range;
    int nbx, nby, nblocks;
    int margin;              // bm_range + radius * ps_range
    int tbw;                 // work-group width in reference blocks (16 or 32)
    int tile_w, tile_h, tile_wq, tile_hp;
    int tile_hp2;            // mixed kernel: FP16 pair words per tile column
    uint32_t* groups;        // [frames][nblocks][8] (may be null when counting)
    unsigned long long* ncand;  // non-null: count in-plane candidates
    int cost;                // block distance, 0 .. kNumCosts - 1 (the mixed kernel: 0 only)
    int kt;                  // mixed kernel: keys kept per temporal frame (4, 8 or 16)
    int mx_spatial;          // mixed kernel: spatial search on the matrix cores where eligible
    int strip;               // mixed kernel: spatial search by strips shared between lanes (bs 2 or 4)
    int obw;                 // blocks owned per work-group row: tbw, or strip_owned_width with strips
    // mixed kernel at radius > 0: the whole search in one launch (0), or its
    // spatial part (1, strips only) then its temporal part (2) in two launches
    // handing the spatial best 8 over in slist, [frames][nblocks][16]
    // (8 positions, then the bits of their 8 distances)
    int phase;
    uint32_t* slist;
};
enum { kPhBoth = 0, kPhSpatial, kPhTemporal };

// Blocks a work-group row of tbw lanes owns with the strip search: the last
// 8 / bs - 1 lanes hold only strips of the row's last blocks.
inline int strip_owned_width(int tbw, int bs) { return tbw - 8 / bs + 1; }
// The strip search serves block steps 4 (rows of 16 or 32 lanes) and 2 (rows
// of 16).
inline bool strip_supported(int tbw, int bs) { return bs == 4 || (bs == 2 && tbw == 16); }

struct FilterArgs {
    const float* src;        // noisy clip, frame 0 plane 0
    const float* ref;        // basic-estimate clip (final pass), same layout
    long long frame_stride, plane_stride;
    int num_frames, first_frame;
    int W, H, stride;
    long long plane_size;    // H * stride
    int radius;
    int nblocks;
    float sigma[3];          // scaled thresholds per plane
    float extractor;         // 2^extractor_exp or 0
    const uint32_t* groups;  // [frames][nblocks][8]
    float* res;              // per frame: [planes][2r+1][2][H][stride]
    long long res_frame_stride, res_plane_stride;
    float* corner;           // [frames][planes][2r+1][H][stride]
    long long corner_frame_stride;
    int bs, nbx, nby;
    int tbx, tby, ntx, nty;  // work-group tile of reference blocks, tile grid
    int rm, rw, rh;          // accumulation region: margin, width, height
    int* scratch;            // fixed-point regions (layout by mode, below)
    const float* fx_range;   // fixed-point unit: see fx_exponent (bm3d_device.h)
    int fx_exp;
    int* wrap;               // wrap counts of the region words, corner-map layout (bm3d_device.h)
    int* wrap_flag;          // set when any wrap was counted
    const float* range;      // largest finite |sample| of the call's input (always measured)
    float vbound;            // |one weighted estimate| <= vbound * range (contribution_bound)
    const void* mx_ops;      // mixed filter: per-lane transform operands (build_filter_mx_ops)
    int t2d, t1d;            // FP32 filter: transform kinds along rows / columns and members
    // Accumulation mode.
    //  0 (slots): one evaluation per work-group (frame first_frame + z of the
    //    grid); region of slot z of evaluation f at scratch
    //    [f][nty][ntx][planes][2r+1][rh][rw], weights at corner
    //    [f][planes][2r+1][H][stride] (corner_frame_stride floats per f).
    //  1 (targets): a work-group walks evaluations eval_lo + chunk * z ..
    //    (at most `chunk`, below eval_hi) of its tile, and sums slot z of
    //    evaluation m, `multiplicity` times, into output frame
    //    t = clamp(m - r + z, 0, num_frames - 1) -- the VAggregate sum of
    //    source/source.cpp:816-846 regrouped by the frame each slot covers.
    //    Region of frame t at scratch + (t % ring) * target_stride +
    //    [nty][ntx][planes][rh][rw]; weights at corner + (t % ring) *
    //    corner_frame_stride + [planes][H][stride]. A region whose
    //    contributing evaluations all lie in the work-group's chunk is
    //    stored; the regions of any other frame are added per pixel to
    //    part + (t % ring) * corner_frame_stride + [planes][H][stride]
    //    (64-bit, zeroed by the caller). Group lists of evaluation m at
    //    index m - eval_lo.
    int mode;
    long long* part;
    int eval_lo, eval_hi, chunk;
    int ring;
    long long target_stride;
};

struct FinalizeArgs {
    const int* scratch;
    const float* fx_range;   // fixed-point unit: see fx_exponent (bm3d_device.h)
    int fx_exp;
    const int* wrap;         // wrap counts, corner-map layout; read when *wrap_flag
    const int* wrap_flag;
    long long corner_frame_stride;
    float* res;
    long long res_frame_stride, res_plane_stride;
    int W, H, stride;
    long long plane_size;
    int nplanes, slots;
    int bs, tbx, tby, ntx, nty, rm, rw, rh;
    int accumulate;          // add to the stack's existing values
    int chroma;
    float sigma[3];
};

struct DenArgs {
    const float* corner;
    long long corner_frame_stride;
    float* res;
    long long res_frame_stride, res_plane_stride;
    int W, H, stride;
    long long plane_size;
    int nplanes, slots;
};

struct DivideArgs {
    const float* res;        // r = 0 stacks [frames][planes][1][2][H][stride]
    long long res_frame_stride, res_plane_stride;
    float* dst;
    long long dst_frame_stride, dst_plane_stride;
    int dst_stride;
    int W, H, stride;
    long long plane_size;
    int nplanes;
};

// Output rows per work-group of bm3d_vagg_targets (64 columns wide), and
// output frames per launch (VTargetArgs::partial).
constexpr int kVaggRows = 32;
constexpr int kVaggMaxFrames = 256;

// VAggregate from target-mode accumulation (bm3d_vagg_targets): output
// frame n = the regions' integer sum of frame n (times the unit) over the
// 8 x 8 box sum of its corner weight map. The stacks are never written.
struct VTargetArgs {
    const int* scratch;      // [ring][nty][ntx][planes][rh][rw]
    long long target_stride; // ints per frame
    const float* corner;     // [ring][planes][H][stride]
    long long corner_frame;  // floats per frame
    int ring;                // frame n at index n % ring
    const float* fx_range;   // fixed-point unit: see fx_exponent (bm3d_device.h)
    int fx_exp;
    const int* wrap;         // wrap counts, corner-map layout; read when *wrap_flag
    const int* wrap_flag;
    // frames first_frame + f whose contributors spanned several runs (bit f,
    // f < 256): their sums are the 64-bit per-pixel values at part + (frame %
    // ring) * corner_frame + [planes][H][stride], not the regions
    const long long* part;
    unsigned partial[8];
    float* dst;              // output frames [first_frame, first_frame + count)
    long long dst_frame_stride, dst_plane_stride;
    int dst_stride;
    int W, H, stride;
    long long plane_size;
    int nplanes;             // planes of the evaluation (1, or 3 with chroma)
    unsigned plane_mask;     // planes to produce (bit p)
    int first_frame;
    int bs, tbx, tby, ntx, nty, rm, rw, rh;
};

struct VAggArgs {
    const float* stacks;     // stacks of frames [stack_first, ...)
    long long stack_frame_stride, stack_plane_stride;
    int stack_first;
    float* dst;              // output frames [first_frame, first_frame + count)
    long long dst_frame_stride, dst_plane_stride;
    int dst_stride;
    int W, H, stride;
    long long plane_size;
    int nplanes;             // planes in each stack (processed planes only)
    int radius, num_frames, first_frame;
};

// Block size tbw x tby (a.tbw blocks wide, tby rows).
hipError_t launch_match(const MatchArgs& a, int tby, size_t lds_bytes, int frames, hipStream_t s);
// Multi-job forms: `jobs` and `grids` are device arrays of njobs entries;
// `wgs` work-groups in all; every job shares the launch's block size
// (threads: tbw x tby for matching, 64 x waves for filtering) and kernel
// variant; lds_bytes covers the largest job.
hipError_t launch_match_multi(const MatchArgs* jobs, const JobGrid* grids, int njobs, int wgs, int threads,
                              size_t lds_bytes, int cost, hipStream_t s);
// (every job of one launch has the same MatchArgs::kt and phase, and
// `variant` is dx_spatial_variant of each job)
hipError_t launch_match_dx_multi(const MatchArgs* jobs, const JobGrid* grids, int njobs, int wgs, int threads,
                                 size_t lds_bytes, int kt, int variant, hipStream_t s, int phase = 0);
// generic: the FP32 filter variant for transform kinds other than the DCT.
hipError_t launch_filter_multi(const FilterArgs* jobs, const JobGrid* grids, int njobs, int wgs, bool final_,
                               bool chroma, bool generic, int waves, size_t lds_bytes, hipStream_t s);
// any_extractor: some job has extractor_exp set (launches the 64-bit-add
// variant beside the other; bm3d_filter_mx.hip).
hipError_t launch_filter_mx_multi(const FilterArgs* jobs, const JobGrid* grids, int njobs, int wgs, bool final_,
                                  bool chroma, int waves, size_t lds_bytes, hipStream_t s, bool any_extractor);
hipError_t launch_vagg_targets_multi(const VTargetArgs* jobs, const JobGrid* grids, int njobs, int wgs,
                                     hipStream_t s);
// Largest |sample| per job: grid of job j = (min(H, 64), frames * nplanes).
struct AbsmaxArgs {
    const float* clip;
    long long frame_stride, plane_stride;
    int nplanes, W, H, stride;
    float* out;
};
hipError_t launch_absmax_multi(const AbsmaxArgs* jobs, const JobGrid* grids, int njobs, int wgs, hipStream_t s);
hipError_t set_match_lds_limit(size_t bytes);
// Mixed-precision matching (bm3d_match_dx.hip): same launch geometry as
// launch_match; LDS bytes as returned by dx_lds_bytes.
hipError_t launch_match_dx(const MatchArgs& a, int tby, size_t lds_bytes, int frames, hipStream_t s);
hipError_t set_match_dx_lds_limit(size_t bytes);
// launch_match_dx through the multi-job kernels, one job whose arguments a
// small kernel first stores in `table` (device memory of kMatchTableBytes):
// reading the arguments from memory instead of the kernel-argument segment,
// the compiler keeps fewer of them in scalar registers (the temporal search
// spills 76 instead of 172 B per lane). Not for counting launches.
constexpr size_t kMatchTableGridOffset = (sizeof(MatchArgs) + 63) / 64 * 64;
constexpr size_t kMatchTableBytes = kMatchTableGridOffset + 64;
hipError_t launch_match_dx_table(const MatchArgs& a, int tby, size_t lds_bytes, int frames, void* table,
                                 hipStream_t s);
// Default keys kept per temporal frame: 8 (16 for ps_num above 4).
int dx_temporal_keys(int ps_num);
// The mixed matcher's kernel variants, by spatial search: per-lane scans
// (block steps dividing 8, and the others), the matrix-core search, and the
// strip searches (block step 4 with rows of 16 or 32 lanes, block step 2).
// Each variant keeps the per-lane scan for the tiles its search does not
// cover.
enum { kSpLane = 0, kSpOther, kSpMx, kSpStrip4, kSpStrip4W, kSpStrip2 };
bool dx_block_step_divides_8(int block_step);
int dx_spatial_variant(const MatchArgs& a);
// Grid (ntx, nty, grid_z): grid_z = evaluations (mode 0) or chunks (mode 1);
// `waves` waves per work-group (1..kMaxFilterWaves).
constexpr int kMaxFilterWaves = 4;
constexpr int kMaxMxFilterWaves = 8;
hipError_t launch_filter(const FilterArgs& a, bool final_, bool chroma, int grid_z, int waves, size_t lds_bytes,
                         hipStream_t s);
hipError_t set_filter_lds_limit(size_t bytes);
// Mixed-precision filter (bm3d_filter_mx.hip): same arguments, tiles and
// outputs as launch_filter; `waves` waves per work-group (1..kMaxMxFilterWaves).
constexpr int kMxOpVectors = 17;  // uint4 per lane of FilterArgs::mx_ops
// kMxOpVectors * 64 * 4 words, host; transform kinds t2d (rows and columns)
// and t1d (members).
void build_filter_mx_ops(uint32_t* out, int t2d, int t1d);
hipError_t launch_filter_mx(const FilterArgs& a, bool final_, bool chroma, int grid_z, int waves, size_t lds_bytes,
                            hipStream_t s);
hipError_t set_filter_mx_lds_limit(size_t bytes);
hipError_t launch_finalize(const FinalizeArgs& a, int count, hipStream_t s);
// Zeroes `count` runs of `n` floats, `step` floats apart; when `flag` is
// given, only if *flag is non-zero (read on the device).
hipError_t launch_zero(float* p, size_t n, size_t step, int count, hipStream_t s, const int* flag = nullptr);
// Bound B on one weighted estimate of a group, |w * est| <= B * M for inputs
// of magnitude <= M (hard thresholding, or Wiener when final), for the
// transform pair (t2d, t1d); host.
float contribution_bound(int t2d, int t1d, bool final_);
// Largest finite |sample| of planes 0 .. nplanes-1 of `frames` frames from
// clip, max-ed into *out (bit pattern compare; *out must start at zero).
hipError_t launch_absmax(const float* clip, long long frame_stride, long long plane_stride, int nplanes, int W,
                         int H, int stride, int frames, float* out, hipStream_t s);
hipError_t launch_den_box(const DenArgs& a, int count, hipStream_t s);
hipError_t launch_divide(const DivideArgs& a, int count, hipStream_t s);
hipError_t launch_vaggregate(const VAggArgs& a, int count, hipStream_t s);
hipError_t launch_vagg_targets(const VTargetArgs& a, int count, hipStream_t s);

}  // namespace bm3d
