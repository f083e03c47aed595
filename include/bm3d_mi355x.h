/*
 * bm3d_mi355x -- C ABI of the MI355X (gfx950) BM3D / V-BM3D denoiser.
 *
 * Drop-in replacement for the device seam of the bm3dcuda VapourSynth
 * plugin. In the reference, the plugin layer (source/source.cpp) talks to
 * the GPU through one C++ function,
 *
 *   cudaGraphExec_t get_graphexec(float *d_res, float *d_src, float *h_res,
 *       int width, int height, int stride, float sigma, int block_step,
 *       int bm_range, int radius, int ps_num, int ps_range, bool chroma,
 *       float sigma_u, float sigma_v, bool final_, float extractor)
 *                                         (source/kernel.cu:40-53, 759-864)
 *
 * which bakes one plane group's parameters into a graph, plus two host loops
 * around it: the r = 0 division `Aggregation` (source/source.cpp:170-188)
 * and `VAggregateGetFrame` (source/source.cpp:746-862). This header replaces
 * all three with device-resident entry points:
 *
 *   bm3d_plan_create / bm3d_plan_destroy   <- get_graphexec (parameters baked
 *                                             into a plan, per plane group)
 *   bm3d_execute                           <- graph launch (kernel node) for
 *                                             any number of output frames
 *   bm3d_execute_window                    <- graph launch on the reference's
 *                                             packed d_src layout
 *   bm3d_aggregate                         <- Aggregation (r = 0)
 *   bm3d_vaggregate                        <- VAggregateGetFrame (r > 0)
 *
 * and the plugin-level filters of source/source.cpp:457-989 (BM3D,
 * VAggregate, BM3Dv2 argument handling) are provided by bm3d_filter_*.
 *
 * Conventions: plain pointers and sizes; all float buffers are device
 * memory (hipMalloc) unless stated; `stream` is a hipStream_t passed as
 * void* (NULL = default stream); calls are asynchronous on that stream.
 * Every entry point returns BM3D_OK or a BM3D_E* code and records a message
 * readable with bm3d_last_error() on the calling thread. No C++ exception
 * crosses this boundary.
 */
#ifndef BM3D_MI355X_H
#define BM3D_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

#define BM3D_OK 0
#define BM3D_EINVAL 1       /* invalid argument (message says which)       */
#define BM3D_EHIP 2         /* HIP runtime error                            */
#define BM3D_EUNSUPPORTED 3 /* valid for the reference, not supported here  */

#define BM3D_PRECISION_FP32 0  /* all FP32 on VALU                          */
#define BM3D_PRECISION_MIXED 1 /* FP16 block-matching pre-scoring, exact
                                  FP32 rescoring of the best few            */

/* Block distance (the bm_error_s option of the NVRTC variant,
 * rtc_source/kernel.hpp:363-504; parsing rtc_source/source.cpp:854-875). */
#define BM3D_BM_SSD 0       /* sum of squared differences (the default)    */
#define BM3D_BM_SAD 1       /* sum of absolute differences                 */
#define BM3D_BM_ZSSD 2      /* SSD after removing the block means          */
#define BM3D_BM_ZSAD 3      /* SAD after removing the block means          */
#define BM3D_BM_SSD_NORM 4  /* 2 - 2 <c, n> / (|c| |n| + FLT_EPSILON)     */

/* 8-point transforms, each scaled to 4 x its orthonormal form (the
 * transform_2d_s / transform_1d_s options, rtc_source/kernel.hpp:103-361;
 * parsing rtc_source/source.cpp:877-912). */
#define BM3D_TRANSFORM_DCT 0      /* DCT-II / DCT-III (the default)        */
#define BM3D_TRANSFORM_HAAR 1     /* Haar, with the reference's inverse    */
#define BM3D_TRANSFORM_WHT 2      /* Walsh-Hadamard, sequency order        */
#define BM3D_TRANSFORM_BIOR1_5 3  /* biorthogonal spline 1.5               */

typedef struct bm3d_plan bm3d_plan;

/* Parameters of one plane group: one plane, or the three planes of a
 * YUV444 clip with chroma = 1 (block matching on plane 0). Field meanings
 * and defaults follow the BM3D filter arguments (source/source.cpp:
 * 503-629); sigma is in user units and is scaled here exactly as
 * source/source.cpp:520-523. */
typedef struct bm3d_desc {
    int width, height;   /* plane size, >= 8 x 8 */
    int stride;          /* row pitch in floats of every plane buffer   */
    float sigma[3];      /* [0]: plane 0; [1], [2]: chroma planes      */
    int block_step;      /* 1..8 */
    int bm_range;        /* >= 1 */
    int radius;          /* 0 = BM3D, > 0 = V-BM3D */
    int ps_num;          /* 1..8 */
    int ps_range;        /* >= 1 */
    int chroma;          /* 0 or 1 */
    int final_;          /* 0: basic (hard threshold); 1: final (Wiener, needs ref) */
    int extractor_exp;   /* 0 = off, else pre-rounding with 2^extractor_exp */
    int precision;       /* BM3D_PRECISION_* */
    /* Options of the NVRTC variant (0 = the bm3dcuda behaviour). The mixed
     * precision matcher serves BM3D_BM_SSD only; other distances run the FP32
     * matcher. */
    int bm_error;        /* BM3D_BM_* */
    int transform_2d;    /* BM3D_TRANSFORM_*: along block rows and columns */
    int transform_1d;    /* BM3D_TRANSFORM_*: along the group's members    */
    /* Precision MIXED: candidates kept per temporal frame for exact
     * rescoring. 0 = 8 (16 for ps_num above 4); else 4, 8 or 16, at least
     * 2 ps_num. Fewer keys: faster matching, group lists further from the
     * FP32 path's (4 at ps_num 2: 0.0014% of the lists of a 1080p test clip
     * differ, not only at ties; 8: none). No counterpart in the reference. */
    int mixed_temporal_keys;
} bm3d_desc;

/* A clip in device memory: frame f, plane p starts at
 * data + f * frame_stride + p * plane_stride; rows are desc.stride apart. */
typedef struct bm3d_clip {
    const float *data;
    long long frame_stride;
    long long plane_stride;
    int num_frames;
} bm3d_clip;

/* Creates a plan on the current HIP device. Validates the parameters with
 * the reference's messages (source/source.cpp:476-611). Any search reach the
 * reference accepts is planned: when bm_range + radius * ps_range outgrows
 * the LDS tiles, the filter adds the members beyond its accumulation margin
 * with global atomics (bm3d_execute_vagg and batches then refuse the plan),
 * and the matcher reads the planes from global memory. */
int bm3d_plan_create(const bm3d_desc *desc, bm3d_plan **plan);
void bm3d_plan_destroy(bm3d_plan *plan);

/* Floats in one output stack: planes * (2r+1) * 2 * height * stride
 * (the reference's d_res shape, source/kernel.cu:41-42). */
long long bm3d_stack_size(const bm3d_plan *plan);

/* Evaluates BM3D for output frames first_frame .. first_frame+count-1 of
 * `src` (ref: the basic estimate, required iff desc.final_). Frame n's
 * temporal slot z is frame clamp(n - r + z, 0, num_frames - 1)
 * (source/source.cpp:240-264). Output frame i's stack starts at
 * d_res + i * res_frame_stride; its plane p at + p * res_plane_stride; slot
 * z of a plane holds `height` rows of the weighted sum, then `height` rows of
 * the weight (source/source.cpp:362-366), rows desc.stride floats apart.
 * A stride of 0 means packed (bm3d_stack_size() per frame,
 * (2r+1) * 2 * height * stride per plane). ref must share src's strides.
 * The stacks are fully overwritten. */
int bm3d_execute(bm3d_plan *plan, float *d_res, long long res_frame_stride, long long res_plane_stride,
                 const bm3d_clip *src, const bm3d_clip *ref, int first_frame, int count, void *stream);

/* BM3D followed by VAggregate (the BM3Dv2 composition, source/source.cpp:
 * 944-988) for output frames first_frame .. first_frame+count-1 (for
 * radius 0, BM3D's aggregated output, as bm3d_execute + bm3d_aggregate):
 * dst frame i, plane p at dst + i * dst_frame_stride + p * dst_plane_stride,
 * rows dst_stride floats apart (0 = packed, width floats). The same values
 * as bm3d_execute into stacks followed by bm3d_vaggregate (bit for bit when
 * extractor_exp is set), computed without writing the stacks: each output
 * reads the accumulation of the 2r + 1 evaluations that cover it. Planes
 * skipped by chroma (sigma < FLT_EPSILON) are not written. Returns
 * BM3D_EUNSUPPORTED when the plan's accumulation tiles cannot hold every
 * group member (very large bm_range + radius * ps_range); use bm3d_execute
 * and bm3d_vaggregate then. No counterpart entry point in the reference. */
int bm3d_execute_vagg(bm3d_plan *plan, float *dst, long long dst_frame_stride, long long dst_plane_stride,
                      int dst_stride, const bm3d_clip *src, const bm3d_clip *ref, int first_frame, int count,
                      void *stream);

/* bm3d_execute without allocating device memory: no counterpart in the
 * reference, for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it). Same arguments,
 * checks and stacks as bm3d_execute, bit for bit when extractor_exp is set.
 * The group lists, weight maps and accumulation regions live in a static
 * 64 MiB device array that each GPU holds from the library's first use of
 * it until the process ends, instead of the plan's buffers: one evaluation
 * needs about 25 bytes per plane pixel and slot (a 1920x1080 grey plane at
 * radius 0 fits, with block_step 4); plans that need more than 64 MiB for
 * one evaluation get BM3D_EINVAL, naming the size. Every call on one GPU
 * shares that array: calls that overlap on a GPU (other streams, other
 * threads) race and must be serialised by the caller. */
int bm3d_execute_noalloc(bm3d_plan *plan, float *d_res, long long res_frame_stride, long long res_plane_stride,
                         const bm3d_clip *src, const bm3d_clip *ref, int first_frame, int count, void *stream);

/* One window in the reference's packed input layout
 * d_src = [(final ? 2 : 1)][planes][2r+1][height][stride], ref first
 * (source/kernel.cu:43-44). Writes one packed stack to d_res. */
int bm3d_execute_window(bm3d_plan *plan, float *d_res, const float *d_src, void *stream);

/* r = 0 aggregation, dst = num / den per pixel (source/source.cpp:170-188)
 * for `count` frames of `nplanes` planes. res: stacks with one slot per
 * plane (strides as in bm3d_execute; 0 = packed). */
int bm3d_aggregate(float *dst, long long dst_frame_stride, long long dst_plane_stride, int dst_stride,
                   const float *res, long long res_frame_stride, long long res_plane_stride,
                   int width, int height, int stride, int nplanes, int count, void *stream);

/* VAggregate (source/source.cpp:746-862) for output frames
 * first_frame .. first_frame+count-1 of a num_frames-frame clip.
 * stacks: the stacks of frames stack_first, stack_first+1, ... (strides as
 * in bm3d_execute; 0 = packed), covering clamp(n - r, 0, N-1) ..
 * clamp(n + r, 0, N-1) for every output frame n. */
int bm3d_vaggregate(float *dst, long long dst_frame_stride, long long dst_plane_stride, int dst_stride,
                    const float *stacks, long long stack_frame_stride, long long stack_plane_stride,
                    int stack_first, int width, int height, int stride, int nplanes, int radius,
                    int num_frames, int first_frame, int count, void *stream);

/* Block matching only (no counterpart entry point in the reference; it
 * exposes the group lists the kernel of source/kernel.cu:432-647 builds in
 * registers). Writes, for each output frame, nblocks x 8 packed positions
 * x | y << 14 | z << 27, blocks in raster order of the reference-block grid
 * (cpu_source/source.cpp:787-791). nblocks via bm3d_num_blocks(). */
int bm3d_match(bm3d_plan *plan, unsigned int *d_groups, const bm3d_clip *src, const bm3d_clip *ref,
               int first_frame, int count, void *stream);
long long bm3d_num_blocks(const bm3d_plan *plan);

/* Block-matching kernel of the plan: 1 for the mixed-precision kernel
 * (precision MIXED: FP16 pre-scoring, exact FP32 rescoring of the best few)
 * with a per-lane spatial search, 3 for the same kernel with its spatial
 * search by strips shared between neighbouring lanes (the default at block
 * step 4, and at block step 2 with 16-lane rows; at radius > 0 it runs as two
 * launches, the spatial search by strips and then the temporal search with
 * one lane per block, handing 64 B per block over in a buffer of the plan;
 * bm3d_execute_noalloc runs both in one launch), 2 for the same kernel with the spatial search's cross
 * terms on the matrix cores where the work-group tile allows (opt-in: slower
 * on MI355X), 0 for the FP32 kernel (precision FP32, or a search window too
 * large for the mixed kernel's 12-bit scan indices). No counterpart in the
 * reference. */
int bm3d_plan_matching_kernel(const bm3d_plan *plan);

/* Filtering kernel of the plan: 1 for the mixed-precision kernel (precision
 * MIXED: the 3-D DCT pair as FP16 matrix products with FP32 accumulation;
 * shrinkage, weights and aggregation in FP32), 0 for the FP32 kernel. No
 * counterpart in the reference. */
int bm3d_plan_filter_kernel(const bm3d_plan *plan);

/* Selects the plan's kernels, for testing and measurement (no counterpart in
 * the reference): `matching` and `filter` each 0 (FP32) or 1 (mixed
 * precision); `matching` 2 is the mixed kernel with its spatial search's
 * cross terms on the matrix cores, 3 the mixed kernel with its strip search
 * (where the block step and tile width allow it; the per-lane search
 * elsewhere). Mixed kernels need a plan created with precision MIXED
 * (BM3D_EINVAL otherwise); the mixed matching kernel also needs search
 * windows within its limits (BM3D_EUNSUPPORTED otherwise). */
int bm3d_plan_select_kernels(bm3d_plan *plan, int matching, int filter);

/* Launch shapes of a plan (no counterpart in the reference: the reference
 * bakes one shape per GPU generation into its kernel, source/kernel.cu:
 * 368-374, 823-828). Every shape gives the same results; only the speed
 * differs. bm3d_plan_create picks defaults; bm3d_plan_autotune measures
 * candidates on the caller's data. */
typedef struct bm3d_tuning {
    int match_rows;     /* rows of reference blocks per matching work-group: 4, 8 or 16 */
    int filter_tile_x;  /* reference blocks per filtering work-group tile: 1..32 each, */
    int filter_tile_y;  /*   at most 256 in all                                         */
    int filter_waves;   /* waves (64 lanes) per filtering work-group: 1..8 (the FP32
                           filter uses at most 4) */
    int batch_frames;   /* evaluations (output frames) per launch: 1..256 */
    int chunks;         /* bm3d_execute_vagg: runs of evaluations per filtering tile and
                           launch, 1..batch_frames (more runs: more work-groups, more
                           regions summed with atomics) */
    int match_width;    /* reference blocks per row of a matching work-group: 16 or 32
                           (match_width x match_rows <= 256) */
} bm3d_tuning;

int bm3d_plan_get_tuning(const bm3d_plan *plan, bm3d_tuning *out);
/* Applies `t` (BM3D_EINVAL outside the ranges above, BM3D_EUNSUPPORTED when
 * its work-group tiles do not fit the device's LDS). */
int bm3d_plan_set_tuning(bm3d_plan *plan, const bm3d_tuning *t);

/* Times candidate launch shapes on frames first_frame .. first_frame +
 * count - 1 of src (and ref for a final-pass plan) with
 * bm3d_execute_vagg (radius > 0) or bm3d_execute (radius 0), applies the
 * fastest and records it in a process-wide cache keyed by the plan's
 * parameters and plane size, so later plans with the same key start from it.
 * When the environment variable BM3D_TUNE_CACHE names a file, the cache is
 * read from and written to it (one line per key). Synchronises `stream`.
 * The outputs written while tuning go to scratch, not to the caller. */
int bm3d_plan_autotune(bm3d_plan *plan, const bm3d_clip *src, const bm3d_clip *ref, int first_frame, int count,
                       void *stream);

/* Kernel timing (no counterpart in the reference). While enabled, every
 * launch of the block-matching (kernel 0) and filtering (kernel 1) kernels
 * is bracketed by HIP events on its stream. bm3d_profile_read waits for the
 * recorded events and returns the summed milliseconds and launch count of
 * one kernel since the last reset. */
int bm3d_profile_enable(int on);
int bm3d_profile_reset(void);
int bm3d_profile_read(int kernel, double *total_ms, long long *launches);

/* Exact number of block-matching candidates evaluated (inside the plane)
 * for the given frames, counted by the matching kernel itself. Used to
 * state algorithmic work; no counterpart in the reference. */
int bm3d_count_candidates(bm3d_plan *plan, const bm3d_clip *src, const bm3d_clip *ref, int first_frame,
                          int count, unsigned long long *ncand, void *stream);

/* ---- batches of clips ------------------------------------------------------
 * Many clips of different sizes, radii and parameters denoised together. Each
 * stage (fixed-point range, block matching, filtering, VAggregate) runs as
 * one launch over every clip of a wave of clips, the waves bounded by a
 * device-memory budget; bm3d_batch_create captures the whole sequence, both
 * passes of every clip, in one HIP graph and bm3d_batch_execute replays it.
 * No counterpart in the reference, whose plugin launches one graph per frame
 * of one clip (source/source.cpp:299-441, source/kernel.cu:759-864). */

/* One plane group of one clip (one plane, or the three planes of a YUV444
 * clip with desc.chroma = 1). Output semantics are BM3Dv2's
 * (source/source.cpp:911-989): BM3D followed by VAggregate for radius > 0,
 * BM3D's aggregated output for radius 0; planes not processed (sigma below
 * FLT_EPSILON, source/source.cpp:916-930) are copied from src. */
typedef struct bm3d_batch_job {
    bm3d_desc desc;   /* parameters; desc.final_ is ignored */
    int passes;       /* 1: out = BM3Dv2(src);
                         2: basic = BM3Dv2(src), out = BM3Dv2(src, ref = basic) */
    bm3d_clip src;    /* device clip, rows desc.stride floats apart */
    float *basic;     /* passes == 2: device frames of the basic estimate, with
                         src's frame and plane strides */
    float *out;       /* device output frames, src's frame and plane strides */
} bm3d_batch_job;

typedef struct bm3d_batch bm3d_batch;

/* Plans every job on the current device (the reference's argument checks
 * and messages, as bm3d_plan_create), packs the jobs into waves whose
 * scratch stays within scratch_budget bytes (0: 16 GiB; a job larger than
 * the budget gets a wave of its own), and captures the graph. The jobs'
 * buffers must stay valid until bm3d_batch_destroy. */
int bm3d_batch_create(const bm3d_batch_job *jobs, int njobs, long long scratch_budget, bm3d_batch **batch);
/* Replays the batch's graph on `stream` (asynchronous). */
int bm3d_batch_execute(bm3d_batch *batch, void *stream);
/* Launch shapes for every job (batch_frames and chunks are not used: a job's
 * frames are evaluated in one run); the graph is captured again. */
int bm3d_batch_set_tuning(bm3d_batch *batch, const bm3d_tuning *t);
int bm3d_batch_get_tuning(const bm3d_batch *batch, bm3d_tuning *t);
/* Times candidate launch shapes by replaying the batch (which writes the
 * jobs' outputs) and keeps the fastest. Synchronises `stream`. */
int bm3d_batch_autotune(bm3d_batch *batch, void *stream);
/* Waves, kernel launches in the graph, and device bytes of scratch. */
int bm3d_batch_info(const bm3d_batch *batch, int *waves, int *launches, long long *scratch_bytes);
void bm3d_batch_destroy(bm3d_batch *batch);

/* ---- plugin-level filters ------------------------------------------------
 * The three filters the bm3dcuda plugin registers (source/source.cpp:
 * 991-1027), with its argument semantics: optional per-plane arrays whose
 * missing entries repeat the previous one, the same defaults, checks and
 * error messages (BM3DCreate, source/source.cpp:457-720; VAggregateCreate
 * 880-909; BM3Dv2Create 911-989), and the same frame semantics (BM3DGetFrame
 * 206-441, VAggregateGetFrame 746-862). Frames live in device memory; a
 * get_frames call produces any range of output frames of a whole clip. */

#define BM3D_GRAY 1
#define BM3D_RGB 2
#define BM3D_YUV 3

typedef struct bm3d_video_info {
    int color_family;    /* BM3D_GRAY, BM3D_RGB or BM3D_YUV */
    int sample_type;     /* 0 integer, 1 float */
    int bits_per_sample;
    int subsampling_w, subsampling_h;
    int num_planes;
    int width, height;   /* 0 = variable (rejected, as in the reference) */
    int num_frames;
} bm3d_video_info;

/* Arguments of BM3D / BM3Dv2 ("clip:clip;ref:clip:opt;sigma:float[]:opt;
 * block_step:int[]:opt;bm_range:int[]:opt;radius:int:opt;ps_num:int[]:opt;
 * ps_range:int[]:opt;chroma:int:opt;device_id:int:opt;fast:int:opt;
 * extractor_exp:int:opt;zero_init:int:opt;", source/source.cpp:1001-1016).
 * n_<name> is the number of values given (0 = key absent). `precision` has
 * no counterpart in the reference (BM3D_PRECISION_*, default FP32). */
typedef struct bm3d_args {
    int n_sigma;          double sigma[3];
    int n_block_step;     long long block_step[3];
    int n_bm_range;       long long bm_range[3];
    int n_radius;         long long radius;
    int n_ps_num;         long long ps_num[3];
    int n_ps_range;       long long ps_range[3];
    int n_chroma;         long long chroma;
    int n_device_id;      long long device_id;
    int n_fast;           long long fast;
    int n_extractor_exp;  long long extractor_exp;
    int n_zero_init;      long long zero_init;
    int n_precision;      long long precision;
    /* The NVRTC variant's string options (rtc_source/source.cpp:854-912):
     * per plane, case-insensitive, missing entries repeat the previous one;
     * bm_error_s in {"ssd", "sad", "zssd", "zsad", "ssd/norm"} (default
     * "ssd"), transform_*_s in {"dct", "haar", "wht", "bior1.5"} (default
     * "dct"). With chroma, plane 0's entries apply to all planes. */
    int n_bm_error_s;     const char *bm_error_s[3];
    int n_transform_2d_s; const char *transform_2d_s[3];
    int n_transform_1d_s; const char *transform_1d_s[3];
} bm3d_args;

/* A clip's planes in device memory: frame f of plane p starts at
 * data[p] + f * frame_stride[p]; rows are stride[p] floats apart. */
typedef struct bm3d_frames {
    float *data[3];
    long long frame_stride[3];
    int stride[3];
} bm3d_frames;

typedef struct bm3d_filter bm3d_filter;

/* BM3D (BM3DCreate). ref: video info of "ref", or NULL when not given. */
int bm3d_filter_create(const bm3d_args *args, const bm3d_video_info *clip, const bm3d_video_info *ref,
                       bm3d_filter **filter);
/* VAggregate (VAggregateCreate): clip = a BM3D output with radius > 0,
 * src = the source clip, planes = indices of the planes to aggregate. */
int bm3d_vaggregate_create(const long long *planes, int n_planes, const bm3d_video_info *clip,
                           const bm3d_video_info *src, bm3d_filter **filter);
/* BM3Dv2 (BM3Dv2Create): BM3D, then VAggregate when radius > 0; the source
 * clip itself when no plane is processed. */
int bm3dv2_filter_create(const bm3d_args *args, const bm3d_video_info *clip, const bm3d_video_info *ref,
                         bm3d_filter **filter);

/* Output video info (BM3DInit: height x 2(2r+1) for radius > 0), the
 * temporal radius and the processed planes (the BM3D_V_radius and
 * BM3D_V_process frame properties, source/source.cpp:428-435). */
int bm3d_filter_info(const bm3d_filter *filter, bm3d_video_info *out, int *radius, int process[3]);

/* Produces output frames first .. first+count-1. clip and ref (BM3D,
 * BM3Dv2) or clip and src (VAggregate; pass src as `ref`) hold all frames
 * of their clips; out receives frames at their own indices. */
int bm3d_filter_get_frames(bm3d_filter *filter, const bm3d_frames *clip, const bm3d_frames *ref,
                           const bm3d_frames *out, int first, int count, void *stream);
void bm3d_filter_destroy(bm3d_filter *filter);

/* ---- frames in host memory (the VapourSynth GetFrame route) -------------
 * Replaces the reference's per-frame staging (source/source.cpp:296-426:
 * vs_bitblt of every window slot into pinned memory, a graph of HtoD,
 * kernel and DtoH, a stream sync, vs_bitblt out), with `fast` selecting
 * kFast = 4 resource sets (source/source.cpp:62, 636-713).
 *
 * One frame: plane p's rows at data[p], stride[p] floats apart, in any host
 * memory (VSFrameRef planes: getReadPtr / getWritePtr). */
typedef struct bm3d_host_frame {
    float *data[3];
    int stride[3];
} bm3d_host_frame;

/* Input frames lo .. hi that output frames first .. first+count-1 read
 * (clamped to the clip): BM3D r, BM3Dv2 2r, VAggregate r frames either
 * side. */
int bm3d_filter_host_range(const bm3d_filter *filter, int first, int count, int *lo, int *hi);

/* bm3d_filter_get_frames on host frames. clip[i] and ref[i] are frame lo + i
 * for i = 0 .. hi - lo (bm3d_filter_host_range); ref is the basic estimate
 * (BM3D, BM3Dv2 final) or the source clip (VAggregate, which reads only
 * frames first .. first+count-1 of it; other entries may be zero). out[i]
 * receives output frame first + i. Synchronous: returns once out is
 * written.
 *
 * Each input frame goes to the device once per call, and only the output
 * frames come back. For BM3Dv2 the per-frame stacks never cross PCIe. Host
 * rows are staged through pinned buffers by a pool of host threads. Ranges
 * of more than a few frames run in chunks whose upload, compute and
 * download overlap on separate streams. Calls from several threads run
 * concurrently on up to 4 resource sets, or 1 when the filter was created
 * with fast = 0, each with its own plans and buffers. A call holds its
 * input range (count + 4r frames) on the device. */
int bm3d_filter_get_frames_host(bm3d_filter *filter, const bm3d_host_frame *clip, const bm3d_host_frame *ref,
                                int lo, bm3d_host_frame *out, int first, int count);

/* ---- failure injection (tests) -------------------------------------------
 * No counterpart in the reference, which reports runtime failures through
 * setError / setFilterError (source/source.cpp:54-59, 305-314, 225-231);
 * this hook lets a test force those failures. Arms a one-shot countdown for
 * one kind of runtime event: device allocations (hipMalloc), pinned host
 * allocations (the host-frame route's staging buffers) or kernel launches.
 * The event that finds the countdown at zero (after = 0: the next one) fails
 * with hipErrorOutOfMemory / hipErrorLaunchFailure without reaching the
 * runtime, and the call that made it returns BM3D_EHIP with a message naming
 * it; later calls work normally. after < 0 disarms the site. Process-wide:
 * events on every thread count. */
#define BM3D_INJECT_DEVICE_ALLOC 1
#define BM3D_INJECT_HOST_ALLOC 2
#define BM3D_INJECT_LAUNCH 3
int bm3d_debug_inject_failure(int site, long long after);

/* Message of the last failing call on this thread ("" if none). */
const char *bm3d_last_error(void);

/* Library version string. */
const char *bm3d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* BM3D_MI355X_H */
