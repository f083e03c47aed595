// Launch layer: the extern "C" entry points of include/bm3d_mi355x.h.
//
// Replaces get_graphexec (source/kernel.cu:759-864): instead of a graph with
// baked HtoD / memset / kernel / DtoH nodes per frame, a plan holds the
// parameters and the device scratch (group lists, corner-weight maps), and
// bm3d_execute runs zero -> match -> filter -> den-box on device-resident
// clips for a whole batch of output frames per launch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "bm3d_inject.h"
#include "bm3d_internal.h"
#include "bm3d_kernels.h"
#include "bm3d_mi355x.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

// Device memory of bm3d_execute_noalloc: one array per GPU, held until the
// process ends, shared by every call on that GPU (which is why such calls
// must not overlap there). Host code uses its address from
// hipGetSymbolAddress, never the name.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int bm3d::fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

void bm3d::clear_error() { g_last_error.clear(); }

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

// Default evaluations per launch.
constexpr int kDefaultBatchFrames = 32;

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

namespace {

int count_positions(int n, int bs) {
    int c = 0;
    for (int v = 0; v < n - 8 + bs; v += bs) ++c;
    return c;
}

// Device memory a plan's batch may take for its own buffers (group lists,
// corner maps, regions); larger clips are processed in more launches.
constexpr long long kScratchBudget = 8LL << 30;

// Evaluations per launch: the tuning's batch, at most n, and at most what
// the scratch budget allows at per_frame bytes each (after `fixed` bytes).
int batch_limit(const bm3d_plan* p, int n, long long per_frame, long long fixed = 0) {
    int b = std::min(n, p->tune.batch_frames);
    const long long cap = (kScratchBudget - fixed) / std::max(per_frame, 1LL);
    if (cap < b) b = (int)cap;
    return std::max(b, 1);
}

// Raises the dynamic-LDS limit of a kernel family (0 matching, 1 mixed
// matching, 2 filtering, 3 mixed filtering) to at least `bytes`. The limit
// is process-wide and only ever raised, so plans of different sizes coexist.
hipError_t raise_lds_limit(int family, size_t bytes) {
    static std::mutex mu;
    static size_t cur[4] = {65536, 65536, 65536, 65536};
    if (bytes <= 65536) return hipSuccess;
    std::lock_guard<std::mutex> g(mu);
    if (bytes <= cur[family]) return hipSuccess;
    hipError_t e = hipSuccess;
    switch (family) {
        case 0: e = bm3d::set_match_lds_limit(bytes); break;
        case 1: e = bm3d::set_match_dx_lds_limit(bytes); break;
        case 2: e = bm3d::set_filter_lds_limit(bytes); break;
        default: e = bm3d::set_filter_mx_lds_limit(bytes); break;
    }
    if (e == hipSuccess) cur[family] = bytes;
    return e;
}

bm3d_tuning default_tuning(const bm3d_plan* p) {
    bm3d_tuning t{};
    t.match_rows = 16;
    t.match_width = 16;
    t.filter_tile_x = 8;
    t.filter_tile_y = 8;
    // the autotuner's choice for a 30-frame 1080p clip on both paths: four
    // evaluation runs per launch give the filter 4x the work-groups; the
    // mixed filter's copy of its transform operands takes 17 KB of LDS per
    // work-group, so it runs 8 waves per work-group for 4 waves per SIMD
    t.filter_waves = p->d.precision == BM3D_PRECISION_MIXED ? 8 : 4;
    t.batch_frames = kDefaultBatchFrames;
    t.chunks = 4;
    return t;
}

const char* tuning_range_error(const bm3d_tuning& t) {
    if (t.match_rows != 4 && t.match_rows != 8 && t.match_rows != 16) return "match_rows must be 4, 8 or 16";
    if (t.match_width != 16 && t.match_width != 32) return "match_width must be 16 or 32";
    if (t.match_width * t.match_rows > 256) return "match_width x match_rows must be at most 256";
    if (t.filter_tile_x < 1 || t.filter_tile_x > 32 || t.filter_tile_y < 1 || t.filter_tile_y > 32 ||
        t.filter_tile_x * t.filter_tile_y > 256)
        return "filter tile must be 1..32 blocks per side and at most 256 blocks";
    if (t.filter_waves < 1 || t.filter_waves > 8) return "filter_waves must be in range [1, 8]";
    if (t.batch_frames < 1 || t.batch_frames > 256) return "batch_frames must be in range [1, 256]";
    if (t.chunks < 1 || t.chunks > t.batch_frames) return "chunks must be in range [1, batch_frames]";
    return nullptr;
}

// Process-wide cache of tuned launch shapes, keyed by the parameters and
// plane size that set a plan's work (optionally backed by the file named in
// BM3D_TUNE_CACHE: lines "<key> : <6 tuning fields>").
struct TuneCache {
    std::mutex mu;
    bool loaded = false;
    std::vector<std::pair<std::string, bm3d_tuning>> entries;
};
TuneCache& tune_cache() {
    static TuneCache c;
    return c;
}

std::string tune_key(const bm3d_plan* p) {
    const bm3d_desc& d = p->d;
    char buf[192];
    std::snprintf(buf, sizeof buf, "w%d h%d s%d bs%d R%d r%d pn%d pr%d c%d f%d p%d e%d t%d%d k%d", d.width,
                  d.height, d.stride, d.block_step, d.bm_range, d.radius, d.ps_num, d.ps_range, d.chroma, d.final_,
                  d.precision, d.bm_error, d.transform_2d, d.transform_1d, bm3d::plan_temporal_keys(p));
    return buf;
}

void load_tune_file(TuneCache& c) {
    if (c.loaded) return;
    c.loaded = true;
    const char* path = std::getenv("BM3D_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        char* colon = std::strchr(line, ':');
        if (!colon) continue;
        std::string key(line, colon - line);
        while (!key.empty() && key.back() == ' ') key.pop_back();
        bm3d_tuning t{};
        if (std::sscanf(colon + 1, "%d %d %d %d %d %d %d", &t.match_rows, &t.filter_tile_x, &t.filter_tile_y,
                        &t.filter_waves, &t.batch_frames, &t.chunks, &t.match_width) == 7 &&
            !tuning_range_error(t))
            c.entries.emplace_back(key, t);
    }
    std::fclose(f);
}

bool cached_tuning(const bm3d_plan* p, bm3d_tuning* out) {
    TuneCache& c = tune_cache();
    std::lock_guard<std::mutex> g(c.mu);
    load_tune_file(c);
    const std::string key = tune_key(p);
    for (auto it = c.entries.rbegin(); it != c.entries.rend(); ++it)
        if (it->first == key) {
            *out = it->second;
            return true;
        }
    return false;
}

void store_tuning(const bm3d_plan* p, const bm3d_tuning& t) {
    TuneCache& c = tune_cache();
    std::lock_guard<std::mutex> g(c.mu);
    load_tune_file(c);
    const std::string key = tune_key(p);
    bool found = false;
    for (auto& e : c.entries)
        if (e.first == key) {
            e.second = t;
            found = true;
        }
    if (!found) c.entries.emplace_back(key, t);
    const char* path = std::getenv("BM3D_TUNE_CACHE");
    if (!path || !*path) return;
    FILE* f = std::fopen(path, "w");
    if (!f) return;
    for (auto& e : c.entries)
        std::fprintf(f, "%s : %d %d %d %d %d %d %d\n", e.first.c_str(), e.second.match_rows, e.second.filter_tile_x,
                     e.second.filter_tile_y, e.second.filter_waves, e.second.batch_frames, e.second.chunks,
                     e.second.match_width);
    std::fclose(f);
}

// Matching tile for a work-group of width x rows reference blocks; false
// when its LDS does not fit the device.
bool match_geometry(bm3d_plan* p, int rows, int width) {
    const int bs = p->d.block_step;
    const int tile_w = (width - 1) * bs + 8 + 2 * p->margin;
    const int tile_h = (rows - 1) * bs + 8 + 2 * p->margin;
    const int tile_wq = (tile_w + bs - 1) / bs;
    const int tile_hp = tile_h | 1;  // odd column pitch: lanes of a row hit distinct banks
    const size_t bytes = (size_t)bs * tile_wq * tile_hp * sizeof(float);
    if (bytes > (size_t)p->max_lds) return false;
    p->tbw = width;
    p->tby = rows;
    p->tile_w = tile_w;
    p->tile_h = tile_h;
    p->tile_wq = tile_wq;
    p->tile_hp = tile_hp;
    p->lds_bytes = bytes;
    return true;
}

// Mixed-precision matching kernel (bm3d_match_dx.hip): the FP32 kernel's
// tile, stored as two planes of FP16 pairs (hp2 words per column, odd so
// that the lanes of a row hit distinct banks). Scan indices are packed in
// 12 bits.
void dx_geometry(bm3d_plan* p) {
    const bm3d_desc& d = p->d;
    p->dx_ok = false;
    if (d.precision != BM3D_PRECISION_MIXED || d.bm_error != BM3D_BM_SSD || p->lds_bytes == 0) return;
    const int nWs = 2 * d.bm_range + 1, nWt = 2 * d.ps_range + 1;
    if (nWs * nWs > 4096 || (d.radius > 0 && d.ps_num * nWt * nWt > 4096)) return;
    p->tile_hp2 = ((p->tile_h + 1) / 2 + 1) | 1;
    const size_t bytes = (size_t)2 * d.block_step * p->tile_wq * p->tile_hp2 * 4;
    // + the kernel's static arrays (11 KB: the spatial list, the matrix-core
    // search's column sums)
    if (bytes + 12288 > (size_t)p->max_lds) return;
    p->dx_lds_bytes = bytes;
    p->dx_ok = true;
}

// Filter tiling: tbx x tby reference blocks per work-group, accumulation
// margin rm; false when the LDS regions (2r + 1 of them) do not fit.
bool filter_geometry(bm3d_plan* p, int tbx, int tby, int rm) {
    const int bs = p->d.block_step;
    const int rw = (tbx - 1) * bs + 8 + 2 * rm;
    const int rh = (tby - 1) * bs + 8 + 2 * rm;
    const size_t bytes = (size_t)p->tw * rw * rh * sizeof(int);
    // + the kernels' static slot table, and the mixed filter's copy of its
    // transform operands
    const size_t fixed = 1024 + (p->d.precision == BM3D_PRECISION_MIXED ? bm3d::kMxOpVectors * 64 * 16 : 0);
    if (bytes + fixed > (size_t)p->max_lds) return false;
    p->ftbx = tbx; p->ftby = tby; p->frm = rm; p->frw = rw; p->frh = rh;
    p->filter_lds_bytes = bytes;
    p->fntx = (p->nbx + tbx - 1) / tbx;
    p->fnty = (p->nby + tby - 1) / tby;
    p->region_ints = (long long)p->fntx * p->fnty * p->nplanes * rw * rh;
    return true;
}

// Applies a tuning's tiles: the requested ones at the full search reach
// when they fit, else the largest default tile that fits, else a 4 x 4 tile
// with a reduced margin (members beyond it go to global atomics, and
// bm3d_execute_vagg is unavailable). A matching tile whose union window
// does not fit even 4 rows of blocks makes the matcher read the planes from
// global memory (no LDS tile; same results, slower).
int apply_tuning(bm3d_plan* p, const bm3d_tuning& t) {
    if (p->pinned)
        return fail(BM3D_EINVAL, "BM3D: the plan is held by a captured graph; its launch shapes cannot change");
    if (!match_geometry(p, t.match_rows, t.match_width) && !match_geometry(p, 4, 16)) {
        p->tbw = t.match_width;
        p->tby = t.match_rows;
        p->tile_w = p->tile_h = p->tile_wq = p->tile_hp = 0;
        p->lds_bytes = 0;
    }
    dx_geometry(p);
    if (!p->dx_ok) p->use_dx = false;
    bool ok = filter_geometry(p, t.filter_tile_x, t.filter_tile_y, p->margin);
    const int fallback[][2] = {{8, 8}, {8, 4}, {4, 4}};
    for (auto& f : fallback)
        if (!ok) ok = filter_geometry(p, f[0], f[1], p->margin);
    for (int rm = p->margin - 1; !ok && rm >= 0; --rm) ok = filter_geometry(p, 4, 4, rm);
    if (!ok) return fail(BM3D_EUNSUPPORTED, "BM3D: no filter tiling fits the device's LDS");
    hipError_t e = raise_lds_limit(0, p->lds_bytes);
    if (e == hipSuccess && p->dx_ok) e = raise_lds_limit(1, p->dx_lds_bytes);
    if (e == hipSuccess) e = raise_lds_limit(2, p->filter_lds_bytes);
    // (only mixed-precision plans can run the mixed filter, whose static LDS
    // leaves less room for the regions)
    if (e == hipSuccess && p->d.precision == BM3D_PRECISION_MIXED) e = raise_lds_limit(3, p->filter_lds_bytes);
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute");
    p->tune = t;
    p->tune.match_rows = p->tby;
    p->tune.match_width = p->tbw;
    p->tune.filter_tile_x = p->ftbx;
    p->tune.filter_tile_y = p->ftby;
    return BM3D_OK;
}

// Grows the plan's device buffers to hold at least the given numbers of
// group-list entries, corner-map floats (and as many wrap counts), region
// ints and 64-bit partial sums. A call on a capturing stream pins the plan: its buffers must already
// be large enough (an eager call sizes them), and from then on they never
// move, since the captured graph holds their addresses.
int ensure_buffers(bm3d_plan* p, size_t groups, size_t corner, size_t scratch, hipStream_t stream,
                   size_t part = 0) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    BM3D_HIP(hipStreamIsCapturing(stream, &cs));
    const bool grow = groups > p->groups_cap || corner > p->corner_cap || scratch > p->scratch_cap ||
                      part > p->part_cap;
    if (cs != hipStreamCaptureStatusNone) p->pinned = true;
    if (grow && p->pinned)
        return fail(BM3D_EINVAL, "BM3D: the plan's buffers are too small for this call and are held by a captured "
                                 "graph (size them with an eager call of the largest range before capturing)");
    if (groups > p->groups_cap) {
        if (p->groups) (void)hipFree(p->groups);
        p->groups = nullptr;
        p->groups_cap = 0;
        BM3D_HIP(bm3d::malloc_device(&p->groups, groups * sizeof(uint32_t)));
        p->groups_cap = groups;
    }
    if (corner > p->corner_cap) {
        if (p->corner) (void)hipFree(p->corner);
        if (p->wrap) (void)hipFree(p->wrap);
        p->corner = nullptr;
        p->wrap = nullptr;
        p->corner_cap = 0;
        BM3D_HIP(bm3d::malloc_device(&p->corner, corner * sizeof(float)));
        BM3D_HIP(bm3d::malloc_device(&p->wrap, corner * sizeof(int)));
        // counts start at zero: readers use them whenever the flag is set
        BM3D_HIP(hipMemset(p->wrap, 0, corner * sizeof(int)));
        p->corner_cap = corner;
    }
    if (scratch > p->scratch_cap) {
        if (p->scratch) (void)hipFree(p->scratch);
        p->scratch = nullptr;
        p->scratch_cap = 0;
        BM3D_HIP(bm3d::malloc_device(&p->scratch, scratch * sizeof(int)));
        p->scratch_cap = scratch;
    }
    if (part > p->part_cap) {
        if (p->part) (void)hipFree(p->part);
        p->part = nullptr;
        p->part_cap = 0;
        BM3D_HIP(bm3d::malloc_device(&p->part, part * sizeof(long long)));
        p->part_cap = part;
    }
    return BM3D_OK;
}

// Block matching for frames first .. first + count - 1 into `groups` (or
// counting candidates into `ncand`): the mixed-precision kernel when the
// plan selected it, else the FP32 kernel.
bm3d::MatchArgs match_args(const bm3d_plan* p, uint32_t* groups, unsigned long long* ncand, const bm3d_clip* src,
                           const bm3d_clip* ref, int first_frame) {
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
    m.tbw = p->tbw;
    m.tile_w = p->tile_w; m.tile_h = p->tile_h; m.tile_wq = p->tile_wq; m.tile_hp = p->tile_hp;
    m.groups = groups;
    m.ncand = ncand;
    m.tile_hp2 = p->tile_hp2;
    m.cost = d.bm_error;
    m.kt = bm3d::plan_temporal_keys(p);
    m.mx_spatial = p->mx_spatial ? 1 : 0;
    m.strip = bm3d::plan_uses_strip(p) ? 1 : 0;
    m.obw = m.strip ? bm3d::strip_owned_width(p->tbw, d.block_step) : p->tbw;
    return m;
}

// Frames per pair of launches of the two-launch strip search: their
// hand-over lists (64 bytes per block) stay within 256 MB.
int split_frames(const bm3d_plan* p, int count) {
    const long long per_frame = (long long)p->nblocks * 16 * sizeof(uint32_t);
    return (int)std::max(1LL, std::min<long long>(count, (256LL << 20) / per_frame));
}

// The hand-over lists of `frames` frames of the two-launch strip search,
// grown if needed; null (*e unchanged) when the buffer is too small and may
// not move (a capturing stream or a plan held by a graph): the search then
// runs in one launch, with the same results.
uint32_t* split_lists(bm3d_plan* p, int frames, hipStream_t stream, hipError_t* e, bool may_alloc) {
    const size_t need = (size_t)frames * p->nblocks * 16;
    if (need <= p->slist_cap) return p->slist;
    if (!may_alloc) return nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    *e = hipStreamIsCapturing(stream, &cs);
    if (*e != hipSuccess || cs != hipStreamCaptureStatusNone || p->pinned) return nullptr;
    if (p->slist) (void)hipFree(p->slist);
    p->slist = nullptr;
    p->slist_cap = 0;
    *e = bm3d::malloc_device(&p->slist, need * sizeof(uint32_t));
    if (*e != hipSuccess) return nullptr;
    p->slist_cap = need;
    return p->slist;
}

hipError_t launch_matching(bm3d_plan* p, uint32_t* groups, unsigned long long* ncand, const bm3d_clip* src,
                           const bm3d_clip* ref, int first_frame, int count, hipStream_t stream,
                           bool may_alloc = true) {
    const bm3d::MatchArgs m = match_args(p, groups, ncand, src, ref, first_frame);
    if (!p->use_dx) return bm3d::launch_match(m, p->tby, p->lds_bytes, count, stream);
    if (m.strip && p->d.radius > 0 && count > 0 && may_alloc) {
        // the spatial search by strips, then the temporal search with one
        // lane per block (the strips' last lanes own no block and would idle
        // through it)
        const int chunk = split_frames(p, count);
        hipError_t e = hipSuccess;
        if (uint32_t* sl = split_lists(p, chunk, stream, &e, may_alloc)) {
            for (int c0 = 0; c0 < count && e == hipSuccess; c0 += chunk) {
                const int nf = std::min(chunk, count - c0);
                bm3d::MatchArgs ms = m;
                ms.first_frame = first_frame + c0;
                ms.slist = sl;
                ms.phase = bm3d::kPhSpatial;
                bm3d::MatchArgs mt = ms;
                mt.phase = bm3d::kPhTemporal;
                mt.strip = 0;
                mt.obw = mt.tbw;
                mt.groups = groups ? groups + (size_t)c0 * p->nblocks * 8 : nullptr;
                // through the launch table (fewer scalar registers); the
                // counting launches take the plain kernels
                if (ncand) {
                    e = bm3d::launch_match_dx(ms, p->tby, p->dx_lds_bytes, nf, stream);
                    if (e == hipSuccess) e = bm3d::launch_match_dx(mt, p->tby, p->dx_lds_bytes, nf, stream);
                } else {
                    e = bm3d::launch_match_dx_table(ms, p->tby, p->dx_lds_bytes, nf, p->match_table, stream);
                    if (e == hipSuccess)
                        e = bm3d::launch_match_dx_table(mt, p->tby, p->dx_lds_bytes, nf, p->match_table, stream);
                }
            }
            return e;
        }
        if (e != hipSuccess) return e;
    }
    return bm3d::launch_match_dx(m, p->tby, p->dx_lds_bytes, count, stream);
}

// Range pointer of the fixed-point unit (null: the unit is fixed).
const float* fx_range(const bm3d_plan* p) { return p->d.extractor_exp ? nullptr : p->range; }

// Measures the largest finite |sample| of the processed planes of frames
// lo .. hi of src into p->range: the fixed-point unit of every evaluation
// and aggregation of the call follows from it (without extractor_exp), and
// so does the filters' choice of 64-bit adds (always).
hipError_t measure_range(bm3d_plan* p, const bm3d_clip* src, int lo, int hi, hipStream_t stream) {
    hipError_t e = bm3d::launch_zero(p->range, 1, 0, 1, stream);
    if (e != hipSuccess) return e;
    return bm3d::launch_absmax(src->data + (size_t)lo * src->frame_stride, src->frame_stride, src->plane_stride,
                               p->nplanes, p->d.width, p->d.height, p->d.stride, hi - lo + 1, p->range, stream);
}

// Filter arguments common to both accumulation modes.
bm3d::FilterArgs filter_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref) {
    const bm3d_desc& d = p->d;
    bm3d::FilterArgs fa{};
    fa.src = src->data;
    fa.ref = d.final_ ? ref->data : nullptr;
    fa.frame_stride = src->frame_stride;
    fa.plane_stride = src->plane_stride;
    fa.num_frames = src->num_frames;
    fa.W = d.width; fa.H = d.height; fa.stride = d.stride;
    fa.plane_size = p->plane_size;
    fa.radius = d.radius;
    fa.nblocks = p->nblocks;
    for (int i = 0; i < 3; ++i) fa.sigma[i] = p->sigma[i];
    fa.extractor = p->extractor;
    fa.groups = p->groups;
    fa.corner = p->corner;
    fa.bs = d.block_step; fa.nbx = p->nbx; fa.nby = p->nby;
    fa.tbx = p->ftbx; fa.tby = p->ftby; fa.ntx = p->fntx; fa.nty = p->fnty;
    fa.rm = p->frm; fa.rw = p->frw; fa.rh = p->frh;
    fa.scratch = p->scratch;
    fa.fx_range = fx_range(p);
    fa.fx_exp = p->fx_exp;
    fa.wrap = p->wrap;
    fa.wrap_flag = p->wrap_flag;
    fa.range = p->range;
    fa.vbound = p->vbound;
    fa.mx_ops = p->mx_ops;
    fa.t2d = d.transform_2d;
    fa.t1d = d.transform_1d;
    return fa;
}

hipError_t launch_filtering(const bm3d_plan* p, const bm3d::FilterArgs& fa, int grid_z, hipStream_t stream) {
    ProfScope ps(1, stream);
    const bool fin = p->d.final_ != 0, chroma = p->d.chroma != 0;
    if (p->use_mx_filter)
        return bm3d::launch_filter_mx(fa, fin, chroma, grid_z, std::min(p->tune.filter_waves, bm3d::kMaxMxFilterWaves),
                                      p->filter_lds_bytes, stream);
    return bm3d::launch_filter(fa, fin, chroma, grid_z, std::min(p->tune.filter_waves, bm3d::kMaxFilterWaves),
                               p->filter_lds_bytes, stream);
}

// VAggregate-from-targets arguments over the plan's buffers (ring frames);
// the destination fields are the caller's.
bm3d::VTargetArgs vtarget_args(const bm3d_plan* p, int ring) {
    const bm3d_desc& d = p->d;
    bm3d::VTargetArgs va{};
    va.scratch = p->scratch;
    va.target_stride = p->region_ints;
    va.corner = p->corner;
    va.corner_frame = (long long)p->nplanes * p->plane_size;
    va.ring = ring;
    va.fx_range = fx_range(p);
    va.fx_exp = p->fx_exp;
    va.wrap = p->wrap;
    va.wrap_flag = p->wrap_flag;
    va.W = d.width; va.H = d.height; va.stride = d.stride;
    va.plane_size = p->plane_size;
    va.nplanes = p->nplanes;
    va.plane_mask = 0;
    for (int i = 0; i < p->nplanes; ++i)
        if (!(d.chroma && p->sigma[i] < std::numeric_limits<float>::epsilon())) va.plane_mask |= 1u << i;
    va.bs = d.block_step; va.tbx = p->ftbx; va.tby = p->ftby; va.ntx = p->fntx; va.nty = p->fnty;
    va.rm = p->frm; va.rw = p->frw; va.rh = p->frh;
    return va;
}

// Slot-mode evaluation of frames n0 .. n0 + nf - 1 (stacks path): group
// lists, slot regions and corner maps per evaluation.
// The device buffers a slot-mode evaluation batch works in: the plan's own
// (bm3d_execute), or carved from the static scratch (bm3d_execute_noalloc).
struct SlotBufs {
    uint32_t* groups;
    float* corner;
    int* wrap;
    int* regions;
    bool wrap_always_zero;  // zero the wrap counts every batch (else only once the plan counted a wrap)
    bool may_alloc;         // matching may allocate its own buffers (not in bm3d_execute_noalloc)
};

SlotBufs plan_slot_bufs(const bm3d_plan* p) { return {p->groups, p->corner, p->wrap, p->scratch, false, true}; }

hipError_t evaluate_slots(bm3d_plan* p, const SlotBufs& b, const bm3d_clip* src, const bm3d_clip* ref, int n0,
                          int nf, float* res, long long res_frame_stride, long long res_plane_stride,
                          hipStream_t stream) {
    const long long corner_eval = (long long)p->nplanes * p->tw * p->plane_size;
    hipError_t e = bm3d::launch_zero(b.corner, (size_t)nf * corner_eval, 0, 1, stream);
    if (e == hipSuccess)
        e = bm3d::launch_zero(reinterpret_cast<float*>(b.wrap), (size_t)nf * corner_eval, 0, 1, stream,
                              b.wrap_always_zero ? nullptr : p->wrap_flag);
    if (e != hipSuccess) return e;
    {
        ProfScope ps(0, stream);
        e = launch_matching(p, b.groups, nullptr, src, ref, n0, nf, stream, b.may_alloc);
    }
    if (e != hipSuccess) return e;
    bm3d::FilterArgs fa = filter_args(p, src, ref);
    fa.groups = b.groups;
    fa.corner = b.corner;
    fa.wrap = b.wrap;
    fa.scratch = b.regions;
    fa.mode = 0;
    fa.first_frame = n0;
    fa.res = res;
    fa.res_frame_stride = res_frame_stride;
    fa.res_plane_stride = res_plane_stride;
    fa.corner_frame_stride = corner_eval;
    return launch_filtering(p, fa, nf, stream);
}

// Checks of bm3d_execute's arguments; fills in packed strides.
int check_execute(bm3d_plan* p, float* d_res, long long& res_frame_stride, long long& res_plane_stride,
                  const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count) {
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
    int dev = 0;
    BM3D_HIP(hipGetDevice(&dev));
    if (dev != p->device) return fail(BM3D_EINVAL, "BM3D: plan used on a different device than it was created on");
    return BM3D_OK;
}

// The stacks of output frames first_frame .. first_frame + count - 1, in
// evaluation batches of `batch` frames over the buffers b.
int run_slots(bm3d_plan* p, const SlotBufs& b, int batch, float* d_res, long long res_frame_stride,
              long long res_plane_stride, const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count,
              hipStream_t stream) {
    const bm3d_desc& d = p->d;
    const long long plane_stack = (long long)p->tw * 2 * p->plane_size;
    const long long corner_eval = (long long)p->nplanes * p->tw * p->plane_size;
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
        BM3D_HIP(evaluate_slots(p, b, src, ref, n0, nf, res, res_frame_stride, res_plane_stride, stream));

        bm3d::FinalizeArgs fz{};
        fz.scratch = b.regions;
        fz.fx_range = fx_range(p);
        fz.fx_exp = p->fx_exp;
        fz.wrap = b.wrap;
        fz.wrap_flag = p->wrap_flag;
        fz.corner_frame_stride = corner_eval;
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
        da.corner = b.corner;
        da.corner_frame_stride = corner_eval;
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

// Address of the current device's `scratch` array (hipGetSymbolAddress once
// per device, with that device current).
int noalloc_scratch(void** out) {
    static std::mutex mu;
    static void* ptrs[64];
    int dev = 0;
    BM3D_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(BM3D_EUNSUPPORTED, "BM3D: bm3d_execute_noalloc serves devices 0..63");
    std::lock_guard<std::mutex> g(mu);
    if (!ptrs[dev]) {
        void* ptr = nullptr;
        BM3D_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        ptrs[dev] = ptr;
    }
    *out = ptrs[dev];
    return BM3D_OK;
}

// Evaluations that can add to output frame t in target mode:
// contributors_lo(t) .. contributors_hi(t) (a slot of evaluation m covers
// frame clamp(m - r + z), z in [0, 2r]).
int contributors_lo(int t, int r) { return std::max(t - r, 0); }
int contributors_hi(int t, int r, int N) { return std::min(t + r, N - 1); }

}  // namespace

namespace bm3d {

int plan_apply_tuning(bm3d_plan* p, const bm3d_tuning& t) { return apply_tuning(p, t); }
const char* plan_tuning_error(const bm3d_tuning& t) { return tuning_range_error(t); }
int plan_temporal_keys(const bm3d_plan* p) {
    return p->d.mixed_temporal_keys ? p->d.mixed_temporal_keys : dx_temporal_keys(p->d.ps_num);
}
bool plan_uses_strip(const bm3d_plan* p) {
    return p->use_dx && p->strip && !p->mx_spatial && p->lds_bytes && strip_supported(p->tbw, p->d.block_step);
}
// A batch's matching arguments describe the strip search wherever the plan
// uses it (at radius > 0 the batch turns them into its two launches, as
// launch_matching does).
bool plan_batch_strip(const bm3d_plan* p) { return plan_uses_strip(p); }
int plan_match_owned_width(const bm3d_plan* p) {
    return plan_batch_strip(p) ? strip_owned_width(p->tbw, p->d.block_step) : p->tbw;
}
int plan_dx_variant(const bm3d_plan* p) {
    MatchArgs m{};
    m.bs = p->d.block_step;
    m.tbw = p->tbw;
    m.strip = plan_batch_strip(p) ? 1 : 0;
    m.mx_spatial = p->mx_spatial ? 1 : 0;
    return dx_spatial_variant(m);
}
MatchArgs plan_match_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref) {
    MatchArgs m = match_args(p, nullptr, nullptr, src, ref, 0);
    m.strip = plan_batch_strip(p) ? 1 : 0;
    m.obw = plan_match_owned_width(p);
    return m;
}
FilterArgs plan_filter_args(const bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref) {
    return filter_args(p, src, ref);
}
VTargetArgs plan_vtarget_args(const bm3d_plan* p) { return vtarget_args(p, 1); }

}  // namespace bm3d

namespace bm3d {

namespace {
std::atomic<long long> g_inject[kInjectSites] = {-1, -1, -1, -1};
}

bool inject_now(int site) {
    std::atomic<long long>& c = g_inject[site];
    long long v = c.load(std::memory_order_relaxed);
    while (v >= 0) {
        if (c.compare_exchange_weak(v, v - 1, std::memory_order_relaxed)) return v == 0;
    }
    return false;
}

}  // namespace bm3d

extern "C" {

int bm3d_debug_inject_failure(int site, long long after) {
    g_last_error.clear();
    if (site < bm3d::kInjectDeviceAlloc || site >= bm3d::kInjectSites)
        return fail(BM3D_EINVAL, "BM3D: unknown failure injection site");
    bm3d::g_inject[site].store(after < 0 ? -1 : after, std::memory_order_relaxed);
    return BM3D_OK;
}

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
    if (d.bm_error < 0 || d.bm_error >= bm3d::kNumCosts) return fail(BM3D_EINVAL, "BM3D: unknown bm_error");
    if (d.transform_2d < 0 || d.transform_2d >= bm3d::kNumTransforms || d.transform_1d < 0 ||
        d.transform_1d >= bm3d::kNumTransforms)
        return fail(BM3D_EINVAL, "BM3D: unknown transform");
    if (d.mixed_temporal_keys != 0 &&
        ((d.mixed_temporal_keys != 4 && d.mixed_temporal_keys != 8 && d.mixed_temporal_keys != 16) ||
         d.mixed_temporal_keys < 2 * d.ps_num))
        return fail(BM3D_EINVAL, "BM3D: mixed_temporal_keys must be 0, or 4, 8 or 16 and at least 2 ps_num");

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

    hipError_t e = hipGetDevice(&p->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&p->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device);
    if (e != hipSuccess) {
        delete p;
        return hip_fail(e, "device set-up");
    }
    bm3d_tuning t = default_tuning(p);
    (void)cached_tuning(p, &t);
    if (int rc = apply_tuning(p, t)) {
        delete p;
        return rc;
    }
    p->use_dx = p->dx_ok;
    // the strip search: about a fifth fewer instructions in the spatial search
    p->strip = true;
    p->vbound = bm3d::contribution_bound(d.transform_2d, d.transform_1d, p->d.final_ != 0);
    // the range, then the wrap flag; from byte 256 the matching launch table
    e = bm3d::malloc_device(&p->range, 256 + bm3d::kMatchTableBytes);
    if (e == hipSuccess) e = hipMemset(p->range, 0, 256);
    if (e != hipSuccess) {
        bm3d_plan_destroy(p);
        return hip_fail(e, "hipMalloc");
    }
    p->wrap_flag = reinterpret_cast<int*>(p->range) + 1;
    p->match_table = reinterpret_cast<char*>(p->range) + 256;

    if (d.precision == BM3D_PRECISION_MIXED) {
        std::vector<uint32_t> ops((size_t)bm3d::kMxOpVectors * 64 * 4);
        bm3d::build_filter_mx_ops(ops.data(), d.transform_2d, d.transform_1d);
        const size_t bytes = ops.size() * sizeof(uint32_t);
        e = bm3d::malloc_device(&p->mx_ops, bytes);
        if (e == hipSuccess) e = hipMemcpy(p->mx_ops, ops.data(), bytes, hipMemcpyHostToDevice);
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
    if (p->slist) (void)hipFree(p->slist);
    if (p->corner) (void)hipFree(p->corner);
    if (p->wrap) (void)hipFree(p->wrap);
    if (p->part) (void)hipFree(p->part);
    if (p->scratch) (void)hipFree(p->scratch);
    if (p->mx_ops) (void)hipFree(p->mx_ops);
    if (p->range) (void)hipFree(p->range);
    if (p->side) (void)hipStreamDestroy(p->side);
    if (p->fork) (void)hipEventDestroy(p->fork);
    if (p->join) (void)hipEventDestroy(p->join);
    delete p;
}

long long bm3d_stack_size(const bm3d_plan* p) { return p ? p->stack_size : 0; }

int bm3d_execute_noalloc(bm3d_plan* p, float* d_res, long long res_frame_stride, long long res_plane_stride,
                         const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count, void* stream_) {
    g_last_error.clear();
    if (int rc = check_execute(p, d_res, res_frame_stride, res_plane_stride, src, ref, first_frame, count)) return rc;
    if (count == 0) return BM3D_OK;
    // one evaluation's buffers: group lists, corner weights, wrap counts, regions
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t corner_eval = (size_t)p->nplanes * p->tw * p->plane_size;
    const size_t groups_b = up((size_t)p->nblocks * 8 * 4), corner_b = up(corner_eval * 4);
    const size_t regions_b = up((size_t)p->tw * p->region_ints * 4);
    const size_t per_eval = groups_b + 2 * corner_b + regions_b;
    if (per_eval > kNoAllocScratchBytes)
        return fail(BM3D_EINVAL, "BM3D: bm3d_execute_noalloc: one evaluation of this plan needs " +
                                     std::to_string(per_eval) + " bytes of scratch, more than its " +
                                     std::to_string(kNoAllocScratchBytes) + " (use bm3d_execute)");
    void* base = nullptr;
    if (int rc = noalloc_scratch(&base)) return rc;
    const int batch = (int)std::min<size_t>((size_t)count, kNoAllocScratchBytes / per_eval);
    char* c = static_cast<char*>(base);
    SlotBufs b;
    b.groups = reinterpret_cast<uint32_t*>(c);
    c += groups_b * batch;
    b.corner = reinterpret_cast<float*>(c);
    c += corner_b * batch;
    b.wrap = reinterpret_cast<int*>(c);
    c += corner_b * batch;
    b.regions = reinterpret_cast<int*>(c);
    b.wrap_always_zero = true;  // the array keeps other calls' counts
    b.may_alloc = false;        // matching in one launch (same results)
    return run_slots(p, b, batch, d_res, res_frame_stride, res_plane_stride, src, ref, first_frame, count,
                     reinterpret_cast<hipStream_t>(stream_));
}

int bm3d_execute(bm3d_plan* p, float* d_res, long long res_frame_stride, long long res_plane_stride,
                 const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count, void* stream_) {
    g_last_error.clear();
    if (int rc = check_execute(p, d_res, res_frame_stride, res_plane_stride, src, ref, first_frame, count)) return rc;
    if (count == 0) return BM3D_OK;
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    const long long corner_eval = (long long)p->nplanes * p->tw * p->plane_size;
    const long long scratch_eval = (long long)p->tw * p->region_ints;
    const int batch = batch_limit(p, count, corner_eval * 8 + scratch_eval * 4 + (long long)p->nblocks * 32);
    if (int rc = ensure_buffers(p, (size_t)batch * p->nblocks * 8, (size_t)batch * corner_eval,
                                (size_t)batch * scratch_eval, stream))
        return rc;
    return run_slots(p, plan_slot_bufs(p), batch, d_res, res_frame_stride, res_plane_stride, src, ref, first_frame,
                     count, stream);
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

int bm3d_plan_matching_kernel(const bm3d_plan* p) {
    if (!p || !p->use_dx) return 0;
    return p->mx_spatial ? 2 : bm3d::plan_uses_strip(p) ? 3 : 1;
}

int bm3d_plan_filter_kernel(const bm3d_plan* p) { return p && p->use_mx_filter ? 1 : 0; }

int bm3d_plan_select_kernels(bm3d_plan* p, int matching, int filter) {
    g_last_error.clear();
    if (!p) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (matching < 0 || matching > 3 || filter < 0 || filter > 1)
        return fail(BM3D_EINVAL, "BM3D: kernel selector must be 0 (FP32) or 1 (mixed precision), or for the mixed "
                                 "matcher 2 (spatial search on the matrix cores) or 3 (spatial search by strips)");
    if ((matching || filter) && p->d.precision != BM3D_PRECISION_MIXED)
        return fail(BM3D_EINVAL, "BM3D: mixed-precision kernels need a plan created with precision MIXED");
    if (matching && p->d.bm_error != BM3D_BM_SSD)
        return fail(BM3D_EUNSUPPORTED, "BM3D: the mixed-precision matching kernel computes the ssd distance only");
    if (matching && !p->dx_ok)
        return fail(BM3D_EUNSUPPORTED, "BM3D: search windows too large for the mixed-precision matching kernel");
    p->use_dx = matching != 0;
    p->mx_spatial = matching == 2;
    p->strip = matching == 3;
    p->use_mx_filter = filter != 0;
    return BM3D_OK;
}

int bm3d_plan_get_tuning(const bm3d_plan* p, bm3d_tuning* out) {
    g_last_error.clear();
    if (!p || !out) return fail(BM3D_EINVAL, "BM3D: null argument");
    *out = p->tune;
    return BM3D_OK;
}

int bm3d_plan_set_tuning(bm3d_plan* p, const bm3d_tuning* t) {
    g_last_error.clear();
    if (!p || !t) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (const char* why = tuning_range_error(*t)) return fail(BM3D_EINVAL, std::string("BM3D: tuning: ") + why);
    const bm3d_tuning old = p->tune;
    if (int rc = apply_tuning(p, *t)) {
        const std::string msg = g_last_error;
        (void)apply_tuning(p, old);
        return fail(rc, msg);
    }
    return BM3D_OK;
}

int bm3d_plan_autotune(bm3d_plan* p, const bm3d_clip* src, const bm3d_clip* ref, int first_frame, int count,
                       void* stream_) {
    g_last_error.clear();
    if (!p || !src || !src->data) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (p->d.final_ && (!ref || !ref->data)) return fail(BM3D_EINVAL, "BM3D: final pass needs \"ref\"");
    if (src->num_frames <= 0 || count <= 0 || first_frame < 0 || first_frame + count > src->num_frames)
        return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    if (p->pinned)
        return fail(BM3D_EINVAL, "BM3D: the plan is held by a captured graph; its launch shapes cannot change");
    hipStream_t stream = reinterpret_cast<hipStream_t>(stream_);
    const bm3d_desc& d = p->d;
    const bool fused = d.radius > 0 && p->frm >= p->margin;
    const size_t out_floats = fused ? (size_t)count * p->nplanes * d.height * d.width : (size_t)count * p->stack_size;
    float* out = nullptr;
    BM3D_HIP(bm3d::malloc_device(&out, out_floats * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    int rc = he == hipSuccess ? BM3D_OK : hip_fail(he, "hipEventCreate");
    auto run = [&]() -> int {
        if (fused) return bm3d_execute_vagg(p, out, 0, 0, 0, src, ref, first_frame, count, stream_);
        return bm3d_execute(p, out, 0, 0, src, ref, first_frame, count, stream_);
    };
    // milliseconds of one run with tuning t (warm), or a negative value when
    // t does not apply
    auto time_one = [&](const bm3d_tuning& t) -> double {
        if (tuning_range_error(t) || apply_tuning(p, t) != BM3D_OK) return -1.0;
        if (run() != BM3D_OK) return -1.0;
        float best = -1.0f;
        for (int rep = 0; rep < 2; ++rep) {
            if (hipEventRecord(e0, stream) != hipSuccess || run() != BM3D_OK ||
                hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)
                return -1.0;
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.0;
            best = best < 0.0f ? ms : std::min(best, ms);
        }
        return best;
    };
    bm3d_tuning best = p->tune;
    double best_ms = rc == BM3D_OK ? time_one(best) : -1.0;
    if (best_ms < 0.0 && rc == BM3D_OK)
        rc = fail(BM3D_EHIP, "BM3D: autotune: the plan's launch shapes failed: " + g_last_error);
    auto consider = [&](const bm3d_tuning& t) {
        const double ms = time_one(t);
        if (ms >= 0.0 && ms < best_ms) {
            best_ms = ms;
            best = t;
        }
    };
    if (rc == BM3D_OK) {
        // coordinate descent: matching rows, then filter tile and waves,
        // then batch size and evaluation runs
        const int shapes[][2] = {{16, 16}, {16, 8}, {16, 4}, {32, 8}, {32, 4}};
        for (auto& sh : shapes) {
            bm3d_tuning t = best;
            t.match_width = sh[0];
            t.match_rows = sh[1];
            consider(t);
        }
        const int tiles[][3] = {{8, 8, 4}, {8, 8, 8}, {16, 8, 4}, {16, 8, 8}, {8, 16, 8}, {16, 16, 8},
                                {8, 4, 4}, {8, 4, 2}, {4, 4, 2}};
        const bm3d_tuning base = best;
        for (auto& f : tiles) {
            bm3d_tuning t = base;
            t.filter_tile_x = f[0];
            t.filter_tile_y = f[1];
            t.filter_waves = f[2];
            consider(t);
        }
        const bm3d_tuning base2 = best;
        for (int b : {8, 16, 32, 64, 128}) {
            for (int c : {1, 2, 4}) {
                bm3d_tuning t = base2;
                t.batch_frames = b;
                t.chunks = std::min(c, b);
                if (c > 1 && !fused) continue;
                consider(t);
            }
        }
        (void)apply_tuning(p, best);
        store_tuning(p, p->tune);
        g_last_error.clear();
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    return rc;
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
    BM3D_HIP(bm3d::malloc_device(&d_cnt, sizeof(unsigned long long)));
    hipError_t e = bm3d::launch_zero(reinterpret_cast<float*>(d_cnt), 2, 0, 1, st);  // the counter's 64 bits
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
    // output frames the evaluations add to
    const int tlo_all = std::max(mlo - r, 0), thi_all = std::min(mhi + r, N - 1);
    const long long corner_target = (long long)p->nplanes * p->plane_size;
    // per resident output frame: corner map, wrap counts, 64-bit partial sums, regions
    const long long target_bytes = corner_target * 16 + p->region_ints * 4;
    const int batch = batch_limit(p, mhi - mlo + 1, target_bytes + (long long)p->nblocks * 32, 2 * r * target_bytes);
    // resident output frames: those a launch adds to (its evaluations +- r)
    const int ring = std::min(batch + 2 * r, thi_all - tlo_all + 1);
    if (int rc = ensure_buffers(p, (size_t)batch * p->nblocks * 8, (size_t)ring * corner_target,
                                (size_t)ring * p->region_ints, stream, (size_t)ring * corner_target))
        return rc;

    bm3d::FilterArgs fa = filter_args(p, src, ref);
    fa.mode = 1;
    fa.part = p->part;
    fa.ring = ring;
    fa.target_stride = p->region_ints;
    fa.corner_frame_stride = corner_target;

    bm3d::VTargetArgs va = vtarget_args(p, ring);
    va.dst_frame_stride = dst_frame_stride;
    va.dst_plane_stride = dst_plane_stride;
    va.dst_stride = dst_stride;
    va.part = p->part;
    // output frames whose contributors span several runs (64-bit sums)
    std::vector<char> partial(thi_all - tlo_all + 1, 0);

    // the input range and the first launch's zero fills run on the side
    // stream, beside the first matching launch (which needs neither); the
    // first filtering launch waits for them
    // (created on first use outside a graph capture, since batches never
    // need it; a call captured before that runs on one stream)
    if (!p->side) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        BM3D_HIP(hipStreamIsCapturing(stream, &cs));
        if (cs == hipStreamCaptureStatusNone) {
            BM3D_HIP(hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking));
            BM3D_HIP(hipEventCreateWithFlags(&p->fork, hipEventDisableTiming));
            BM3D_HIP(hipEventCreateWithFlags(&p->join, hipEventDisableTiming));
        }
    }
    hipStream_t side = stream;
    if (p->side && p->fork && p->join) {
        side = p->side;
        BM3D_HIP(hipEventRecord(p->fork, stream));
        BM3D_HIP(hipStreamWaitEvent(side, p->fork, 0));
    }
    BM3D_HIP(measure_range(p, src, std::max(mlo - r, 0), std::min(mhi + r, N - 1), side));
    int next_out = first_frame;
    int touched_hi = tlo_all - 1;  // highest output frame a previous launch added to
    for (int b0 = mlo; b0 <= mhi;) {
        const bool first_launch = b0 == mlo;
        hipStream_t zs = first_launch ? side : stream;
        const int b1 = std::min(b0 + batch, mhi + 1);
        const int nf = b1 - b0;
        const int chunk = (nf + p->tune.chunks - 1) / std::max(1, p->tune.chunks);
        const int nchunks = (nf + chunk - 1) / chunk;
        // output frames first added to by this launch start from zero: the
        // corner maps always (float atomics), the regions when they are
        // summed atomically (contributions from more than one run)
        const int tnew_lo = std::max(std::max(b0 - r, 0), touched_hi + 1);
        const int tnew_hi = std::min(b1 - 1 + r, N - 1);
        for (int t = tnew_lo; t <= tnew_hi;) {
            const int slot = t % ring;
            const int run = std::min(tnew_hi - t + 1, ring - slot);
            BM3D_HIP(bm3d::launch_zero(p->corner + (size_t)slot * corner_target, (size_t)run * corner_target, 0, 1,
                                       zs));
            BM3D_HIP(bm3d::launch_zero(reinterpret_cast<float*>(p->wrap) + (size_t)slot * corner_target,
                                       (size_t)run * corner_target, 0, 1, zs, p->wrap_flag));
            t += run;
        }
        for (int t = tnew_lo; t <= tnew_hi; ++t) {
            const int clo = contributors_lo(t, r), chi = contributors_hi(t, r, N);
            const int k = (clo - b0) / chunk;
            const bool whole = clo >= b0 && chi < b1 && chi < b0 + (k + 1) * chunk;
            partial[t - tlo_all] = !whole;
            if (!whole)
                BM3D_HIP(bm3d::launch_zero(reinterpret_cast<float*>(p->part + (size_t)(t % ring) * corner_target),
                                           (size_t)corner_target * 2, 0, 1, zs));
        }
        touched_hi = std::max(touched_hi, tnew_hi);
        const bool joined = first_launch && side != stream;
        if (joined) BM3D_HIP(hipEventRecord(p->join, side));
        {
            ProfScope ps(0, stream);
            BM3D_HIP(launch_matching(p, p->groups, nullptr, src, ref, b0, nf, stream));
        }
        if (joined) BM3D_HIP(hipStreamWaitEvent(stream, p->join, 0));
        fa.eval_lo = b0;
        fa.eval_hi = b1;
        fa.chunk = chunk;
        BM3D_HIP(launch_filtering(p, fa, nchunks, stream));
        // outputs whose evaluations are all done
        const int out_hi = (b1 - 1 == N - 1) ? first_frame + count - 1
                                              : std::min(first_frame + count - 1, b1 - 1 - r);
        while (out_hi >= next_out) {
            const int n = std::min(out_hi - next_out + 1, bm3d::kVaggMaxFrames);
            va.first_frame = next_out;
            va.dst = dst + (size_t)(next_out - first_frame) * dst_frame_stride;
            std::memset(va.partial, 0, sizeof va.partial);
            for (int f = 0; f < n; ++f)
                if (partial[next_out + f - tlo_all]) va.partial[f >> 5] |= 1u << (f & 31);
            BM3D_HIP(bm3d::launch_vagg_targets(va, n, stream));
            next_out += n;
        }
        b0 = b1;
    }
    if (next_out != first_frame + count) return fail(BM3D_EHIP, "BM3D: internal frame schedule error");
    return BM3D_OK;
}

int bm3d_execute_window(bm3d_plan* p, float* d_res, const float* d_src, void* stream) {
    g_last_error.clear();
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
