// Host-frame path of the plugin-level filters (bm3d_filter_get_frames_host):
// the route a VapourSynth GetFrame takes when its frames live in host memory.
//
// The reference stages every GetFrame through pinned memory on one of kFast
// resource sets (source/source.cpp:296-426). It bitblts each of the 2r + 1
// window slots (twice in the final pass), runs a graph of HtoD, kernel and
// DtoH of the whole stack, syncs the stream and bitblts the stack out;
// VAggregate then runs on the host. A call here instead:
//   * uploads each input frame of its range (lo .. hi) once, through pinned
//     buffers that a pool of host threads fills row block by row block;
//   * runs the device path (bm3d_filter_get_frames) on the device copies;
//   * downloads only the output frames (for BM3Dv2 the denoised frames, never
//     the stacks);
//   * splits ranges of more than kChunk output frames into chunks whose
//     upload, compute and download overlap on three streams (events order
//     them; pinned buffers are double-buffered);
//   * runs on one of up to kLanes resource sets (the reference's fast =
//     kFast = 4; 1 with fast = 0). Each set is an independent copy of the
//     filter with its own plans, streams and buffers, so concurrent callers
//     overlap their copies and kernels.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bm3d_inject.h"
#include "bm3d_internal.h"

using bm3d::fail;

namespace bm3d {
namespace {

constexpr int kLanes = 4;    // source/source.cpp:62 (kFast)
constexpr int kChunk = 16;   // output frames per pipelined chunk

#define BM3D_HOST_HIP(expr)                                                                          \
    do {                                                                                             \
        hipError_t _e = (expr);                                                                      \
        if (_e != hipSuccess)                                                                        \
            return fail(BM3D_EHIP, std::string("BM3D: '") + #expr + "' failed: " + hipGetErrorString(_e)); \
    } while (0)

// ------------------------------------------------------------ host copies

struct Copy2D {
    char* dst;
    size_t dpitch;
    const char* src;
    size_t spitch;
    size_t width;  // bytes per row
    int rows;
};

// A process-wide pool of host threads for pageable <-> pinned row copies.
// run() splits its copies into pieces of about 1 MiB, lets the workers and
// the calling thread take pieces, and returns when all are done.
class CopyPool {
public:
    static CopyPool& get() {
        static CopyPool pool;
        return pool;
    }

    void run(const std::vector<Copy2D>& copies) {
        auto b = std::make_unique<Batch>();
        for (const Copy2D& c : copies) {
            if (c.rows <= 0 || c.width == 0) continue;
            const int per = (int)std::max<size_t>(1, (size_t(1) << 20) / c.width);
            for (int r = 0; r < c.rows; r += per) {
                Copy2D p = c;
                p.dst += (size_t)r * c.dpitch;
                p.src += (size_t)r * c.spitch;
                p.rows = std::min(per, c.rows - r);
                b->pieces.push_back(p);
            }
        }
        if (b->pieces.empty()) return;
        if (b->pieces.size() == 1 || threads_.empty()) {
            for (const Copy2D& p : b->pieces) copy(p);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            queue_.push_back(b.get());
            ++b->users;
        }
        cv_.notify_all();
        work(b.get());
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [&] { return b->done == b->pieces.size() && b->users == 0; });
        auto it = std::find(queue_.begin(), queue_.end(), b.get());
        if (it != queue_.end()) queue_.erase(it);
    }

private:
    struct Batch {
        std::vector<Copy2D> pieces;
        std::atomic<size_t> next{0};
        size_t done = 0;  // guarded by mu_
        int users = 0;    // threads working on it, guarded by mu_
    };

    CopyPool() {
        const unsigned hc = std::thread::hardware_concurrency();
        const unsigned n = std::min(8u, std::max(2u, hc / 2));
        for (unsigned i = 0; i < n; ++i) threads_.emplace_back([this] { worker(); });
    }
    ~CopyPool() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : threads_) t.join();
    }

    static void copy(const Copy2D& c) {
        if (c.dpitch == c.width && c.spitch == c.width) {
            std::memcpy(c.dst, c.src, c.width * (size_t)c.rows);
            return;
        }
        for (int r = 0; r < c.rows; ++r) std::memcpy(c.dst + (size_t)r * c.dpitch, c.src + (size_t)r * c.spitch, c.width);
    }

    // Takes pieces of b until none is left; the caller holds a use of b.
    void work(Batch* b) {
        size_t n = 0;
        for (size_t i; (i = b->next.fetch_add(1)) < b->pieces.size(); ++n) copy(b->pieces[i]);
        std::lock_guard<std::mutex> g(mu_);
        b->done += n;
        --b->users;
        if (b->done == b->pieces.size() && b->users == 0) done_cv_.notify_all();
    }

    void worker() {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
            if (stop_) return;
            Batch* b = queue_.front();
            if (b->next.load() >= b->pieces.size()) {
                queue_.pop_front();  // exhausted: its owner finishes it
                continue;
            }
            ++b->users;
            lk.unlock();
            work(b);
            lk.lock();
        }
    }

    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<Batch*> queue_;
    std::vector<std::thread> threads_;
    bool stop_ = false;
};

// -------------------------------------------------------------- lanes

struct Lane {
    bm3d_filter* f = nullptr;  // this set's own copy of the filter
    hipStream_t up = nullptr, comp = nullptr, down = nullptr;
    hipEvent_t up_done[2]{}, computed[2]{}, down_done[2]{};
    float* d_in = nullptr;
    size_t d_in_cap = 0;  // floats
    float* d_out = nullptr;
    size_t d_out_cap = 0;
    float* h_in[2]{};
    float* h_out[2]{};
    size_t h_in_cap = 0, h_out_cap = 0;  // floats per buffer

    ~Lane() {
        if (up) (void)hipStreamSynchronize(up);
        if (comp) (void)hipStreamSynchronize(comp);
        if (down) (void)hipStreamSynchronize(down);
        for (int i = 0; i < 2; ++i) {
            if (up_done[i]) (void)hipEventDestroy(up_done[i]);
            if (computed[i]) (void)hipEventDestroy(computed[i]);
            if (down_done[i]) (void)hipEventDestroy(down_done[i]);
            if (h_in[i]) (void)hipHostFree(h_in[i]);
            if (h_out[i]) (void)hipHostFree(h_out[i]);
        }
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        if (up) (void)hipStreamDestroy(up);
        if (comp) (void)hipStreamDestroy(comp);
        if (down) (void)hipStreamDestroy(down);
        bm3d_filter_destroy(f);
    }
};

int new_lane(const bm3d_filter* f, std::unique_ptr<Lane>* out) {
    auto l = std::make_unique<Lane>();
    if (int rc = clone_filter(f, &l->f)) return rc;
    BM3D_HOST_HIP(hipStreamCreateWithFlags(&l->up, hipStreamNonBlocking));
    BM3D_HOST_HIP(hipStreamCreateWithFlags(&l->comp, hipStreamNonBlocking));
    BM3D_HOST_HIP(hipStreamCreateWithFlags(&l->down, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        BM3D_HOST_HIP(hipEventCreateWithFlags(&l->up_done[i], hipEventDisableTiming));
        BM3D_HOST_HIP(hipEventCreateWithFlags(&l->computed[i], hipEventDisableTiming));
        BM3D_HOST_HIP(hipEventCreateWithFlags(&l->down_done[i], hipEventDisableTiming));
    }
    *out = std::move(l);
    return BM3D_OK;
}

int grow_device(float** p, size_t* cap, size_t floats) {
    if (floats <= *cap) return BM3D_OK;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    BM3D_HOST_HIP(bm3d::malloc_device(p, floats * sizeof(float)));
    *cap = floats;
    return BM3D_OK;
}

int grow_pinned(float* (&p)[2], size_t* cap, size_t floats) {
    if (floats <= *cap) return BM3D_OK;
    for (auto& q : p) {
        if (q) (void)hipHostFree(q);
        q = nullptr;
    }
    *cap = 0;
    for (auto& q : p) BM3D_HOST_HIP(bm3d::malloc_pinned(&q, floats * sizeof(float)));
    *cap = floats;
    return BM3D_OK;
}

int plane_w(const bm3d_video_info& v, int p) { return p == 0 ? v.width : v.width >> v.subsampling_w; }
int plane_h(const bm3d_video_info& v, int p) { return p == 0 ? v.height : v.height >> v.subsampling_h; }
int padded(int w) { return (w + 63) & ~63; }

// Frames reach: output frames [first, first + count) read input frames
// [first - reach, first + count - 1 + reach] (clamped).
int reach(const bm3d_filter* f) {
    if (f->kind == Kind::kBM3Dv2) return f->passthrough ? 0 : 2 * f->radius;
    return f->radius;
}

// Device layout of one clip's planes over `frames` consecutive frames:
// plane p is a block of frames x h[p] x s[p] floats (chroma planes, all of
// one size, are then equally spaced, as the chroma plans need).
struct Layout {
    int np = 1;
    int w[3]{}, h[3]{}, s[3]{};
    size_t frame[3]{};  // floats per frame per plane
    size_t per_frame = 0;
    void init(const bm3d_video_info& v) {
        np = std::min(v.num_planes, 3);
        per_frame = 0;
        for (int p = 0; p < np; ++p) {
            w[p] = plane_w(v, p);
            h[p] = plane_h(v, p);
            s[p] = padded(w[p]);
            frame[p] = (size_t)h[p] * s[p];
            per_frame += frame[p];
        }
    }
};

// bm3d_frames over a device block holding frames base .. base + frames - 1:
// data[p] addresses frame 0, so frame n is at data[p] + n * frame_stride[p].
bm3d_frames device_frames(const Layout& L, float* block, int base, int frames) {
    bm3d_frames d{};
    size_t off = 0;
    for (int p = 0; p < L.np; ++p) {
        const uintptr_t start = reinterpret_cast<uintptr_t>(block + off);
        d.data[p] = reinterpret_cast<float*>(start - (uintptr_t)base * L.frame[p] * sizeof(float));
        d.frame_stride[p] = (long long)L.frame[p];
        d.stride[p] = L.s[p];
        off += (size_t)frames * L.frame[p];
    }
    return d;
}

// Adds the copies that stage host frames n0 .. n1 of one clip (host[i] is
// frame host_lo + i) into pinned memory at `pin`, as plane blocks of
// n1 - n0 + 1 frames. Returns the floats of pinned memory they fill.
size_t stage_in(const Layout& L, const bm3d_host_frame* host, int host_lo, int n0, int n1, float* pin,
                std::vector<Copy2D>& copies) {
    const int n = n1 - n0 + 1;
    size_t off = 0;
    for (int p = 0; p < L.np; ++p) {
        for (int k = 0; k < n; ++k) {
            const bm3d_host_frame& hf = host[n0 + k - host_lo];
            copies.push_back({reinterpret_cast<char*>(pin + off + (size_t)k * L.frame[p]), (size_t)L.s[p] * 4,
                              reinterpret_cast<const char*>(hf.data[p]), (size_t)hf.stride[p] * 4, (size_t)L.w[p] * 4,
                              L.h[p]});
        }
        off += (size_t)n * L.frame[p];
    }
    return off;
}

// Queues the HtoD copies of staged frames n0 .. n1 (see stage_in).
hipError_t upload(const Layout& L, int n0, int n1, const float* pin, float* block, int base, int frames,
                  hipStream_t s) {
    const int n = n1 - n0 + 1;
    size_t pin_off = 0, dev_off = 0;
    for (int p = 0; p < L.np; ++p) {
        const size_t bytes = (size_t)n * L.frame[p] * sizeof(float);
        hipError_t e = hipMemcpyAsync(block + dev_off + (size_t)(n0 - base) * L.frame[p], pin + pin_off, bytes,
                                      hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        pin_off += (size_t)n * L.frame[p];
        dev_off += (size_t)frames * L.frame[p];
    }
    return hipSuccess;
}

int validate_frames(const Layout& L, const bm3d_host_frame* fr, int n, const char* what) {
    for (int i = 0; i < n; ++i)
        for (int p = 0; p < L.np; ++p)
            if (!fr[i].data[p] || fr[i].stride[p] < L.w[p])
                return fail(BM3D_EINVAL, std::string("BM3D: ") + what + " host frame " + std::to_string(i) +
                                             " plane " + std::to_string(p) + " is null or its stride is too small");
    return BM3D_OK;
}

}  // namespace

struct HostPool {
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<Lane*> idle;
    int building = 0;  // lanes being created
    int max_lanes = kLanes;
};

void destroy_host_pool(HostPool* pool) { delete pool; }

namespace {

// Runs one call on lane l (the lane's filter is a copy of f).
int run_call(const bm3d_filter* f, Lane& l, const bm3d_host_frame* clip, const bm3d_host_frame* ref, int lo, int hi,
             bm3d_host_frame* out, int first, int count) {
    const bool vagg = f->kind == Kind::kVAggregate;
    const bool has_ref = vagg || (f->kind == Kind::kBM3D ? f->final_ : (f->bm3d && f->bm3d->final_));
    Layout Lc, Lr, Lo;
    Lc.init(vagg ? f->recipe.clip : f->vi);  // VAggregate reads BM3D stacks
    Lr.init(f->vi);
    Lo.init(f->out_vi);
    const int nin = hi - lo + 1;
    // ref frames kept on the device: the input range, or (VAggregate: src)
    // the output range
    const int rlo = vagg ? first : lo, rn = vagg ? count : nin;
    if (int rc = validate_frames(Lc, clip, nin, "clip")) return rc;
    if (has_ref) {
        if (!ref) return fail(BM3D_EINVAL, vagg ? "VAggregate: \"src\" frames missing" : "BM3D: \"ref\" frames missing");
        if (int rc = validate_frames(Lr, ref + (rlo - lo), rn, "ref")) return rc;
    }
    if (int rc = validate_frames(Lo, out, count, "out")) return rc;

    const size_t in_clip = (size_t)nin * Lc.per_frame, in_ref = has_ref ? (size_t)rn * Lr.per_frame : 0;
    if (int rc = grow_device(&l.d_in, &l.d_in_cap, in_clip + in_ref)) return rc;
    if (int rc = grow_device(&l.d_out, &l.d_out_cap, (size_t)count * Lo.per_frame)) return rc;
    float* d_clip = l.d_in;
    float* d_ref = l.d_in + in_clip;

    // chunks of output frames and the input frames each adds
    const int nchunks = (count + kChunk - 1) / kChunk;
    const int csize = (count + nchunks - 1) / nchunks;
    struct Chunk {
        int f0, n;        // output frames
        int c0, c1;       // new clip frames (c0 > c1: none)
        int r0, r1;       // new ref frames
    };
    std::vector<Chunk> ch;
    int clip_hi = lo - 1, ref_hi = rlo - 1;
    size_t pin_in = 0, pin_out = 0;
    for (int c = 0; c < nchunks; ++c) {
        Chunk k{};
        k.f0 = first + c * csize;
        k.n = std::min(csize, first + count - k.f0);
        int clo, chi;
        bm3d_filter_host_range(f, k.f0, k.n, &clo, &chi);
        k.c0 = clip_hi + 1;
        k.c1 = chi;
        clip_hi = std::max(clip_hi, chi);
        const int rhi = vagg ? k.f0 + k.n - 1 : chi;
        k.r0 = ref_hi + 1;
        k.r1 = has_ref ? rhi : ref_hi;
        ref_hi = std::max(ref_hi, k.r1);
        const size_t need = (size_t)std::max(0, k.c1 - k.c0 + 1) * Lc.per_frame +
                            (size_t)std::max(0, k.r1 - k.r0 + 1) * Lr.per_frame;
        pin_in = std::max(pin_in, need);
        pin_out = std::max(pin_out, (size_t)k.n * Lo.per_frame);
        ch.push_back(k);
    }
    if (int rc = grow_pinned(l.h_in, &l.h_in_cap, pin_in)) return rc;
    if (int rc = grow_pinned(l.h_out, &l.h_out_cap, pin_out)) return rc;

    const bm3d_frames dclip = device_frames(Lc, d_clip, lo, nin);
    const bm3d_frames dref = has_ref ? device_frames(Lr, d_ref, rlo, rn) : bm3d_frames{};
    const bm3d_frames dout = device_frames(Lo, l.d_out, first, count);

    // downloads chunk c's outputs into pinned set c % 2, then to the caller
    auto finish = [&](int c) -> int {
        const int set = c % 2;
        BM3D_HOST_HIP(hipEventSynchronize(l.down_done[set]));
        std::vector<Copy2D> copies;
        size_t off = 0;
        for (int p = 0; p < Lo.np; ++p) {
            for (int k = 0; k < ch[c].n; ++k) {
                bm3d_host_frame& hf = out[ch[c].f0 - first + k];
                copies.push_back({reinterpret_cast<char*>(hf.data[p]), (size_t)hf.stride[p] * 4,
                                  reinterpret_cast<const char*>(l.h_out[set] + off + (size_t)k * Lo.frame[p]),
                                  (size_t)Lo.s[p] * 4, (size_t)Lo.w[p] * 4, Lo.h[p]});
            }
            off += (size_t)ch[c].n * Lo.frame[p];
        }
        CopyPool::get().run(copies);
        return BM3D_OK;
    };

    for (int c = 0; c < nchunks; ++c) {
        const int set = c % 2;
        const Chunk& k = ch[c];
        // stage and upload this chunk's new input frames (the pinned set was
        // last used by chunk c - 2, whose upload must be complete)
        BM3D_HOST_HIP(hipEventSynchronize(l.up_done[set]));
        std::vector<Copy2D> copies;
        float* pin = l.h_in[set];
        size_t used_clip = 0;
        if (k.c1 >= k.c0) used_clip = stage_in(Lc, clip, lo, k.c0, k.c1, pin, copies);
        if (has_ref && k.r1 >= k.r0) (void)stage_in(Lr, ref, lo, k.r0, k.r1, pin + used_clip, copies);
        CopyPool::get().run(copies);
        if (k.c1 >= k.c0) BM3D_HOST_HIP(upload(Lc, k.c0, k.c1, pin, d_clip, lo, nin, l.up));
        if (has_ref && k.r1 >= k.r0) BM3D_HOST_HIP(upload(Lr, k.r0, k.r1, pin + used_clip, d_ref, rlo, rn, l.up));
        BM3D_HOST_HIP(hipEventRecord(l.up_done[set], l.up));
        // compute once every frame it reads is on the device
        BM3D_HOST_HIP(hipStreamWaitEvent(l.comp, l.up_done[set], 0));
        if (int rc = bm3d_filter_get_frames(l.f, &dclip, has_ref ? &dref : nullptr, &dout, k.f0, k.n, l.comp)) {
            (void)hipStreamSynchronize(l.comp);
            return rc;
        }
        BM3D_HOST_HIP(hipEventRecord(l.computed[set], l.comp));
        // download into pinned set c % 2 (chunk c - 2 has been copied out)
        BM3D_HOST_HIP(hipStreamWaitEvent(l.down, l.computed[set], 0));
        size_t off = 0, dev_off = 0;
        for (int p = 0; p < Lo.np; ++p) {
            BM3D_HOST_HIP(hipMemcpyAsync(l.h_out[set] + off, l.d_out + dev_off + (size_t)(k.f0 - first) * Lo.frame[p],
                                         (size_t)k.n * Lo.frame[p] * sizeof(float), hipMemcpyDeviceToHost, l.down));
            off += (size_t)k.n * Lo.frame[p];
            dev_off += (size_t)count * Lo.frame[p];
        }
        BM3D_HOST_HIP(hipEventRecord(l.down_done[set], l.down));
        if (c > 0)
            if (int rc = finish(c - 1)) return rc;
    }
    return finish(nchunks - 1);
}

int acquire_lane(bm3d_filter* f, Lane** out) {
    {
        static std::mutex create_mu;
        std::lock_guard<std::mutex> g(create_mu);
        if (!f->host) {
            f->host = new HostPool;
            f->host->max_lanes = f->fast ? kLanes : 1;
        }
    }
    HostPool& P = *f->host;
    std::unique_lock<std::mutex> lk(P.mu);
    for (;;) {
        if (!P.idle.empty()) {
            *out = P.idle.back();
            P.idle.pop_back();
            return BM3D_OK;
        }
        if ((int)P.lanes.size() + P.building < P.max_lanes) {
            ++P.building;  // build outside the lock
            lk.unlock();
            std::unique_ptr<Lane> l;
            const int rc = new_lane(f, &l);
            lk.lock();
            --P.building;
            if (rc != BM3D_OK) {
                P.cv.notify_all();
                return rc;
            }
            P.lanes.push_back(std::move(l));
            *out = P.lanes.back().get();
            return BM3D_OK;
        }
        P.cv.wait(lk);
    }
}

void release_lane(bm3d_filter* f, Lane* l) {
    {
        std::lock_guard<std::mutex> g(f->host->mu);
        f->host->idle.push_back(l);
    }
    f->host->cv.notify_one();
}

}  // namespace
}  // namespace bm3d

extern "C" {

int bm3d_filter_host_range(const bm3d_filter* f, int first, int count, int* lo, int* hi) {
    if (!f || !lo || !hi) return fail(BM3D_EINVAL, "BM3D: null argument");
    if (first < 0 || count <= 0 || first + count > f->num_frames)
        return fail(BM3D_EINVAL, "BM3D: frame range out of bounds");
    const int r = bm3d::reach(f);
    *lo = std::max(first - r, 0);
    *hi = std::min(first + count - 1 + r, f->num_frames - 1);
    return BM3D_OK;
}

int bm3d_filter_get_frames_host(bm3d_filter* f, const bm3d_host_frame* clip, const bm3d_host_frame* ref, int lo,
                                bm3d_host_frame* out, int first, int count) {
    if (!f || !clip || !out) return fail(BM3D_EINVAL, "BM3D: null argument");
    int want_lo = 0, hi = 0;
    if (int rc = bm3d_filter_host_range(f, first, count, &want_lo, &hi)) return rc;
    if (lo != want_lo) return fail(BM3D_EINVAL, "BM3D: host frames must start at frame lo of bm3d_filter_host_range");
    if (f->kind == bm3d::Kind::kBM3Dv2 && f->passthrough) {  // the source clip itself (source/source.cpp:932-942)
        bm3d::Layout L;
        L.init(f->vi);
        if (int rc = bm3d::validate_frames(L, clip, count, "clip")) return rc;
        if (int rc = bm3d::validate_frames(L, out, count, "out")) return rc;
        std::vector<bm3d::Copy2D> copies;
        for (int k = 0; k < count; ++k)
            for (int p = 0; p < L.np; ++p)
                copies.push_back({reinterpret_cast<char*>(out[k].data[p]), (size_t)out[k].stride[p] * 4,
                                  reinterpret_cast<const char*>(clip[k].data[p]), (size_t)clip[k].stride[p] * 4,
                                  (size_t)L.w[p] * 4, L.h[p]});
        bm3d::CopyPool::get().run(copies);
        return BM3D_OK;
    }
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    const int dev = f->kind == bm3d::Kind::kBM3Dv2 && f->bm3d ? f->bm3d->device : f->device;
    if (hipSetDevice(dev) != hipSuccess) return fail(BM3D_EHIP, "BM3D: hipSetDevice failed");
    bm3d::Lane* lane = nullptr;
    int rc = bm3d::acquire_lane(f, &lane);
    if (rc == BM3D_OK) {
        rc = bm3d::run_call(f, *lane, clip, ref, lo, hi, out, first, count);
        if (rc != BM3D_OK) {
            (void)hipStreamSynchronize(lane->up);
            (void)hipStreamSynchronize(lane->comp);
            (void)hipStreamSynchronize(lane->down);
        }
        bm3d::release_lane(f, lane);
    }
    if (prev >= 0) (void)hipSetDevice(prev);
    return rc;
}

}  // extern "C"
