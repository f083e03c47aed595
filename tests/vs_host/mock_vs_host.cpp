// A minimal VapourSynth API-3 host for testing the plugin entry: it loads a
// plugin, records its registrations, builds source clips from raw float
// files, invokes the registered functions with typed argument maps, and
// pulls frames through the filter graph (arInitial, then the requested
// frames, then arAllFramesReady) from several threads at once. Nodes and
// frames are reference counted as in VapourSynth, so a sanitizer build of
// host and plugin sees every leak and double free.
//
// Commands, one per line on stdin:
//   plugin <path.so>
//   source <name> <GRAYS|GRAY8|GRAYH|YUV444PS|YUV420PS|RGBS> <W> <H> <N> <raw>
//   invoke <name> <function> key=i:1,2 key=f:1.5 key=s:dct,haar key=n:<node> ...
//   fetch <name> <raw out> <threads>
//   props <name> <frame>
//   inject <library.so> <site> <after>
//   free <name>
// Raw clips are [frame][plane][rows][cols] float32 without padding.
#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "vs_api3.h"

struct Prop {
    char type = 'u';  // i, f, s, c (node), v (frame)
    std::vector<int64_t> i;
    std::vector<double> f;
    std::vector<std::string> s;
    std::vector<VSNodeRef*> n;
    std::vector<const VSFrameRef*> v;
};

struct VSMap {
    std::map<std::string, Prop> props;
    std::string error;
    bool has_error = false;
    VSMap() = default;
    VSMap(const VSMap&) = delete;
    ~VSMap();
};

struct VSCore {
    int unused = 0;
};

struct VSFrameRef {
    std::atomic<int> refs{1};
    const VSFormat* fmt = nullptr;
    int w[3] = {0, 0, 0}, h[3] = {0, 0, 0}, stride[3] = {0, 0, 0};
    uint8_t* data[3] = {nullptr, nullptr, nullptr};
    VSMap props;
    ~VSFrameRef() {
        for (auto* p : data) std::free(p);
    }
};

struct Node {
    std::atomic<int> refs{1};
    VSVideoInfo vi{};
    std::vector<const VSFrameRef*> frames;  // source clips
    VSFilterGetFrame get = nullptr;          // filters
    VSFilterFree free_fn = nullptr;
    void* inst = nullptr;
    std::string name;
};

struct VSNodeRef {
    Node* node;
};

struct VSFrameContext {
    std::vector<std::pair<Node*, int>> requests;
    std::map<std::pair<Node*, int>, const VSFrameRef*> ready;
    std::string error;
};

namespace {

VSCore g_core;
VSAPI g_api{};

const VSFormat kFormats[] = {
    {"GRAYS", 1, cmGray, stFloat, 32, 4, 0, 0, 1},    {"GRAY8", 2, cmGray, stInteger, 8, 1, 0, 0, 1},
    {"GRAYH", 3, cmGray, stFloat, 16, 2, 0, 0, 1},    {"YUV444PS", 4, cmYUV, stFloat, 32, 4, 0, 0, 3},
    {"YUV420PS", 5, cmYUV, stFloat, 32, 4, 1, 1, 3},  {"RGBS", 6, cmRGB, stFloat, 32, 4, 0, 0, 3},
};

const VSFormat* find_format(const std::string& name) {
    for (const auto& f : kFormats)
        if (name == f.name) return &f;
    return nullptr;
}

// ---------------------------------------------------------------- nodes / frames
void node_release(Node* n) {
    if (--n->refs == 0) {
        if (n->free_fn) n->free_fn(n->inst, &g_core, &g_api);
        for (auto* f : n->frames) g_api.freeFrame(f);
        delete n;
    }
}

VSNodeRef* new_ref(Node* n) {
    ++n->refs;
    return new VSNodeRef{n};
}

VSFrameRef* new_frame(const VSFormat* fmt, int width, int height) {
    auto* f = new VSFrameRef;
    f->fmt = fmt;
    for (int p = 0; p < fmt->numPlanes; ++p) {
        f->w[p] = p ? width >> fmt->subSamplingW : width;
        f->h[p] = p ? height >> fmt->subSamplingH : height;
        // a row pitch wider than the row, as VapourSynth's aligned frames have
        f->stride[p] = ((f->w[p] * fmt->bytesPerSample + 63) / 64) * 64 + 64;
        const size_t bytes = (size_t)f->stride[p] * f->h[p];
        f->data[p] = static_cast<uint8_t*>(std::aligned_alloc(64, bytes));
        std::memset(f->data[p], 0xff, bytes);  // NaN until written
    }
    return f;
}

void copy_props(const VSMap* src, VSMap* dst) {
    for (const auto& kv : src->props) {
        Prop p = kv.second;
        for (auto*& r : p.n) r = new_ref(r->node);
        for (auto*& v : p.v) v = g_api.cloneFrameRef(v);
        dst->props[kv.first] = std::move(p);
    }
}

const VSFrameRef* evaluate(Node* node, int n, std::string* err) {
    n = std::clamp(n, 0, node->vi.numFrames - 1);
    if (!node->get) return g_api.cloneFrameRef(node->frames[n]);
    VSFrameContext ctx;
    void* frame_data = nullptr;
    const VSFrameRef* f = node->get(n, arInitial, &node->inst, &frame_data, &ctx, &g_core, &g_api);
    if (f) g_api.freeFrame(f);
    for (size_t i = 0; i < ctx.requests.size() && ctx.error.empty(); ++i) {
        const auto req = ctx.requests[i];
        if (ctx.ready.count(req)) continue;
        const VSFrameRef* r = evaluate(req.first, req.second, err);
        if (!r) ctx.error = *err;
        else ctx.ready[req] = r;
    }
    f = nullptr;
    if (ctx.error.empty()) f = node->get(n, arAllFramesReady, &node->inst, &frame_data, &ctx, &g_core, &g_api);
    for (auto& kv : ctx.ready) g_api.freeFrame(kv.second);
    if (!ctx.error.empty() || !f) {
        if (f) g_api.freeFrame(f);
        *err = ctx.error.empty() ? "filter returned no frame" : ctx.error;
        return nullptr;
    }
    return f;
}

// ---------------------------------------------------------------- VSAPI
const VSFrameRef* VS_CC clone_frame(const VSFrameRef* f) {
    ++const_cast<VSFrameRef*>(f)->refs;
    return f;
}
VSNodeRef* VS_CC clone_node(VSNodeRef* r) { return new_ref(r->node); }
void VS_CC free_frame(const VSFrameRef* f) {
    if (f && --const_cast<VSFrameRef*>(f)->refs == 0) delete f;
}
void VS_CC free_node(VSNodeRef* r) {
    if (!r) return;
    node_release(r->node);
    delete r;
}
VSFrameRef* VS_CC new_video_frame(const VSFormat* fmt, int width, int height, const VSFrameRef* prop_src, VSCore*) {
    VSFrameRef* f = new_frame(fmt, width, height);
    if (prop_src) copy_props(&prop_src->props, &f->props);
    return f;
}
void VS_CC create_filter(const VSMap* in, VSMap* out, const char* name, VSFilterInit init, VSFilterGetFrame get,
                         VSFilterFree free_fn, int, int, void* inst, VSCore* core) {
    Node* n = new Node;
    n->name = name;
    n->get = get;
    n->free_fn = free_fn;
    n->inst = inst;
    init(const_cast<VSMap*>(in), out, &n->inst, reinterpret_cast<VSNode*>(n), core, &g_api);
    VSNodeRef r{n};
    g_api.propSetNode(out, "clip", &r, paReplace);
    node_release(n);
}
void VS_CC set_error(VSMap* m, const char* msg) {
    m->error = msg ? msg : "";
    m->has_error = true;
}
const char* VS_CC get_error(const VSMap* m) { return m->has_error ? m->error.c_str() : nullptr; }
void VS_CC set_filter_error(const char* msg, VSFrameContext* ctx) { ctx->error = msg ? msg : "unknown error"; }
const VSFrameRef* VS_CC get_frame_filter(int n, VSNodeRef* r, VSFrameContext* ctx) {
    n = std::clamp(n, 0, r->node->vi.numFrames - 1);
    auto it = ctx->ready.find({r->node, n});
    return it == ctx->ready.end() ? nullptr : clone_frame(it->second);
}
void VS_CC request_frame_filter(int n, VSNodeRef* r, VSFrameContext* ctx) {
    n = std::clamp(n, 0, r->node->vi.numFrames - 1);
    ctx->requests.emplace_back(r->node, n);
}
int VS_CC get_stride(const VSFrameRef* f, int p) { return f->stride[p]; }
const uint8_t* VS_CC get_read_ptr(const VSFrameRef* f, int p) { return f->data[p]; }
uint8_t* VS_CC get_write_ptr(VSFrameRef* f, int p) { return f->data[p]; }
VSMap* VS_CC create_map() { return new VSMap; }
void VS_CC free_map(VSMap* m) { delete m; }
const VSVideoInfo* VS_CC get_video_info(VSNodeRef* r) { return &r->node->vi; }
void VS_CC set_video_info(const VSVideoInfo* vi, int, VSNode* node) { reinterpret_cast<Node*>(node)->vi = *vi; }
const VSFormat* VS_CC get_frame_format(const VSFrameRef* f) { return f->fmt; }
int VS_CC get_frame_width(const VSFrameRef* f, int p) { return f->w[p]; }
int VS_CC get_frame_height(const VSFrameRef* f, int p) { return f->h[p]; }
const VSMap* VS_CC get_props_ro(const VSFrameRef* f) { return &f->props; }
VSMap* VS_CC get_props_rw(VSFrameRef* f) { return &f->props; }

const Prop* find(const VSMap* m, const char* key, int* error) {
    auto it = m->props.find(key);
    if (it == m->props.end()) {
        if (error) *error = peUnset;
        else std::abort();  // the API aborts on an unchecked missing key
        return nullptr;
    }
    if (error) *error = 0;
    return &it->second;
}
int VS_CC num_elements(const VSMap* m, const char* key) {
    auto it = m->props.find(key);
    if (it == m->props.end()) return -1;
    const Prop& p = it->second;
    return (int)(p.type == 'i' ? p.i.size() : p.type == 'f' ? p.f.size() : p.type == 's' ? p.s.size()
                                           : p.type == 'c' ? p.n.size() : p.v.size());
}
char VS_CC get_type(const VSMap* m, const char* key) {
    auto it = m->props.find(key);
    return it == m->props.end() ? 'u' : it->second.type;
}
template <typename T>
bool index_ok(const std::vector<T>& v, int index, int* error) {
    if (index >= 0 && index < (int)v.size()) return true;
    if (error) *error = peIndex;
    else std::abort();
    return false;
}
int64_t VS_CC get_int(const VSMap* m, const char* key, int index, int* error) {
    const Prop* p = find(m, key, error);
    if (!p) return 0;
    if (p->type != 'i') {
        if (error) *error = peType;
        return 0;
    }
    return index_ok(p->i, index, error) ? p->i[index] : 0;
}
double VS_CC get_float(const VSMap* m, const char* key, int index, int* error) {
    const Prop* p = find(m, key, error);
    if (!p) return 0;
    if (p->type != 'f') {
        if (error) *error = peType;
        return 0;
    }
    return index_ok(p->f, index, error) ? p->f[index] : 0;
}
const char* VS_CC get_data(const VSMap* m, const char* key, int index, int* error) {
    const Prop* p = find(m, key, error);
    if (!p) return nullptr;
    if (p->type != 's') {
        if (error) *error = peType;
        return nullptr;
    }
    return index_ok(p->s, index, error) ? p->s[index].c_str() : nullptr;
}
VSNodeRef* VS_CC get_node(const VSMap* m, const char* key, int index, int* error) {
    const Prop* p = find(m, key, error);
    if (!p) return nullptr;
    if (p->type != 'c') {
        if (error) *error = peType;
        return nullptr;
    }
    return index_ok(p->n, index, error) ? new_ref(p->n[index]->node) : nullptr;
}
const int64_t* VS_CC get_int_array(const VSMap* m, const char* key, int* error) {
    const Prop* p = find(m, key, error);
    return p && p->type == 'i' ? p->i.data() : nullptr;
}
Prop& slot(VSMap* m, const char* key, char type, int append) {
    Prop& p = m->props[key];
    if (append != paAppend || p.type != type) {
        for (auto* r : p.n) free_node(r);
        for (auto* v : p.v) free_frame(v);
        p = Prop{};
        p.type = type;
    }
    return p;
}
int VS_CC set_int(VSMap* m, const char* key, int64_t v, int append) {
    slot(m, key, 'i', append).i.push_back(v);
    return 0;
}
int VS_CC set_float(VSMap* m, const char* key, double v, int append) {
    slot(m, key, 'f', append).f.push_back(v);
    return 0;
}
int VS_CC set_data(VSMap* m, const char* key, const char* d, int size, int append) {
    slot(m, key, 's', append).s.emplace_back(d, size < 0 ? std::strlen(d) : (size_t)size);
    return 0;
}
int VS_CC set_node(VSMap* m, const char* key, VSNodeRef* r, int append) {
    slot(m, key, 'c', append).n.push_back(new_ref(r->node));
    return 0;
}
int VS_CC set_int_array(VSMap* m, const char* key, const int64_t* v, int size) {
    Prop& p = slot(m, key, 'i', paReplace);
    p.i.assign(v, v + size);
    return 0;
}

// ---------------------------------------------------------------- registry
struct Function {
    std::string args;
    VSPublicFunction fn;
    void* data;
};
std::map<std::string, Function> g_functions;

void VS_CC config_plugin(const char* id, const char* ns, const char* name, int api, int readonly, VSPlugin*) {
    std::printf("config %s %s %d %d %s\n", id, ns, api, readonly, name);
}
void VS_CC register_function(const char* name, const char* args, VSPublicFunction fn, void* data, VSPlugin*) {
    g_functions[name] = {args, fn, data};
    std::printf("register %s %s\n", name, args);
}

void init_api() {
    g_api.cloneFrameRef = clone_frame;
    g_api.cloneNodeRef = clone_node;
    g_api.freeFrame = free_frame;
    g_api.freeNode = free_node;
    g_api.newVideoFrame = new_video_frame;
    g_api.createFilter = create_filter;
    g_api.setError = set_error;
    g_api.getError = get_error;
    g_api.setFilterError = set_filter_error;
    g_api.getFrameFilter = get_frame_filter;
    g_api.requestFrameFilter = request_frame_filter;
    g_api.getStride = get_stride;
    g_api.getReadPtr = get_read_ptr;
    g_api.getWritePtr = get_write_ptr;
    g_api.createMap = create_map;
    g_api.freeMap = free_map;
    g_api.getVideoInfo = get_video_info;
    g_api.setVideoInfo = set_video_info;
    g_api.getFrameFormat = get_frame_format;
    g_api.getFrameWidth = get_frame_width;
    g_api.getFrameHeight = get_frame_height;
    g_api.getFramePropsRO = get_props_ro;
    g_api.getFramePropsRW = get_props_rw;
    g_api.propNumElements = num_elements;
    g_api.propGetType = get_type;
    g_api.propGetInt = get_int;
    g_api.propGetFloat = get_float;
    g_api.propGetData = get_data;
    g_api.propGetNode = get_node;
    g_api.propSetInt = set_int;
    g_api.propSetFloat = set_float;
    g_api.propSetData = set_data;
    g_api.propSetNode = set_node;
    g_api.propGetIntArray = get_int_array;
    g_api.propSetIntArray = set_int_array;
}

std::map<std::string, VSNodeRef*> g_nodes;

std::vector<std::string> split(const std::string& s, char sep) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, sep)) out.push_back(item);
    return out;
}

bool cmd_source(std::istringstream& in) {
    std::string name, fmt_name, path;
    int W, H, N;
    in >> name >> fmt_name >> W >> H >> N >> path;
    const VSFormat* fmt = find_format(fmt_name);
    if (!fmt) return false;
    std::ifstream raw(path, std::ios::binary);
    Node* n = new Node;
    n->name = "source";
    n->vi = {fmt, 30000, 1001, W, H, N, 0};
    for (int k = 0; k < N; ++k) {
        VSFrameRef* f = new_frame(fmt, W, H);
        set_int(&f->props, "_SourceFrame", k, paReplace);
        for (int p = 0; p < fmt->numPlanes; ++p) {
            const size_t row = (size_t)f->w[p] * fmt->bytesPerSample;
            for (int y = 0; y < f->h[p]; ++y) raw.read(reinterpret_cast<char*>(f->data[p] + (size_t)y * f->stride[p]), row);
        }
        n->frames.push_back(f);
    }
    if (!raw) {
        node_release(n);
        return false;
    }
    g_nodes[name] = new VSNodeRef{n};
    std::printf("ok %s %d %d %d\n", name.c_str(), W, H, N);
    return true;
}

void cmd_invoke(std::istringstream& in) {
    std::string name, fn;
    in >> name >> fn;
    VSMap args;
    std::string kv;
    while (in >> kv) {
        const size_t eq = kv.find('=');
        const std::string key = kv.substr(0, eq), val = kv.substr(eq + 1);
        const char t = val[0];
        for (const std::string& item : split(val.substr(2), ',')) {
            if (t == 'i') set_int(&args, key.c_str(), std::stoll(item), paAppend);
            else if (t == 'f') set_float(&args, key.c_str(), std::stod(item), paAppend);
            else if (t == 's') set_data(&args, key.c_str(), item.c_str(), -1, paAppend);
            else if (t == 'n') set_node(&args, key.c_str(), g_nodes.at(item), paAppend);
        }
    }
    auto it = g_functions.find(fn);
    if (it == g_functions.end()) {
        std::printf("error %s no such function\n", fn.c_str());
        return;
    }
    VSMap out;
    it->second.fn(&args, &out, it->second.data, &g_core, &g_api);
    if (const char* e = get_error(&out)) {
        std::printf("error %s %s\n", fn.c_str(), e);
        return;
    }
    int err = 0;
    VSNodeRef* r = get_node(&out, "clip", 0, &err);
    if (err) {
        std::printf("error %s no clip returned\n", fn.c_str());
        return;
    }
    if (g_nodes.count(name)) free_node(g_nodes[name]);
    g_nodes[name] = r;
    const VSVideoInfo& vi = r->node->vi;
    std::printf("ok %s %d %d %d %s\n", name.c_str(), vi.width, vi.height, vi.numFrames, r->node->name.c_str());
}

void cmd_fetch(std::istringstream& in) {
    std::string name, path;
    int threads = 1;
    in >> name >> path >> threads;
    Node* node = g_nodes.at(name)->node;
    const VSVideoInfo vi = node->vi;
    const VSFormat* fmt = vi.format;
    size_t frame_bytes = 0;
    for (int p = 0; p < fmt->numPlanes; ++p)
        frame_bytes += (size_t)(p ? vi.width >> fmt->subSamplingW : vi.width) *
                       (p ? vi.height >> fmt->subSamplingH : vi.height) * fmt->bytesPerSample;
    std::vector<char> out(frame_bytes * vi.numFrames);
    std::mutex mu;
    std::vector<std::string> errors;
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int k = next++; k < vi.numFrames; k = next++) {
            std::string err;
            const VSFrameRef* f = evaluate(node, k, &err);
            if (!f) {
                std::lock_guard<std::mutex> g(mu);
                errors.push_back(std::to_string(k) + " " + err);
                continue;
            }
            char* dst = out.data() + frame_bytes * k;
            for (int p = 0; p < fmt->numPlanes; ++p) {
                const size_t row = (size_t)f->w[p] * fmt->bytesPerSample;
                for (int y = 0; y < f->h[p]; ++y, dst += row) std::memcpy(dst, f->data[p] + (size_t)y * f->stride[p], row);
            }
            free_frame(f);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    if (!errors.empty()) {
        std::sort(errors.begin(), errors.end());
        std::printf("fetch-error %s\n", errors.front().c_str());
        return;
    }
    std::ofstream(path, std::ios::binary).write(out.data(), (std::streamsize)out.size());
    std::printf("fetched %s %d\n", name.c_str(), vi.numFrames);
}

void cmd_props(std::istringstream& in) {
    std::string name;
    int k = 0;
    in >> name >> k;
    std::string err;
    const VSFrameRef* f = evaluate(g_nodes.at(name)->node, k, &err);
    if (!f) {
        std::printf("fetch-error %d %s\n", k, err.c_str());
        return;
    }
    for (const auto& kv : f->props.props) {
        std::printf("prop %s %c", kv.first.c_str(), kv.second.type);
        for (auto v : kv.second.i) std::printf(" %lld", (long long)v);
        for (auto v : kv.second.f) std::printf(" %.17g", v);
        std::printf("\n");
    }
    free_frame(f);
}

}  // namespace

VSMap::~VSMap() {
    for (auto& kv : props) {
        for (auto* r : kv.second.n) free_node(r);
        for (auto* v : kv.second.v) free_frame(v);
    }
}

int main() {
    init_api();
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        if (!(in >> cmd)) continue;
        if (cmd == "plugin") {
            std::string path;
            in >> path;
            void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
            auto init = h ? reinterpret_cast<VSInitPlugin>(dlsym(h, "VapourSynthPluginInit")) : nullptr;
            if (!init) {
                std::printf("error plugin %s\n", dlerror());
                return 2;
            }
            init(config_plugin, register_function, nullptr);
        } else if (cmd == "source") {
            if (!cmd_source(in)) {
                std::printf("error source\n");
                return 2;
            }
        } else if (cmd == "invoke") {
            cmd_invoke(in);
        } else if (cmd == "fetch") {
            cmd_fetch(in);
        } else if (cmd == "props") {
            cmd_props(in);
        } else if (cmd == "inject") {
            // inject <library.so> <site> <after>: bm3d_debug_inject_failure of
            // the library the plugin loaded
            std::string path;
            int site = 0;
            long long after = 0;
            in >> path >> site >> after;
            void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_NOLOAD);
            using Fn = int (*)(int, long long);
            auto fn = h ? reinterpret_cast<Fn>(dlsym(h, "bm3d_debug_inject_failure")) : nullptr;
            if (!fn || fn(site, after) != 0) {
                std::printf("error inject\n");
                return 2;
            }
            dlclose(h);
        } else if (cmd == "free") {
            std::string name;
            in >> name;
            free_node(g_nodes.at(name));
            g_nodes.erase(name);
        } else {
            std::printf("error unknown command %s\n", cmd.c_str());
        }
        std::fflush(stdout);
    }
    for (auto& kv : g_nodes) free_node(kv.second);
    g_nodes.clear();
    std::printf("done\n");
    return 0;
}
