// mh_api.cpp - the C ABI of libmarmel_hip (declared in include/marmel_hip.h).
//
// mh_scene_create converts the host's flattened scene into the device layout
// (mh_scene_dev.h) and uploads it once; the render / query entry points launch
// the kernels on the scene's own stream, time them with HIP events recorded
// on that stream, and copy results back only when the caller asks.
#include "marmel_hip.h"
#include "mh_build.h"
#include "mh_flare.h"
#include "mh_kernels.h"
#include "mh_layout.h"
#include "mh_pack.h"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <algorithm>
#include <vector>

using namespace mh;

namespace {

thread_local std::string g_err;
thread_local mh_timing g_timing;
thread_local double g_seconds = 0.0;
thread_local mh_kernel_times g_ktimes;
// primary samples of the last render: total and the largest per pixel
thread_local unsigned long long g_total_samples = 0;
thread_local int g_max_samples = 0;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

} // namespace

int mh::layout_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

#define HIP_TRY(expr)                                                                                 \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess)                                                                         \
            return fail(100 + (int)e_, "%s failed: %s", #expr, hipGetErrorString(e_));               \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Event pairs recorded around kernel launches. A pair is read back only when
// its slot is reused (N launches later) or at the end of the render, so the
// host never waits on the launches it has just queued.
class EventClock : public KernelClock {
public:
    static const int N = 1024;
    double ms[8] = {};
    long long launches[8] = {};

    ~EventClock() override
    {
        for (auto e : a_) (void)hipEventDestroy(e);
        for (auto e : b_) (void)hipEventDestroy(e);
    }
    hipError_t start()
    {
        if (a_.empty()) {
            a_.assign(N, nullptr);
            b_.assign(N, nullptr);
            kind_.assign(N, -1);
            for (int i = 0; i < N; ++i) {
                hipError_t e = hipEventCreate(&a_[i]);
                if (e == hipSuccess) e = hipEventCreate(&b_[i]);
                if (e != hipSuccess) return e;
            }
        }
        for (int i = 0; i < N; ++i) kind_[i] = -1;
        for (int k = 0; k < 8; ++k) {
            ms[k] = 0.0;
            launches[k] = 0;
        }
        next_ = 0;
        return hipSuccess;
    }
    hipError_t begin(int kind, hipStream_t st) override
    {
        (void)kind;
        int i = (int)(next_ % N);
        hipError_t e = harvest(i);
        if (e != hipSuccess) return e;
        return hipEventRecord(a_[i], st);
    }
    hipError_t end(int kind, hipStream_t st) override
    {
        int i = (int)(next_ % N);
        kind_[i] = kind;
        ++next_;
        return hipEventRecord(b_[i], st);
    }
    hipError_t drain()
    {
        for (int i = 0; i < (int)kind_.size(); ++i) {
            hipError_t e = harvest(i);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

private:
    std::vector<hipEvent_t> a_, b_;
    std::vector<int> kind_;
    long long next_ = 0;
    hipError_t harvest(int i)
    {
        int k = kind_[i];
        if (k < 0) return hipSuccess;
        hipError_t e = hipEventSynchronize(b_[i]);
        if (e != hipSuccess) return e;
        float t = 0.f;
        e = hipEventElapsedTime(&t, a_[i], b_[i]);
        if (e != hipSuccess) return e;
        ms[k & 7] += t;
        launches[k & 7] += 1;
        kind_[i] = -1;
        return hipSuccess;
    }
};

} // namespace

struct mh_scene {
    int device = 0;                // the device current at creation (owns stream and memory)
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<DevBuf> bufs;
    size_t device_bytes = 0;
    SceneView<float> v32;
    SceneView<double> v64;
    // framebuffer / scratch (grown on demand)
    float* fb_linear = nullptr;
    unsigned char* fb_rgb = nullptr;
    size_t fb_pixels = 0;
    int fb_w = 0, fb_h = 0;
    unsigned long long* d_rays = nullptr;
    int* d_overflow = nullptr;
    void* wstack = nullptr;
    size_t wstack_bytes = 0;
    void* qbuf = nullptr;
    size_t qbuf_bytes = 0;
    // wavefront state (grown on demand) and the captured batch graph
    void* wf_mem = nullptr;
    size_t wf_bytes = 0;
    int wf_pool = 0;
    size_t wf_npix = 0;
    WfState ws{};
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    std::vector<unsigned char> graph_key;
    EventClock clock;
    // material tables, the spheres that may be emissive lights and each
    // material's primitive bounds (for updates)
    void* d_m32 = nullptr;
    void* d_m64 = nullptr;
    std::vector<LightSphere> spheres;
    std::vector<double> mat_box;
    // adjoint buffers
    float* d_dimg = nullptr;
    size_t dimg_n = 0;
    double* d_grad = nullptr;
    size_t grad_n = 0;
    void* d_frozen = nullptr;     // frozen-sampling material table (fp32)
    size_t frozen_bytes = 0;
    // fog as given, and whether a material is water (marched water segments)
    mh_fog fog_host{};
    int has_water = 0;
    // the tree: its statistics, and for a device build the binary nodes and
    // leaf order (device, owned through bufs)
    mh_bvh_info bvh{};
    const mh_bvh_node* tree_nodes = nullptr;
    const int32_t* tree_leaf = nullptr;
    // adaptive Whitted state (grown on demand): sums, counts, active list,
    // per-pass list lengths and the resolve statistics in one allocation
    void* ad_mem = nullptr;
    size_t ad_bytes = 0;
    // flare scratch for rows too wide for LDS
    void* flare_mem = nullptr;
    size_t flare_bytes = 0;
    // per-sample sums of the last recording render and the view they belong
    // to (dropped by any material / sky / fog update)
    float4* rec = nullptr;
    size_t rec_n = 0;
    bool rec_valid = false;
    mh_camera rec_cam{};
    int rec_key[5] = {0, 0, 0, 0, 0};
};

namespace {

int upload(mh_scene* s, const void* src, size_t bytes, void** dst)
{
    *dst = nullptr;
    if (bytes == 0) return 0;
    void* p = nullptr;
    HIP_TRY(hipMalloc(&p, bytes));
    s->bufs.push_back({p, bytes});
    s->device_bytes += bytes;
    HIP_TRY(hipMemcpyAsync(p, src, bytes, hipMemcpyHostToDevice, s->stream));
    *dst = p;
    return 0;
}

int validate_sky(const mh_sky& k)
{
    if (k.cloud_thickness > 1e-4) return fail(4, "volumetric clouds (cloud_thickness %g) are outside the GPU path", k.cloud_thickness);
    if (k.star_intensity > 0.0 || k.nebula_intensity > 0.0 || k.galaxy_intensity > 0.0)
        return fail(4, "stars / nebula / galaxy sky layers are outside the GPU path");
    if (k.fog_density != 0.0) return fail(5, "mh_sky.fog_density must be 0: fog is mh_scene_soa.fog");
    return 0;
}

int validate_fog(const mh_fog& f)
{
    double v[10] = {f.density, f.color.x, f.color.y, f.color.z, f.height, f.height_falloff, f.inscatter_strength,
                    f.sun_anisotropy, f.noise_scale, f.noise_amount};
    for (double x : v)
        if (!std::isfinite(x)) return fail(5, "fog parameter is not finite");
    if (f.density < 0.0) return fail(5, "negative fog density %g", f.density);
    return 0;
}

int validate_materials(const mh_material* mats, int n)
{
    for (int i = 0; i < n; ++i) {
        const mh_material& m = mats[i];
        if (m.bump_strength > 1e-6) return fail(4, "material %d: bump mapping is outside the GPU path", i);
        if (m.texture_kind != MH_TEXTURE_NONE && m.texture_kind != MH_TEXTURE_CHECKER &&
            m.texture_kind != MH_TEXTURE_UV_CHECKER)
            return fail(4, "material %d: texture kind %d is outside the GPU path", i, m.texture_kind);
    }
    return 0;
}

// with_tree false: the tree fields are not part of the scene (device build)
int validate(const mh_scene_soa* soa, bool with_tree = true)
{
    if (soa->abi_version != MH_ABI_VERSION) return fail(5, "abi_version %d != %d", soa->abi_version, MH_ABI_VERSION);
    if (soa->n_prims < 0 || soa->n_materials < 0 || soa->n_nodes < 0 || soa->n_leaf_prims < 0 || soa->n_planes < 0)
        return fail(5, "negative count in scene");
    if ((soa->n_prims && !soa->prims) || (soa->n_materials && !soa->materials) || (soa->n_nodes && !soa->nodes) ||
        (soa->n_leaf_prims && !soa->leaf_prims) || (soa->n_planes && !soa->planes))
        return fail(1, "NULL array with non-zero count");
    int rc = validate_sky(soa->sky);
    if (rc) return rc;
    rc = validate_fog(soa->fog);
    if (rc) return rc;
    rc = validate_materials(soa->materials, soa->n_materials);
    if (rc) return rc;
    int n_planes = 0;
    for (int i = 0; i < soa->n_prims; ++i) {
        const mh_prim& p = soa->prims[i];
        if (p.kind < MH_PRIM_SPHERE || p.kind > MH_PRIM_CYLINDER)
            return fail(4, "prim %d: kind %d is outside the GPU path", i, p.kind);
        if (p.kind == MH_PRIM_PLANE) ++n_planes;
        if ((p.flags & ~(MH_PRIM_SMOOTH | MH_PRIM_UV)) != 0) return fail(5, "prim %d: bad flags", i);
    }
    if (!with_tree) {
        if (soa->n_prims - n_planes > (1 << 27)) return fail(5, "too many primitives (%d)", soa->n_prims - n_planes);
        return 0;
    }
    if (n_planes != soa->n_planes) return fail(5, "plane list has %d entries, scene has %d planes", soa->n_planes, n_planes);
    for (int i = 0; i < soa->n_planes; ++i) {
        int pi = soa->planes[i];
        if (pi < 0 || pi >= soa->n_prims || soa->prims[pi].kind != MH_PRIM_PLANE)
            return fail(5, "plane list entry %d is not a plane", i);
    }
    if (soa->n_leaf_prims != soa->n_prims - n_planes)
        return fail(5, "leaf list has %d entries, expected %d", soa->n_leaf_prims, soa->n_prims - n_planes);
    if (soa->n_leaf_prims > (1 << 27)) return fail(5, "too many primitives (%d)", soa->n_leaf_prims);
    if ((soa->n_leaf_prims > 0) != (soa->n_nodes > 0)) return fail(5, "BVH / primitive count mismatch");
    for (int i = 0; i < soa->n_leaf_prims; ++i) {
        int pi = soa->leaf_prims[i];
        if (pi < 0 || pi >= soa->n_prims || soa->prims[pi].kind == MH_PRIM_PLANE)
            return fail(5, "leaf list entry %d is invalid", i);
    }
    return 0;
}

int ensure_fb(mh_scene* s, int w, int h)
{
    size_t px = (size_t)w * (size_t)h;
    if (px > s->fb_pixels) {
        if (s->fb_linear) (void)hipFree(s->fb_linear);
        if (s->fb_rgb) (void)hipFree(s->fb_rgb);
        s->fb_linear = nullptr;
        s->fb_rgb = nullptr;
        s->fb_pixels = 0;
        HIP_TRY(hipMalloc(&s->fb_linear, px * 3 * sizeof(float)));
        HIP_TRY(hipMalloc(&s->fb_rgb, px * 3));
        s->fb_pixels = px;
    }
    s->fb_w = w;
    s->fb_h = h;
    return 0;
}

int ensure_query(mh_scene* s, size_t bytes)
{
    if (bytes > s->qbuf_bytes) {
        if (s->qbuf) (void)hipFree(s->qbuf);
        s->qbuf = nullptr;
        s->qbuf_bytes = 0;
        HIP_TRY(hipMalloc(&s->qbuf, bytes));
        s->qbuf_bytes = bytes;
    }
    return 0;
}

// Carve the wavefront state for `pool` paths and `npix` pixels out of one
// allocation (`jid`: per-slot job ids of the batched engine).
int alloc_wf(int pool, size_t npix, bool jid, void** mem, size_t* bytes, WfState* w)
{
    size_t P = (size_t)pool;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t sz[] = {al(P * 16), al(P * 16), al(P * 16), al(P * 16), al(P * 8), al(P * 9 * 16), al(P * 4),
                   al(P * 4),  al(P * 4),  al(P * 4),  al(P * 4),  al(P * 9 * 16), al(P * 9 * 16), al(64),
                   al(npix * 3 * 4), al(64), jid ? al(P * 4) : 0, jid ? al(P * 4) : 0, al(P * 9 * 4),
                   al(P * 9 * 4)};
    const int nbuf = (int)(sizeof sz / sizeof sz[0]);
    size_t total = 0;
    for (size_t v : sz) total += v;
    HIP_TRY(hipMalloc(mem, total));
    *bytes = total;
    char* p = (char*)*mem;
    void* ptrs[nbuf];
    for (int i = 0; i < nbuf; ++i) {
        ptrs[i] = sz[i] ? p : nullptr;
        p += sz[i];
    }
    w->p0 = (float4*)ptrs[0];
    w->p1 = (float4*)ptrs[1];
    w->p2 = (float4*)ptrs[2];
    w->p3 = (float4*)ptrs[3];
    w->hit = (float2*)ptrs[4];
    w->pend = (float4*)ptrs[5];
    w->vis = (uint32_t*)ptrs[6];
    w->npend = (uint32_t*)ptrs[7];
    w->q[0] = (uint32_t*)ptrs[8];
    w->q[1] = (uint32_t*)ptrs[9];
    w->term = (uint32_t*)ptrs[10];
    w->sh0 = (float4*)ptrs[11];
    w->sh1 = (float4*)ptrs[12];
    w->cnt = (uint32_t*)ptrs[13];
    w->acc = (float*)ptrs[14];
    w->batch = (unsigned long long*)ptrs[15];
    w->jid = (uint32_t*)ptrs[16];
    w->freeq = (uint32_t*)ptrs[17];
    w->shq = (uint32_t*)ptrs[18];
    w->slowq = (uint32_t*)ptrs[19];
    return 0;
}

int ensure_wf(mh_scene* s, int pool, size_t npix)
{
    if (pool <= s->wf_pool && npix <= s->wf_npix) return 0;
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    if (s->graph_def) (void)hipGraphDestroy(s->graph_def);
    s->graph = nullptr;
    s->graph_def = nullptr;
    s->graph_key.clear();
    if (s->wf_mem) (void)hipFree(s->wf_mem);
    s->wf_mem = nullptr;
    s->wf_bytes = 0;
    s->wf_pool = 0;
    s->wf_npix = 0;
    int rc = alloc_wf(pool, npix, false, &s->wf_mem, &s->wf_bytes, &s->ws);
    if (rc) return rc;
    s->wf_pool = pool;
    s->wf_npix = npix;
    return 0;
}

// The scene's views with the requested BVH width (0 = default).
template <typename R> SceneView<R> with_width(const SceneView<R>& v, int width)
{
    SceneView<R> o = v;
    if (width == 2) o.bvh4 = 0;
    return o;
}

int check_width(int w, const SceneView<float>& v)
{
    if (w != 0 && w != 2 && w != 4) return fail(4, "BVH width %d is not built (2 or 4)", w);
    if (w == 4 && !v.bvh4 && v.nodes4 == nullptr && v.root_ref != INT_MIN && v.n_nodes > 0)
        return fail(4, "this scene has no 4-wide tree (too deep for the traversal stack)");
    return 0;
}

float ms_between(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) ms = 0.f;
    return ms;
}

} // namespace


// Batched renderer: its own stream, path state, tables and graph.
struct mh_batch {
    hipStream_t stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    void* wf_mem = nullptr;
    size_t wf_bytes = 0;
    int wf_pool = 0;
    WfState ws{};
    void* tab = nullptr;          // scene table then job table
    size_t tab_bytes = 0;
    unsigned long long* acc64 = nullptr;
    size_t acc_pix = 0;
    float* fb_linear = nullptr;
    unsigned char* fb_rgb = nullptr;
    size_t fb_pix = 0;
    unsigned long long* d_rays = nullptr;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
    std::vector<unsigned char> graph_key;
    EventClock clock;
    std::vector<std::pair<std::vector<long long>, mh_tune>> tuned;  // autotune cache
    mh_tune last{};
    unsigned int* poll = nullptr;   // pinned: queue length + cursor, read between graph replays
    hipGraphExec_t sgraph = nullptr;  // two captured regeneration steps
    hipGraph_t sgraph_def = nullptr;
    std::vector<unsigned char> sgraph_key;
    BUnit* units = nullptr;         // batched megakernel work units (device)
    size_t units_n = 0;
    int count_on = 0;                       // count traversal work (WF_COUNT launches)
    unsigned long long* visits = nullptr;   // device counters, mh_traversal_counts order
    mh_traversal_counts counts{};           // of the last render
};

namespace {

void drop_graph(mh_batch* b)
{
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    if (b->graph_def) (void)hipGraphDestroy(b->graph_def);
    b->graph = nullptr;
    b->graph_def = nullptr;
    b->graph_key.clear();
    if (b->sgraph) (void)hipGraphExecDestroy(b->sgraph);
    if (b->sgraph_def) (void)hipGraphDestroy(b->sgraph_def);
    b->sgraph = nullptr;
    b->sgraph_def = nullptr;
    b->sgraph_key.clear();
}

// Path regeneration: steps until every path of the workload (or the first
// `limit` ones) has ended. The host replays pairs of steps (the queues
// ping-pong) and between rounds reads the queue length and cursor back.
int stream_run(mh_batch* b, WfMulti mj, int feat, const mh_tune& t, long long limit)
{
    if (limit >= 0 && limit < mj.total) mj.total = limit;
    mj.stream = 1;
    HIP_TRY(wf_stream_init(b->ws, t.pool, b->stream));
    long long min_pairs = (mj.total + 2LL * t.pool - 1) / (2LL * t.pool);
    long long round = min_pairs < 1 ? 1 : min_pairs;
    bool graph = t.use_graph && !t.kernel_clock;
    if (graph) {
        std::vector<unsigned char> key(sizeof mj + sizeof b->ws + 4 * sizeof(int));
        int extra[4] = {t.grid_blocks, t.block, feat, t.pool};
        memcpy(key.data(), &mj, sizeof mj);
        memcpy(key.data() + sizeof mj, &b->ws, sizeof b->ws);
        memcpy(key.data() + sizeof mj + sizeof b->ws, extra, sizeof extra);
        if (!b->sgraph || key != b->sgraph_key) {
            if (b->sgraph) (void)hipGraphExecDestroy(b->sgraph);
            if (b->sgraph_def) (void)hipGraphDestroy(b->sgraph_def);
            b->sgraph = nullptr;
            b->sgraph_def = nullptr;
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = wf_stream_step(mj, b->ws, 0, b->d_rays, t.grid_blocks, t.block, feat, b->stream);
            if (ce == hipSuccess)
                ce = wf_stream_step(mj, b->ws, 1, b->d_rays, t.grid_blocks, t.block, feat, b->stream);
            hipError_t ee = hipStreamEndCapture(b->stream, &b->sgraph_def);
            HIP_TRY(ce);
            HIP_TRY(ee);
            HIP_TRY(hipGraphInstantiate(&b->sgraph, b->sgraph_def, nullptr, nullptr, 0));
            b->sgraph_key = key;
        }
    }
    if (t.kernel_clock) HIP_TRY(b->clock.start());
    for (int guard = 0; guard < 1 << 20; ++guard) {
        for (long long i = 0; i < round; ++i) {
            if (graph) {
                HIP_TRY(hipGraphLaunch(b->sgraph, b->stream));
            } else {
                KernelClock* kc = t.kernel_clock ? &b->clock : nullptr;
                HIP_TRY(wf_stream_step(mj, b->ws, 0, b->d_rays, t.grid_blocks, t.block, feat, b->stream, kc));
                HIP_TRY(wf_stream_step(mj, b->ws, 1, b->d_rays, t.grid_blocks, t.block, feat, b->stream, kc));
            }
        }
        // paths queued for the next step (cnt[0]) and the cursor
        HIP_TRY(hipMemcpyAsync(b->poll, b->ws.cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, b->stream));
        HIP_TRY(hipMemcpyAsync(b->poll + 2, b->ws.batch, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
        unsigned long long cursor;
        memcpy(&cursor, b->poll + 2, sizeof cursor);
        if (b->poll[0] == 0 && (long long)cursor >= mj.total) return 0;
        round = 4;
    }
    return fail(6, "path regeneration did not finish");
}

template <typename T> int grow(T** p, size_t* have, size_t need, size_t elem)
{
    if (need <= *have) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    HIP_TRY(hipMalloc((void**)p, need * elem));
    *have = need;
    return 0;
}

// Run the batches of one render: `limit` paths (all when < 0) with settings t.
int batch_run(mh_batch* b, const WfMulti& mj, int max_depth, int feat, const mh_tune& t, long long limit)
{
    if (!t.fixed_batches) return stream_run(b, mj, feat, t, limit);  // path regeneration (default)
    long long total = limit >= 0 && limit < mj.total ? limit : mj.total;
    long long nb = (total + t.pool - 1) / t.pool;
    if (t.kernel_clock) {
        HIP_TRY(b->clock.start());
        for (long long i = 0; i < nb; ++i)
            HIP_TRY(wf_multi_batch(mj, b->ws, t.pool, max_depth, b->d_rays, t.grid_blocks, t.block, feat,
                                   b->stream, &b->clock));
    } else if (t.use_graph) {
        std::vector<unsigned char> key(sizeof mj + sizeof b->ws + 5 * sizeof(int));
        int extra[5] = {t.pool, max_depth, t.grid_blocks, t.block, feat};
        memcpy(key.data(), &mj, sizeof mj);
        memcpy(key.data() + sizeof mj, &b->ws, sizeof b->ws);
        memcpy(key.data() + sizeof mj + sizeof b->ws, extra, sizeof extra);
        if (!b->graph || key != b->graph_key) {
            drop_graph(b);
            HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = wf_multi_batch(mj, b->ws, t.pool, max_depth, b->d_rays, t.grid_blocks, t.block, feat,
                                           b->stream);
            hipError_t ee = hipStreamEndCapture(b->stream, &b->graph_def);
            HIP_TRY(ce);
            HIP_TRY(ee);
            HIP_TRY(hipGraphInstantiate(&b->graph, b->graph_def, nullptr, nullptr, 0));
            b->graph_key = key;
        }
        for (long long i = 0; i < nb; ++i) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
    } else {
        for (long long i = 0; i < nb; ++i)
            HIP_TRY(wf_multi_batch(mj, b->ws, t.pool, max_depth, b->d_rays, t.grid_blocks, t.block, feat, b->stream));
    }
    return 0;
}

int ensure_batch_state(mh_batch* b, int pool)
{
    if (pool <= b->wf_pool) return 0;
    drop_graph(b);
    if (b->wf_mem) (void)hipFree(b->wf_mem);
    b->wf_mem = nullptr;
    b->wf_bytes = 0;
    b->wf_pool = 0;
    int rc = alloc_wf(pool, 1, true, &b->wf_mem, &b->wf_bytes, &b->ws);
    if (rc) return rc;
    b->wf_pool = pool;
    return 0;
}

// Paths per ms of `trial` paths of the workload under settings t.
int batch_trial(mh_batch* b, const WfMulti& mj, int max_depth, int feat, mh_tune t, long long trial, double* rate)
{
    int rc = ensure_batch_state(b, t.pool);
    if (rc) return rc;
    t.use_graph = 0;
    t.kernel_clock = 0;
    HIP_TRY(wf_multi_begin(b->ws, mj, b->stream));
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    rc = batch_run(b, mj, max_depth, feat, t, trial);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    float ms = ms_between(b->ev[0], b->ev[1]);
    long long n = trial < mj.total ? trial : mj.total;
    long long done = (n + t.pool - 1) / t.pool * (long long)t.pool;  // whole batches ran
    if (done > mj.total) done = mj.total;
    *rate = ms > 0.f ? (double)done / ms : 0.0;
    return 0;
}

// Coordinate search by the path rate of a fixed trial of the workload's first
// paths: path regeneration or fixed batches jointly with the pool size (the
// better mode depends on how the workload's depths and scene sizes mix, and
// each mode has its own best pool), then grid blocks, then block size.
int autotune(mh_batch* b, const WfMulti& mj, int max_depth, int feat, mh_tune* t)
{
    // every candidate runs the same first paths (as many as the largest pool
    // holds), so that rates compare the same work: the workload's first
    // jobs can be much cheaper or dearer than the rest
    long long trial = std::min<long long>(mj.total, 1LL << 26);
    double best = -1.0;
    mh_tune cur = *t;
    // pools up to the 2^26 slots the shadow tags address (35 GB of state:
    // HBM holds it); a pool's trial covers at least the pool
    const int pools[] = {1 << 19, 1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25, 1 << 26};
    int pool_best = cur.pool, mode_best = 0;
    for (int mode = 0; mode < 2; ++mode)
        for (int p : pools) {
            if (p > 2 * mj.total && p != pools[0]) break;  // larger pools cannot fill up
            mh_tune c = cur;
            c.pool = p;
            c.fixed_batches = mode;
            double r;
            int rc = batch_trial(b, mj, max_depth, feat, c, std::max<long long>(trial, p), &r);
            if (rc == 100 + (int)hipErrorOutOfMemory && p != pools[0]) {
                // the device cannot hold this pool next to what else it holds:
                // the largest pool that fitted is the ceiling
                (void)hipGetLastError();
                g_err.clear();
                break;
            }
            if (rc) return rc;
            if (r > best) {
                best = r;
                pool_best = p;
                mode_best = mode;
            }
        }
    cur.pool = pool_best;
    cur.fixed_batches = mode_best;
    best = -1.0;
    const int grids[] = {1024, 2048, 4096, 8192, 16384};
    int grid_best = cur.grid_blocks;
    for (int g : grids) {
        mh_tune c = cur;
        c.grid_blocks = g;
        double r;
        int rc = batch_trial(b, mj, max_depth, feat, c, std::max<long long>(trial, cur.pool), &r);
        if (rc) return rc;
        if (r > best) {
            best = r;
            grid_best = g;
        }
    }
    cur.grid_blocks = grid_best;
    best = -1.0;
    const int blocks[] = {128, 256};
    int block_best = cur.block;
    for (int bl : blocks) {
        mh_tune c = cur;
        c.block = bl;
        double r;
        int rc = batch_trial(b, mj, max_depth, feat, c, std::max<long long>(trial, cur.pool), &r);
        if (rc) return rc;
        if (r > best) {
            best = r;
            block_best = bl;
        }
    }
    cur.block = block_best;
    *t = cur;
    return 0;
}

} // namespace

extern "C" {

int mh_device_info(char* name, size_t name_len, int* cu_count, size_t* hbm_bytes)
{
    hipDeviceProp_t p;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    if (name && name_len) {
        snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = p.totalGlobalMem;
    return 0;
}

namespace {

mh_scene* new_scene()
{
    mh_scene* s = new mh_scene();
    if (hipGetDevice(&s->device) != hipSuccess) {
        fail(101, "hipGetDevice failed");
        mh_scene_destroy(s);
        return nullptr;
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        fail(101, "hipStreamCreate failed");
        mh_scene_destroy(s);
        return nullptr;
    }
    for (auto& e : s->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            fail(101, "hipEventCreate failed");
            mh_scene_destroy(s);
            return nullptr;
        }
    return s;
}

// Materials, lights, counters and the two precision views: the part of scene
// creation that does not depend on where the tree was built.
int finish_scene(mh_scene* s, const mh_scene_soa* soa, Layout lay, const std::vector<DMat<float>>& m32,
                 const std::vector<DMat<double>>& m64)
{
    void *d_m32, *d_m64;
    int rc;
    if ((rc = upload(s, m32.data(), m32.size() * sizeof(DMat<float>), &d_m32)) ||
        (rc = upload(s, m64.data(), m64.size() * sizeof(DMat<double>), &d_m64)))
        return rc;
    if (hipMalloc(&s->d_rays, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&s->d_overflow, sizeof(int)) != hipSuccess)
        return fail(102, "hipMalloc failed for counters");
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(103, "upload failed");

    s->d_m32 = d_m32;
    s->d_m64 = d_m64;
    lay.m32 = d_m32;
    lay.m64 = d_m64;
    s->spheres = light_spheres(soa);
    DLight<float> l32[8];
    DLight<double> l64[8];
    int nl = collect_lights(s->spheres, soa->materials, soa->n_materials, l32, l64);
    int has_transmissive = any_transmissive(soa->materials, soa->n_materials);
    s->fog_host = soa->fog;
    s->has_water = any_water(soa->materials, soa->n_materials);
    s->mat_box = material_boxes(soa);
    float tlo[3], thi[3];
    transmissive_box(s->mat_box, soa->materials, soa->n_materials, tlo, thi);
    const int split = trans_split_pays(s->mat_box, tlo, thi);
    fill_view(s->v32, lay, soa, s->has_water, l32, nl, has_transmissive, tlo, thi, split);
    fill_view(s->v64, lay, soa, s->has_water, l64, nl, has_transmissive, tlo, thi, split);
    return 0;
}

} // namespace

int mh_scene_create(const mh_scene_soa* soa, mh_scene** out)
{
    if (!soa || !out) return fail(1, "NULL argument");
    *out = nullptr;
    int rc = validate(soa);
    if (rc) return rc;
    HostLayout hl;
    rc = build_host_layout(soa, &hl);
    if (rc) return rc;

    mh_scene* s = new_scene();
    if (!s) return 101;
    auto bail = [&](int code) {
        mh_scene_destroy(s);
        return code;
    };
    Layout lay;
    lay.root_ref = hl.root_ref;
    lay.root4 = hl.root4;
    lay.bvh4 = hl.bvh4;
    lay.n_nodes = (int)hl.nodes.size();
    lay.n_planes = soa->n_planes;
    void *d_nodes, *d_nodes4, *d_p32, *d_p64, *d_pl32, *d_pl64, *d_plidx, *d_a32, *d_a64;
    if ((rc = upload(s, hl.nodes.data(), hl.nodes.size() * sizeof(DNode), &d_nodes)) ||
        (rc = upload(s, hl.nodes4.data(), hl.nodes4.size() * sizeof(DNode4), &d_nodes4)) ||
        (rc = upload(s, hl.p32.data(), hl.p32.size() * sizeof(P32), &d_p32)) ||
        (rc = upload(s, hl.p64.data(), hl.p64.size() * sizeof(P64), &d_p64)) ||
        (rc = upload(s, hl.pl32.data(), hl.pl32.size() * sizeof(P32), &d_pl32)) ||
        (rc = upload(s, hl.pl64.data(), hl.pl64.size() * sizeof(P64), &d_pl64)) ||
        (rc = upload(s, soa->planes, (size_t)soa->n_planes * sizeof(int), &d_plidx)) ||
        (rc = upload(s, hl.a32.data(), hl.a32.size() * sizeof(DAttr<float>), &d_a32)) ||
        (rc = upload(s, hl.a64.data(), hl.a64.size() * sizeof(DAttr<double>), &d_a64)))
        return bail(rc);
    lay.nodes = d_nodes;
    lay.nodes4 = d_nodes4;
    lay.p32 = d_p32;
    lay.p64 = d_p64;
    lay.pl32 = d_pl32;
    lay.pl64 = d_pl64;
    lay.plidx = d_plidx;
    lay.a32 = d_a32;
    lay.a64 = d_a64;
    s->bvh.n_nodes = soa->n_nodes;
    s->bvh.n_leaf = soa->n_leaf_prims;
    s->bvh.max_depth = 0;
    s->bvh.width4 = lay.bvh4 ? 1 : 0;
    s->bvh.build_ms = 0.0;
    s->bvh.device_built = 0;
    if ((rc = finish_scene(s, soa, lay, hl.m32, hl.m64))) return bail(rc);
    *out = s;
    return 0;
}

int mh_scene_create_gpu_bvh(const mh_scene_soa* soa, int max_leaf, int bins, mh_scene** out)
{
    if (!soa || !out) return fail(1, "NULL argument");
    *out = nullptr;
    // the tree fields are not read: validate the rest
    mh_scene_soa v = *soa;
    v.n_nodes = 0;
    v.nodes = nullptr;
    v.n_leaf_prims = 0;
    v.leaf_prims = nullptr;
    v.n_planes = 0;
    v.planes = nullptr;
    int rc = validate(&v, false);
    if (rc) return rc;
    if (max_leaf < 0 || max_leaf > 8 || bins < 0 || bins > 64)
        return fail(5, "BVH build options out of range (max_leaf %d, bins %d)", max_leaf, bins);

    mh_scene* s = new_scene();
    if (!s) return 101;
    auto bail = [&](int code) {
        mh_scene_destroy(s);
        return code;
    };
    bool any_attr = has_attrs(soa);
    void *d_prims = nullptr, *d_attrs = nullptr;
    auto free_in = [&]() {
        if (d_prims) (void)hipFree(d_prims);
        if (d_attrs) (void)hipFree(d_attrs);
    };
    size_t pbytes = sizeof(mh_prim) * (size_t)soa->n_prims;
    if (pbytes && hipMalloc(&d_prims, pbytes) != hipSuccess) return bail(fail(102, "hipMalloc failed for prims"));
    if (soa->attrs && soa->n_prims > 0 &&
        hipMalloc(&d_attrs, sizeof(mh_prim_attr) * (size_t)soa->n_prims) != hipSuccess) {
        free_in();
        return bail(fail(102, "hipMalloc failed for attributes"));
    }
    GpuBvh g;
    int depth_err = 0;
    hipError_t he = hipSuccess;
    if (pbytes) he = hipMemcpyAsync(d_prims, soa->prims, pbytes, hipMemcpyHostToDevice, s->stream);
    if (he == hipSuccess && d_attrs)
        he = hipMemcpyAsync(d_attrs, soa->attrs, sizeof(mh_prim_attr) * (size_t)soa->n_prims, hipMemcpyHostToDevice,
                            s->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(s->stream);
    if (he == hipSuccess) he = hipEventRecord(s->ev[0], s->stream);
    int brc = he == hipSuccess ? gpu_build_scene((const mh_prim*)d_prims, (const mh_prim_attr*)d_attrs, soa->n_prims,
                                                 any_attr, max_leaf, bins, s->stream, &g, &depth_err)
                               : (int)he;
    if (brc == 0) brc = (int)hipEventRecord(s->ev[1], s->stream);
    if (brc == 0) brc = (int)hipEventSynchronize(s->ev[1]);
    float ms = 0.f;
    if (brc == 0) (void)hipEventElapsedTime(&ms, s->ev[0], s->ev[1]);
    free_in();
    for (void* p : g.owned) s->bufs.push_back({p, 0});
    s->device_bytes += g.bytes;
    if (brc != 0) return bail(fail(100 + brc, "device BVH build failed: %s", hipGetErrorString((hipError_t)brc)));
    if (depth_err) return bail(fail(5, "BVH depth %d exceeds the GPU traversal bound (60)", depth_err));

    Layout lay;
    lay.nodes = g.dnodes;
    lay.n_nodes = g.n_inner;
    lay.root_ref = g.root_ref;
    bool use4 = g.n_nodes > 0 && 3 * g.depth4 + 1 <= MH_STACK;
    lay.bvh4 = use4 && g.n_nodes4 > 0;
    lay.nodes4 = g.dnodes4;
    lay.root4 = lay.bvh4 ? g.root4 : g.root_ref;
    lay.p32 = g.p32;
    lay.p64 = g.p64;
    lay.pl32 = g.pl32;
    lay.pl64 = g.pl64;
    lay.plidx = g.planes;
    lay.a32 = g.a32;
    lay.a64 = g.a64;
    lay.n_planes = g.n_planes;
    s->bvh.n_nodes = g.n_nodes;
    s->bvh.n_leaf = g.n_leaf;
    s->bvh.max_depth = g.max_depth;
    s->bvh.width4 = lay.bvh4 ? 1 : 0;
    s->bvh.build_ms = ms;
    s->bvh.device_built = 1;
    s->tree_nodes = g.nodes;
    s->tree_leaf = g.leaf_prims;
    std::vector<DMat<float>> m32(soa->n_materials);
    std::vector<DMat<double>> m64(soa->n_materials);
    for (int i = 0; i < soa->n_materials; ++i) {
        m32[i] = conv_mat<float>(soa->materials[i]);
        m64[i] = conv_mat<double>(soa->materials[i]);
    }
    if ((rc = finish_scene(s, soa, lay, m32, m64))) return bail(rc);
    *out = s;
    return 0;
}

int mh_scene_bvh_info(mh_scene* s, mh_bvh_info* info)
{
    if (!s || !info) return fail(1, "NULL argument");
    *info = s->bvh;
    return 0;
}

int mh_scene_bvh_export(mh_scene* s, mh_bvh_node* nodes, int32_t* leaf_prims)
{
    if (!s || !nodes || !leaf_prims) return fail(1, "NULL argument");
    if (!s->bvh.device_built) return fail(4, "the scene's tree was built on the host (it holds the tree)");
    HIP_TRY(hipMemcpyAsync(nodes, s->tree_nodes, sizeof(mh_bvh_node) * (size_t)s->bvh.n_nodes, hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipMemcpyAsync(leaf_prims, s->tree_leaf, sizeof(int32_t) * (size_t)s->bvh.n_leaf, hipMemcpyDeviceToHost,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

void mh_scene_destroy(mh_scene* s)
{
    if (!s) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& b : s->bufs) (void)hipFree(b.p);
    if (s->fb_linear) (void)hipFree(s->fb_linear);
    if (s->fb_rgb) (void)hipFree(s->fb_rgb);
    if (s->d_rays) (void)hipFree(s->d_rays);
    if (s->d_overflow) (void)hipFree(s->d_overflow);
    if (s->wstack) (void)hipFree(s->wstack);
    if (s->qbuf) (void)hipFree(s->qbuf);
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    if (s->graph_def) (void)hipGraphDestroy(s->graph_def);
    if (s->wf_mem) (void)hipFree(s->wf_mem);
    if (s->d_dimg) (void)hipFree(s->d_dimg);
    if (s->d_grad) (void)hipFree(s->d_grad);
    if (s->d_frozen) (void)hipFree(s->d_frozen);
    if (s->ad_mem) (void)hipFree(s->ad_mem);
    if (s->flare_mem) (void)hipFree(s->flare_mem);
    if (s->rec) (void)hipFree(s->rec);
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// Scene parameters changed: the captured wavefront graph and the recorded
// per-sample sums no longer describe the scene.
static void drop_scene_graph(mh_scene* s)
{
    s->rec_valid = false;
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    if (s->graph_def) (void)hipGraphDestroy(s->graph_def);
    s->graph = nullptr;
    s->graph_def = nullptr;
    s->graph_key.clear();
}

int mh_scene_set_materials(mh_scene* s, const mh_material* mats, int n)
{
    if (!s || (!mats && n > 0)) return fail(1, "NULL argument");
    if (n != s->v32.n_mats) return fail(5, "material count %d != the scene's %d", n, s->v32.n_mats);
    int rc = validate_materials(mats, n);
    if (rc) return rc;
    std::vector<DMat<float>> m32((size_t)n);
    std::vector<DMat<double>> m64((size_t)n);
    for (int i = 0; i < n; ++i) {
        m32[i] = conv_mat<float>(mats[i]);
        m64[i] = conv_mat<double>(mats[i]);
    }
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(s->d_m32, m32.data(), m32.size() * sizeof(DMat<float>), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipMemcpyAsync(s->d_m64, m64.data(), m64.size() * sizeof(DMat<double>), hipMemcpyHostToDevice, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    DLight<float> l32[8];
    DLight<double> l64[8];
    int nl = collect_lights(s->spheres, mats, n, l32, l64);
    int tr = any_transmissive(mats, n);
    s->v32.n_lights = s->v64.n_lights = nl;
    for (int i = 0; i < 8; ++i) {
        s->v32.lights[i] = l32[i];
        s->v64.lights[i] = l64[i];
    }
    s->v32.has_transmissive = s->v64.has_transmissive = tr;
    float tlo[3], thi[3];
    transmissive_box(s->mat_box, mats, n, tlo, thi);
    s->v32.tlo = V3<float>{tlo[0], tlo[1], tlo[2]};
    s->v32.thi = V3<float>{thi[0], thi[1], thi[2]};
    s->v64.tlo = V3<double>{tlo[0], tlo[1], tlo[2]};
    s->v64.thi = V3<double>{thi[0], thi[1], thi[2]};
    s->v32.trans_split = s->v64.trans_split = trans_split_pays(s->mat_box, tlo, thi);
    s->has_water = any_water(mats, n);
    s->v32.fog_kernel = s->v64.fog_kernel = needs_fog_kernel(s->fog_host, s->has_water);
    drop_scene_graph(s);  // captured launches hold the scene view by value
    return 0;
}

int mh_scene_set_sky(mh_scene* s, const mh_sky* sky)
{
    if (!s || !sky) return fail(1, "NULL argument");
    int rc = validate_sky(*sky);
    if (rc) return rc;
    s->v32.sky = conv_sky<float>(*sky);
    s->v64.sky = conv_sky<double>(*sky);
    drop_scene_graph(s);
    return 0;
}

int mh_scene_set_fog(mh_scene* s, const mh_fog* fog)
{
    if (!s || !fog) return fail(1, "NULL argument");
    int rc = validate_fog(*fog);
    if (rc) return rc;
    s->fog_host = *fog;
    s->v32.fog = conv_fog<float>(*fog);
    s->v64.fog = conv_fog<double>(*fog);
    s->v32.fog_kernel = s->v64.fog_kernel = needs_fog_kernel(*fog, s->has_water);
    drop_scene_graph(s);
    return 0;
}

int mh_render_frozen(mh_scene* s, const mh_camera* cam, const mh_render_params* p, const mh_material* frozen, int n,
                     unsigned char* rgb_out, float* linear_out)
{
    if (!s || !cam || !p || !frozen) return fail(1, "NULL argument");
    if (p->width <= 0 || p->height <= 0) return fail(2, "non-positive image size %dx%d", p->width, p->height);
    if (p->spp < 1 || p->max_depth < 0 || p->sample_offset < 0)
        return fail(3, "bad spp %d / depth %d / offset %d", p->spp, p->max_depth, p->sample_offset);
    if ((long long)p->width * p->height > (1LL << 28)) return fail(2, "image too large");
    if (p->mode != MH_MODE_PATHTRACE || p->precision != MH_PREC_FP32)
        return fail(4, "frozen sampling renders the fp32 path tracer only");
    if (n != s->v32.n_mats) return fail(5, "frozen material count %d != the scene's %d", n, s->v32.n_mats);
    int rc = validate_materials(frozen, n);
    if (rc) return rc;
    std::vector<DMat<float>> f((size_t)n);
    for (int i = 0; i < n; ++i) f[i] = conv_mat<float>(frozen[i]);
    size_t fb = f.size() * sizeof(DMat<float>);
    if (fb > s->frozen_bytes) {
        if (s->d_frozen) (void)hipFree(s->d_frozen);
        s->d_frozen = nullptr;
        s->frozen_bytes = 0;
        HIP_TRY(hipMalloc(&s->d_frozen, fb > 0 ? fb : 1));
        s->frozen_bytes = fb;
    }
    rc = ensure_fb(s, p->width, p->height);
    if (rc) return rc;
    if (fb) HIP_TRY(hipMemcpyAsync(s->d_frozen, f.data(), fb, hipMemcpyHostToDevice, s->stream));
    LaunchOut lo{s->fb_linear, s->fb_rgb, s->d_rays, s->d_overflow};
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    HIP_TRY(launch_render_frozen(s->v32, conv_cam<float>(*cam), p->width, p->height, p->spp, p->sample_offset,
                                 p->max_depth, s->d_frozen, lo, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    size_t n3 = (size_t)p->width * p->height * 3;
    if (rgb_out) HIP_TRY(hipMemcpyAsync(rgb_out, s->fb_rgb, n3, hipMemcpyDeviceToHost, s->stream));
    if (linear_out) HIP_TRY(hipMemcpyAsync(linear_out, s->fb_linear, n3 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = g_timing.total_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.samples = (int64_t)p->width * p->height * p->spp;
    g_seconds = g_timing.kernel_ms * 1e-3;
    return 0;
}

int mh_adjoint(mh_scene* s, const mh_camera* cam, const mh_render_params* p, const float* d_image, float* grad)
{
    if (!s || !cam || !p || !d_image || !grad) return fail(1, "NULL argument");
    if (p->width <= 0 || p->height <= 0) return fail(2, "non-positive image size %dx%d", p->width, p->height);
    if (p->spp < 1 || p->max_depth < 0 || p->sample_offset < 0)
        return fail(3, "bad spp %d / depth %d / offset %d", p->spp, p->max_depth, p->sample_offset);
    if ((long long)p->width * p->height > (1LL << 28)) return fail(2, "image too large");
    if (p->mode != MH_MODE_PATHTRACE) return fail(4, "the adjoint differentiates the path tracer only");
    if (p->precision != MH_PREC_FP32) return fail(4, "the adjoint computes in fp32 only");
    size_t n3 = (size_t)p->width * p->height * 3;
    size_t ng = (size_t)s->v32.n_mats * 8 + 3;
    int rc = grow(&s->d_dimg, &s->dimg_n, n3, sizeof(float));
    if (rc) return rc;
    rc = grow(&s->d_grad, &s->grad_n, ng, sizeof(double));
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(s->d_dimg, d_image, n3 * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipMemsetAsync(s->d_grad, 0, ng * sizeof(double), s->stream));
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    int key[5] = {p->width, p->height, p->spp, p->sample_offset, p->max_depth};
    bool use_rec = s->rec_valid && !memcmp(key, s->rec_key, sizeof key) && !memcmp(cam, &s->rec_cam, sizeof *cam);
    HIP_TRY(launch_adjoint(s->v32, conv_cam<float>(*cam), p->width, p->height, p->spp, p->sample_offset, p->max_depth,
                           s->d_dimg, s->d_grad, use_rec ? s->rec : nullptr, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    std::vector<double> g(ng);
    HIP_TRY(hipMemcpyAsync(g.data(), s->d_grad, ng * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    for (size_t i = 0; i < ng; ++i) grad[i] = (float)g[i];
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = g_timing.total_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.samples = (int64_t)p->width * p->height * p->spp;
    g_seconds = g_timing.kernel_ms * 1e-3;
    return 0;
}

int mh_render_record(mh_scene* s, const mh_camera* cam, const mh_render_params* p, unsigned char* rgb_out,
                     float* linear_out)
{
    if (!s || !cam || !p) return fail(1, "NULL argument");
    if (p->width <= 0 || p->height <= 0) return fail(2, "non-positive image size %dx%d", p->width, p->height);
    if (p->spp < 1 || p->max_depth < 0 || p->sample_offset < 0)
        return fail(3, "bad spp %d / depth %d / offset %d", p->spp, p->max_depth, p->sample_offset);
    if ((long long)p->width * p->height > (1LL << 28)) return fail(2, "image too large");
    if (p->mode != MH_MODE_PATHTRACE) return fail(4, "the recording render is the path tracer");
    if (p->precision != MH_PREC_FP32) return fail(4, "the recording render computes in fp32 only");
    size_t nrec = (size_t)p->width * p->height * (size_t)p->spp;
    if (nrec * sizeof(float4) > ((size_t)16 << 30)) return fail(5, "record of %zu samples exceeds 16 GiB", nrec);
    s->rec_valid = false;
    int rc = grow(&s->rec, &s->rec_n, nrec, sizeof(float4));
    if (rc) return rc;
    rc = ensure_fb(s, p->width, p->height);
    if (rc) return rc;
    LaunchOut lo{s->fb_linear, s->fb_rgb, s->d_rays, s->d_overflow};
    HIP_TRY(hipMemsetAsync(s->d_rays, 0, sizeof(unsigned long long), s->stream));
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    HIP_TRY(launch_render_record(s->v32, conv_cam<float>(*cam), p->width, p->height, p->spp, p->sample_offset,
                                 p->max_depth, s->rec, lo, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    size_t n3 = (size_t)p->width * p->height * 3;
    if (rgb_out) HIP_TRY(hipMemcpyAsync(rgb_out, s->fb_rgb, n3, hipMemcpyDeviceToHost, s->stream));
    if (linear_out) HIP_TRY(hipMemcpyAsync(linear_out, s->fb_linear, n3 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int key[5] = {p->width, p->height, p->spp, p->sample_offset, p->max_depth};
    memcpy(s->rec_key, key, sizeof key);
    s->rec_cam = *cam;
    s->rec_valid = true;
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = g_timing.total_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.samples = (int64_t)p->width * p->height * p->spp;
    g_seconds = g_timing.kernel_ms * 1e-3;
    g_total_samples = (unsigned long long)g_timing.samples;
    g_max_samples = p->spp;
    return 0;
}

size_t mh_scene_device_bytes(const mh_scene* s) { return s ? s->device_bytes : 0; }

int mh_trace_closest(mh_scene* s, const mh_ray* rays, int n, double tmin, double tmax, int precision, mh_hit* out)
{
    if (!s || (!rays && n > 0) || (!out && n > 0)) return fail(1, "NULL argument");
    if (n <= 0) return 0;
    size_t rb = (size_t)n * sizeof(mh_ray), hb = (size_t)n * sizeof(mh_hit);
    int rc = ensure_query(s, rb + hb);
    if (rc) return rc;
    mh_ray* d_r = (mh_ray*)s->qbuf;
    mh_hit* d_h = (mh_hit*)((char*)s->qbuf + rb);
    HIP_TRY(hipMemcpyAsync(d_r, rays, rb, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    if (precision == MH_PREC_FP64) HIP_TRY(launch_closest<double>(s->v64, d_r, n, tmin, tmax, d_h, s->stream));
    else HIP_TRY(launch_closest<float>(s->v32, d_r, n, tmin, tmax, d_h, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    HIP_TRY(hipMemcpyAsync(out, d_h, hb, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = g_timing.total_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.rays = n;
    return 0;
}

int mh_trace_any(mh_scene* s, const mh_ray* rays, int n, double tmin, double tmax, int precision, uint8_t* out)
{
    if (!s || (!rays && n > 0) || (!out && n > 0)) return fail(1, "NULL argument");
    if (n <= 0) return 0;
    size_t rb = (size_t)n * sizeof(mh_ray);
    int rc = ensure_query(s, rb + (size_t)n);
    if (rc) return rc;
    mh_ray* d_r = (mh_ray*)s->qbuf;
    uint8_t* d_o = (uint8_t*)s->qbuf + rb;
    HIP_TRY(hipMemcpyAsync(d_r, rays, rb, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    if (precision == MH_PREC_FP64) HIP_TRY(launch_any<double>(s->v64, d_r, n, tmin, tmax, d_o, s->stream));
    else HIP_TRY(launch_any<float>(s->v32, d_r, n, tmin, tmax, d_o, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    HIP_TRY(hipMemcpyAsync(out, d_o, (size_t)n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = g_timing.total_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.rays = n;
    return 0;
}

// Explicit Whitted work stacks of a persistent grid of `threads` lanes
// (at most 256 Ki, at most one per pixel): 17 pending rays per level suffice.
// MARMEL_WHITTED_STACK_CAP (a test hook) lowers the capacity, so that the
// kernels' own bound check (every push tests it first) reports an overflow.
static int ensure_wstack(mh_scene* s, int w, int h, int depth, bool f64, int* cap_out, int* threads_out)
{
    int cap = 17 * (depth + 2) + 4;
    if (const char* e = getenv("MARMEL_WHITTED_STACK_CAP")) {
        int c = atoi(e);
        if (c >= 1 && c < cap) cap = c;
    }
    int threads = 256 * 1024;
    size_t px = (size_t)w * h;
    if ((size_t)threads > px) threads = (int)((px + 255) / 256 * 256);
    size_t item = f64 ? whitted_item_bytes<double>() : whitted_item_bytes<float>();
    size_t need = item * (size_t)cap * (size_t)threads;
    if (need > s->wstack_bytes) {
        if (s->wstack) (void)hipFree(s->wstack);
        s->wstack = nullptr;
        s->wstack_bytes = 0;
        HIP_TRY(hipMalloc(&s->wstack, need));
        s->wstack_bytes = need;
    }
    *cap_out = cap;
    *threads_out = threads;
    return 0;
}

int mh_render(mh_scene* s, const mh_camera* cam, const mh_render_params* p, unsigned char* rgb_out, float* linear_out)
{
    if (!s || !cam || !p) return fail(1, "NULL argument");
    if (p->width <= 0 || p->height <= 0) return fail(2, "non-positive image size %dx%d", p->width, p->height);
    if (p->spp < 1 || p->max_depth < 0) return fail(3, "bad spp %d / depth %d", p->spp, p->max_depth);
    if ((long long)p->width * p->height > (1LL << 28)) return fail(2, "image too large");
    if (p->mode != MH_MODE_PATHTRACE && p->mode != MH_MODE_WHITTED) return fail(5, "unknown mode %d", p->mode);
    if (p->sample_offset < 0) return fail(3, "negative sample offset");
    if (int wrc = check_width(p->bvh_width, s->v32)) return wrc;
    SceneView<float> v32 = with_width(s->v32, p->bvh_width);
    SceneView<double> v64 = with_width(s->v64, p->bvh_width);
    int rc = ensure_fb(s, p->width, p->height);
    if (rc) return rc;
    LaunchOut lo{s->fb_linear, s->fb_rgb, s->d_rays, s->d_overflow};
    HIP_TRY(hipMemsetAsync(s->d_rays, 0, sizeof(unsigned long long), s->stream));
    HIP_TRY(hipMemsetAsync(s->d_overflow, 0, sizeof(int), s->stream));
    bool f64 = p->precision == MH_PREC_FP64;
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    if (p->mode == MH_MODE_PATHTRACE && p->engine == MH_ENGINE_WAVEFRONT) {
        if (f64) return fail(5, "the wavefront engine computes in fp32 only");
        int npix = p->width * p->height;
        // a batch holds at most one sample of each pixel, or whole samples
        // (pool a multiple of npix, summed in sample order at the batch end:
        // the same sums). 1080p terrain, 64 spp: 1 sample per batch 327, 4
        // samples 451, 8 samples 486, 16 samples 506 Msamples/s; the default
        // (8) keeps a scene's path state near 10 GB
        long long want = p->pool > 0 ? p->pool : std::min<long long>(1LL << 24, 8LL * npix);
        const long long all = (long long)npix * p->spp;
        if (want > all) want = all;
        if (want > (1LL << 26)) want = 1LL << 26;  // the shadow tags address 2^26 slots
        int pool = want >= 2LL * npix ? (int)(want / npix * npix) : (int)(want < npix ? want : npix);
        rc = ensure_wf(s, pool, (size_t)npix);
        if (rc) return rc;
        int grid = p->grid_blocks > 0 ? p->grid_blocks : 4096;  // 1080p terrain: 2048 -> 9 % slower, 8192 -> 0.5 %
        CamView<float> cvw = conv_cam<float>(*cam);
        WfJob job{p->width, p->height, p->spp, p->sample_offset, p->max_depth, npix,
                  (long long)npix * p->spp, pool};
        long long nb = (job.total + pool - 1) / pool;
        HIP_TRY(wf_enqueue_begin(s->ws, npix, s->stream));
        if (p->kernel_clock) {
            HIP_TRY(s->clock.start());
            for (long long b = 0; b < nb; ++b)
                HIP_TRY(wf_enqueue_batch(v32, cvw, job, s->ws, s->d_rays, grid, s->stream, &s->clock));
        } else if (p->use_graph) {
            int gk[2] = {grid, v32.bvh4};
            std::vector<unsigned char> key(sizeof job + sizeof cvw + sizeof gk);
            memcpy(key.data(), &job, sizeof job);
            memcpy(key.data() + sizeof job, &cvw, sizeof cvw);
            memcpy(key.data() + sizeof job + sizeof cvw, gk, sizeof gk);
            if (!s->graph || key != s->graph_key) {
                if (s->graph) (void)hipGraphExecDestroy(s->graph);
                if (s->graph_def) (void)hipGraphDestroy(s->graph_def);
                s->graph = nullptr;
                s->graph_def = nullptr;
                HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
                hipError_t ce = wf_enqueue_batch(v32, cvw, job, s->ws, s->d_rays, grid, s->stream);
                hipError_t ee = hipStreamEndCapture(s->stream, &s->graph_def);
                HIP_TRY(ce);
                HIP_TRY(ee);
                HIP_TRY(hipGraphInstantiate(&s->graph, s->graph_def, nullptr, nullptr, 0));
                s->graph_key = key;
            }
            for (long long b = 0; b < nb; ++b) HIP_TRY(hipGraphLaunch(s->graph, s->stream));
        } else {
            for (long long b = 0; b < nb; ++b)
                HIP_TRY(wf_enqueue_batch(v32, cvw, job, s->ws, s->d_rays, grid, s->stream));
        }
        HIP_TRY(wf_enqueue_end(s->ws, npix, p->spp, lo, s->stream));
    } else if (p->mode == MH_MODE_PATHTRACE) {
        if (f64)
            HIP_TRY(launch_pathtrace<double>(v64, conv_cam<double>(*cam), p->width, p->height, p->spp,
                                             p->sample_offset, p->max_depth, lo, 256, s->stream));
        else
            HIP_TRY(launch_pathtrace<float>(v32, conv_cam<float>(*cam), p->width, p->height, p->spp,
                                            p->sample_offset, p->max_depth, lo, 256, s->stream));
    } else {
        int cap = 0, threads = 0;
        rc = ensure_wstack(s, p->width, p->height, p->max_depth, f64, &cap, &threads);
        if (rc) return rc;
        if (f64)
            HIP_TRY(launch_whitted<double>(v64, conv_cam<double>(*cam), p->width, p->height, p->spp,
                                           p->sample_offset, p->max_depth, lo, s->wstack, cap, threads, s->stream));
        else
            HIP_TRY(launch_whitted<float>(v32, conv_cam<float>(*cam), p->width, p->height, p->spp,
                                          p->sample_offset, p->max_depth, lo, s->wstack, cap, threads, s->stream));
    }
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    size_t n3 = (size_t)p->width * p->height * 3;
    if (rgb_out) HIP_TRY(hipMemcpyAsync(rgb_out, s->fb_rgb, n3, hipMemcpyDeviceToHost, s->stream));
    if (linear_out) HIP_TRY(hipMemcpyAsync(linear_out, s->fb_linear, n3 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    unsigned long long rays = 0;
    int ovf = 0;
    HIP_TRY(hipMemcpyAsync(&rays, s->d_rays, sizeof rays, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(&ovf, s->d_overflow, sizeof ovf, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipEventRecord(s->ev[2], s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.total_ms = ms_between(s->ev[0], s->ev[2]);
    g_timing.samples = (int64_t)p->width * p->height * p->spp;
    g_timing.rays = (int64_t)rays;
    g_seconds = g_timing.kernel_ms * 1e-3;
    g_total_samples = (unsigned long long)p->width * (unsigned long long)p->height * (unsigned long long)p->spp;
    g_max_samples = p->spp;
    memset(&g_ktimes, 0, sizeof g_ktimes);
    if (p->kernel_clock && p->mode == MH_MODE_PATHTRACE && p->engine == MH_ENGINE_WAVEFRONT) {
        HIP_TRY(s->clock.drain());
        for (int k = 0; k < 8; ++k) {
            g_ktimes.ms[k] = s->clock.ms[k];
            g_ktimes.launches[k] = s->clock.launches[k];
        }
    }
    if (ovf) return fail(6, "Whitted work stack overflow");
    return 0;
}

} // extern "C"

// ---- render without allocation ---------------------------------------------
//
// The scratch of mh_render_noalloc is one file-scope device array: the ray
// counter and overflow flag in its first 256 bytes, then a linear
// framebuffer of MH_NOALLOC_MAX_PIXELS pixels. Each GPU holds its own copy;
// host code reaches it only through the address hipGetSymbolAddress gives
// with that GPU current (the array's name is a host-side stand-in).
namespace {

constexpr size_t NOALLOC_HEAD = 256;
constexpr size_t NOALLOC_BYTES = NOALLOC_HEAD + (size_t)MH_NOALLOC_MAX_PIXELS * 3 * sizeof(float);
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

std::mutex g_scratch_mu;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// Clears the head of the scratch (a fill kernel rather than a memset, so a
// caller's stream capture records no memset node).
__global__ void k_noalloc_clear(unsigned long long* head)
{
    if (threadIdx.x < NOALLOC_HEAD / sizeof(unsigned long long)) head[threadIdx.x] = 0ull;
}

// Device address of `dev`'s scratch, looked up once per device with that
// device current.
int noalloc_scratch(int dev, void** out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(5, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = 0;
        HIP_TRY(hipGetDevice(&cur));
        if (cur != dev) HIP_TRY(hipSetDevice(dev));
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != dev) HIP_TRY(hipSetDevice(cur));
        HIP_TRY(e);
        if (!ptr) return fail(100, "hipGetSymbolAddress gave no address for the scratch");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

} // namespace

extern "C" {

int mh_render_noalloc(mh_scene* s, const mh_camera* cam, const mh_render_params* p, unsigned char* d_rgb_out,
                      float* d_linear_out)
{
    if (!s || !cam || !p || !d_rgb_out) return fail(1, "NULL argument");
    if (p->width <= 0 || p->height <= 0) return fail(2, "non-positive image size %dx%d", p->width, p->height);
    if (p->spp < 1 || p->max_depth < 0) return fail(3, "bad spp %d / depth %d", p->spp, p->max_depth);
    if ((long long)p->width * p->height > (long long)MH_NOALLOC_MAX_PIXELS)
        return fail(2, "image of %lld pixels exceeds the no-allocation limit of %d pixels",
                    (long long)p->width * p->height, MH_NOALLOC_MAX_PIXELS);
    if (p->sample_offset < 0) return fail(3, "negative sample offset");
    if (p->mode != MH_MODE_PATHTRACE || p->engine != MH_ENGINE_MEGAKERNEL)
        return fail(4, "mh_render_noalloc runs the path-tracing megakernel only");
    if (int wrc = check_width(p->bvh_width, s->v32)) return wrc;
    void* scr = nullptr;
    if (int rc = noalloc_scratch(s->device, &scr)) return rc;
    unsigned long long* d_rays = (unsigned long long*)scr;
    int* d_ovf = (int*)((char*)scr + sizeof(unsigned long long));
    float* lin = d_linear_out ? d_linear_out : (float*)((char*)scr + NOALLOC_HEAD);
    LaunchOut lo{lin, d_rgb_out, d_rays, d_ovf};
    hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, s->stream, (unsigned long long*)scr);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    if (p->precision == MH_PREC_FP64)
        HIP_TRY(launch_pathtrace<double>(with_width(s->v64, p->bvh_width), conv_cam<double>(*cam), p->width,
                                         p->height, p->spp, p->sample_offset, p->max_depth, lo, 256, s->stream));
    else
        HIP_TRY(launch_pathtrace<float>(with_width(s->v32, p->bvh_width), conv_cam<float>(*cam), p->width,
                                        p->height, p->spp, p->sample_offset, p->max_depth, lo, 256, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    unsigned long long rays = 0;
    HIP_TRY(hipMemcpyAsync(&rays, d_rays, sizeof rays, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = g_timing.total_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.samples = (int64_t)p->width * p->height * p->spp;
    g_timing.rays = (int64_t)rays;
    g_seconds = g_timing.kernel_ms * 1e-3;
    g_total_samples = (unsigned long long)p->width * (unsigned long long)p->height * (unsigned long long)p->spp;
    g_max_samples = p->spp;
    memset(&g_ktimes, 0, sizeof g_ktimes);
    return 0;
}

int mh_framebuffer_download(mh_scene* s, unsigned char* rgb_out, float* linear_out)
{
    if (!s) return fail(1, "NULL argument");
    if (!s->fb_rgb) return fail(5, "no framebuffer yet");
    size_t n3 = (size_t)s->fb_w * s->fb_h * 3;
    if (rgb_out) HIP_TRY(hipMemcpyAsync(rgb_out, s->fb_rgb, n3, hipMemcpyDeviceToHost, s->stream));
    if (linear_out) HIP_TRY(hipMemcpyAsync(linear_out, s->fb_linear, n3 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

static int render_drop_in(mh_scene* s, const mh_camera* cam, int w, int h, int spp, int depth, unsigned char* rgb,
                          int mode)
{
    if (!s || !cam || !rgb) return fail(1, "NULL argument");
    if (w <= 0 || h <= 0) return fail(2, "non-positive image size");
    if (spp < 1 || depth < 0) return fail(3, "bad spp / depth");
    mh_render_params p;
    memset(&p, 0, sizeof p);
    p.width = w;
    p.height = h;
    p.spp = spp;
    p.max_depth = depth;
    p.mode = mode;
    // Whitted is the byte-parity renderer: run it in fp64. The path tracer
    // is a Monte Carlo estimate: fp32 keeps it within its own noise.
    p.precision = mode == MH_MODE_WHITTED ? MH_PREC_FP64 : MH_PREC_FP32;
    return mh_render(s, cam, &p, rgb, nullptr);
}

int mh_pathtrace_render(mh_scene* s, const mh_camera* cam, int width, int height, int spp, int max_depth,
                        unsigned char* rgb_out)
{
    return render_drop_in(s, cam, width, height, spp, max_depth, rgb_out, MH_MODE_PATHTRACE);
}

int mh_render_image(mh_scene* s, const mh_camera* cam, int width, int height, int spp, int max_depth,
                    unsigned char* rgb_out)
{
    return render_drop_in(s, cam, width, height, spp, max_depth, rgb_out, MH_MODE_WHITTED);
}

// Adaptive Whitted (reference src/render.c:1337-1496 and render_image_ex,
// src/render.c:1627-1680): validation and defaults as the reference's, then
// the whole pass sequence on the scene's stream without host round trips.
static int render_adaptive_impl(mh_scene* s, const mh_camera* cam, int w, int h, const mh_render_ex_params* p,
                                int precision, unsigned char* rgb_out, float* linear_out, int32_t* nsamp_out)
{
    if (!p || !s || !cam) return fail(1, "NULL argument");
    if (w <= 0 || h <= 0) return fail(2, "non-positive image size %dx%d", w, h);
    if (p->samples_per_pixel < 1 || p->max_depth < 0)
        return fail(3, "bad spp %d / depth %d", p->samples_per_pixel, p->max_depth);
    if ((long long)w * h > (1LL << 28)) return fail(2, "image too large");
    if (precision != MH_PREC_FP32 && precision != MH_PREC_FP64) return fail(5, "unknown precision %d", precision);
    int n0 = p->samples_per_pixel;
    if (!p->adaptive) {
        mh_render_params rp;
        memset(&rp, 0, sizeof rp);
        rp.width = w;
        rp.height = h;
        rp.spp = n0;
        rp.max_depth = p->max_depth;
        rp.mode = MH_MODE_WHITTED;
        rp.precision = precision;
        int rc = mh_render(s, cam, &rp, rgb_out, linear_out);
        if (rc == 0 && nsamp_out)
            for (size_t i = 0; i < (size_t)w * h; ++i) nsamp_out[i] = n0;
        return rc;
    }
    long long nmax = p->adaptive_max_spp;
    if (nmax <= 0) nmax = 4LL * n0;
    if (nmax < n0) nmax = n0;
    if (nmax > INT_MAX) nmax = INT_MAX;
    double tau = p->adaptive_tau;
    if (!(tau > 0.0)) tau = 0.02;
    long long max_passes = 1;
    if (nmax > n0) max_passes += (nmax - n0 + n0 - 1) / n0;
    if (max_passes > (1 << 20)) return fail(3, "adaptive cap %lld over base %d needs too many passes", nmax, n0);
    bool f64 = precision == MH_PREC_FP64;
    int rc = ensure_fb(s, w, h);
    if (rc) return rc;
    int cap = 0, threads = 0;
    rc = ensure_wstack(s, w, h, p->max_depth, f64, &cap, &threads);
    if (rc) return rc;
    size_t npix = (size_t)w * h;
    size_t o_ns = 5 * npix * sizeof(double);
    size_t o_list = o_ns + npix * sizeof(int);
    size_t o_cnt = o_list + npix * sizeof(int);
    size_t o_tot = (o_cnt + (size_t)max_passes * sizeof(int) + 7) & ~(size_t)7;
    size_t need = o_tot + 16;
    if (need > s->ad_bytes) {
        if (s->ad_mem) (void)hipFree(s->ad_mem);
        s->ad_mem = nullptr;
        s->ad_bytes = 0;
        HIP_TRY(hipMalloc(&s->ad_mem, need));
        s->ad_bytes = need;
    }
    char* base = (char*)s->ad_mem;
    AdaptBuf ab{(double*)base, (int*)(base + o_ns), (int*)(base + o_list), (int*)(base + o_cnt),
                (unsigned long long*)(base + o_tot), (int*)(base + o_tot + 8)};
    LaunchOut lo{s->fb_linear, s->fb_rgb, s->d_rays, s->d_overflow};
    HIP_TRY(hipMemsetAsync(s->d_rays, 0, sizeof(unsigned long long), s->stream));
    HIP_TRY(hipMemsetAsync(s->d_overflow, 0, sizeof(int), s->stream));
    HIP_TRY(hipEventRecord(s->ev[0], s->stream));
    // sums and counts start at zero; the list needs no clearing
    HIP_TRY(hipMemsetAsync(base, 0, o_list, s->stream));
    HIP_TRY(hipMemsetAsync(base + o_cnt, 0, need - o_cnt, s->stream));
    SceneView<float> v32 = with_width(s->v32, 0);
    SceneView<double> v64 = with_width(s->v64, 0);
    if (f64)
        HIP_TRY(launch_whitted_adaptive<double>(v64, conv_cam<double>(*cam), w, h, n0, (int)nmax, tau, p->max_depth,
                                                ab, (int)max_passes, lo, s->wstack, cap, threads, s->stream));
    else
        HIP_TRY(launch_whitted_adaptive<float>(v32, conv_cam<float>(*cam), w, h, n0, (int)nmax, tau, p->max_depth,
                                               ab, (int)max_passes, lo, s->wstack, cap, threads, s->stream));
    HIP_TRY(hipEventRecord(s->ev[1], s->stream));
    size_t n3 = npix * 3;
    if (rgb_out) HIP_TRY(hipMemcpyAsync(rgb_out, s->fb_rgb, n3, hipMemcpyDeviceToHost, s->stream));
    if (linear_out) HIP_TRY(hipMemcpyAsync(linear_out, s->fb_linear, n3 * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if (nsamp_out) HIP_TRY(hipMemcpyAsync(nsamp_out, ab.ns, npix * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    unsigned long long stats[2] = {0, 0}, rays = 0;
    int ovf = 0;
    HIP_TRY(hipMemcpyAsync(stats, ab.total, 16, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(&rays, s->d_rays, sizeof rays, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipMemcpyAsync(&ovf, s->d_overflow, sizeof ovf, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipEventRecord(s->ev[2], s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    int mx = 0;
    memcpy(&mx, &stats[1], sizeof mx);
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = ms_between(s->ev[0], s->ev[1]);
    g_timing.total_ms = ms_between(s->ev[0], s->ev[2]);
    g_timing.samples = (int64_t)stats[0];
    g_timing.rays = (int64_t)rays;
    g_seconds = g_timing.kernel_ms * 1e-3;
    memset(&g_ktimes, 0, sizeof g_ktimes);
    g_total_samples = stats[0];
    g_max_samples = mx;
    if (ovf) return fail(6, "Whitted work stack overflow");
    return 0;
}

int mh_render_image_ex(mh_scene* s, const mh_camera* cam, int width, int height, const mh_render_ex_params* params,
                       unsigned char* rgb_out)
{
    if (!params) return fail(1, "NULL argument");
    if (!s || !cam || !rgb_out) return fail(1, "NULL argument");
    return render_adaptive_impl(s, cam, width, height, params, MH_PREC_FP64, rgb_out, nullptr, nullptr);
}

int mh_render_adaptive(mh_scene* s, const mh_camera* cam, int width, int height, const mh_render_ex_params* params,
                       int precision, unsigned char* rgb_out, float* linear_out, int32_t* nsamp_out)
{
    return render_adaptive_impl(s, cam, width, height, params, precision, rgb_out, linear_out, nsamp_out);
}

int mh_last_sample_stats(unsigned long long* total, int* max_per_pixel)
{
    if (total) *total = g_total_samples;
    if (max_per_pixel) *max_per_pixel = g_max_samples;
    return 0;
}

// ---- flare post-pass (reference src/flare.c:60-173) ------------------------

// Per-call constants as the reference derives them; 1 = disabled (nothing
// to do). alpha = expf(-1 / radius) is the host's expf, like the reference's.
static int flare_prepare(const mh_flare_params* p, int w, FlareK* k)
{
    if (!p->enabled || p->intensity <= 0.0) return 1;
    double th = p->threshold < 0.0 ? 0.0 : (p->threshold > 0.999 ? 0.999 : p->threshold);
    double len = p->streak_length < 0.005 ? 0.005 : (p->streak_length > 1.0 ? 1.0 : p->streak_length);
    mh_vec3 t = p->tint;
    if (t.x * t.x + t.y * t.y + t.z * t.z < 1e-6) t = mh_vec3{0.15, 0.45, 1.0};
    float r1 = (float)(len * (double)w * 0.12);
    if (r1 < 2.0f) r1 = 2.0f;
    float r2 = (float)(len * (double)w);
    if (r2 < 6.0f) r2 = 6.0f;
    k->intensity = p->intensity;
    k->thresh = th;
    k->tint[0] = t.x;
    k->tint[1] = t.y;
    k->tint[2] = t.z;
    k->a1 = expf(-1.0f / r1);
    k->a2 = expf(-1.0f / r2);
    k->n1 = 0.45f * 0.045f;
    k->n2 = 0.55f * 0.045f;
    return 0;
}

// Rows in LDS when 15 float planes fit the device's per-workgroup limit, else
// in a scratch slice per workgroup (grown on demand in *mem).
static int flare_run(unsigned char* d_rgb, int w, int h, const FlareK& k, void** mem, size_t* bytes, hipStream_t st)
{
    int dev = 0, lim = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    float* scr = nullptr;
    int grid = h < 4096 ? h : 4096;
    if (flare_lds_bytes(w) > (size_t)lim) {
        grid = h < 512 ? h : 512;
        size_t need = (size_t)grid * flare_lds_bytes(w);
        if (need > *bytes) {
            if (*mem) (void)hipFree(*mem);
            *mem = nullptr;
            *bytes = 0;
            HIP_TRY(hipMalloc(mem, need));
            *bytes = need;
        }
        scr = (float*)*mem;
    }
    HIP_TRY(launch_flare(d_rgb, w, h, k, scr, grid, st));
    return 0;
}

namespace {
// Device image, scratch and stream of mh_flare_apply, one set per device
// (kept until the process ends); one caller at a time.
struct FlareCtx {
    hipStream_t stream = nullptr;
    void* img = nullptr;
    size_t img_bytes = 0;
    void* scr = nullptr;
    size_t scr_bytes = 0;
};
std::mutex g_flare_mu;
std::map<int, FlareCtx> g_flare;
} // namespace

int mh_flare_apply(unsigned char* rgb, int width, int height, const mh_flare_params* params)
{
    if (!rgb || width <= 0 || height <= 0 || !params) return fail(1, "invalid flare arguments");
    FlareK k;
    if (flare_prepare(params, width, &k)) return 0;
    if ((long long)width * height > (1LL << 28)) return fail(2, "image too large");
    std::lock_guard<std::mutex> lock(g_flare_mu);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    FlareCtx& c = g_flare[dev];
    if (!c.stream) HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    size_t n = (size_t)width * height * 3;
    if (n > c.img_bytes) {
        if (c.img) (void)hipFree(c.img);
        c.img = nullptr;
        c.img_bytes = 0;
        HIP_TRY(hipMalloc(&c.img, n));
        c.img_bytes = n;
    }
    unsigned char* d = (unsigned char*)c.img;
    HIP_TRY(hipMemcpyAsync(d, rgb, n, hipMemcpyHostToDevice, c.stream));
    int rc = flare_run(d, width, height, k, &c.scr, &c.scr_bytes, c.stream);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(rgb, d, n, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return 0;
}

int mh_scene_flare(mh_scene* s, const mh_flare_params* params)
{
    if (!s || !params) return fail(1, "NULL argument");
    if (!s->fb_rgb || s->fb_w <= 0 || s->fb_h <= 0) return fail(5, "no framebuffer yet");
    FlareK k;
    if (flare_prepare(params, s->fb_w, &k)) return 0;
    int rc = flare_run(s->fb_rgb, s->fb_w, s->fb_h, k, &s->flare_mem, &s->flare_bytes, s->stream);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return 0;
}

int mh_batch_create(mh_batch** out)
{
    if (!out) return fail(1, "NULL argument");
    *out = nullptr;
    mh_batch* b = new mh_batch();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        mh_batch_destroy(b);
        return fail(101, "hipStreamCreate failed");
    }
    for (auto& e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            mh_batch_destroy(b);
            return fail(101, "hipEventCreate failed");
        }
    if (hipHostMalloc((void**)&b->poll, 16, hipHostMallocDefault) != hipSuccess) {
        b->poll = nullptr;
        mh_batch_destroy(b);
        return fail(102, "hipHostMalloc failed");
    }
    if (hipMalloc(&b->d_rays, sizeof(unsigned long long)) != hipSuccess) {
        mh_batch_destroy(b);
        return fail(102, "hipMalloc failed");
    }
    *out = b;
    return 0;
}

void mh_batch_destroy(mh_batch* b)
{
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->wf_mem) (void)hipFree(b->wf_mem);
    if (b->tab) (void)hipFree(b->tab);
    if (b->acc64) (void)hipFree(b->acc64);
    if (b->fb_linear) (void)hipFree(b->fb_linear);
    if (b->fb_rgb) (void)hipFree(b->fb_rgb);
    if (b->d_rays) (void)hipFree(b->d_rays);
    if (b->poll) (void)hipHostFree(b->poll);
    for (auto& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->units) (void)hipFree(b->units);
    if (b->visits) (void)hipFree(b->visits);
    delete b;
}

// Samples per batched-megakernel unit.
#ifndef MH_BATCH_CHUNK
#define MH_BATCH_CHUNK 16  // 32-scene batch: 2 -> 941, 4 -> 973, 16 -> 998, 64 -> 1000 Msamples/s
#endif

// Batched megakernel units, chunk-major (the first samples of every job
// first), so a prefix of the list is a sample of the whole workload.
static void make_units(const std::vector<BJob>& bj, int chunk, std::vector<BUnit>* out)
{
    out->clear();
    int max_chunks = 0;
    for (const auto& j : bj) max_chunks = std::max(max_chunks, (j.spp + chunk - 1) / chunk);
    for (int c = 0; c < max_chunks; ++c)
        for (size_t ji = 0; ji < bj.size(); ++ji) {
            const BJob& j = bj[ji];
            if (c * chunk >= j.spp) continue;
            int tiles = ((j.w + 15) / 16) * ((j.h + 15) / 16);
            int se = std::min(j.spp, (c + 1) * chunk);
            for (int t = 0; t < tiles; ++t) out->push_back(BUnit{(int)ji, t, c * chunk, se});
        }
}

// The first `n` units (all when n < 0) on the batched megakernel.
static int mega_run(mh_batch* b, const WfMulti& mj, int feat, int n)
{
    int total = (int)b->units_n;
    if (n < 0 || n > total) n = total;
    HIP_TRY(batch_mega(mj, b->units, n, (feat & WF_FOG) != 0, b->d_rays, b->stream));
    return 0;
}

int mh_render_batch(mh_batch* b, mh_scene* const* scenes, int n_scenes, const mh_job* jobs, int n_jobs,
                    const mh_tune* tune, unsigned char* rgb_out, float* linear_out)
{
    if (!b || !scenes || !jobs) return fail(1, "NULL argument");
    if (n_scenes < 1 || n_jobs < 1) return fail(2, "empty batch (%d scenes, %d jobs)", n_scenes, n_jobs);
    for (int i = 0; i < n_scenes; ++i)
        if (!scenes[i]) return fail(1, "NULL scene %d", i);
    mh_tune t{};
    if (tune) t = *tune;
    if (t.bvh_width != 0 && t.bvh_width != 2 && t.bvh_width != 4)
        return fail(4, "BVH node width %d is not built (2 or 4)", t.bvh_width);
    // shadow-ray tags hold the slot in 28 bits
    if (t.pool < 0 || t.pool > (1 << 26) || t.grid_blocks < 0 ||
        (t.block != 0 && t.block != 64 && t.block != 128 && t.block != 256))
        return fail(4, "bad launch settings (pool %d, grid %d, block %d)", t.pool, t.grid_blocks, t.block);
    if (t.engine != MH_BATCH_WAVEFRONT && t.engine != MH_BATCH_MEGAKERNEL)
        return fail(4, "unknown batch engine %d", t.engine);

    std::vector<SceneView<float>> views((size_t)n_scenes);
    int feat = b->count_on ? WF_COUNT : 0;
    for (int i = 0; i < n_scenes; ++i) {
        if (scenes[i]->v32.has_transmissive) {
            feat |= WF_TRANS;
            // the queue is split only when every transmissive scene's box
            // is small (SceneView::trans_split); test hook:
            // MARMEL_SHADOW_SPLIT=0 / 1 never / always splits
            const char* e = std::getenv("MARMEL_SHADOW_SPLIT");
            const bool force = e && (e[0] == '0' || e[0] == '1') && e[1] == 0;
            if (force ? e[0] == '0' : !scenes[i]->v32.trans_split) feat |= WF_NOSPLIT;
        }
        if (scenes[i]->v32.fog_kernel) feat |= WF_FOG;
    }
    // width 4: every scene whose 4-wide tree was built; width 2: binary
    // (the kernels a launch runs depend on it: WF_WIDE)
    auto set_width = [&](int w) {
        feat &= ~WF_WIDE;
        for (int i = 0; i < n_scenes; ++i) {
            views[i] = with_width(scenes[i]->v32, w == 2 ? 2 : 0);
            if (views[i].bvh4) feat |= WF_WIDE;
        }
    };
    set_width(t.bvh_width);
    std::vector<BJob> bj((size_t)n_jobs);
    long long begin = 0, pix = 0;
    int max_depth = 0;
    for (int i = 0; i < n_jobs; ++i) {
        const mh_job& j = jobs[i];
        if (j.scene < 0 || j.scene >= n_scenes) return fail(5, "job %d: scene %d out of range", i, j.scene);
        if (j.width <= 0 || j.height <= 0) return fail(2, "job %d: non-positive image size %dx%d", i, j.width, j.height);
        if ((long long)j.width * j.height > (1LL << 28)) return fail(2, "job %d: image too large", i);
        if (j.spp < 1 || j.max_depth < 0 || j.sample_offset < 0)
            return fail(3, "job %d: bad spp %d / depth %d / offset %d", i, j.spp, j.max_depth, j.sample_offset);
        if (j.max_depth > 254) return fail(4, "job %d: depth %d above the batched engine's 254", i, j.max_depth);
        if (j.spp > MH_BATCH_MAX_SPP)
            return fail(4, "job %d: %d samples per pixel above the batched engine's %d per call (render more "
                           "samples as calls with increasing sample_offset)", i, j.spp, MH_BATCH_MAX_SPP);
        BJob& d = bj[i];
        d.scene = j.scene;
        d.w = j.width;
        d.h = j.height;
        d.spp = j.spp;
        d.s0 = j.sample_offset;
        d.depth = j.max_depth;
        d.npix = j.width * j.height;
        d.begin = begin;
        d.acc_off = pix;
        d.cam = conv_cam<float>(j.cam);
        begin += (long long)d.npix * j.spp;
        pix += d.npix;
        if (j.max_depth > max_depth) max_depth = j.max_depth;
    }
    if (pix * 3 > (1LL << 31)) return fail(2, "batch output too large (%lld pixels)", pix);

    // tables, accumulators, framebuffer
    size_t vb = (views.size() * sizeof(SceneView<float>) + 255) & ~(size_t)255;
    size_t need = vb + bj.size() * sizeof(BJob);
    if (need > b->tab_bytes) {
        drop_graph(b);
        if (b->tab) (void)hipFree(b->tab);
        b->tab = nullptr;
        b->tab_bytes = 0;
        HIP_TRY(hipMalloc(&b->tab, need));
        b->tab_bytes = need;
    }
    size_t old_acc = b->acc_pix;
    int rc = grow(&b->acc64, &b->acc_pix, (size_t)pix * 3, sizeof(unsigned long long));
    if (rc) return rc;
    if (b->acc_pix != old_acc) drop_graph(b);
    if ((size_t)pix > b->fb_pix) {
        if (b->fb_linear) (void)hipFree(b->fb_linear);
        if (b->fb_rgb) (void)hipFree(b->fb_rgb);
        b->fb_linear = nullptr;
        b->fb_rgb = nullptr;
        b->fb_pix = 0;
        HIP_TRY(hipMalloc(&b->fb_linear, (size_t)pix * 3 * sizeof(float)));
        HIP_TRY(hipMalloc(&b->fb_rgb, (size_t)pix * 3));
        b->fb_pix = (size_t)pix;
    }
    HIP_TRY(hipMemcpyAsync(b->tab, views.data(), views.size() * sizeof(SceneView<float>), hipMemcpyHostToDevice,
                           b->stream));
    HIP_TRY(hipMemcpyAsync((char*)b->tab + vb, bj.data(), bj.size() * sizeof(BJob), hipMemcpyHostToDevice, b->stream));
    WfMulti mj{};
    mj.scenes = (const SceneView<float>*)b->tab;
    mj.jobs = (const BJob*)((char*)b->tab + vb);
    mj.n_jobs = n_jobs;
    mj.total = begin;
    mj.total_pix = pix;
    mj.acc64 = b->acc64;
    mj.visits = b->visits;
    mj.sh_stride = 1;
    for (const auto& v : views) mj.sh_stride = std::max(mj.sh_stride, v.n_lights + 1);
    // every job on one scene: the kernels read its view at a launch-uniform
    // index (scalar loads) instead of looking up each path's job
    mj.one_scene = bj[0].scene;
    for (const BJob& d : bj)
        if (d.scene != mj.one_scene) mj.one_scene = -1;

    if (t.pool == 0) t.pool = 1 << 21;
    if (t.grid_blocks == 0) t.grid_blocks = 4096;
    if (t.block == 0) t.block = 256;
    // batched megakernel units (MH_BATCH_CHUNK samples of one 16x16 tile of one job)
    bool want_mega = t.engine == MH_BATCH_MEGAKERNEL || t.autotune;
    if (want_mega) {
        std::vector<BUnit> units;
        make_units(bj, MH_BATCH_CHUNK, &units);
        if (units.size() > (size_t)INT_MAX) return fail(2, "batch has too many work units");
        if (units.size() > b->units_n) {
            if (b->units) (void)hipFree(b->units);
            b->units = nullptr;
            b->units_n = 0;
            HIP_TRY(hipMalloc(&b->units, units.size() * sizeof(BUnit)));
        }
        b->units_n = units.size();
        HIP_TRY(hipMemcpyAsync(b->units, units.data(), units.size() * sizeof(BUnit), hipMemcpyHostToDevice,
                               b->stream));
        HIP_TRY(hipStreamSynchronize(b->stream));
    }
    if (t.autotune) {
        std::vector<long long> key = {n_jobs, begin, pix, max_depth};
        for (int i = 0; i < n_scenes; ++i) key.push_back((long long)(uintptr_t)scenes[i]);
        bool hit = false;
        for (auto& kv : b->tuned)
            if (kv.first == key) {
                t.pool = kv.second.pool;
                t.grid_blocks = kv.second.grid_blocks;
                t.block = kv.second.block;
                t.bvh_width = kv.second.bvh_width;
                t.fixed_batches = kv.second.fixed_batches;
                t.engine = kv.second.engine;
                hit = true;
            }
        if (!hit) {
            rc = autotune(b, mj, max_depth, feat, &t);
            if (rc) return rc;
            // node width: time both layouts at the chosen launch settings
            if (t.bvh_width == 0) {
                long long trial = std::max<long long>(std::min<long long>(mj.total, 1LL << 23), t.pool);
                double best = -1.0;
                int wbest = 4;
                for (int w : {2, 4}) {
                    set_width(w);
                    HIP_TRY(hipMemcpyAsync(b->tab, views.data(), views.size() * sizeof(SceneView<float>),
                                           hipMemcpyHostToDevice, b->stream));
                    double r;
                    rc = batch_trial(b, mj, max_depth, feat, t, trial, &r);
                    if (rc) return rc;
                    if (r > best) {
                        best = r;
                        wbest = w;
                    }
                }
                t.bvh_width = wbest;
            }
            // engine: the tuned wavefront batches against the batched
            // megakernel, each over the whole workload once
            set_width(t.bvh_width);
            HIP_TRY(hipMemcpyAsync(b->tab, views.data(), views.size() * sizeof(SceneView<float>),
                                   hipMemcpyHostToDevice, b->stream));
            double r_wf;
            mh_tune tw = t;
            tw.engine = MH_BATCH_WAVEFRONT;
            rc = batch_trial(b, mj, max_depth, feat, tw, mj.total, &r_wf);
            if (rc) return rc;
            // the megakernel on its own best layout: the 4-wide tree where built
            set_width(4);
            HIP_TRY(hipMemcpyAsync(b->tab, views.data(), views.size() * sizeof(SceneView<float>),
                                   hipMemcpyHostToDevice, b->stream));
            HIP_TRY(wf_multi_begin(b->ws, mj, b->stream));
            HIP_TRY(hipEventRecord(b->ev[0], b->stream));
            rc = mega_run(b, mj, feat, -1);
            if (rc) return rc;
            HIP_TRY(hipEventRecord(b->ev[1], b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
            float ms = ms_between(b->ev[0], b->ev[1]);
            double r_mk = ms > 0.f ? (double)mj.total / ms : 0.0;
            if (r_mk > r_wf) {
                t.engine = MH_BATCH_MEGAKERNEL;
                t.bvh_width = 4;
            } else {
                t.engine = MH_BATCH_WAVEFRONT;
            }
            b->tuned.push_back({key, t});
        }
        set_width(t.bvh_width);
        HIP_TRY(hipMemcpyAsync(b->tab, views.data(), views.size() * sizeof(SceneView<float>), hipMemcpyHostToDevice,
                               b->stream));
    }
    if (t.bvh_width == 0) t.bvh_width = 4;
    if (t.engine == MH_BATCH_WAVEFRONT) {
        rc = ensure_batch_state(b, t.pool);
        if (rc) return rc;
    }

    LaunchOut lo{b->fb_linear, b->fb_rgb, b->d_rays, nullptr};
    HIP_TRY(hipMemsetAsync(b->d_rays, 0, sizeof(unsigned long long), b->stream));
    if (b->count_on) HIP_TRY(hipMemsetAsync(b->visits, 0, sizeof(mh_traversal_counts), b->stream));
    HIP_TRY(wf_multi_begin(b->ws, mj, b->stream));
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    rc = t.engine == MH_BATCH_MEGAKERNEL ? mega_run(b, mj, feat, -1) : batch_run(b, mj, max_depth, feat, t, -1);
    if (rc) return rc;
    HIP_TRY(wf_multi_end(mj, lo, b->stream));
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    size_t n3 = (size_t)pix * 3;
    if (rgb_out) HIP_TRY(hipMemcpyAsync(rgb_out, b->fb_rgb, n3, hipMemcpyDeviceToHost, b->stream));
    if (linear_out) HIP_TRY(hipMemcpyAsync(linear_out, b->fb_linear, n3 * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    unsigned long long rays = 0;
    HIP_TRY(hipMemcpyAsync(&rays, b->d_rays, sizeof rays, hipMemcpyDeviceToHost, b->stream));
    mh_traversal_counts counts{};
    if (b->count_on)
        HIP_TRY(hipMemcpyAsync(&counts, b->visits, sizeof counts, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipEventRecord(b->ev[2], b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->counts = counts;
    memset(&g_timing, 0, sizeof g_timing);
    g_timing.kernel_ms = ms_between(b->ev[0], b->ev[1]);
    g_timing.total_ms = ms_between(b->ev[0], b->ev[2]);
    g_timing.samples = begin;
    g_timing.rays = (int64_t)rays;
    g_seconds = g_timing.kernel_ms * 1e-3;
    memset(&g_ktimes, 0, sizeof g_ktimes);
    if (t.kernel_clock) {
        HIP_TRY(b->clock.drain());
        for (int k = 0; k < 8; ++k) {
            g_ktimes.ms[k] = b->clock.ms[k];
            g_ktimes.launches[k] = b->clock.launches[k];
        }
    }
    b->last = t;
    return 0;
}

int mh_batch_last_tune(const mh_batch* b, mh_tune* out)
{
    if (!b || !out) return fail(1, "NULL argument");
    *out = b->last;
    return 0;
}

int mh_batch_set_counting(mh_batch* b, int on)
{
    if (!b) return fail(1, "NULL argument");
    if (on && !b->visits) HIP_TRY(hipMalloc(&b->visits, sizeof(mh_traversal_counts)));
    b->count_on = on ? 1 : 0;
    return 0;
}

int mh_batch_last_counts(const mh_batch* b, mh_traversal_counts* out)
{
    if (!b || !out) return fail(1, "NULL argument");
    *out = b->counts;
    return 0;
}

int mh_last_timing(mh_timing* out)
{
    if (!out) return fail(1, "NULL argument");
    *out = g_timing;
    return 0;
}

int mh_last_kernel_times(mh_kernel_times* out)
{
    if (!out) return fail(1, "NULL argument");
    *out = g_ktimes;
    return 0;
}

double mh_last_seconds(void) { return g_seconds; }

const char* mh_last_error(void) { return g_err.c_str(); }

} // extern "C"
