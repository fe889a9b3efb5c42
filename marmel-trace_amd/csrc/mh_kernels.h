// mh_kernels.h - host-callable launchers of the render / query kernels.
// Each launcher is instantiated once per arithmetic type, in its own
// translation unit (mh_kernels_f32.hip is built with FMA contraction,
// mh_kernels_f64.hip without, so the double path rounds like the reference).
#pragma once

#include "marmel_hip.h"
#include "mh_scene_dev.h"

namespace mh {

template <typename R> struct WItem;

struct LaunchOut {
    float* linear;               // w*h*3 mean radiance (device)
    unsigned char* rgb;          // w*h*3 tone-mapped bytes (device)
    unsigned long long* rays;    // ray-query counter (device)
    int* overflow;               // Whitted stack overflow flag (device)
};

template <typename R>
hipError_t launch_pathtrace(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int spp, int s0, int depth,
                            const LaunchOut& out, int block, hipStream_t st);

// Path-replay adjoint (fp32): accumulates dLoss/dtheta into grad
// (n_mats * 8 + 3 doubles: per material albedo rgb, roughness, metallic,
// emissive rgb; then sun colour rgb) from dimg = dLoss/d(pixel mean).
// rec: the per-sample sums of a recording render of the same view and
// parameters (each path is then replayed once), or null (traced twice).
hipError_t launch_adjoint(const SceneView<float>& sc, const CamView<float>& cam, int w, int h, int spp, int s0,
                          int depth, const float* dimg, double* grad, const float4* rec, hipStream_t st);
// fp32 path tracer in the adjoint's code that records each sample's raw sum
// (one float4 per sample and pixel, sample-major) for launch_adjoint.
hipError_t launch_render_record(const SceneView<float>& sc, const CamView<float>& cam, int w, int h, int spp, int s0,
                                int depth, float4* rec, const LaunchOut& out, hipStream_t st);

// fp32 path tracer whose sampling decisions use the device material table
// `frozen` (DMat<float>[n_mats]) and BSDF values the scene's materials.
hipError_t launch_render_frozen(const SceneView<float>& sc, const CamView<float>& cam, int w, int h, int spp, int s0,
                                int depth, const void* frozen, const LaunchOut& out, hipStream_t st);

// ---- wavefront engine (fp32) ---------------------------------------------

// Device state of the wavefront renderer for `pool` paths in flight.
struct WfState {
    float4* p0;       // origin.xyz, seed (bits)
    float4* p1;       // direction.xyz, meta (bits): bounce | last_spec << 8 | (medium + 1) << 9
    float4* p2;       // throughput.xyz, pixel (bits)
    float4* p3;       // radiance.xyz, sample index (bits)
    float2* hit;      // t, leaf record (bits)
    float4* pend;     // 9 pending NEE estimates per path (rgb)
    uint32_t* vis;    // visibility bits of the pending estimates
    uint32_t* npend;  // pending estimates: lights in bits 0..3, sun in bit 4
    uint32_t* q[2];   // extension queues (ping-pong)
    uint32_t* term;   // paths that ended this bounce
    float4* sh0;      // shadow rays cast this bounce, dense (count in cnt[2]): origin.xyz, tmax
    float4* sh1;      // direction.xyz, (slot << 4 | estimate) bits
    uint32_t* shq;    // batched engine: scene index of each shadow ray
    uint32_t* cnt;    // [0..1] ext queue lengths, [2] shadow, [3] term, [4] free slots, [5] slow list
    float* acc;       // per-pixel radiance sums (w*h*3)
    unsigned long long* batch;  // [0] first global path index of the current batch
    uint32_t* jid;    // batched engine: job of each path slot
    uint32_t* freeq;  // batched engine: free path slots (count in cnt[4])
    uint32_t* slowq;  // shadow-queue positions of rays that can meet a transmissive surface (count in cnt[5])
};

// One render job of the batched engine: a camera view of one scene.
struct BJob {
    int scene;              // index into the scene table
    int w, h, spp, s0, depth;
    int npix;
    long long begin;        // first global path index (paths ordered job, sample, pixel)
    long long acc_off;      // first pixel of this job in the accumulators / output
    CamView<float> cam;
};

// Batched engine tables (device pointers). Single-scene launches pass zeros.
struct WfMulti {
    const SceneView<float>* scenes;
    const BJob* jobs;
    int n_jobs;
    long long total;             // paths over all jobs
    long long total_pix;         // pixels over all jobs
    unsigned long long* acc64;   // fixed-point (2^-24) per-pixel sums, 3 per pixel
    int stream;                  // path regeneration: per-slot bounce, free-slot list
    int sh_stride;               // shadow-queue blocks per bounce: max lights over the scenes + 1 (sun)
    int one_scene;               // >= 0: every job renders scenes[one_scene] (WF_ONE kernels); -1: per job
    unsigned long long* visits;  // WF_COUNT launches: traversal counts, MH_V_* (device)
};

// Batched megakernel work unit: one 16x16 tile of a job, samples
// [s_begin, s_end) of it.
struct BUnit {
    int job, tile, s_begin, s_end;
};
// Batched megakernel: every unit in one launch (sums into mj.acc64;
// resolve with wf_multi_end). fog: some scene has fog or marched water.
hipError_t batch_mega(const WfMulti& mj, const BUnit* units, int n_units, bool fog, unsigned long long* rays,
                      hipStream_t st);

struct WfJob {
    int w, h, spp, s0, depth;
    int npix;
    long long total;  // npix * spp
    int pool;         // batch size: at most npix (one sample of each pixel), or a multiple of npix
                      // (whole samples, added to the pixel sums in sample order at the batch end)
};

// Bytes of device state per path slot (state + pending + queues + shadow rays).
size_t wf_bytes_per_path();

// Optional per-launch timing: begin/end are called around each kernel launch
// on the launch stream (direct launches only, never during graph capture).
struct KernelClock {
    virtual ~KernelClock() {}
    virtual hipError_t begin(int kind, hipStream_t st) = 0;
    virtual hipError_t end(int kind, hipStream_t st) = 0;
};

// Record the kernels of one batch (gen, depth+1 bounces, advance) on `st`.
hipError_t wf_enqueue_batch(const SceneView<float>& sc, const CamView<float>& cam, const WfJob& job,
                            const WfState& ws, unsigned long long* rays, int grid, hipStream_t st,
                            KernelClock* kc = nullptr);
hipError_t wf_enqueue_begin(const WfState& ws, int npix, hipStream_t st);
hipError_t wf_enqueue_end(const WfState& ws, int npix, int spp, const LaunchOut& out, hipStream_t st);

// Kernel variants a batch needs (bit set): some scene has a transmissive
// material (shadow rays walk through it) / has fog or marched water / is
// traversed on its 4-wide tree.
// WF_NOSPLIT: a transmissive launch sends every shadow ray through the
// closest-hit steps (k_wf_shadow<., true>) instead of splitting the queue by
// the transmissive box (some scene's box is large, SceneView::trans_split;
// the same image either way).
enum { WF_TRANS = 1, WF_FOG = 2, WF_WIDE = 4, WF_COUNT = 8, WF_NOSPLIT = 16 };
// Launch modes of the wavefront kernels. WF_SINGLE: one scene and one job,
// both kernel arguments. The batched engine's job table: WF_ONE when every
// job renders the same scene (read at a launch-uniform index, so its fields
// are scalar loads), WF_TABLE when each path slot looks its job's scene up.
enum WfMode { WF_SINGLE = 0, WF_ONE = 1, WF_TABLE = 2 };
// WF_COUNT: the extension and any-hit shadow kernels count their traversal
// work into WfMulti::visits (order of mh_traversal_counts; batched engine).

// Batched engine: one batch of `pool` paths drawn from all jobs, bounce
// sequence to the deepest job's depth. `feat`: WF_* bits over the batch's scenes.
hipError_t wf_multi_begin(const WfState& ws, const WfMulti& mj, hipStream_t st);
hipError_t wf_multi_batch(const WfMulti& mj, const WfState& ws, int pool, int max_depth,
                          unsigned long long* rays, int grid, int block, int feat, hipStream_t st,
                          KernelClock* kc = nullptr);
hipError_t wf_multi_end(const WfMulti& mj, const LaunchOut& out, hipStream_t st);

// Path regeneration: init (all slots free), and one step = refill + extend +
// shade + shadow + finalize on extension queue `which`.
hipError_t wf_stream_init(const WfState& ws, int pool, hipStream_t st);
hipError_t wf_stream_step(const WfMulti& mj, const WfState& ws, int which, unsigned long long* rays, int grid,
                          int block, int feat, hipStream_t st, KernelClock* kc = nullptr);

// Whitted: persistent grid of `threads` lanes, each with `cap` stack items
// of `item_bytes` at stack + lane * cap * item_bytes.
template <typename R>
hipError_t launch_whitted(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int spp, int s0, int depth,
                          const LaunchOut& out, void* stack, int cap, int threads, hipStream_t st);

template <typename R> size_t whitted_item_bytes();
// Adaptive Whitted state: per-pixel sums in five planes of w*h doubles
// (R, G, B, luminance, luminance^2), sample counts, the active list.
struct AdaptBuf {
    double* sum;
    int* ns;
    int* list;
    int* count;                 // [pass] length of that pass's active list (zeroed)
    unsigned long long* total;  // sum of the per-pixel sample counts (zeroed)
    int* max_n;                 // largest per-pixel sample count (zeroed)
};
// Pass 0 (n0 samples everywhere), then max_passes - 1 refinement passes of
// flag + sample, then the resolve; no host synchronisation between passes
// (a pass whose list is empty does nothing).
template <typename R>
hipError_t launch_whitted_adaptive(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int n0, int nmax,
                                   double tau, int depth, const AdaptBuf& ab, int max_passes, const LaunchOut& out,
                                   void* stack, int cap, int threads, hipStream_t st);


template <typename R>
hipError_t launch_closest(const SceneView<R>& sc, const mh_ray* rays, int n, double tmin, double tmax, mh_hit* out,
                          hipStream_t st);

template <typename R>
hipError_t launch_any(const SceneView<R>& sc, const mh_ray* rays, int n, double tmin, double tmax, uint8_t* out,
                      hipStream_t st);

} // namespace mh
