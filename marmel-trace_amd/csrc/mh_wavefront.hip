// mh_wavefront.hip - host-side enqueue of the wavefront kernels (fp32).
#include "mh_wavefront.h"

namespace mh {

size_t wf_bytes_per_path()
{
    // p0..p3 + hit + 9 pending + vis + npend + 2 queues + term + 9 shadow rays + 9 shadow queue entries
    // + 9 slow-list entries
    return 4 * sizeof(float4) + sizeof(float2) + 9 * sizeof(float4) + 2 * sizeof(uint32_t) +
           3 * sizeof(uint32_t) + 9 * 2 * sizeof(float4) + 9 * sizeof(uint32_t) + 9 * sizeof(uint32_t);
}

hipError_t wf_enqueue_begin(const WfState& ws, int npix, hipStream_t st)
{
    int blocks = (npix * 3 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_wf_clear, dim3(blocks), dim3(256), 0, st, ws, npix);
    return hipGetLastError();
}

// The extension kernel: dynamic fetch (on the tree width each scene view
// selects).
template <int MODE> static auto extend_kernel(int feat)
{
    if (MH_EXTEND_REFILL > 0) {
        if constexpr (MODE != WF_SINGLE)
            if (feat & WF_COUNT) return k_wf_extend_refill<MODE, true>;
        return k_wf_extend_refill<MODE, false>;
    }
    return k_wf_extend<MODE>;
}

// The any-hit shadow kernel (no transmissive material in the launch).
template <int MODE> static auto shadow_any(int feat)
{
    if (MH_SHADOW_REFILL > 0) {
        if constexpr (MODE != WF_SINGLE)
            if (feat & WF_COUNT) return k_wf_shadow_refill<MODE, true>;
        return k_wf_shadow_refill<MODE, false>;
    }
    return k_wf_shadow<MODE, false>;
}

// Transmissive launches: 1 = the dynamic-fetch any-hit kernel takes every
// shadow ray whose segment misses the transmissive surfaces' box and lists
// the others for the closest-hit steps of k_wf_shadow<., true, true>; 0 =
// k_wf_shadow<., true> takes every ray through the closest-hit steps.
#ifndef MH_SHADOW_SPLIT
#define MH_SHADOW_SPLIT 1
#endif
static bool split_trans(int feat) { return (feat & WF_TRANS) && !(feat & WF_NOSPLIT) && MH_SHADOW_SPLIT && MH_SHADOW_REFILL > 0; }
template <int MODE> static auto shadow_trans(int feat)
{
    if (split_trans(feat)) {
        if constexpr (MODE != WF_SINGLE)
            if (feat & WF_COUNT) return k_wf_shadow_refill<MODE, true, true>;
        return k_wf_shadow_refill<MODE, false, true>;
    }
    return k_wf_shadow<MODE, true>;
}

// Workgroups of the slow-list kernel: the queue is split only where few rays
// meet the transmissive box, so the list is short, and a full persistent
// grid of workgroups that find nothing to do costs ~0.1 ms per launch.
static dim3 list_grid(int grid) { return dim3(grid < 1024 ? grid : 1024); }

// Dynamic LDS of the shade kernel: the shadow rays of each thread, staged
// until their queue positions are known (9 light blocks + the origin).
static size_t shade_lds(int block) { return (size_t)10 * (size_t)block * sizeof(float4); }

// Launch wrapped in the optional clock.
#define MH_TIMED(kind, launch)                                                                            \
    do {                                                                                                  \
        if (kc) {                                                                                         \
            hipError_t e_ = kc->begin(kind, st);                                                          \
            if (e_ != hipSuccess) return e_;                                                              \
        }                                                                                                 \
        launch;                                                                                           \
        if (kc) {                                                                                         \
            hipError_t e_ = kc->end(kind, st);                                                            \
            if (e_ != hipSuccess) return e_;                                                              \
        }                                                                                                 \
    } while (0)

// The bounce sequence of one batch, shared by both engines.
template <int MODE>
static hipError_t enqueue_bounces(const SceneView<float>& sc, const CamView<float>& cam, const WfJob& job,
                                  const WfState& ws, const WfMulti& mj, unsigned long long* rays, int grid, int block,
                                  int feat, hipStream_t st, KernelClock* kc)
{
    const bool split = split_trans(feat);
    auto shadow = (feat & WF_TRANS) ? shadow_trans<MODE>(feat) : shadow_any<MODE>(feat);
    auto shade = (feat & WF_FOG) ? k_wf_shade<MODE, true> : k_wf_shade<MODE, false>;
    // single-scene batches of several whole samples sum at the batch end
    const int defer = MODE == WF_SINGLE && job.pool > job.npix ? 1 : 0;
    dim3 G(grid), B(block);
    MH_TIMED(MH_K_GEN, hipLaunchKernelGGL(k_wf_gen<MODE>, G, B, 0, st, sc, cam, job, ws, mj));
    for (int b = 0; b <= job.depth; ++b) {
        int which = b & 1;
        MH_TIMED(MH_K_OTHER, hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, st, ws.cnt, which ^ 1));
        MH_TIMED(MH_K_EXTEND, hipLaunchKernelGGL(extend_kernel<MODE>(feat), G, B, 0, st, sc, ws, which, mj));
        MH_TIMED(MH_K_SHADE, hipLaunchKernelGGL(shade, G, B, shade_lds(block), st, sc, ws, job, which, b, rays, mj));
        MH_TIMED(MH_K_SHADOW, hipLaunchKernelGGL(shadow, G, B, 0, st, sc, ws, rays, mj));
        if (split)
            MH_TIMED(MH_K_SHADOW, hipLaunchKernelGGL((k_wf_shadow<MODE, true, true>), list_grid(grid), B, 0, st, sc, ws,
                                                     rays, mj));
        MH_TIMED(MH_K_FINALIZE, hipLaunchKernelGGL(k_wf_finalize<MODE>, G, B, 0, st, ws, mj, defer));
    }
    // batched engine, fixed batches / single-scene batches of whole samples:
    // the batch's pixel sums (k_wf_finalize left each path's radiance in its
    // slot)
    if (MODE != WF_SINGLE)
        MH_TIMED(MH_K_FINALIZE, hipLaunchKernelGGL(k_wf_batch_sum, G, B, 0, st, ws, mj, job.pool));
    else if (defer)
        MH_TIMED(MH_K_FINALIZE, hipLaunchKernelGGL(k_wf_single_sum, G, B, 0, st, ws, job));
    MH_TIMED(MH_K_OTHER, hipLaunchKernelGGL(k_wf_next_batch, dim3(1), dim3(64), 0, st, ws, job.pool));
    return hipGetLastError();
}

hipError_t wf_enqueue_batch(const SceneView<float>& sc, const CamView<float>& cam, const WfJob& job,
                            const WfState& ws, unsigned long long* rays, int grid, hipStream_t st, KernelClock* kc)
{
    WfMulti none{};
    int feat = (sc.has_transmissive ? WF_TRANS : 0) | (sc.has_transmissive && !sc.trans_split ? WF_NOSPLIT : 0) |
               (sc.fog_kernel ? WF_FOG : 0) | (sc.bvh4 ? WF_WIDE : 0);
    return enqueue_bounces<WF_SINGLE>(sc, cam, job, ws, none, rays, grid, 256, feat, st, kc);
}

hipError_t wf_multi_begin(const WfState& ws, const WfMulti& mj, hipStream_t st)
{
    long long n = mj.total_pix * 3;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_wf_multi_clear, dim3(blocks), dim3(256), 0, st, ws, mj);
    return hipGetLastError();
}

hipError_t wf_multi_batch(const WfMulti& mj, const WfState& ws, int pool, int max_depth, unsigned long long* rays,
                          int grid, int block, int feat, hipStream_t st, KernelClock* kc)
{
    SceneView<float> sc{};
    CamView<float> cam{};
    WfJob job{};
    job.depth = max_depth;
    job.pool = pool;
    if (mj.one_scene >= 0) return enqueue_bounces<WF_ONE>(sc, cam, job, ws, mj, rays, grid, block, feat, st, kc);
    return enqueue_bounces<WF_TABLE>(sc, cam, job, ws, mj, rays, grid, block, feat, st, kc);
}

hipError_t wf_stream_init(const WfState& ws, int pool, hipStream_t st)
{
    int blocks = (pool + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_wf_stream_init, dim3(blocks), dim3(256), 0, st, ws, pool);
    return hipGetLastError();
}

template <int MODE>
static hipError_t stream_step(const WfMulti& mj, const WfState& ws, int which, unsigned long long* rays, int grid,
                              int block, int feat, hipStream_t st, KernelClock* kc)
{
    dim3 G(grid), B(block);
    SceneView<float> sc{};
    WfJob job{};
    const bool split = split_trans(feat);
    auto shadow = (feat & WF_TRANS) ? shadow_trans<MODE>(feat) : shadow_any<MODE>(feat);
    auto shade = (feat & WF_FOG) ? k_wf_shade<MODE, true> : k_wf_shade<MODE, false>;
    MH_TIMED(MH_K_GEN, hipLaunchKernelGGL(k_wf_refill<MODE>, G, B, 0, st, ws, mj, which));
    MH_TIMED(MH_K_OTHER, hipLaunchKernelGGL(k_wf_refill_commit, dim3(1), dim3(64), 0, st, ws, mj));
    MH_TIMED(MH_K_OTHER, hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, st, ws.cnt, which ^ 1));
    MH_TIMED(MH_K_EXTEND, hipLaunchKernelGGL(extend_kernel<MODE>(feat), G, B, 0, st, sc, ws, which, mj));
    MH_TIMED(MH_K_SHADE, hipLaunchKernelGGL(shade, G, B, shade_lds(block), st, sc, ws, job, which, 0, rays, mj));
    MH_TIMED(MH_K_SHADOW, hipLaunchKernelGGL(shadow, G, B, 0, st, sc, ws, rays, mj));
    if (split)
        MH_TIMED(MH_K_SHADOW, hipLaunchKernelGGL((k_wf_shadow<MODE, true, true>), list_grid(grid), B, 0, st, sc, ws,
                                                 rays, mj));
    MH_TIMED(MH_K_FINALIZE, hipLaunchKernelGGL(k_wf_finalize<MODE>, G, B, 0, st, ws, mj, 0));
    return hipGetLastError();
}

hipError_t wf_stream_step(const WfMulti& mj, const WfState& ws, int which, unsigned long long* rays, int grid,
                          int block, int feat, hipStream_t st, KernelClock* kc)
{
    if (mj.one_scene >= 0) return stream_step<WF_ONE>(mj, ws, which, rays, grid, block, feat, st, kc);
    return stream_step<WF_TABLE>(mj, ws, which, rays, grid, block, feat, st, kc);
}

hipError_t batch_mega(const WfMulti& mj, const BUnit* units, int n_units, bool fog, unsigned long long* rays,
                      hipStream_t st)
{
    if (n_units <= 0) return hipSuccess;
    auto kern = fog ? k_batch_mega<true> : k_batch_mega<false>;
    hipLaunchKernelGGL(kern, dim3(n_units), dim3(256), 0, st, mj, units, n_units, rays);
    return hipGetLastError();
}

hipError_t wf_multi_end(const WfMulti& mj, const LaunchOut& out, hipStream_t st)
{
    long long n = mj.total_pix;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_wf_multi_resolve, dim3(blocks), dim3(256), 0, st, mj, out);
    return hipGetLastError();
}

hipError_t wf_enqueue_end(const WfState& ws, int npix, int spp, const LaunchOut& out, hipStream_t st)
{
    int blocks = (npix + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_wf_resolve, dim3(blocks), dim3(256), 0, st, ws, npix, spp, out);
    return hipGetLastError();
}

} // namespace mh
