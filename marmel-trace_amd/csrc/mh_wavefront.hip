// mh_wavefront.hip - host-side enqueue of the wavefront kernels (fp32).
#include "mh_wavefront.h"

namespace mh {

size_t wf_bytes_per_path()
{
    // p0..p3 + hit + 9 pending + vis + npend + 2 queues + term + 9 shadow rays
    return 4 * sizeof(float4) + sizeof(float2) + 9 * sizeof(float4) + 2 * sizeof(uint32_t) +
           3 * sizeof(uint32_t) + 9 * 2 * sizeof(float4);
}

hipError_t wf_enqueue_begin(const WfState& ws, int npix, hipStream_t st)
{
    int blocks = (npix * 3 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_wf_clear, dim3(blocks), dim3(256), 0, st, ws, npix);
    return hipGetLastError();
}

hipError_t wf_enqueue_batch(const SceneView<float>& sc, const CamView<float>& cam, const WfJob& job,
                            const WfState& ws, unsigned long long* rays, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(k_wf_gen, dim3(grid), dim3(256), 0, st, sc, cam, job, ws);
    for (int b = 0; b <= job.depth; ++b) {
        int which = b & 1;
        hipLaunchKernelGGL(k_wf_reset, dim3(1), dim3(64), 0, st, ws.cnt, which ^ 1);
        hipLaunchKernelGGL(k_wf_extend, dim3(grid), dim3(256), 0, st, sc, ws, which);
        hipLaunchKernelGGL(k_wf_shade, dim3(grid), dim3(256), 0, st, sc, ws, job, which, b, rays);
        hipLaunchKernelGGL(k_wf_shadow, dim3(grid), dim3(256), 0, st, sc, ws, rays);
        hipLaunchKernelGGL(k_wf_finalize, dim3(grid), dim3(256), 0, st, ws);
    }
    hipLaunchKernelGGL(k_wf_next_batch, dim3(1), dim3(64), 0, st, ws, job.pool);
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
