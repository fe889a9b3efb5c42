// mh_flare.hip - anamorphic streak flare post-pass on an 8-bit RGB image
// (reference src/flare.c:60-173). Built with -ffp-contract=off: every
// product and sum rounds as the reference's scalar C does, so the bytes
// equal the reference's.
//
// One wavefront per scanline. Phase 1 (all lanes, over x): highlight energy
// of each pixel in double -> three float source planes. Phase 2 (12 lanes):
// the first-order recurrences acc = src + alpha * acc, one lane each per
// colour channel, scale (core and tail radius) and direction (forward and
// backward); their rounding fixes the result, so each runs in order, with
// its loads and stores blocked through registers so LDS latency is paid
// once per block of 16. Phase 3 (all lanes): per pixel and channel the two
// scales' (forward + backward - src) * weight * 0.045 are added to zero in
// the reference's order, composited and clamped.
// The 15 planes of a row (60 bytes per pixel) live in LDS when they fit,
// else in a global scratch slice per workgroup.
#include "mh_flare.h"

#include <hip/hip_runtime.h>

namespace mh {

// n steps of acc = src + alpha * acc along src[0], src[step], ... (step = 1
// forward, -1 backward), writing each acc to the same position of out.
__device__ __forceinline__ void recur(const float* src, float* out, int n, int step, float alpha)
{
    constexpr int B = 16;
    float acc = 0.0f;
    int i = 0;
    for (; i + B <= n; i += B) {
        float v[B];
#pragma unroll
        for (int k = 0; k < B; ++k) v[k] = src[(i + k) * step];
#pragma unroll
        for (int k = 0; k < B; ++k) {
            acc = v[k] + alpha * acc;
            v[k] = acc;
        }
#pragma unroll
        for (int k = 0; k < B; ++k) out[(i + k) * step] = v[k];
    }
    for (; i < n; ++i) {
        acc = src[i * step] + alpha * acc;
        out[i * step] = acc;
    }
}

template <bool LDS>
__global__ __launch_bounds__(64) void k_flare(unsigned char* rgb, int w, int h, FlareK k, float* scratch)
{
    extern __shared__ float lds[];
    int lane = threadIdx.x;
    size_t wz = (size_t)w;
    // planes: src[3], then per (channel, scale, direction) the recurrence
    float* buf = LDS ? lds : scratch + (size_t)blockIdx.x * 15 * wz;
    for (int y = blockIdx.x; y < h; y += gridDim.x) {
        unsigned char* row = rgb + (size_t)y * wz * 3;
        int any = 0;
        for (int x = lane; x < w; x += 64) {
            double c0 = (double)row[3 * x] / 255.0;
            double c1 = (double)row[3 * x + 1] / 255.0;
            double c2 = (double)row[3 * x + 2] / 255.0;
            double lum = 0.2126 * c0 + 0.7152 * c1 + 0.0722 * c2;
            double mx = c0 > c1 ? (c0 > c2 ? c0 : c2) : (c1 > c2 ? c1 : c2);
            double m = 0.5 * (lum + mx);
            float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
            if (m > k.thresh) {
                double e = (m - k.thresh) / (1.0 - k.thresh) * k.intensity;
                s0 = (float)((0.25 * c0 + 0.75 * k.tint[0]) * e);
                s1 = (float)((0.25 * c1 + 0.75 * k.tint[1]) * e);
                s2 = (float)((0.25 * c2 + 0.75 * k.tint[2]) * e);
                any = 1;
            }
            buf[x] = s0;
            buf[wz + x] = s1;
            buf[2 * wz + x] = s2;
        }
        if (__ballot(any) == 0) continue;   // the row has no highlight: unchanged
        __syncthreads();
        if (lane < 12) {
            int c = lane >> 2, scale = (lane >> 1) & 1, bwd = lane & 1;
            const float* src = buf + (size_t)c * wz;
            float* out = buf + (size_t)(3 + lane) * wz;
            int o = bwd ? w - 1 : 0;
            recur(src + o, out + o, w, bwd ? -1 : 1, scale ? k.a2 : k.a1);
        }
        __syncthreads();
        for (int x = lane; x < w; x += 64) {
            for (int c = 0; c < 3; ++c) {
                const float* q = buf + (size_t)(3 + 4 * c) * wz + x;
                float sv = buf[(size_t)c * wz + x];
                float v1 = (q[0] + q[wz] - sv) * k.n1;
                float v2 = (q[2 * wz] + q[3 * wz] - sv) * k.n2;
                float d = 0.0f;
                d += v1;
                d += v2;
                int v = (int)((float)row[3 * x + c] + d * 255.0f + 0.5f);
                row[3 * x + c] = (unsigned char)(v > 255 ? 255 : (v < 0 ? 0 : v));
            }
        }
        __syncthreads();
    }
}

hipError_t launch_flare(unsigned char* rgb, int w, int h, const FlareK& k, float* scratch, int grid, hipStream_t st)
{
    if (scratch) {
        hipLaunchKernelGGL(k_flare<false>, dim3(grid), dim3(64), 0, st, rgb, w, h, k, scratch);
    } else {
        size_t lds = flare_lds_bytes(w);
        if (lds > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute((const void*)k_flare<true>,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_flare<true>, dim3(grid), dim3(64), lds, st, rgb, w, h, k, nullptr);
    }
    return hipGetLastError();
}

} // namespace mh
