// mh_flare.hip - anamorphic streak flare post-pass on an 8-bit RGB image
// (reference src/flare.c:60-173). Built with -ffp-contract=off: every
// product and sum rounds as the reference's scalar C does, so the bytes
// equal the reference's.
//
// One wavefront per scanline. Phase 1 (all lanes, over x): highlight energy
// of each pixel in double -> three float source planes. Phase 2 (lanes 0-2,
// one colour channel each): the forward recurrence acc = src + alpha * acc
// into a buffer, then the backward one combined on the fly into the streak
// accumulator, first at the core radius, then at the tail radius; these are
// first-order recurrences whose rounding fixes the result, so they run in
// order, one lane per channel. Phase 3 (all lanes): composite and clamp.
// The 9 planes of a row live in LDS when they fit (36 bytes per pixel), else
// in a global scratch slice per workgroup.
#include "mh_flare.h"

#include <hip/hip_runtime.h>

namespace mh {

__device__ __forceinline__ void streak(const float* src, float* dst, float* fwd, int n, float alpha, float norm)
{
    float acc = 0.0f;
#pragma unroll 8
    for (int x = 0; x < n; ++x) {
        acc = src[x] + alpha * acc;
        fwd[x] = acc;
    }
    acc = 0.0f;
#pragma unroll 8
    for (int x = n - 1; x >= 0; --x) {
        acc = src[x] + alpha * acc;
        dst[x] += (fwd[x] + acc - src[x]) * norm;
    }
}

template <bool LDS>
__global__ __launch_bounds__(64) void k_flare(unsigned char* rgb, int w, int h, FlareK k, float* scratch)
{
    extern __shared__ float lds[];
    int lane = threadIdx.x;
    size_t wz = (size_t)w;
    float* buf = LDS ? lds : scratch + (size_t)blockIdx.x * 9 * wz;
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
            buf[3 * wz + x] = 0.0f;
            buf[4 * wz + x] = 0.0f;
            buf[5 * wz + x] = 0.0f;
        }
        if (__ballot(any) == 0) continue;   // the row has no highlight: unchanged
        __syncthreads();
        if (lane < 3) {
            const float* src = buf + (size_t)lane * wz;
            float* dst = buf + (size_t)(3 + lane) * wz;
            float* fwd = buf + (size_t)(6 + lane) * wz;
            streak(src, dst, fwd, w, k.a1, k.n1);
            streak(src, dst, fwd, w, k.a2, k.n2);
        }
        __syncthreads();
        for (int x = lane; x < w; x += 64) {
            for (int c = 0; c < 3; ++c) {
                int v = (int)((float)row[3 * x + c] + buf[(size_t)(3 + c) * wz + x] * 255.0f + 0.5f);
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
