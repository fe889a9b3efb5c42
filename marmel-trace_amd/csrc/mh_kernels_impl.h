// mh_kernels_impl.h - kernel bodies and launchers, included by the two
// per-precision translation units.
#pragma once

#include "mh_kernels.h"
#include "mh_render.h"

namespace mh {

// Path index within a sample -> pixel (y * w + x), visiting 8x8 tiles row by
// row and each tile row by row (edge tiles narrower / shorter). A bijection
// on [0, w*h): the estimator per (pixel, sample) is unchanged, but a wave of
// 64 consecutive paths covers one tile, so its camera rays and their first
// shadow rays stay coherent in the tree.
__device__ __forceinline__ uint32_t tile_pixel(uint32_t i, uint32_t w, uint32_t h)
{
    uint32_t ty = i / (8u * w), off = i - ty * 8u * w;
    uint32_t th = h - ty * 8u < 8u ? h - ty * 8u : 8u;
    uint32_t nfull = w / 8u, tx = off / (8u * th);
    uint32_t px, py;
    if (tx < nfull) {
        uint32_t loc = off - tx * 8u * th;
        px = tx * 8u + (loc & 7u);
        py = ty * 8u + (loc >> 3);
    } else {
        uint32_t pw = w - nfull * 8u, loc = off - nfull * 8u * th;
        px = nfull * 8u + loc % pw;
        py = ty * 8u + loc / pw;
    }
    return py * w + px;
}


__device__ __forceinline__ void count_rays(unsigned long long* ctr, int nq)
{
    // wave-level reduction, one atomic per wave
    unsigned long long v = (unsigned long long)nq;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(ctr, v);
}

// One thread per pixel, 16x16-pixel tiles per 256-thread block; the thread
// loops over the pixel's samples in order (same summation order as the
// reference pixel loop), then writes the mean and its 8-bit tone map.
template <typename R, bool FOG>
__global__ __launch_bounds__(256) void k_pathtrace(SceneView<R> sc, CamView<R> cam, int w, int h, int spp, int s0,
                                                   int depth, int tiles_x, int ntiles, LaunchOut out)
{
#if MH_MEGA_LDS
    MH_STACK_LDS(ts, MH_MEGA_LDS, 256)
#else
    MH_STACK_PRIV(ts)
#endif
    // tiles in launch order: workgroups are dealt round-robin over the 8
    // XCDs, so neighbouring tiles run on different XCDs at the same time
    // (measured 61 % faster on the 1080p terrain than giving each XCD a
    // contiguous band of tiles)
    int tile = blockIdx.x;
    int nq = 0;
    if (tile < ntiles) {
        int tx = tile % tiles_x, ty = tile / tiles_x;
        int x = tx * 16 + (threadIdx.x & 15);
        int y = ty * 16 + (threadIdx.x >> 4);
        if (x < w && y < h) {
            V3<R> acc = pt_pixel<FOG>(sc, ts, cam, x, y, w, h, spp, s0, depth, &nq);
            size_t p = ((size_t)y * (size_t)w + (size_t)x) * 3;
            out.linear[p] = (float)acc.x;
            out.linear[p + 1] = (float)acc.y;
            out.linear[p + 2] = (float)acc.z;
            out.rgb[p] = to_byte(acc.x);
            out.rgb[p + 1] = to_byte(acc.y);
            out.rgb[p + 2] = to_byte(acc.z);
        }
    }
    count_rays(out.rays, nq);
}

template <typename R>
__global__ __launch_bounds__(256) void k_whitted(SceneView<R> sc, CamView<R> cam, int w, int h, int spp, int s0,
                                                 int depth, LaunchOut out, WItem<R>* stack, int cap)
{
    MH_STACK_PRIV(ts)
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    int nlanes = gridDim.x * blockDim.x;
    WItem<R>* stk = stack + (size_t)lane * (size_t)cap;
    int nq = 0, ovf = 0;
    int g = (int)(msqrt((double)spp) + 0.5);
    bool strat = g > 0 && g * g == spp;
    for (int i = lane; i < w * h; i += nlanes) {
        int pix = (int)tile_pixel((uint32_t)i, (uint32_t)w, (uint32_t)h);
        int x = pix % w, y = pix / w;
        V3<R> acc = whitted_pixel(sc, ts, cam, x, y, w, h, spp, s0, depth, g, strat, stk, cap, &ovf, &nq);
        size_t p = ((size_t)y * (size_t)w + (size_t)x) * 3;
        out.linear[p] = (float)acc.x;
        out.linear[p + 1] = (float)acc.y;
        out.linear[p + 2] = (float)acc.z;
        out.rgb[p] = to_byte(acc.x);
        out.rgb[p + 1] = to_byte(acc.y);
        out.rgb[p + 2] = to_byte(acc.z);
    }
    if (ovf) atomicOr(out.overflow, 1);
    count_rays(out.rays, nq);
}

// ---- adaptive Whitted ------------------------------------------------------
// One refinement pass: each listed pixel (every pixel when `count` is null)
// takes up to n0 more samples, continuing its global sample index. Sample s
// uses the fixed render's offset (stratified when n0 is a perfect square)
// for s < n0 and uniform jitter after; the pass's sums are formed serially in
// sample order and then added to the pixel's running sums, so the result is
// independent of the launch shape. Follows reference src/render.c:1151-1233.
template <typename R>
__global__ __launch_bounds__(256) void k_adapt_sample(SceneView<R> sc, CamView<R> cam, int w, int h, int n0, int nmax,
                                                      int depth, AdaptBuf ab, const int* count, LaunchOut out,
                                                      WItem<R>* stack, int cap)
{
    MH_STACK_PRIV(ts)
    int lane = blockIdx.x * blockDim.x + threadIdx.x;
    int nlanes = gridDim.x * blockDim.x;
    WItem<R>* stk = stack + (size_t)lane * (size_t)cap;
    int nq = 0, ovf = 0;
    int g = (int)(msqrt((double)n0) + 0.5);
    bool strat = g > 0 && g * g == n0;
    size_t npix = (size_t)w * (size_t)h;
    int items = count ? *count : (int)npix;
    for (int i = lane; i < items; i += nlanes) {
        int p = count ? ab.list[i] : (int)tile_pixel((uint32_t)i, (uint32_t)w, (uint32_t)h);
        int x = p % w, y = p / w;
        int n = ab.ns[p];
        int b = nmax - n < n0 ? nmax - n : n0;
        if (b <= 0) continue;
        double sr = 0.0, sg = 0.0, sb = 0.0, sy = 0.0, sy2 = 0.0;
        for (int k = 0; k < b; ++k) {
            int s = n + k;
            R jx = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 0));
            R jy = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 1));
            R du = jx, dv = jy;
            if (strat && s < n0) {
                du = ((R)(s % g) + jx) / (R)g;
                dv = ((R)(s / g) + jy) / (R)g;
            }
            R uu = ((R)x + du) / (R)w;
            R vv = R(1) - ((R)y + dv) / (R)h;
            R l1 = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 2));
            R l2 = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 3));
            V3<R> o, d;
            camera_ray(cam, uu, vv, l1, l2, &o, &d);
            V3<R> c = whitted_sample(sc, ts, o, d, depth, hash3((uint32_t)x, (uint32_t)y, (uint32_t)s), stk, cap,
                                     &ovf, &nq);
            double cr = (double)c.x, cg = (double)c.y, cb = (double)c.z;
            sr += cr;
            sg += cg;
            sb += cb;
            double Y = 0.2126 * cr + 0.7152 * cg + 0.0722 * cb;
            sy += Y;
            sy2 += Y * Y;
        }
        ab.sum[p] += sr;
        ab.sum[npix + p] += sg;
        ab.sum[2 * npix + p] += sb;
        ab.sum[3 * npix + p] += sy;
        ab.sum[4 * npix + p] += sy2;
        ab.ns[p] = n + b;
    }
    if (ovf) atomicOr(out.overflow, 1);
    count_rays(out.rays, nq);
}

// Active set of the next pass from each pixel's own running statistics:
// refine while n < nmax and (n < 2 or the relative standard error of the
// mean luminance exceeds tau). Appends the pixels to ab.list and their number
// to *count: one atomic per workgroup and round of 1024 pixels (an offset
// shared through LDS), as a counter hit by every wave serialises.
// Follows reference src/render.c:1330-1361.
template <typename R>
__global__ __launch_bounds__(256) void k_adapt_flag(AdaptBuf ab, int w, int h, int nmax, double tau, int* count)
{
    __shared__ int wave_n[4][4];
    __shared__ int base_at;
    int npix = w * h;
    int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int base = blockIdx.x * 1024; base < npix; base += gridDim.x * 1024) {
        int flag[4];
        unsigned long long m[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // pixels in 8x8-tile order: the list, and so each refinement
            // pass's waves, keep neighbouring pixels together
            int li = base + r * 256 + threadIdx.x;
            int p = li < npix ? (int)tile_pixel((uint32_t)li, (uint32_t)w, (uint32_t)h) : npix;
            int f = 0;
            if (li < npix) {
                int n = ab.ns[p];
                if (n < nmax) {
                    if (n < 2) {
                        f = 1;
                    } else {
                        double nf = (double)n;
                        double sy = ab.sum[3 * (size_t)npix + p], sy2 = ab.sum[4 * (size_t)npix + p];
                        double ybar = sy / nf;
                        double var = (sy2 - sy * ybar) / (nf - 1.0);
                        if (var < 0.0) var = 0.0;
                        double e = sqrt(var / nf) / (ybar + 1e-4);
                        f = e > tau;
                    }
                }
            }
            flag[r] = f;
            m[r] = __ballot(f);
            if (lane == 0) wave_n[r][wave] = __popcll(m[r]);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int tot = 0;
            for (int r = 0; r < 4; ++r)
                for (int q = 0; q < 4; ++q) tot += wave_n[r][q];
            base_at = tot ? atomicAdd(count, tot) : 0;
        }
        __syncthreads();
        int at = base_at;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            int before = 0;
            for (int rr = 0; rr < r; ++rr)
                for (int q = 0; q < 4; ++q) before += wave_n[rr][q];
            for (int q = 0; q < wave; ++q) before += wave_n[r][q];
            if (flag[r]) {
                int li = base + r * 256 + threadIdx.x;
                ab.list[at + before + __popcll(m[r] & ((1ull << lane) - 1ull))] =
                    (int)tile_pixel((uint32_t)li, (uint32_t)w, (uint32_t)h);
            }
        }
        __syncthreads();  // wave_n and base_at are reused by the next round
    }
}

// Mean of each pixel's samples -> linear float and 8-bit output; the sum and
// maximum of the per-pixel sample counts, reduced per workgroup before one
// atomic each (reference src/render.c:1474-1494).
template <typename R>
__global__ __launch_bounds__(256) void k_adapt_resolve(AdaptBuf ab, int npix, LaunchOut out)
{
    __shared__ long long s_tot[4];
    __shared__ int s_max[4];
    long long tot = 0;
    int mx = 0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        int n = ab.ns[p];
        double inv = 1.0 / (double)n;
        size_t o = (size_t)p * 3;
        for (int c = 0; c < 3; ++c) {
            double m = ab.sum[(size_t)c * (size_t)npix + p] * inv;
            out.linear[o + c] = (float)m;
            out.rgb[o + c] = to_byte(m);
        }
        tot += n;
        mx = n > mx ? n : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
        tot += __shfl_xor(tot, off);
        int o = __shfl_xor(mx, off);
        mx = o > mx ? o : mx;
    }
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_tot[wave] = tot;
        s_max[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        int m = 0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
            t += s_tot[q];
            m = s_max[q] > m ? s_max[q] : m;
        }
        if (t) {
            atomicAdd(ab.total, (unsigned long long)t);
            atomicMax(ab.max_n, m);
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_closest(SceneView<R> sc, const mh_ray* rays, int n, double tmin, double tmax,
                                                 mh_hit* out)
{
    MH_STACK_PRIV(ts)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mh_ray r = rays[i];
    V3<R> o = mk((R)r.ox, (R)r.oy, (R)r.oz);
    V3<R> d = unit_dir(mk((R)r.dx, (R)r.dy, (R)r.dz));
    DHit<R> h;
    mh_hit res;
    res.reserved = 0;
    if (intersect(sc, ts, o, d, (R)tmin, (R)tmax, &h)) {
        res.t = (double)h.t;
        res.prim_index = h.prim;
        res.material = h.mat;
        res.front_face = h.front;
        res.nx = (double)h.n.x;
        res.ny = (double)h.n.y;
        res.nz = (double)h.n.z;
        res.u = (double)h.u;
        res.v = (double)h.v;
    } else {
        res.t = 0.0;
        res.prim_index = -1;
        res.material = 0;
        res.front_face = 0;
        res.nx = res.ny = res.nz = res.u = res.v = 0.0;
    }
    out[i] = res;
}

template <typename R>
__global__ __launch_bounds__(256) void k_any(SceneView<R> sc, const mh_ray* rays, int n, double tmin, double tmax,
                                             uint8_t* out)
{
    MH_STACK_PRIV(ts)
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    mh_ray r = rays[i];
    V3<R> o = mk((R)r.ox, (R)r.oy, (R)r.oz);
    V3<R> d = unit_dir(mk((R)r.dx, (R)r.dy, (R)r.dz));
    out[i] = any_hit(sc, ts, o, d, (R)tmin, (R)tmax) ? 1 : 0;
}

template <typename R>
hipError_t launch_pathtrace(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int spp, int s0, int depth,
                            const LaunchOut& out, int block, hipStream_t st)
{
    (void)block;
    int tiles_x = (w + 15) / 16, tiles_y = (h + 15) / 16;
    int ntiles = tiles_x * tiles_y;
    int nblocks = (ntiles + 7) / 8 * 8;
    // the fog variant only for scenes that need it: the fog code raises the
    // kernel's register count
    auto kern = sc.fog_kernel ? k_pathtrace<R, true> : k_pathtrace<R, false>;
    hipLaunchKernelGGL(kern, dim3(nblocks), dim3(256), 0, st, sc, cam, w, h, spp, s0, depth, tiles_x, ntiles, out);
    return hipGetLastError();
}

template <typename R> size_t whitted_item_bytes() { return sizeof(WItem<R>); }

template <typename R>
hipError_t launch_whitted(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int spp, int s0, int depth,
                          const LaunchOut& out, void* stack, int cap, int threads, hipStream_t st)
{
    int nblocks = threads / 256;
    hipLaunchKernelGGL(k_whitted<R>, dim3(nblocks), dim3(256), 0, st, sc, cam, w, h, spp, s0, depth, out,
                       (WItem<R>*)stack, cap);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_whitted_adaptive(const SceneView<R>& sc, const CamView<R>& cam, int w, int h, int n0, int nmax,
                                   double tau, int depth, const AdaptBuf& ab, int max_passes, const LaunchOut& out,
                                   void* stack, int cap, int threads, hipStream_t st)
{
    int npix = w * h;
    int nblocks = threads / 256;
    int fblocks = (npix + 1023) / 1024;  // flag: 1024 pixels per workgroup and round
    if (fblocks > 1024) fblocks = 1024;
    int rblocks = (npix + 255) / 256;    // resolve: grid-strided, one atomic pair per workgroup
    if (rblocks > 1024) rblocks = 1024;
    hipLaunchKernelGGL(k_adapt_sample<R>, dim3(nblocks), dim3(256), 0, st, sc, cam, w, h, n0, nmax, depth, ab,
                       (const int*)nullptr, out, (WItem<R>*)stack, cap);
    for (int pass = 1; pass < max_passes; ++pass) {
        hipLaunchKernelGGL(k_adapt_flag<R>, dim3(fblocks), dim3(256), 0, st, ab, w, h, nmax, tau, ab.count + pass);
        hipLaunchKernelGGL(k_adapt_sample<R>, dim3(nblocks), dim3(256), 0, st, sc, cam, w, h, n0, nmax, depth, ab,
                           (const int*)(ab.count + pass), out, (WItem<R>*)stack, cap);
    }
    hipLaunchKernelGGL(k_adapt_resolve<R>, dim3(rblocks), dim3(256), 0, st, ab, npix, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_closest(const SceneView<R>& sc, const mh_ray* rays, int n, double tmin, double tmax, mh_hit* out,
                          hipStream_t st)
{
    hipLaunchKernelGGL(k_closest<R>, dim3((n + 255) / 256), dim3(256), 0, st, sc, rays, n, tmin, tmax, out);
    return hipGetLastError();
}

template <typename R>
hipError_t launch_any(const SceneView<R>& sc, const mh_ray* rays, int n, double tmin, double tmax, uint8_t* out,
                      hipStream_t st)
{
    hipLaunchKernelGGL(k_any<R>, dim3((n + 255) / 256), dim3(256), 0, st, sc, rays, n, tmin, tmax, out);
    return hipGetLastError();
}

} // namespace mh

#define MH_INSTANTIATE(R)                                                                                          \
    template hipError_t mh::launch_pathtrace<R>(const SceneView<R>&, const CamView<R>&, int, int, int, int, int,  \
                                                const LaunchOut&, int, hipStream_t);                               \
    template hipError_t mh::launch_whitted<R>(const SceneView<R>&, const CamView<R>&, int, int, int, int, int,    \
                                              const LaunchOut&, void*, int, int, hipStream_t);                     \
    template size_t mh::whitted_item_bytes<R>();                                                                   \
    template hipError_t mh::launch_whitted_adaptive<R>(const SceneView<R>&, const CamView<R>&, int, int, int, int, \
                                                       double, int, const AdaptBuf&, int, const LaunchOut&, void*, \
                                                       int, int, hipStream_t);                                     \
    template hipError_t mh::launch_closest<R>(const SceneView<R>&, const mh_ray*, int, double, double, mh_hit*,    \
                                              hipStream_t);                                                        \
    template hipError_t mh::launch_any<R>(const SceneView<R>&, const mh_ray*, int, double, double, uint8_t*,       \
                                          hipStream_t);
