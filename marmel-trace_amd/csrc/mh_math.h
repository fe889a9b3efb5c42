// mh_math.h - device vector math, integer-hash PRNG and noise for the
// marmel HIP path tracer. Templated on the arithmetic type R (float for the
// throughput path, double for the parity path).
//
// Operation order follows the reference vector helpers (src/vec3.h, vec3.c)
// so that the double instantiation, compiled without FMA contraction,
// performs the same IEEE operations as the reference.
#pragma once

#if defined(MH_HOST_BUILD)
#include "hd_compat.h"  // CPU build of the render code for sanitizer runs (tests/hostdev)
#include <cassert>
// bounds checks that only the CPU build compiles (the GPU code masks these
// indices and relies on the bounds proved at upload)
#define MH_HOST_ASSERT(c) assert(c)
#else
#include <hip/hip_runtime.h>
#define MH_HOST_ASSERT(c) ((void)0)
#endif
#include <stdint.h>

namespace mh {

template <typename R> struct V3 {
    R x, y, z;
};

template <typename R> __device__ __forceinline__ V3<R> mk(R x, R y, R z) { return V3<R>{x, y, z}; }
template <typename R> __device__ __forceinline__ V3<R> operator+(V3<R> a, V3<R> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <typename R> __device__ __forceinline__ V3<R> operator-(V3<R> a, V3<R> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <typename R> __device__ __forceinline__ V3<R> operator-(V3<R> a) { return mk(-a.x, -a.y, -a.z); }
template <typename R> __device__ __forceinline__ V3<R> operator*(V3<R> a, R s) { return mk(a.x * s, a.y * s, a.z * s); }
template <typename R> __device__ __forceinline__ V3<R> vmul(V3<R> a, V3<R> b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
template <typename R> __device__ __forceinline__ R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> __device__ __forceinline__ V3<R> cross(V3<R> a, V3<R> b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <typename R> __device__ __forceinline__ R len2(V3<R> a) { return dot(a, a); }
template <typename R> __device__ __forceinline__ V3<R> lerp(V3<R> a, V3<R> b, R t)
{
    return mk(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}
template <typename R> __device__ __forceinline__ V3<R> reflect(V3<R> i, V3<R> n)
{
    R d = dot(i, n);
    return mk(i.x - R(2) * d * n.x, i.y - R(2) * d * n.y, i.z - R(2) * d * n.z);
}
// Hit point o + t d, each product rounded before the sum (never a fused
// multiply-add), as the reference's double code rounds it. On a surface
// that lies on a coordinate plane (a ground plane y = 0, a box face) the
// rounded sum is then exactly the plane's coordinate for almost every ray,
// as in the reference; a fused o.y + t d.y lands on either side of it half
// the time, which flips floor() in the 3D checker texture.
template <typename R> __device__ __forceinline__ V3<R> at(V3<R> o, V3<R> d, R t)
{
#pragma clang fp contract(off)
    return mk(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z);
}
template <typename R> __device__ __forceinline__ bool finite3(V3<R> v)
{
    return isfinite(v.x) && isfinite(v.y) && isfinite(v.z);
}
template <typename R> __device__ __forceinline__ V3<R> zero3() { return mk(R(0), R(0), R(0)); }

// Zero vector when |a|^2 <= 1e-12, else a * (1/|a|).
template <typename R> __device__ __forceinline__ V3<R> normalize(V3<R> a)
{
    R l2 = len2(a);
    if (l2 <= R(1e-12)) return zero3<R>();
    return a * (R(1) / sqrt(l2));
}

// Accurate scalar math for both precisions (the fp32 instantiation resolves
// to the float overloads).
__device__ __forceinline__ float msqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double msqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float mpow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double mpow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float mexp(float a) { return expf(a); }
__device__ __forceinline__ double mexp(double a) { return exp(a); }
__device__ __forceinline__ float mfloor(float a) { return floorf(a); }
__device__ __forceinline__ double mfloor(double a) { return floor(a); }
__device__ __forceinline__ float matan2(float a, float b) { return atan2f(a, b); }
__device__ __forceinline__ double matan2(double a, double b) { return atan2(a, b); }
__device__ __forceinline__ float masin(float a) { return asinf(a); }
__device__ __forceinline__ double masin(double a) { return asin(a); }
__device__ __forceinline__ float mtan(float a) { return tanf(a); }
__device__ __forceinline__ double mtan(double a) { return tan(a); }
__device__ __forceinline__ void msincos(float a, float* s, float* c) { sincosf(a, s, c); }
__device__ __forceinline__ void msincos(double a, double* s, double* c) { sincos(a, s, c); }
__device__ __forceinline__ float mcos(float a) { return cosf(a); }
__device__ __forceinline__ double mcos(double a) { return cos(a); }
template <typename R> __device__ __forceinline__ R mfabs(R a) { return a < R(0) ? -a : a; }
// fmin/fmax with the C99 NaN rule (return the non-NaN operand).
__device__ __forceinline__ float mfmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double mfmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float mfmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double mfmax(double a, double b) { return fmax(a, b); }

template <typename R> struct Consts;
template <> struct Consts<float> {
    static constexpr float pi = 3.14159265358979323846f;
    static constexpr float umax = 0.99999994f; // largest float below 1
};
template <> struct Consts<double> {
    static constexpr double pi = 3.14159265358979323846;
    static constexpr double umax = 0.9999999999999999;
};

// ---- integer-hash PRNG (reference src/pathtrace.c:92-117) -------------

__device__ __forceinline__ uint32_t hash_u32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__device__ __forceinline__ uint32_t hash3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t h = 0x9e3779b9u ^ hash_u32(a + 0x85ebca6bu);
    h = hash_u32(h + b);
    h ^= hash_u32(c + 0xc2b2ae35u);
    return hash_u32(h);
}

// 24 bits -> [0,1): exact in both float and double.
template <typename R> __device__ __forceinline__ R rand01(uint32_t a, uint32_t b, uint32_t c)
{
    return R(hash3(a, b, c) >> 8) * R(1.0 / 16777216.0);
}

// ---- 2D gradient noise and fBm (reference src/noise.c:31-61,158-187,272-294)

__device__ __forceinline__ uint32_t lattice_hash(int x, int y, uint32_t seed)
{
    uint32_t h = seed + 0x9e3779b9u;
    h ^= (uint32_t)x * 0x85ebca6bu;
    h ^= (uint32_t)y * 0xc2b2ae35u;
    h ^= 0u * 0x27d4eb2fu;
    return hash_u32(h);
}

template <typename R> __device__ __forceinline__ R fade5(R t)
{
    return t * t * t * (t * (t * R(6) - R(15)) + R(10));
}
template <typename R> __device__ __forceinline__ R mixs(R a, R b, R t) { return a + t * (b - a); }
template <typename R> __device__ __forceinline__ R clampr(R v, R lo, R hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Corner gradient g(angle) . (dx, dy) with angle = hash / (2^32-1) * 2 pi.
template <typename R> __device__ __forceinline__ R corner(int ix, int iy, uint32_t seed, R dx, R dy)
{
    R a = (R)lattice_hash(ix, iy, seed) / (R)4294967295.0 * R(2) * Consts<R>::pi;
    R s, c;
    msincos(a, &s, &c);
    return c * dx + s * dy;
}

template <typename R> __device__ R perlin2(R x, R y, uint32_t seed)
{
    R fx = mfloor(x), fy = mfloor(y);
    int ix = (int)fx, iy = (int)fy;
    R rx = x - fx, ry = y - fy;
    R u = fade5(rx), v = fade5(ry);
    R d00 = corner<R>(ix, iy, seed, rx, ry);
    R d10 = corner<R>(ix + 1, iy, seed, rx - R(1), ry);
    R d01 = corner<R>(ix, iy + 1, seed, rx, ry - R(1));
    R d11 = corner<R>(ix + 1, iy + 1, seed, rx - R(1), ry - R(1));
    R val = mixs(mixs(d00, d10, u), mixs(d01, d11, u), v);
    return clampr(val * R(1.4142135623730951), R(-1), R(1));
}

template <typename R> __device__ R fbm2(R x, R y, int octaves, uint32_t seed)
{
    int n = octaves <= 0 ? 1 : octaves;
    R amp = R(1), freq = R(1), sum = R(0), norm = R(0);
    for (int i = 0; i < n; ++i) {
        sum += amp * perlin2<R>(x * freq, y * freq, seed);
        norm += amp;
        amp *= R(0.5);
        freq *= R(2);
    }
    return clampr(sum / norm, R(-1), R(1));
}

// ---- 3D gradient noise and fBm (reference src/noise.c:189-253, 296-317)

__device__ __forceinline__ uint32_t lattice_hash3(int x, int y, int z, uint32_t seed)
{
    uint32_t h = seed + 0x9e3779b9u;
    h ^= (uint32_t)x * 0x85ebca6bu;
    h ^= (uint32_t)y * 0xc2b2ae35u;
    h ^= (uint32_t)z * 0x27d4eb2fu;
    return hash_u32(h);
}

template <typename R> __device__ __forceinline__ R lattice_signed3(int x, int y, int z, uint32_t seed)
{
    return (R)lattice_hash3(x, y, z, seed) / (R)4294967295.0 * R(2) - R(1);
}

// Per corner a gradient of three signed lattice values (seed, seed ^
// 0x68bc21eb, seed ^ 0x02e5be93), normalised (the x axis when degenerate),
// dotted with the corner offset; quintic trilinear blend, x then y then z;
// scaled by 2/sqrt(3) and clamped.
template <typename R> __device__ R perlin3(R x, R y, R z, uint32_t seed)
{
    R fx = mfloor(x), fy = mfloor(y), fz = mfloor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    R rx = x - fx, ry = y - fy, rz = z - fz;
    R dv[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        R gx = lattice_signed3<R>(ix + ox, iy + oy, iz + oz, seed);
        R gy = lattice_signed3<R>(ix + ox, iy + oy, iz + oz, seed ^ 0x68bc21ebu);
        R gz = lattice_signed3<R>(ix + ox, iy + oy, iz + oz, seed ^ 0x02e5be93u);
        R len = msqrt(gx * gx + gy * gy + gz * gz);
        if (len < R(1e-12)) {
            gx = R(1);
            gy = R(0);
            gz = R(0);
        } else {
            gx /= len;
            gy /= len;
            gz /= len;
        }
        dv[c] = gx * (rx - (R)ox) + gy * (ry - (R)oy) + gz * (rz - (R)oz);
    }
    R u = fade5(rx), v = fade5(ry), w = fade5(rz);
    R a = mixs(mixs(dv[0], dv[1], u), mixs(dv[2], dv[3], u), v);
    R b = mixs(mixs(dv[4], dv[5], u), mixs(dv[6], dv[7], u), v);
    return clampr(mixs(a, b, w) * R(1.1547005383792515), R(-1), R(1));
}

// 3 octaves, lacunarity 2, gain 0.5 (the fog turbulence).
template <typename R> __device__ R fbm3(R x, R y, R z, uint32_t seed)
{
    R amp = R(1), freq = R(1), sum = R(0), norm = R(0);
    for (int i = 0; i < 3; ++i) {
        sum += amp * perlin3<R>(x * freq, y * freq, z * freq, seed);
        norm += amp;
        amp *= R(0.5);
        freq *= R(2);
    }
    return clampr(sum / norm, R(-1), R(1));
}

} // namespace mh
