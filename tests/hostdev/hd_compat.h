// hd_compat.h - the few HIP device-side names the render headers use, for a
// host (CPU) build of that code (tests/hostdev, test infrastructure only):
// function qualifiers, the vector record types and bit casts. The render
// headers include this instead of <hip/hip_runtime.h> when MH_HOST_BUILD is
// defined; nothing in the library is built this way.
#pragma once

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <math.h>  // the float overloads of sqrt, floor, ... in the global namespace, as HIP has them

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))

struct float2 { float x, y; };
struct float4 { float x, y, z, w; };
struct int4 { int x, y, z, w; };
struct uint4 { unsigned x, y, z, w; };

inline float2 make_float2(float x, float y) { return float2{x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int4 make_int4(int x, int y, int z, int w) { return int4{x, y, z, w}; }

inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
inline float __uint_as_float(unsigned i) { float f; std::memcpy(&f, &i, 4); return f; }

using std::isfinite;
