// mh_pack.h - conversion of the ABI primitive / attribute records into the
// device records (mh_scene_dev.h). Used on the host (scene upload of a host
// tree) and on the device (scenes whose tree is built on the GPU).
#pragma once

#include "marmel_hip.h"
#include "mh_scene_dev.h"

namespace mh {

template <typename R> __host__ __device__ inline V3<R> cv(const mh_vec3& v) { return V3<R>{(R)v.x, (R)v.y, (R)v.z}; }

__host__ __device__ inline float int_bits_as_float(int v)
{
    union {
        int i;
        float f;
    } u;
    u.i = v;
    return u.f;
}

// fp32 record (layout in mh_scene_dev.h); triangle edges are formed from the
// double vertices, then rounded.
__host__ __device__ inline P32 pack32(const mh_prim& q)
{
    P32 r;
    float kf = int_bits_as_float(q.kind | (q.flags << 8));
    float mf = int_bits_as_float(q.material);
    r.a = make_float4((float)q.p0.x, (float)q.p0.y, (float)q.p0.z, kf);
    r.b = make_float4(0.f, 0.f, 0.f, 0.f);
    r.c = make_float4(0.f, 0.f, 0.f, mf);
    switch (q.kind) {
    case MH_PRIM_TRIANGLE:
        r.b = make_float4((float)(q.p1.x - q.p0.x), (float)(q.p1.y - q.p0.y), (float)(q.p1.z - q.p0.z), 0.f);
        r.c = make_float4((float)(q.p2.x - q.p0.x), (float)(q.p2.y - q.p0.y), (float)(q.p2.z - q.p0.z), mf);
        break;
    case MH_PRIM_SPHERE:
        r.b = make_float4((float)q.r0, 0.f, 0.f, 0.f);
        break;
    case MH_PRIM_CYLINDER:
        r.b = make_float4((float)q.p1.x, (float)q.p1.y, (float)q.p1.z, 0.f);
        r.c = make_float4((float)q.r0, (float)q.r1, 0.f, mf);
        break;
    default: // box, plane
        r.b = make_float4((float)q.p1.x, (float)q.p1.y, (float)q.p1.z, 0.f);
        break;
    }
    return r;
}

__host__ __device__ inline void set_prim_index32(P32& r, int idx) { r.b.w = int_bits_as_float(idx); }

__host__ __device__ inline P64 pack64(const mh_prim& q, int idx)
{
    P64 r;
    r.p0[0] = q.p0.x; r.p0[1] = q.p0.y; r.p0[2] = q.p0.z;
    r.p1[0] = q.p1.x; r.p1[1] = q.p1.y; r.p1[2] = q.p1.z;
    r.p2[0] = q.p2.x; r.p2[1] = q.p2.y; r.p2[2] = q.p2.z;
    r.r0 = q.r0;
    r.r1 = q.r1;
    r.kind = q.kind;
    r.prim = idx;
    r.material = q.material;
    r.flags = q.flags;
    return r;
}

template <typename R> __host__ __device__ inline DAttr<R> conv_attr(const mh_prim_attr& a)
{
    DAttr<R> d;
    d.n0 = cv<R>(a.n0); d.n1 = cv<R>(a.n1); d.n2 = cv<R>(a.n2);
    d.uv0 = cv<R>(a.uv0); d.uv1 = cv<R>(a.uv1); d.uv2 = cv<R>(a.uv2);
    return d;
}

} // namespace mh
