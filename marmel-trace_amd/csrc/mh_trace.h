// mh_trace.h - ray/primitive intersection and BVH2 traversal on the GPU.
//
// Primitive tests follow reference src/geometry.c (sphere 338-369, plane
// 372-393, box 396-479, triangle 482-538, cylinder 549-699, occlusion
// 1373-1461); traversal follows the nearest-hit semantics of reference
// src/bvh.c:557-712 (plane side list first, then the tree; equal t resolves
// to the lowest prim index) and the any-hit semantics of src/bvh.c:714-813.
//
// The traversal is a stack walk over 64-byte nodes holding both child boxes,
// so one node fetch tests two boxes and pushes the far child only. Nodes are
// culled when their entry distance exceeds the running closest hit (strict >,
// which keeps equal-t ties reachable). The fp32 path widens every slab
// interval by a few ulps (robust traversal), the fp64 path runs the slab test
// in double over fp32 boxes that the host rounded outward, so it never culls a
// node the double-precision reference would visit.
#pragma once

#include "mh_scene_dev.h"

namespace mh {

// fp32 box hits put on the face's plane (1), or left at o + t d (0: the
// adjoint kernels, whose register budget sits at the 256-VGPR ceiling,
// where it costs half their occupancy).
#ifndef MH_SNAP_BOX
#define MH_SNAP_BOX 1
#endif

template <typename R> struct PD {
    int kind, flags, prim, mat;
    V3<R> p0, p1, p2; // fp32: triangles carry (v0, e1, e2)
    R r0, r1;
};

// Load a 16-byte-aligned record through a global-address-space pointer, in
// 16-byte vectors. Scene arrays are device allocations, but a pointer read
// from a device table (the batched engine's scenes) is generic to the
// compiler, which then fetches with flat instructions that also wait on LDS
// traffic (the traversal stack).
// Read of an LDS (__shared__) slot as an LDS access that the compiler may
// neither drop nor merge: a read of one of two places (an LDS slot or a
// private overflow array) would otherwise become a single flat load after
// the branch, which also waits on vector memory.
template <typename T> __device__ __forceinline__ T lds_read(const T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const volatile __attribute__((address_space(3))) T*)p;
#else
    return *p;
#endif
}

typedef unsigned int mh_u4v __attribute__((ext_vector_type(4)));
template <typename T> __device__ __forceinline__ T ldg(const T* p)
{
    static_assert(sizeof(T) % 16 == 0, "16-byte records");
    T r;
#if defined(__HIP_DEVICE_COMPILE__)
    const __attribute__((address_space(1))) mh_u4v* s = (const __attribute__((address_space(1))) mh_u4v*)p;
    mh_u4v v[sizeof(T) / 16];
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 16); ++i) v[i] = s[i];
    __builtin_memcpy(&r, v, sizeof(T));
#else
    r = *p;
#endif
    return r;
}

__device__ __forceinline__ PD<float> load_prim(const P32* p)
{
    P32 r = ldg(p);
    float4 a = r.a, b = r.b, c = r.c;
    PD<float> d;
    int kf = __float_as_int(a.w);
    d.kind = kf & 0xff;
    d.flags = kf >> 8;
    d.prim = __float_as_int(b.w);
    d.mat = __float_as_int(c.w);
    d.p0 = mk(a.x, a.y, a.z);
    d.p1 = mk(b.x, b.y, b.z);
    d.p2 = mk(c.x, c.y, c.z);
    d.r0 = d.kind == K_SPHERE ? b.x : c.x;
    d.r1 = c.y;
    return d;
}

__device__ __forceinline__ PD<double> load_prim(const P64* p)
{
    PD<double> d;
    d.kind = p->kind;
    d.flags = p->flags;
    d.prim = p->prim;
    d.mat = p->material;
    d.p0 = mk(p->p0[0], p->p0[1], p->p0[2]);
    d.p1 = mk(p->p1[0], p->p1[1], p->p1[2]);
    d.p2 = mk(p->p2[0], p->p2[1], p->p2[2]);
    d.r0 = p->r0;
    d.r1 = p->r1;
    return d;
}

template <typename R> struct Eps;
template <> struct Eps<float> {
    static constexpr float bary = 1e-6f;   // closes fp32 cracks along shared edges
};
template <> struct Eps<double> {
    static constexpr double bary = 1e-9;   // reference GEO_BARY_EPS
};

template <typename R> struct DHit {
    R t, u, v;
    V3<R> p, n;
    int front, prim, mat;
};

template <typename R>
__device__ __forceinline__ void store_hit(DHit<R>* h, R t, V3<R> p, V3<R> n, V3<R> d, int mat, R u, R v)
{
    bool front = dot(d, n) < R(0);
    h->t = t;
    h->p = p;
    h->n = front ? n : -n;
    h->front = front ? 1 : 0;
    h->mat = mat;
    h->u = u;
    h->v = v;
}

// ---- sphere -------------------------------------------------------------

template <typename R, bool FULL>
__device__ __forceinline__ bool hit_sphere(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout, DHit<R>* h)
{
    if (p.r0 <= R(0)) return false;
    V3<R> l = o - p.p0;
    R hb = dot(l, d);
    R c = len2(l) - p.r0 * p.r0;
    R disc = hb * hb - c;
    if (disc < R(0)) return false;
    R sq = msqrt(disc);
    R t = -hb - sq;
    if (t < tmin) {
        t = -hb + sq;
        if (t < tmin || t > tmax) return false;
    } else if (t > tmax) {
        return false;
    }
    *tout = t;
    if (FULL) {
        V3<R> pt = at(o, d, t);
        V3<R> n = (pt - p.p0) * (R(1) / p.r0);
        if constexpr (sizeof(R) == 4) {
            // fp32: o + t d is off the sphere by about |o| t ulp, which at a
            // hundred units from the ray origin reaches the 1e-4 offset of the
            // next ray; project the point back onto the sphere
            // |n| = 1 + e with e of a few ulp * |o| t / r: one Newton step of
            // 1 / |n| (1.5 - 0.5 |n|^2) removes e to second order
            n = n * (R(1.5) - R(0.5) * len2(n));
            pt = p.p0 + n * p.r0;
        }
        R phi = matan2(n.z, n.x);
        R th = masin(mfmax(R(-1), mfmin(R(1), n.y)));
        R u = R(0.5) + phi / (R(2) * Consts<R>::pi);
        R v = R(0.5) + th / Consts<R>::pi;
        u = u - mfloor(u);
        v = mfmax(R(0), mfmin(R(1), v));
        store_hit(h, t, pt, n, d, p.mat, u, v);
    }
    return true;
}

// any-hit sphere: nearest root of t^2 + 2 hb t + c in [tmin, tmax]
template <typename R>
__device__ __forceinline__ bool occ_sphere(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax)
{
    if (p.r0 <= R(0)) return false;
    V3<R> l = o - p.p0;
    R hb = dot(l, d);
    R c = len2(l) - p.r0 * p.r0;
    R disc = hb * hb - c;
    if (disc < R(0)) return false;
    R sq = msqrt(disc);
    R t0 = -hb - sq, t1 = -hb + sq;
    if (t0 > t1) { R tt = t0; t0 = t1; t1 = tt; }
    return (t0 >= tmin && t0 <= tmax) || (t1 >= tmin && t1 <= tmax);
}

// ---- plane ----------------------------------------------------------------

template <typename R, bool FULL>
__device__ __forceinline__ bool hit_plane(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout, DHit<R>* h)
{
    V3<R> n = normalize(p.p1);
    R den = dot(d, n);
    if (mfabs(den) < R(1e-12)) return false;
    R t = dot(p.p0 - o, n) / den;
    if (t < tmin || t > tmax) return false;
    *tout = t;
    if (FULL) {
        V3<R> pt = at(o, d, t);
        V3<R> hint = mfabs(n.y) > R(0.9) ? mk(R(0), R(0), R(1)) : mk(R(0), R(1), R(0));
        V3<R> T = normalize(cross(hint, n));
        V3<R> B = cross(n, T);
        V3<R> df = pt - p.p0;
        R u = dot(df, T), v = dot(df, B);
        u = u - mfloor(u);
        v = v - mfloor(v);
        store_hit(h, t, pt, n, d, p.mat, u, v);
    }
    return true;
}

// ---- box ------------------------------------------------------------------

template <typename R, bool FULL>
__device__ bool hit_box(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout, DHit<R>* h)
{
    V3<R> lo = p.p0 - p.p1, hi = p.p0 + p.p1;
    R os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
    R los[3] = {lo.x, lo.y, lo.z}, his[3] = {hi.x, hi.y, hi.z};
    R tn = -INFINITY, tf = INFINITY;
    int ea = -1, es = 0, xa = -1, xs = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R inv = R(1) / ds[i];
        R t1 = (los[i] - os[i]) * inv, t2 = (his[i] - os[i]) * inv;
        R tnear = mfmin(t1, t2), tfar = mfmax(t1, t2);
        if (tnear > tn) { tn = tnear; ea = i; es = inv > R(0) ? -1 : 1; }
        if (tfar < tf) { tf = tfar; xa = i; xs = inv > R(0) ? 1 : -1; }
        if (tf < tn) return false;
    }
    R th;
    int axis, sgn;
    if (tn >= tmin && tn <= tmax) { th = tn; axis = ea; sgn = es; }
    else if (tf >= tmin && tf <= tmax) { th = tf; axis = xa; sgn = xs; }
    else return false;
    *tout = th;
    if (FULL) {
        V3<R> pt = at(o, d, th);
        // fp32: the hit point put on the face's own plane (see hit_sphere)
        constexpr bool snap = sizeof(R) == 4 && MH_SNAP_BOX;
        V3<R> n;
        if (axis < 0) {
            n = -d;
        } else {
            n = zero3<R>();
            if (axis == 0) {
                n.x = (R)sgn;
                if (snap) pt.x = sgn > 0 ? hi.x : lo.x;
            } else if (axis == 1) {
                n.y = (R)sgn;
                if (snap) pt.y = sgn > 0 ? hi.y : lo.y;
            } else {
                n.z = (R)sgn;
                if (snap) pt.z = sgn > 0 ? hi.z : lo.z;
            }
        }
        R u = R(0), v = R(0);
        R dx = his[0] - los[0], dy = his[1] - los[1], dz = his[2] - los[2];
        if (axis == 0) {
            u = dz > R(1e-12) ? (pt.z - los[2]) / dz : R(0);
            v = dy > R(1e-12) ? (pt.y - los[1]) / dy : R(0);
        } else if (axis == 1) {
            u = dx > R(1e-12) ? (pt.x - los[0]) / dx : R(0);
            v = dz > R(1e-12) ? (pt.z - los[2]) / dz : R(0);
        } else {
            u = dx > R(1e-12) ? (pt.x - los[0]) / dx : R(0);
            v = dy > R(1e-12) ? (pt.y - los[1]) / dy : R(0);
        }
        u = mfmax(R(0), mfmin(R(1), u));
        v = mfmax(R(0), mfmin(R(1), v));
        store_hit(h, th, pt, n, d, p.mat, u, v);
    }
    return true;
}

template <typename R>
__device__ bool occ_box(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax)
{
    V3<R> lo = p.p0 - p.p1, hi = p.p0 + p.p1;
    R os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
    R los[3] = {lo.x, lo.y, lo.z}, his[3] = {hi.x, hi.y, hi.z};
    R tn = -INFINITY, tf = INFINITY;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R inv = R(1) / ds[i];
        R t1 = (los[i] - os[i]) * inv, t2 = (his[i] - os[i]) * inv;
        if (t1 > t2) { R tt = t1; t1 = t2; t2 = tt; }
        if (t1 > tn) tn = t1;
        if (t2 < tf) tf = t2;
        if (tn > tf) return false;
    }
    return (tn >= tmin && tn <= tmax) || (tf >= tmin && tf <= tmax);
}

// ---- triangle (Moller-Trumbore, two-sided) ---------------------------------

template <typename R> __device__ __forceinline__ void tri_edges(const PD<R>& p, V3<R>* e1, V3<R>* e2);
template <> __device__ __forceinline__ void tri_edges<float>(const PD<float>& p, V3<float>* e1, V3<float>* e2)
{
    *e1 = p.p1;
    *e2 = p.p2;
}
template <> __device__ __forceinline__ void tri_edges<double>(const PD<double>& p, V3<double>* e1, V3<double>* e2)
{
    *e1 = p.p1 - p.p0;
    *e2 = p.p2 - p.p0;
}

template <typename R, bool FULL>
__device__ __forceinline__ bool hit_tri(const PD<R>& p, const DAttr<R>* attr, V3<R> o, V3<R> d, R tmin, R tmax, R* tout,
                                        DHit<R>* h)
{
    V3<R> e1, e2;
    tri_edges(p, &e1, &e2);
    V3<R> pv = cross(d, e2);
    R det = dot(e1, pv);
    if (mfabs(det) < R(1e-9)) return false;
    R idet = R(1) / det;
    V3<R> tv = o - p.p0;
    R u = dot(tv, pv) * idet;
    if (u < -Eps<R>::bary || u > R(1) + Eps<R>::bary) return false;
    V3<R> qv = cross(tv, e1);
    R v = dot(d, qv) * idet;
    if (v < -Eps<R>::bary || u + v > R(1) + Eps<R>::bary) return false;
    R t = dot(e2, qv) * idet;
    if (t < tmin || t > tmax) return false;
    *tout = t;
    if (FULL) {
        // fp64: o + t d as the reference forms it. fp32: from the barycentric
        // coordinates, accurate to the triangle's own coordinates; o + t d is
        // off the plane by about |o| t ulp, which at a hundred units from the
        // ray origin reaches the 1e-4 offset of the next ray (paths then
        // re-hit their own surface: the far terrain of the 1080p frame came
        // out 1.6 % too dark)
        V3<R> pt;
        if constexpr (sizeof(R) == 4) pt = p.p0 + (e1 * u + e2 * v);
        else pt = at(o, d, t);
        V3<R> n;
        bool have = false;
        if ((p.flags & F_SMOOTH) && attr) {
            R w = R(1) - u - v;
            n = attr->n0 * w + (attr->n1 * u + attr->n2 * v);
            R nl = msqrt(len2(n));
            if (nl > R(1e-12)) {
                n = n * (R(1) / nl);
                have = true;
            }
        }
        if (!have) {
            n = cross(e1, e2);
            R nl = msqrt(len2(n));
            if (nl < R(1e-15)) return false;
            n = n * (R(1) / nl);
        }
        R hu = u, hv = v;
        if ((p.flags & F_UV) && attr) {
            R w = R(1) - u - v;
            hu = w * attr->uv0.x + u * attr->uv1.x + v * attr->uv2.x;
            hv = w * attr->uv0.y + u * attr->uv1.y + v * attr->uv2.y;
        }
        store_hit(h, t, pt, n, d, p.mat, hu, hv);
    }
    return true;
}

// ---- tapered capped cylinder -------------------------------------------------

template <typename R, bool FULL>
__device__ bool hit_cyl(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax, R* tout, DHit<R>* h)
{
    V3<R> A = p.p1 - p.p0;
    R hgt = msqrt(len2(A));
    if (hgt < R(1e-12)) return false;
    V3<R> ax = A * (R(1) / hgt);
    R r0 = p.r0, r1 = p.r1;
    R k = (r1 - r0) / hgt;
    V3<R> d0 = o - p.p0;
    R od = dot(d, ax), d0d = dot(d0, ax);
    V3<R> dp = d - ax * od;
    V3<R> d0p = d0 - ax * d0d;
    R ra = k * od, rb = r0 + k * d0d;
    R qa = len2(dp) - ra * ra;
    R qb = dot(dp, d0p) - rb * ra;
    R qc = len2(d0p) - rb * rb;
    bool found = false;
    int is_cap = 0, cap_id = 0;
    R bt = R(0), bs = R(0);
    V3<R> bn = zero3<R>();
    if (mfabs(qa) > R(1e-15)) {
        R disc = qb * qb - qa * qc;
        if (disc >= R(0)) {
            R sq = msqrt(disc);
            R ro0 = (-qb - sq) / qa, ro1 = (-qb + sq) / qa;
            if (ro0 > ro1) { R tt = ro0; ro0 = ro1; ro1 = tt; }
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                R t = i == 0 ? ro0 : ro1;
                if (t < tmin || t > tmax) continue;
                R s = (d0d + t * od) / hgt;
                if (s < R(0) || s > R(1)) continue;
                if (!found || t < bt) {
                    found = true;
                    bt = t;
                    bs = s;
                    is_cap = 0;
                    if (FULL) {
                        V3<R> pt = at(o, d, t);
                        V3<R> ap = p.p0 + ax * (s * hgt);
                        V3<R> rad = pt - ap;
                        R rl = msqrt(len2(rad));
                        V3<R> rd = rl > R(1e-15) ? rad * (R(1) / rl) : zero3<R>();
                        V3<R> n = normalize(rd * hgt - ax * (r1 - r0));
                        if (len2(n) < R(0.5)) n = rd;
                        bn = n;
                    }
                }
            }
        }
    } else if (mfabs(qb) > R(1e-15)) {
        R t = -qc / (R(2) * qb);
        if (t >= tmin && t <= tmax) {
            R s = (d0d + t * od) / hgt;
            if (s >= R(0) && s <= R(1)) {
                found = true;
                bt = t;
                bs = s;
                is_cap = 0;
                if (FULL) {
                    V3<R> pt = at(o, d, t);
                    V3<R> ap = p.p0 + ax * (s * hgt);
                    V3<R> rd = normalize(pt - ap);
                    V3<R> n = normalize(rd * hgt - ax * (r1 - r0));
                    if (len2(n) < R(0.5)) n = rd;
                    bn = n;
                }
            }
        }
    }
    if (mfabs(od) > R(1e-15)) {
#pragma unroll
        for (int cap = 0; cap < 2; ++cap) {
            R sc = cap == 0 ? R(0) : R(1);
            R rc = cap == 0 ? r0 : r1;
            if (rc <= R(0)) continue;
            V3<R> cc = cap == 0 ? p.p0 : p.p1;
            R t = ((sc * hgt) - d0d) / od;
            if (t < tmin || t > tmax) continue;
            V3<R> pt = at(o, d, t);
            V3<R> rad = pt - cc;
            if (len2(rad) > rc * rc) continue;
            if (!found || t < bt) {
                found = true;
                bt = t;
                bn = cap == 0 ? -ax : ax;
                is_cap = 1;
                cap_id = cap;
            }
        }
    }
    if (!found) return false;
    *tout = bt;
    if (FULL) {
        V3<R> pt = at(o, d, bt);
        R u, v;
        V3<R> hint = mfabs(ax.y) > R(0.9) ? mk(R(0), R(0), R(1)) : mk(R(0), R(1), R(0));
        V3<R> T = normalize(cross(hint, ax));
        V3<R> B = cross(ax, T);
        if (is_cap) {
            V3<R> cc = cap_id == 0 ? p.p0 : p.p1;
            R rc = cap_id == 0 ? r0 : r1;
            V3<R> rad = pt - cc;
            R rx = rc > R(1e-12) ? dot(rad, T) / (R(2) * rc) : R(0);
            R ry = rc > R(1e-12) ? dot(rad, B) / (R(2) * rc) : R(0);
            u = mfmax(R(0), mfmin(R(1), R(0.5) + rx));
            v = mfmax(R(0), mfmin(R(1), R(0.5) + ry));
        } else {
            V3<R> ap = p.p0 + ax * (bs * hgt);
            V3<R> rad = pt - ap;
            R phi = matan2(dot(rad, B), dot(rad, T));
            u = R(0.5) + phi / (R(2) * Consts<R>::pi);
            u = u - mfloor(u);
            v = mfmax(R(0), mfmin(R(1), bs));
        }
        store_hit(h, bt, pt, bn, d, p.mat, u, v);
    }
    return true;
}

// ---- dispatch -------------------------------------------------------------------

template <typename R, bool FULL>
__device__ __forceinline__ bool hit_prim(const PD<R>& p, const DAttr<R>* attr, V3<R> o, V3<R> d, R tmin, R tmax,
                                         R* t, DHit<R>* h)
{
    switch (p.kind) {
    case K_TRIANGLE: return hit_tri<R, FULL>(p, attr, o, d, tmin, tmax, t, h);
    case K_SPHERE: return hit_sphere<R, FULL>(p, o, d, tmin, tmax, t, h);
    case K_BOX: return hit_box<R, FULL>(p, o, d, tmin, tmax, t, h);
    case K_CYLINDER: return hit_cyl<R, FULL>(p, o, d, tmin, tmax, t, h);
    case K_PLANE: return hit_plane<R, FULL>(p, o, d, tmin, tmax, t, h);
    default: return false;
    }
}

template <typename R>
__device__ __forceinline__ bool occ_prim(const PD<R>& p, V3<R> o, V3<R> d, R tmin, R tmax)
{
    R t;
    switch (p.kind) {
    case K_TRIANGLE: return hit_tri<R, false>(p, nullptr, o, d, tmin, tmax, &t, nullptr);
    case K_SPHERE: return occ_sphere<R>(p, o, d, tmin, tmax);
    case K_BOX: return occ_box<R>(p, o, d, tmin, tmax);
    case K_CYLINDER: return hit_cyl<R, false>(p, o, d, tmin, tmax, &t, nullptr);
    case K_PLANE: return hit_plane<R, false>(p, o, d, tmin, tmax, &t, nullptr);
    default: return false;
    }
}

// ---- BVH2 traversal -------------------------------------------------------------


// Traversal stacks. A stack policy is a small per-lane handle type passed by
// value down to the traversal routines; each traversal builds its own
// Lane<KeepT> from it, so the stack pointer lives in that traversal's
// registers. KeepT: entries also carry their entry distance (closest hit culls
// on it); any-hit keeps node references only. Host-built trees have depth
// <= 60, so MH_STACK = 64 entries never overflow; indices are masked so that
// even a malformed tree cannot address outside the arrays.
//
// PrivStack: the whole stack in private (scratch) memory, cached in L1/L2.
// LdsStack<K>: the top K entries of each lane in LDS, lane-interleaved so a
// wave's pushes hit consecutive banks; when full, the OLDEST entry moves to a
// private overflow array.
struct PrivStack {
    template <bool KeepT> struct Lane {
        int sp;
        int pref[MH_STACK];
        float pt[KeepT ? MH_STACK : 1];
        __device__ __forceinline__ explicit Lane(PrivStack) : sp(0) {}
        __device__ __forceinline__ bool empty() const { return sp == 0; }
        __device__ __forceinline__ void push(int r, float t)
        {
            MH_HOST_ASSERT(sp >= 0 && sp < MH_STACK);
            pref[sp & (MH_STACK - 1)] = r;
            if (KeepT) pt[sp & (MH_STACK - 1)] = t;
            ++sp;
        }
        __device__ __forceinline__ int pop(float* t)
        {
            MH_HOST_ASSERT(sp > 0);
            --sp;
            *t = KeepT ? pt[sp & (MH_STACK - 1)] : 0.f;
            return pref[sp & (MH_STACK - 1)];
        }
    };
};

template <int K> struct LdsStack {
    static_assert(K > 0 && (K & (K - 1)) == 0 && K < MH_STACK, "LDS stack depth must be a power of two");
    int* ref;   // this lane's first LDS slot; slot k at ref[k * stride]
    float* tt;  // entry distances, same layout
    int stride;
    template <bool KeepT> struct Lane {
        int* ref;
        float* tt;
        int stride;
        int sp, base, psp;
        int pref[MH_STACK - K];
        float pt[KeepT ? MH_STACK - K : 1];
        __device__ __forceinline__ explicit Lane(LdsStack h)
            : ref(h.ref), tt(h.tt), stride(h.stride), sp(0), base(0), psp(0)
        {
        }
        __device__ __forceinline__ bool empty() const { return sp == 0 && psp == 0; }
        __device__ __forceinline__ void push(int r, float t)
        {
            if (sp == K) {
                MH_HOST_ASSERT(psp < MH_STACK - K);
                int j = psp < MH_STACK - K - 1 ? psp : MH_STACK - K - 1;
                pref[j] = ref[base * stride];
                if (KeepT) pt[j] = tt[base * stride];
                ++psp;
                base = (base + 1) & (K - 1);
                --sp;
            }
            int k = (base + sp) & (K - 1);
            ref[k * stride] = r;
            if (KeepT) tt[k * stride] = t;
            ++sp;
        }
        __device__ __forceinline__ int pop(float* t)
        {
            if (sp > 0) {
                --sp;
                int k = (base + sp) & (K - 1);
                *t = KeepT ? tt[k * stride] : 0.f;
                return ref[k * stride];
            }
            --psp;
            MH_HOST_ASSERT(psp >= 0);
            int j = psp < MH_STACK - K - 1 ? psp : MH_STACK - K - 1;
            *t = KeepT ? pt[j] : 0.f;
            return pref[j];
        }
    };
};

// Per-kernel stack choice (0 = PrivStack, K = LdsStack<K>), measured on the
// 1080p terrain workload: LDS pays off for the closest-hit extend kernel and
// for the any-hit-only shadow kernel (occupancy 8: a private stack there is
// ~130 MB of scratch, beyond L2 and MALL, written back to HBM); the
// transmissive shadow walk and the megakernel, at occupancy 5-6, run faster
// from private memory.
#ifndef MH_EXTEND_LDS
#define MH_EXTEND_LDS 8
#endif
#ifndef MH_SHADOW_LDS
#define MH_SHADOW_LDS 0
#endif
#ifndef MH_SHADOW_ANY_LDS
#define MH_SHADOW_ANY_LDS 16  // any-hit-only shadow kernel (no transmissive material)
#endif
#ifndef MH_MEGA_LDS
#define MH_MEGA_LDS 0
#endif

// Stack declarations at the top of a kernel of `threads` lanes per block.
#define MH_STACK_PRIV(ts) ::mh::PrivStack ts;
#define MH_STACK_LDS(ts, K, threads)                                                                      \
    __shared__ int ts##_lds_ref[(K) * (threads)];                                                       \
    __shared__ float ts##_lds_t[(K) * (threads)];                                                       \
    ::mh::LdsStack<K> ts{ts##_lds_ref + threadIdx.x, ts##_lds_t + threadIdx.x, (int)blockDim.x};
// References only: for kernels whose traversals are all any-hit.
#define MH_STACK_LDS_REFS(ts, K, threads)                                                                 \
    __shared__ int ts##_lds_ref[(K) * (threads)];                                                       \
    ::mh::LdsStack<K> ts{ts##_lds_ref + threadIdx.x, nullptr, (int)blockDim.x};

// Both child slabs of one node. fp32: origin-scaled planes and a widened exit
// (robust traversal); fp64: the reference's slab arithmetic in double.
// 1: the fp32 slab setup guards direction components of magnitude below
// 1e-30 (see Slab<float>::inv). The path-replay adjoint's translation unit
// sets 0: its kernel sits at 251 of 256 registers, and the guard's compare
// and select took it to occupancy 1 (two-pass gradient 55 -> 82 ms); there
// an exact zero needs a uniform sample of exactly 0 (2^-24 per sample) in an
// axis-aligned frame, and the adjoint is compared with the oracle within a
// tolerance only.
#ifndef MH_SLAB_GUARD
#define MH_SLAB_GUARD 1
#endif

template <typename R> struct Slab;

template <> struct Slab<float> {
    float ix, iy, iz, ox, oy, oz;
    // MH_SLAB_GUARD: a direction component of magnitude below 1e-30 (zero
    // included) counts as 1e-30, so its inverse stays finite and
    // plane * i - o * i keeps the sign of plane - o (with i = inf,
    // lo * i - o * i is inf - inf or -inf and would drop a box whose slab
    // holds the origin).
    __device__ __forceinline__ static float inv(float v)
    {
        return MH_SLAB_GUARD ? 1.0f / (fabsf(v) < 1e-30f ? 1e-30f : v) : 1.0f / v;
    }
    __device__ __forceinline__ void init(V3<float> o, V3<float> d)
    {
        ix = inv(d.x); iy = inv(d.y); iz = inv(d.z);
        ox = o.x * ix; oy = o.y * iy; oz = o.z * iz;
    }
    // Offset (0 or 16 bytes) of the near row of an axis in a lo / hi pair of
    // float4 rows: the hi row when the ray runs towards -axis.
    __device__ __forceinline__ static uint32_t near_off(float i) { return i < 0.f ? 16u : 0u; }
    // One box from its near and far planes per axis (the lo / hi planes
    // ordered by the direction's sign): the same decision and entry distance
    // as one(), whose min / max of each axis pair picks the same values.
    __device__ __forceinline__ bool nf(float nx, float fx, float ny, float fy, float nz, float fz, float tmin,
                                       float tmax, float* te = nullptr) const
    {
        float en = fmaxf(fmaxf(fmaf(nx, ix, -ox), fmaf(ny, iy, -oy)), fmaxf(fmaf(nz, iz, -oz), tmin));
        float ex = fminf(fminf(fmaf(fx, ix, -ox), fmaf(fy, iy, -oy)), fminf(fmaf(fz, iz, -oz), tmax));
        if (te) *te = en;
        return en <= ex * 1.0000004f;
    }
    __device__ __forceinline__ void test(const DNode& n, float tmin, float tmax, bool* hl, bool* hr, float* tl,
                                         float* tr) const
    {
        float lx0 = fmaf(n.c0.x, ix, -ox), lx1 = fmaf(n.c0.y, ix, -ox);
        float ly0 = fmaf(n.c0.z, iy, -oy), ly1 = fmaf(n.c0.w, iy, -oy);
        float lz0 = fmaf(n.c1.x, iz, -oz), lz1 = fmaf(n.c1.y, iz, -oz);
        float rx0 = fmaf(n.c1.z, ix, -ox), rx1 = fmaf(n.c1.w, ix, -ox);
        float ry0 = fmaf(n.c2.x, iy, -oy), ry1 = fmaf(n.c2.y, iy, -oy);
        float rz0 = fmaf(n.c2.z, iz, -oz), rz1 = fmaf(n.c2.w, iz, -oz);
        float len = fmaxf(fmaxf(fminf(lx0, lx1), fminf(ly0, ly1)), fmaxf(fminf(lz0, lz1), tmin));
        float lex = fminf(fminf(fmaxf(lx0, lx1), fmaxf(ly0, ly1)), fminf(fmaxf(lz0, lz1), tmax));
        float ren = fmaxf(fmaxf(fminf(rx0, rx1), fminf(ry0, ry1)), fmaxf(fminf(rz0, rz1), tmin));
        float rex = fminf(fminf(fmaxf(rx0, rx1), fmaxf(ry0, ry1)), fminf(fmaxf(rz0, rz1), tmax));
        const float widen = 1.0000004f; // 1 + 2 gamma(3)
        *hl = len <= lex * widen;
        *hr = ren <= rex * widen;
        *tl = len;
        *tr = ren;
    }
    // one box (x0, x1, y0, y1, z0, z1 = min / max per axis)
    __device__ __forceinline__ void one(float x0, float x1, float y0, float y1, float z0, float z1, float tmin,
                                        float tmax, bool* hit, float* te) const
    {
        float ax = fmaf(x0, ix, -ox), bx = fmaf(x1, ix, -ox);
        float ay = fmaf(y0, iy, -oy), by = fmaf(y1, iy, -oy);
        float az = fmaf(z0, iz, -oz), bz = fmaf(z1, iz, -oz);
        float en = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
        float ex = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
        *hit = en <= ex * 1.0000004f;
        *te = en;
    }
};

template <> struct Slab<double> {
    double ix, iy, iz, ox, oy, oz;
    __device__ __forceinline__ void init(V3<double> o, V3<double> d)
    {
        ix = 1.0 / d.x; iy = 1.0 / d.y; iz = 1.0 / d.z;
        ox = o.x; oy = o.y; oz = o.z;
    }
    __device__ __forceinline__ void one(double x0, double x1, double y0, double y1, double z0, double z1,
                                        double tmin, double tmax, bool* hit, float* te) const
    {
        double ax = (x0 - ox) * ix, bx = (x1 - ox) * ix;
        double ay = (y0 - oy) * iy, by = (y1 - oy) * iy;
        double az = (z0 - oz) * iz, bz = (z1 - oz) * iz;
        double en = fmax(fmax(fmin(ax, bx), fmin(ay, by)), fmin(az, bz));
        double ex = fmin(fmin(fmax(ax, bx), fmax(ay, by)), fmax(az, bz));
        *hit = ex >= en && ex >= tmin && en <= tmax;
        double e = en < tmin ? tmin : en;
        // round DOWN to float: culling on a lower bound never drops a node
        float f = (float)e;
        if ((double)f > e) f = nextafterf(f, -INFINITY);
        *te = f;
    }
    __device__ __forceinline__ void test(const DNode& n, double tmin, double tmax, bool* hl, bool* hr, float* tl,
                                         float* tr) const
    {
        one(n.c0.x, n.c0.y, n.c0.z, n.c0.w, n.c1.x, n.c1.y, tmin, tmax, hl, tl);
        one(n.c1.z, n.c1.w, n.c2.x, n.c2.y, n.c2.z, n.c2.w, tmin, tmax, hr, tr);
    }
};

// Whether 4-wide any-hit traversal visits children nearest first (1, the
// default). The unsorted variant (0) measured faster in the megakernel but a
// build with it produced wrong fp64 path-traced images, without a cause found;
// it stays a tuning switch only.
// Whether closest_hit<double> also visits the nearest hit child straight from
// a register, as the fp32 path does (0, the default: push-all in fp64; the
// register-held variant is a test switch, see closest_hit).
#ifndef MH_HELD_FP64
#define MH_HELD_FP64 0
#endif

#ifndef MH_ANY_SORT
#define MH_ANY_SORT 1
#endif

// The hit children of a BVH4 node, nearest first: refs r[0..*nh) with entry
// distances t[0..*nh).
template <typename R>
__device__ __forceinline__ void children4(const Slab<R>& sl, const DNode4& n, R tmin, R tmax, int r[4], float t[4],
                                          int* nh)
{
    const float inf = __builtin_inff();
    int cnt = n.meta.x;
    bool h0, h1, h2, h3;
    float t0, t1, t2, t3;
    sl.one(n.lox.x, n.hix.x, n.loy.x, n.hiy.x, n.loz.x, n.hiz.x, tmin, tmax, &h0, &t0);
    sl.one(n.lox.y, n.hix.y, n.loy.y, n.hiy.y, n.loz.y, n.hiz.y, tmin, tmax, &h1, &t1);
    sl.one(n.lox.z, n.hix.z, n.loy.z, n.hiy.z, n.loz.z, n.hiz.z, tmin, tmax, &h2, &t2);
    sl.one(n.lox.w, n.hix.w, n.loy.w, n.hiy.w, n.loz.w, n.hiz.w, tmin, tmax, &h3, &t3);
    float k[4] = {h0 ? t0 : inf, h1 ? t1 : inf, (h2 && cnt > 2) ? t2 : inf, (h3 && cnt > 3) ? t3 : inf};
    int q[4] = {n.ref.x, n.ref.y, n.ref.z, n.ref.w};
    // sorting network (0,1) (2,3) (0,2) (1,3) (1,2): ascending entry distance
#define MH_CSWAP(a, b)                                                                                    \
    if (k[b] < k[a]) {                                                                                    \
        float tk = k[a];                                                                                  \
        k[a] = k[b];                                                                                      \
        k[b] = tk;                                                                                        \
        int tq = q[a];                                                                                    \
        q[a] = q[b];                                                                                      \
        q[b] = tq;                                                                                        \
    }
    MH_CSWAP(0, 1)
    MH_CSWAP(2, 3)
    MH_CSWAP(0, 2)
    MH_CSWAP(1, 3)
    MH_CSWAP(1, 2)
#undef MH_CSWAP
    int c = 0;
    for (int i = 0; i < 4; ++i) {
        r[i] = q[i];
        t[i] = k[i];
        c += k[i] < inf ? 1 : 0;
    }
    *nh = c;
}

// 1: the fp32 closest-hit and any-hit walks test each child of a 4-wide node
// from its near and far planes only (children4_nf); 0: whole nodes
// (children4).
#ifndef MH_TRACE_NF
#define MH_TRACE_NF 1
#endif

// children4 from the near and far planes of node `ref` (the rows picked by
// address from the ray's direction signs, near-row offsets from
// Slab::near_off): 112 of the node's 128 bytes, two planes per axis.
__device__ __forceinline__ void children4_nf(const Slab<float>& sl, const DNode4* nodes4, int ref, uint32_t nrx,
                                             uint32_t nry, uint32_t nrz, float tmin, float tmax, int r[4],
                                             float t[4], int* nh)
{
    const char* nb = (const char*)nodes4;
    // node arrays stay below 4 GiB: 32-bit byte offsets
    const uint32_t off = (uint32_t)ref << 7;
    const uint32_t vx = off | nrx, vy = off | nry, vz = off | nrz;
    const float4 nx4 = ldg((const float4*)(nb + (size_t)vx));
    const float4 fx4 = ldg((const float4*)(nb + (size_t)(vx ^ 16u)));
    const float4 ny4 = ldg((const float4*)(nb + (size_t)vy + 32));
    const float4 fy4 = ldg((const float4*)(nb + (size_t)(vy ^ 16u) + 32));
    const float4 nz4 = ldg((const float4*)(nb + (size_t)vz + 64));
    const float4 fz4 = ldg((const float4*)(nb + (size_t)(vz ^ 16u) + 64));
    const int4 rf = ldg((const int4*)(nb + (size_t)off + 96));
    const float inf = __builtin_inff();
    float k[4];
    const bool h0 = sl.nf(nx4.x, fx4.x, ny4.x, fy4.x, nz4.x, fz4.x, tmin, tmax, &k[0]);
    const bool h1 = sl.nf(nx4.y, fx4.y, ny4.y, fy4.y, nz4.y, fz4.y, tmin, tmax, &k[1]);
    const bool h2 = sl.nf(nx4.z, fx4.z, ny4.z, fy4.z, nz4.z, fz4.z, tmin, tmax, &k[2]) && rf.z != INT_MIN;
    const bool h3 = sl.nf(nx4.w, fx4.w, ny4.w, fy4.w, nz4.w, fz4.w, tmin, tmax, &k[3]) && rf.w != INT_MIN;
    if (!h0) k[0] = inf;
    if (!h1) k[1] = inf;
    if (!h2) k[2] = inf;
    if (!h3) k[3] = inf;
    int q[4] = {rf.x, rf.y, rf.z, rf.w};
#define MH_CSWAP(a, b)                                                                                    \
    if (k[b] < k[a]) {                                                                                    \
        float tk = k[a];                                                                                  \
        k[a] = k[b];                                                                                      \
        k[b] = tk;                                                                                        \
        int tq = q[a];                                                                                    \
        q[a] = q[b];                                                                                      \
        q[b] = tq;                                                                                        \
    }
    MH_CSWAP(0, 1)
    MH_CSWAP(2, 3)
    MH_CSWAP(0, 2)
    MH_CSWAP(1, 3)
    MH_CSWAP(1, 2)
#undef MH_CSWAP
    int c = 0;
    for (int i = 0; i < 4; ++i) {
        r[i] = q[i];
        t[i] = k[i];
        c += k[i] < inf ? 1 : 0;
    }
    *nh = c;
}

// Closest hit. Returns the leaf-record index of the winner (or -1 - plane
// slot for a plane, or INT_MIN for a miss) and its t / prim index.
// fp32: the nearest hit child is visited next straight from a register, the
// others wait on the stack (same visit order, one stack round trip fewer per
// node: the extend kernel 11 % faster on the 1080p terrain frame). fp64 keeps
// the push-all loop: the register-held variant made the fp64 megakernel's
// images differ from run to run, with no cause found in this source. any_hit
// keeps push-all in both precisions (the shadow kernel ran 4 % slower with
// the register-held child).
template <typename R, typename TS>
__device__ int closest_hit(const SceneView<R>& sc, TS ts, V3<R> o, V3<R> d, R tmin, R tmax, R* best_t,
                           int* best_prim)
{
    int best = INT_MIN;
    int bprim = 0x7fffffff;
    R closest = tmax;
    for (int i = 0; i < sc.n_planes; ++i) {
        PD<R> p = load_prim(&sc.planes[i]);
        R t;
        if (hit_plane<R, false>(p, o, d, tmin, closest, &t, nullptr)) {
            int pi = sc.plane_index[i];
            if (best == INT_MIN || t < closest || (t == closest && pi < bprim)) {
                closest = t;
                bprim = pi;
                best = -1 - i;
            }
        }
    }
    if (sc.root_ref != INT_MIN && sc.bvh4) {
        Slab<R> sl;
        sl.init(o, d);
        typename TS::template Lane<true> st(ts);
        if constexpr (sizeof(R) == 4 || MH_HELD_FP64) {
            st.push(sc.root4, (float)tmin);
            int ref = 0;
            bool held = false;  // the nearest hit child, visited next without a stack round trip
            uint32_t nrx = 0, nry = 0, nrz = 0;
            if constexpr (sizeof(R) == 4 && MH_TRACE_NF) {
                nrx = Slab<float>::near_off(sl.ix);
                nry = Slab<float>::near_off(sl.iy);
                nrz = Slab<float>::near_off(sl.iz);
            }
            for (;;) {
                if (!held) {
                    if (st.empty()) break;
                    float te;
                    ref = st.pop(&te);
                    if (best != INT_MIN && (R)te > closest) continue;
                }
                held = false;
                if (ref >= 0) {
                    int r[4], nh;
                    float t[4];
                    if constexpr (sizeof(R) == 4 && MH_TRACE_NF) {
                        children4_nf(sl, sc.nodes4, ref, nrx, nry, nrz, tmin, closest, r, t, &nh);
                    } else {
                        DNode4 n = ldg(&sc.nodes4[ref]);
                        children4(sl, n, tmin, closest, r, t, &nh);
                    }
                    // push the farthest first; the nearest stays in a register
                    for (int i = 3; i >= 1; --i)
                        if (i < nh) st.push(r[i], t[i]);
                    ref = r[0];
                    held = nh > 0;
                } else {
                    int code = ~ref;
                    int first = code >> 3, cnt = (code & 7) + 1;
                    for (int i = 0; i < cnt; ++i) {
                        PD<R> p = load_prim(&sc.prims[first + i]);
                        R t;
                        if (hit_prim<R, false>(p, nullptr, o, d, tmin, closest, &t, nullptr)) {
                            if (best == INT_MIN || t < closest || (t == closest && p.prim < bprim)) {
                                closest = t;
                                bprim = p.prim;
                                best = first + i;
                            }
                        }
                    }
                }
            }
        } else {
            st.push(sc.root4, (float)tmin);
            while (!st.empty()) {
                float te;
                int ref = st.pop(&te);
                if (best != INT_MIN && (R)te > closest) continue;
                if (ref >= 0) {
                    DNode4 n = ldg(&sc.nodes4[ref]);
                    int r[4], nh;
                    float t[4];
                    children4(sl, n, tmin, closest, r, t, &nh);
                    // push the farthest first: the nearest is popped next
                    for (int i = 3; i >= 0; --i)
                        if (i < nh) st.push(r[i], t[i]);
                } else {
                    int code = ~ref;
                    int first = code >> 3, cnt = (code & 7) + 1;
                    for (int i = 0; i < cnt; ++i) {
                        PD<R> p = load_prim(&sc.prims[first + i]);
                        R t;
                        if (hit_prim<R, false>(p, nullptr, o, d, tmin, closest, &t, nullptr)) {
                            if (best == INT_MIN || t < closest || (t == closest && p.prim < bprim)) {
                                closest = t;
                                bprim = p.prim;
                                best = first + i;
                            }
                        }
                    }
                }
            }
        }
    } else if (sc.root_ref != INT_MIN) {
        Slab<R> sl;
        sl.init(o, d);
        typename TS::template Lane<true> st(ts);
        if constexpr (sizeof(R) == 4 || MH_HELD_FP64) {
            st.push(sc.root_ref, (float)tmin);
            int ref = 0;
            bool held = false;  // the nearer hit child, visited next without a stack round trip
            for (;;) {
                if (!held) {
                    if (st.empty()) break;
                    float te;
                    ref = st.pop(&te);
                    if (best != INT_MIN && (R)te > closest) continue;
                }
                held = false;
                if (ref >= 0) {
                    DNode n = ldg(&sc.nodes[ref]);
                    bool hl, hr;
                    float tl, tr;
                    sl.test(n, tmin, closest, &hl, &hr, &tl, &tr);
                    bool lfirst = hl && (!hr || tl <= tr);
                    if (hl && hr) st.push(lfirst ? n.c3.y : n.c3.x, lfirst ? tr : tl);
                    ref = lfirst ? n.c3.x : n.c3.y;
                    held = hl || hr;
                } else {
                    int code = ~ref;
                    int first = code >> 3, cnt = (code & 7) + 1;
                    for (int i = 0; i < cnt; ++i) {
                        PD<R> p = load_prim(&sc.prims[first + i]);
                        R t;
                        if (hit_prim<R, false>(p, nullptr, o, d, tmin, closest, &t, nullptr)) {
                            if (best == INT_MIN || t < closest || (t == closest && p.prim < bprim)) {
                                closest = t;
                                bprim = p.prim;
                                best = first + i;
                            }
                        }
                    }
                }
            }
        } else {
            st.push(sc.root_ref, (float)tmin);
            while (!st.empty()) {
                float te;
                int ref = st.pop(&te);
                if (best != INT_MIN && (R)te > closest) continue;
                if (ref >= 0) {
                    DNode n = ldg(&sc.nodes[ref]);
                    bool hl, hr;
                    float tl, tr;
                    sl.test(n, tmin, closest, &hl, &hr, &tl, &tr);
                    if (hl && hr) {
                        bool lfirst = tl <= tr;
                        st.push(lfirst ? n.c3.y : n.c3.x, lfirst ? tr : tl);
                        st.push(lfirst ? n.c3.x : n.c3.y, lfirst ? tl : tr);
                    } else if (hl) {
                        st.push(n.c3.x, tl);
                    } else if (hr) {
                        st.push(n.c3.y, tr);
                    }
                } else {
                    int code = ~ref;
                    int first = code >> 3, cnt = (code & 7) + 1;
                    for (int i = 0; i < cnt; ++i) {
                        PD<R> p = load_prim(&sc.prims[first + i]);
                        R t;
                        if (hit_prim<R, false>(p, nullptr, o, d, tmin, closest, &t, nullptr)) {
                            if (best == INT_MIN || t < closest || (t == closest && p.prim < bprim)) {
                                closest = t;
                                bprim = p.prim;
                                best = first + i;
                            }
                        }
                    }
                }
            }
        }
    }
    *best_t = closest;
    *best_prim = bprim;
    return best;
}

// Recompute the full hit record of the winner found by closest_hit.
template <typename R>
__device__ __forceinline__ void finish_hit(const SceneView<R>& sc, int rec, V3<R> o, V3<R> d, R tmin, R tmax, DHit<R>* h)
{
    R t;
    bool ok;
    if (rec < 0) {
        int slot = -1 - rec;
        PD<R> p = load_prim(&sc.planes[slot]);
        ok = hit_plane<R, true>(p, o, d, tmin, tmax, &t, h);
        h->prim = sc.plane_index[slot];
    } else {
        PD<R> p = load_prim(&sc.prims[rec]);
        ok = hit_prim<R, true>(p, sc.attrs ? &sc.attrs[rec] : nullptr, o, d, tmin, tmax, &t, h);
        h->prim = p.prim;
    }
    // the winner's full test repeats the query's, so it passes; were it ever
    // to fail, the record holds no material (callers treat it as no surface)
    // instead of whatever the registers held
    if (!ok) store_hit(h, tmax, o + d * tmax, -d, d, -1, R(0), R(0));
}

template <typename R, typename TS>
__device__ __forceinline__ bool intersect(const SceneView<R>& sc, TS ts, V3<R> o, V3<R> d, R tmin, R tmax,
                                          DHit<R>* h)
{
    R t;
    int prim;
    int rec = closest_hit(sc, ts, o, d, tmin, tmax, &t, &prim);
    if (rec == INT_MIN) return false;
    finish_hit(sc, rec, o, d, tmin, t, h);
    return true;
}

// W = 0: the scene's tree width (sc.bvh4); W = 2: always the binary tree
// (kernels at full occupancy whose per-node work dominates).
template <typename R, typename TS, int W = 0>
__device__ bool any_hit(const SceneView<R>& sc, TS ts, V3<R> o, V3<R> d, R tmin, R tmax)
{
    for (int i = 0; i < sc.n_planes; ++i) {
        PD<R> p = load_prim(&sc.planes[i]);
        if (occ_prim(p, o, d, tmin, tmax)) return true;
    }
    if (sc.root_ref == INT_MIN) return false;
    Slab<R> sl;
    sl.init(o, d);
    typename TS::template Lane<false> st(ts);
    if (W != 2 && sc.bvh4) {
        uint32_t nrx = 0, nry = 0, nrz = 0;
        if constexpr (sizeof(R) == 4 && MH_TRACE_NF && MH_ANY_SORT) {
            nrx = Slab<float>::near_off(sl.ix);
            nry = Slab<float>::near_off(sl.iy);
            nrz = Slab<float>::near_off(sl.iz);
        }
        st.push(sc.root4, 0.f);
        while (!st.empty()) {
            float te;
            int ref = st.pop(&te);
            if (ref >= 0) {
#if MH_ANY_SORT
                int r[4], nh;
                float t[4];
                if constexpr (sizeof(R) == 4 && MH_TRACE_NF) {
                    children4_nf(sl, sc.nodes4, ref, nrx, nry, nrz, tmin, tmax, r, t, &nh);
                } else {
                    DNode4 n = ldg(&sc.nodes4[ref]);
                    children4(sl, n, tmin, tmax, r, t, &nh);
                }
                for (int i = 3; i >= 0; --i)
                    if (i < nh) st.push(r[i], 0.f);
#else
                DNode4 n = ldg(&sc.nodes4[ref]);
                // any hit ends the query: visit order only affects speed
                bool h0, h1, h2, h3;
                float t0, t1, t2, t3;
                int cnt = n.meta.x;
                sl.one(n.lox.w, n.hix.w, n.loy.w, n.hiy.w, n.loz.w, n.hiz.w, tmin, tmax, &h3, &t3);
                sl.one(n.lox.z, n.hix.z, n.loy.z, n.hiy.z, n.loz.z, n.hiz.z, tmin, tmax, &h2, &t2);
                sl.one(n.lox.y, n.hix.y, n.loy.y, n.hiy.y, n.loz.y, n.hiz.y, tmin, tmax, &h1, &t1);
                sl.one(n.lox.x, n.hix.x, n.loy.x, n.hiy.x, n.loz.x, n.hiz.x, tmin, tmax, &h0, &t0);
                if (h3 && cnt > 3) st.push(n.ref.w, 0.f);
                if (h2 && cnt > 2) st.push(n.ref.z, 0.f);
                if (h1) st.push(n.ref.y, 0.f);
                if (h0) st.push(n.ref.x, 0.f);
#endif
            } else {
                int code = ~ref;
                int first = code >> 3, cnt = (code & 7) + 1;
                for (int i = 0; i < cnt; ++i) {
                    PD<R> p = load_prim(&sc.prims[first + i]);
                    if (occ_prim(p, o, d, tmin, tmax)) return true;
                }
            }
        }
        return false;
    }
    st.push(sc.root_ref, 0.f);
    while (!st.empty()) {
        float te;
        int ref = st.pop(&te);
        if (ref >= 0) {
            DNode n = ldg(&sc.nodes[ref]);
            bool hl, hr;
            float tl, tr;
            sl.test(n, tmin, tmax, &hl, &hr, &tl, &tr);
            if (hl && hr) {
                bool lfirst = tl <= tr;
                st.push(lfirst ? n.c3.y : n.c3.x, 0.f);
                st.push(lfirst ? n.c3.x : n.c3.y, 0.f);
            } else if (hl) {
                st.push(n.c3.x, 0.f);
            } else if (hr) {
                st.push(n.c3.y, 0.f);
            }
        } else {
            int code = ~ref;
            int first = code >> 3, cnt = (code & 7) + 1;
            for (int i = 0; i < cnt; ++i) {
                PD<R> p = load_prim(&sc.prims[first + i]);
                if (occ_prim(p, o, d, tmin, tmax)) return true;
            }
        }
    }
    return false;
}

// 1: the bundle walk below takes children nearest first along the axis
// (forest scene, fp64 Whitted 1280x720x4 depth 5: 52.8 ms vs 53.9 in index
// order, the same bytes)
#ifndef MH_BUNDLE_SORT
#define MH_BUNDLE_SORT 1
#endif

// Occlusion of n <= 32 unit-direction rays d[0..n) that leave o within the
// angle `spread` (radians) of the unit axis a, over [tmin, tmax]: bit k of the
// result is set when ray k is blocked, as n separate any_hit queries would
// set it. One walk of the tree, nearest child first along the axis, visits
// every node that the cone around a of half-angle `spread` can meet, and
// each ray still open is tested exactly against the primitives of the leaves
// reached (stopping once all are blocked). The cone test grows each box by
// e = D * spread, D an upper bound (L1) of the distance from o to the box: a
// ray k that meets the box at t <= D passes within t * |d[k] - a| <=
// D * spread of the axis point o + t a, which therefore lies in the grown
// box. Double precision only (the growth must not be lost to rounding of the
// box bounds).
template <typename TS>
__device__ uint32_t any_hit_bundle(const SceneView<double>& sc, TS ts, V3<double> o, const V3<double>* d, int n,
                                   V3<double> a, double spread, double tmin, double tmax)
{
    const uint32_t all = n >= 32 ? 0xffffffffu : (1u << n) - 1u;
    uint32_t occ = 0;
    for (int i = 0; i < sc.n_planes; ++i) {
        PD<double> p = load_prim(&sc.planes[i]);
        for (int k = 0; k < n; ++k)
            if (!((occ >> k) & 1u) && occ_prim(p, o, d[k], tmin, tmax)) occ |= 1u << k;
    }
    if (occ == all || sc.root_ref == INT_MIN) return occ;
    // growth per unit distance, with a margin for the rounding of the
    // directions and of the bound itself
    const double g = spread * 1.000001 + 1e-9;
    Slab<double> sl;
    sl.init(o, a);
    // the grown box's entry distance along the axis in *te (near-first order)
    auto meets = [&](double x0, double x1, double y0, double y1, double z0, double z1, float* te) -> bool {
        const double D = mfabs((x0 + x1) * 0.5 - o.x) + mfabs((y0 + y1) * 0.5 - o.y) +
                         mfabs((z0 + z1) * 0.5 - o.z) + (x1 - x0) * 0.5 + (y1 - y0) * 0.5 + (z1 - z0) * 0.5;
        const double e = D * g;
        bool hit;
        sl.one(x0 - e, x1 + e, y0 - e, y1 + e, z0 - e, z1 + e, tmin, tmax, &hit, te);
        return hit;
    };
    auto leaf = [&](int ref) -> bool {
        const int code = ~ref;
        const int first = code >> 3, cnt = (code & 7) + 1;
        for (int i = 0; i < cnt; ++i) {
            PD<double> p = load_prim(&sc.prims[first + i]);
            for (int k = 0; k < n; ++k)
                if (!((occ >> k) & 1u) && occ_prim(p, o, d[k], tmin, tmax)) occ |= 1u << k;
            if (occ == all) return true;
        }
        return false;
    };
    typename TS::template Lane<false> st(ts);
    if (sc.bvh4) {
        st.push(sc.root4, 0.f);
        while (!st.empty()) {
            float te;
            const int ref = st.pop(&te);
            if (ref >= 0) {
                const DNode4 nd = ldg(&sc.nodes4[ref]);
                const int cnt = nd.meta.x;
                const float inf = __builtin_inff();
                float kk[4];
                int qq[4] = {nd.ref.x, nd.ref.y, nd.ref.z, nd.ref.w};
                if (!meets(nd.lox.x, nd.hix.x, nd.loy.x, nd.hiy.x, nd.loz.x, nd.hiz.x, &kk[0])) kk[0] = inf;
                if (!meets(nd.lox.y, nd.hix.y, nd.loy.y, nd.hiy.y, nd.loz.y, nd.hiz.y, &kk[1])) kk[1] = inf;
                if (!(cnt > 2 && meets(nd.lox.z, nd.hix.z, nd.loy.z, nd.hiy.z, nd.loz.z, nd.hiz.z, &kk[2]))) kk[2] = inf;
                if (!(cnt > 3 && meets(nd.lox.w, nd.hix.w, nd.loy.w, nd.hiy.w, nd.loz.w, nd.hiz.w, &kk[3]))) kk[3] = inf;
                if (MH_BUNDLE_SORT) {
                    // nearest first along the axis: blocked rays end the walk sooner
                    auto cs = [&](int a_, int b_) {
                        if (kk[b_] < kk[a_]) {
                            const float tk_ = kk[a_];
                            kk[a_] = kk[b_];
                            kk[b_] = tk_;
                            const int tq_ = qq[a_];
                            qq[a_] = qq[b_];
                            qq[b_] = tq_;
                        }
                    };
                    cs(0, 1);
                    cs(2, 3);
                    cs(0, 2);
                    cs(1, 3);
                    cs(1, 2);
                }
                for (int i = 3; i >= 0; --i)
                    if (kk[i] < inf) st.push(qq[i], 0.f);
            } else if (leaf(ref)) {
                return occ;
            }
        }
        return occ;
    }
    st.push(sc.root_ref, 0.f);
    while (!st.empty()) {
        float te;
        const int ref = st.pop(&te);
        if (ref >= 0) {
            const DNode nd = ldg(&sc.nodes[ref]);
            float tl, tr;
            const bool hr = meets(nd.c1.z, nd.c1.w, nd.c2.x, nd.c2.y, nd.c2.z, nd.c2.w, &tr);
            const bool hl = meets(nd.c0.x, nd.c0.y, nd.c0.z, nd.c0.w, nd.c1.x, nd.c1.y, &tl);
            if (hl && hr && MH_BUNDLE_SORT && tr < tl) {
                st.push(nd.c3.x, 0.f);
                st.push(nd.c3.y, 0.f);
            } else {
                if (hr) st.push(nd.c3.y, 0.f);
                if (hl) st.push(nd.c3.x, 0.f);
            }
        } else if (leaf(ref)) {
            return occ;
        }
    }
    return occ;
}

// Normalise a direction unless it is already unit within 1e-12 in |d|^2,
// as the reference's bvh_intersect / primitive_intersect do on entry.
template <typename R> __device__ __forceinline__ V3<R> unit_dir(V3<R> d)
{
    R l2 = len2(d);
    if (l2 > R(0) && mfabs(l2 - R(1)) > R(1e-12)) d = normalize(d);
    return d;
}

} // namespace mh
