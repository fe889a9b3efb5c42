// mh_shade.h - materials, sky and sampling on the GPU.
//
// Follows reference src/material.c (shade_local 157-188, shade_pbr 218-290,
// glossy dir 303-349, ambient 351-374, sky_sample 473-601, sky_basis /
// sun disk / sphere cap 1062-1146, water 1161-1267), reference
// src/sampling.c:52-281 (basis, cosine hemisphere, Fresnel, refraction,
// p_spec, GGX) and reference src/texture.c:269-327 (albedo textures).
#pragma once

#include "mh_scene_dev.h"

namespace mh {

template <typename R> __device__ __forceinline__ R clamp01m(R v) // material.c: NaN passes
{
    if (v < R(0)) return R(0);
    if (v > R(1)) return R(1);
    return v;
}
template <typename R> __device__ __forceinline__ R clamp01n(R v) // sampling.c: NaN -> 0
{
    if (!(v > R(0))) return R(0);
    if (v > R(1)) return R(1);
    return v;
}
template <typename R> __device__ __forceinline__ R clamp_n(R v, R lo, R hi)
{
    if (!(v > lo)) return lo;
    if (v > hi) return hi;
    return v;
}
template <typename R> __device__ __forceinline__ R clamp_u(R u)
{
    if (!(u > R(0))) return R(0);
    if (u > Consts<R>::umax) return Consts<R>::umax;
    return u;
}
template <typename R> __device__ __forceinline__ R smooth3(R e0, R e1, R x)
{
    R den = e1 - e0;
    if (den == R(0)) return x < e0 ? R(0) : R(1);
    R t = clamp01m((x - e0) / den);
    return t * t * (R(3) - R(2) * t);
}
template <typename R> __device__ __forceinline__ R powpos(R b, R e) { return b <= R(0) ? R(0) : mpow(b, e); }

// Frame by crossing with the world axis least aligned with n (n unit).
template <typename R> __device__ __forceinline__ void frame(V3<R> n, V3<R>* t, V3<R>* b)
{
    V3<R> axis;
    if (mfabs(n.x) <= mfabs(n.y) && mfabs(n.x) <= mfabs(n.z)) axis = mk(R(1), R(0), R(0));
    else if (mfabs(n.y) <= mfabs(n.z)) axis = mk(R(0), R(1), R(0));
    else axis = mk(R(0), R(0), R(1));
    *t = normalize(cross(axis, n));
    *b = cross(n, *t);
}

// sampling.c basis: re-normalises n first.
template <typename R> __device__ __forceinline__ void basis_n(V3<R> n, V3<R>* t, V3<R>* b)
{
    V3<R> nn = normalize(n);
    if (len2(nn) <= R(0)) nn = mk(R(0), R(1), R(0));
    frame(nn, t, b);
}

template <typename R> __device__ V3<R> cosine_dir(V3<R> n, R u1, R u2)
{
    V3<R> nn = normalize(n);
    if (len2(nn) <= R(0)) nn = mk(R(0), R(1), R(0));
    R r1 = clamp_u(u1), r2 = clamp_u(u2);
    R r = msqrt(r1);
    R phi = R(2) * Consts<R>::pi * r2;
    R s, c;
    msincos(phi, &s, &c);
    R x = r * c, y = r * s;
    R z = msqrt(mfmax(R(0), R(1) - r1));
    V3<R> t, b;
    basis_n(nn, &t, &b);
    return normalize((t * x + b * y) + nn * z);
}

template <typename R> __device__ __forceinline__ R schlick_s(R c, R f0) // sampling.c
{
    R cc = clamp01n(c), f = clamp01n(f0);
    R m = R(1) - cc, m2 = m * m;
    return f + (R(1) - f) * (m2 * m2 * m);
}
template <typename R> __device__ __forceinline__ V3<R> schlick_s3(R c, V3<R> f0)
{
    return mk(schlick_s(c, f0.x), schlick_s(c, f0.y), schlick_s(c, f0.z));
}

template <typename R> __device__ R fresnel_diel(R cos_i, R eta)
{
    R c = clamp01n(cos_i);
    R e = eta > R(0) ? eta : R(0);
    R s2t = e * e * (R(1) - c * c);
    if (s2t >= R(1)) return R(1);
    R ct = msqrt(R(1) - s2t);
    R rs = (e * c - ct) / (e * c + ct);
    R rp = (e * ct - c) / (e * ct + c);
    return clamp01n(R(0.5) * (rs * rs + rp * rp));
}

template <typename R> __device__ bool refract_n(V3<R> incident, V3<R> n, R eta, V3<R>* out)
{
    V3<R> i = normalize(incident), nn = normalize(n);
    if (len2(i) <= R(0) || len2(nn) <= R(0)) return false;
    R e = eta > R(0) ? eta : R(0);
    R ci = -dot(i, nn);
    R s2t = e * e * (R(1) - ci * ci);
    if (s2t >= R(1)) return false;
    R ct = msqrt(R(1) - s2t);
    R k = e * ci - ct;
    *out = normalize(mk(e * i.x + k * nn.x, e * i.y + k * nn.y, e * i.z + k * nn.z));
    return true;
}

template <typename R> __device__ __forceinline__ R p_specular(R metallic, R roughness, R f0)
{
    R m = clamp01n(metallic), r = clamp01n(roughness), f = clamp01n(f0);
    R p = f + (R(1) - f) * m;
    p *= R(1) - R(0.5) * r * (R(1) - m);
    return clamp_n(p, R(1e-3), R(1));
}

template <typename R> __device__ __forceinline__ R d_ggx(R ndh, R a)
{
    R a2 = a * a;
    R d = ndh * ndh * (a2 - R(1)) + R(1);
    return a2 / (Consts<R>::pi * d * d);
}

template <typename R> __device__ V3<R> ggx_half(V3<R> n, R alpha, R u1, R u2)
{
    V3<R> nn = normalize(n);
    if (len2(nn) <= R(0)) nn = mk(R(0), R(1), R(0));
    R a = alpha > R(1e-4) ? alpha : R(1e-4);
    R a2 = a * a;
    R r1 = clamp_u(u1), r2 = clamp_u(u2);
    R ct = msqrt((R(1) - r1) / (R(1) + (a2 - R(1)) * r1));
    R st = msqrt(mfmax(R(0), R(1) - ct * ct));
    R phi = R(2) * Consts<R>::pi * r2;
    R s, c;
    msincos(phi, &s, &c);
    V3<R> t, b;
    basis_n(nn, &t, &b);
    return normalize((t * (st * c) + b * (st * s)) + nn * ct);
}

template <typename R> __device__ __forceinline__ R ggx_pdf(R ndh, R hdv, R alpha)
{
    if (!(hdv > R(0)) || !(ndh > R(0))) return R(0);
    R a = alpha > R(1e-4) ? alpha : R(1e-4);
    R c = ndh > R(1) ? R(1) : ndh;
    return d_ggx(c, a) * c / (R(4) * hdv);
}

template <typename R> __device__ __forceinline__ R rough_alpha(R r)
{
    R c = clamp01m(r);
    R a = c * c;
    return a < R(1e-4) ? R(1e-4) : a;
}

// Blinn-Phong NEE response: albedo*L*NdotL + specular*L*(N.H)^shininess.
template <typename R> __device__ V3<R> shade_local(const DMat<R>& m, V3<R> alb, V3<R> N, V3<R> L, V3<R> V, V3<R> light)
{
    R ndl = dot(N, L);
    if (ndl <= R(0)) return zero3<R>();
    V3<R> diff = vmul(alb, light) * ndl;
    V3<R> hv = L + V;
    V3<R> spec = zero3<R>();
    if (len2(hv) > R(1e-12)) {
        R ndh = dot(N, normalize(hv));
        if (ndh > R(0)) spec = vmul(m.specular, light) * powpos(ndh, m.shininess);
    }
    return diff + spec;
}

// Cook-Torrance NEE response.
template <typename R> __device__ V3<R> shade_pbr(const DMat<R>& m, V3<R> alb, V3<R> N, V3<R> L, V3<R> V, V3<R> light)
{
    R ndl = dot(N, L);
    if (ndl <= R(0)) return zero3<R>();
    R ndv = dot(N, V);
    if (ndv <= R(0)) return zero3<R>();
    R met = clamp01m(m.metallic);
    R alpha = rough_alpha(m.roughness);
    V3<R> hv = L + V;
    R ndh = R(0);
    if (len2(hv) > R(1e-12)) {
        ndh = dot(N, normalize(hv));
        if (ndh < R(0)) ndh = R(0);
    }
    V3<R> f0 = lerp(mk(R(0.04), R(0.04), R(0.04)), alb, met);
    R c = clamp01m(ndv), q = R(1) - c, q2 = q * q, k5 = q2 * q2 * q;
    V3<R> F = mk(f0.x + (R(1) - f0.x) * k5, f0.y + (R(1) - f0.y) * k5, f0.z + (R(1) - f0.z) * k5);
    R D = d_ggx(ndh, alpha);
    R rr = clamp01m(m.roughness);
    R k = (rr + R(1)) * (rr + R(1)) / R(8);
    if (k < R(0)) k = R(0);
    else if (k > R(1)) k = R(1);
    R G = (ndl / (ndl * (R(1) - k) + k)) * (ndv / (ndv * (R(1) - k) + k));
    R den = R(4) * ndl * ndv;
    R ss = den > R(0) ? D * G / den : R(0);
    V3<R> spec = F * ss;
    V3<R> kd = (mk(R(1), R(1), R(1)) - F) * (R(1) - met);
    V3<R> fd = vmul(kd, alb) * (R(1) / Consts<R>::pi);
    return vmul(light, fd + spec) * ndl;
}

template <typename R> __device__ __forceinline__ V3<R> shade(const DMat<R>& m, V3<R> alb, V3<R> N, V3<R> L, V3<R> V, V3<R> c)
{
    return m.pbr ? shade_pbr(m, alb, N, L, V, c) : shade_local(m, alb, N, L, V, c);
}

template <typename R> __device__ V3<R> glossy_dir(const DMat<R>& m, V3<R> N, V3<R> inc, R r1, R r2)
{
    V3<R> mirror = reflect(inc, N);
    if (m.roughness <= R(1e-6)) return mirror;
    R a = rough_alpha(m.roughness);
    R a2 = a * a;
    if (!(r1 > R(0))) r1 = R(0);
    else if (r1 > Consts<R>::umax) r1 = Consts<R>::umax;
    if (!(r2 > R(0))) r2 = R(0);
    else if (r2 > Consts<R>::umax) r2 = Consts<R>::umax;
    R ct = msqrt((R(1) - r1) / (R(1) + (a2 - R(1)) * r1));
    R st = msqrt(R(1) - ct * ct);
    R phi = R(2) * Consts<R>::pi * r2;
    R s, c;
    msincos(phi, &s, &c);
    V3<R> t, b;
    frame(N, &t, &b);
    V3<R> H = normalize((t * (st * c) + b * (st * s)) + N * ct);
    V3<R> Rv = reflect(inc, H);
    if (!(dot(Rv, N) > R(0))) return mirror;
    return normalize(Rv);
}

template <typename R> __device__ __forceinline__ V3<R> ambient(V3<R> alb, V3<R> N, const DSky<R>& sky)
{
    V3<R> ground = sky.horizon * R(0.45);
    R w = clamp01m(R(0.5) + R(0.5) * N.y);
    V3<R> hemi = lerp(ground, sky.zenith, w);
    return vmul(alb, hemi) * R(0.15);
}

// Gradient + sun glow + planar fBm cloud layer.
template <typename R> __device__ V3<R> sky_sample(V3<R> dir, const DSky<R>& sky)
{
    R h = clamp01m(dir.y);
    R t = powpos(h, sky.gamma);
    V3<R> base = lerp(sky.horizon, sky.zenith, t);
    R cs = dot(dir, sky.sun_dir);
    if (cs > R(0)) base = base + sky.sun_color * (powpos(cs, sky.glow_exp) * sky.glow_strength);
    if (dir.y > R(0.03)) {
        R tt = sky.cloud_height / dir.y;
        R px = dir.x * tt, pz = dir.z * tt;
        R raw = fbm2<R>(px * sky.cloud_scale, pz * sky.cloud_scale, sky.octaves, sky.seed);
        R dens = clamp01m(R(0.5) * (raw + R(1)));
        R al = powpos(smooth3(sky.coverage - sky.softness, sky.coverage + sky.softness, dens), sky.sharpness);
        al *= smooth3(R(0), R(0.15), dir.y);
        if (al > R(0)) {
            V3<R> cc = mk(R(0.95), R(0.95), R(0.98)) * (R(0.6) + R(0.4) * al);
            R lit = cs > R(0) ? cs : R(0);
            V3<R> lc = cc + vmul(cc, sky.sun_color) * (R(0.35) * lit);
            base = lerp(base, lc, al);
        }
    }
    if (!(base.x >= R(0))) base.x = R(0);
    if (!(base.y >= R(0))) base.y = R(0);
    if (!(base.z >= R(0))) base.z = R(0);
    return base;
}

template <typename R> __device__ V3<R> sun_disk_dir(V3<R> sun, R radius_deg, R r1, R r2)
{
    if (radius_deg <= R(0)) return sun;
    V3<R> n = normalize(sun);
    if (len2(n) == R(0)) return mk(R(0), R(1), R(0));
    r1 = r1 < R(0) ? R(0) : (r1 > R(1) ? R(1) : r1);
    r2 = r2 < R(0) ? R(0) : (r2 > R(1) ? R(1) : r2);
    R rr = radius_deg * (Consts<R>::pi / R(180));
    R off = msqrt(r1) * rr;
    R th = R(2) * Consts<R>::pi * r2;
    R s, c;
    msincos(th, &s, &c);
    V3<R> t, b;
    frame(n, &t, &b);
    V3<R> radial = t * c + b * s;
    return normalize(n + radial * mtan(off));
}

template <typename R> __device__ V3<R> cone_dir(V3<R> w, R cos_max, R u1, R u2)
{
    w = normalize(w);
    if (len2(w) == R(0)) return mk(R(0), R(1), R(0));
    if (!(cos_max >= R(0))) cos_max = R(0);
    else if (cos_max > R(1)) cos_max = R(1);
    u1 = clamp_u(u1);
    u2 = clamp_u(u2);
    R ca = R(1) - u1 * (R(1) - cos_max);
    R sa = msqrt(R(1) - ca * ca);
    R phi = R(2) * Consts<R>::pi * u2;
    R s, c;
    msincos(phi, &s, &c);
    V3<R> t, b;
    frame(w, &t, &b);
    return normalize((t * (sa * c) + b * (sa * s)) + w * ca);
}

template <typename R> __device__ V3<R> water_normal(R x, R z)
{
    // (ax, az, w = 2 pi / lambda, amp, phase); time = 0
    const R W[4][5] = {
        {R(0.9584), R(0.2855), R(2.0 * 3.14159265358979323846 / 24.0), R(0.1800), R(0.0)},
        {R(0.7071), R(-0.7071), R(2.0 * 3.14159265358979323846 / 13.0), R(0.1000), R(1.7)},
        {R(-0.4061), R(0.9138), R(2.0 * 3.14159265358979323846 / 7.0), R(0.0500), R(3.1)},
        {R(0.2005), R(0.9797), R(2.0 * 3.14159265358979323846 / 4.0), R(0.0250), R(0.6)},
    };
    R gx = R(0), gz = R(0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        R ph = (W[i][0] * x + W[i][1] * z) * W[i][2] + W[i][4] + R(0);
        R c = mcos(ph) * W[i][3] * W[i][2];
        gx += c * W[i][0];
        gz += c * W[i][1];
    }
    const R f = R(0.35), amp = R(0.06), e = R(0.02);
    R hl = fbm2<R>((x - e) * f, z * f, 4, 1337u);
    R hr = fbm2<R>((x + e) * f, z * f, 4, 1337u);
    R hd = fbm2<R>(x * f, (z - e) * f, 4, 1337u);
    R hu = fbm2<R>(x * f, (z + e) * f, 4, 1337u);
    gx += amp * (hr - hl) / (R(2) * e);
    gz += amp * (hu - hd) / (R(2) * e);
    V3<R> n = normalize(mk(-gx, R(1), -gz));
    if (len2(n) < R(1e-12)) return mk(R(0), R(1), R(0));
    if (n.y < R(0)) n = -n;
    return n;
}

template <typename R> __device__ __forceinline__ R transmit(R a, R d)
{
    if (a == R(0)) return R(1);
    R t = mexp(-a * d);
    if (!(t > R(0))) return R(0);
    if (t > R(1)) return R(1);
    return t;
}

// Per-channel transmittance T and deep-colour term (1-T)*deep of a segment.
template <typename R> __device__ __forceinline__ void medium_terms(const DMat<R>& m, R depth, V3<R>* T, V3<R>* deep)
{
    R d = depth > R(0) ? depth : R(0);
    *T = mk(transmit(m.absorption.x, d), transmit(m.absorption.y, d), transmit(m.absorption.z, d));
    *deep = mk(m.deep_color.x * (R(1) - T->x), m.deep_color.y * (R(1) - T->y), m.deep_color.z * (R(1) - T->z));
}

template <typename R> __device__ __forceinline__ V3<R> water_attenuate(const DMat<R>& m, V3<R> inner, R depth)
{
    V3<R> T, deep;
    medium_terms(m, depth, &T, &deep);
    V3<R> o = mk(inner.x * T.x + deep.x, inner.y * T.y + deep.y, inner.z * T.z + deep.z);
    if (!(o.x >= R(0))) o.x = R(0);
    if (!(o.y >= R(0))) o.y = R(0);
    if (!(o.z >= R(0))) o.z = R(0);
    return o;
}

template <typename R> __device__ __forceinline__ V3<R> texture_albedo(const DMat<R>& m, V3<R> p, R u, R v)
{
    if (m.texture_kind == 1) {
        R s = m.texture_scale > R(0) ? m.texture_scale : R(1);
        R is = R(1) / s;
        int px = (int)(((long long)mfloor(p.x * is)) & 1);
        int py = (int)(((long long)mfloor(p.y * is)) & 1);
        int pz = (int)(((long long)mfloor(p.z * is)) & 1);
        return vmul(m.albedo, ((px + py + pz) & 1) ? m.tex_a : m.tex_b);
    }
    if (m.texture_kind == 7) {
        R s = m.texture_scale > R(0) ? m.texture_scale : R(1);
        int pu = (int)(((long long)mfloor(u * s)) & 1);
        int pv = (int)(((long long)mfloor(v * s)) & 1);
        return vmul(m.albedo, ((pu + pv) & 1) ? m.tex_a : m.tex_b);
    }
    return m.albedo;
}

template <typename R> __device__ V3<R> limb_glow(const DMat<R>& m, V3<R> N, V3<R> V, V3<R> sun)
{
    R cv = mfmax(R(0), dot(N, V));
    R limb = mpow(R(1) - cv, R(3.5));
    R sd = dot(N, sun);
    R day = R(0);
    if (sd > R(-0.15)) {
        day = (sd + R(0.15)) / R(0.35);
        if (day > R(1)) day = R(1);
    }
    V3<R> g = m.glow;
    if (sd > R(-0.10) && sd < R(0.25)) {
        R st = R(1) - mfabs(sd - R(0.05)) / R(0.20);
        if (st > R(0)) g = lerp(g, mk(R(1), R(0.45), R(0.15)), st * R(0.6));
    }
    return g * (limb * day * R(2.5));
}

} // namespace mh
