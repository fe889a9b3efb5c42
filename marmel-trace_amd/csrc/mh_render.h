// mh_render.h - GPU path tracer and fixed-spp Whitted renderer.
//
// Path tracer: reference src/pathtrace.c (lobe sampling 251-453, medium
// 220-233, NEE + shadow segments 504-606, radiance 800-1028, pixel loop
// 1047-1096). Whitted: reference src/render.c:507-970. Every random number is
// the same pure function of (pixel, sample, bounce, channel) as in the
// reference, so the GPU and CPU images differ only by floating-point rounding.
#pragma once

#include "mh_shade.h"
#include "mh_trace.h"

namespace mh {

// 1: the double-precision Whitted renderer tests the 16 sun-disk shadow rays
// of a hit as one bundle (any_hit_bundle: one walk, the same occlusion bits);
// 0: sixteen any_hit walks. The 16 rays to an emissive light stay separate
// walks: their cone soon grows wider than small primitives, and a bundle walk
// then reaches many more leaves, each tested against all 16 rays.
#ifndef MH_WHITTED_BUNDLE
#define MH_WHITTED_BUNDLE 1
#endif

// ---- path tracer --------------------------------------------------------------

template <typename R> struct Lobe {
    V3<R> wi, w;
    int refr, spec;
};

template <typename R>
__device__ bool sample_pbr(const DMat<R>& m, V3<R> alb, V3<R> N, V3<R> d, V3<R> V, R u1, R u2, R ul, Lobe<R>* out)
{
    out->spec = 0;
    out->refr = 0;
    R met = clamp01n(m.metallic), rough = clamp01n(m.roughness);
    R alpha = rough_alpha(rough);
    V3<R> f0 = lerp(mk(R(0.04), R(0.04), R(0.04)), alb, met);
    R lum = R(0.2126) * f0.x + R(0.7152) * f0.y + R(0.0722) * f0.z;
    lum = lum > R(0) ? lum : R(0);
    R ps = p_specular(met, rough, lum);
    R ndv = dot(N, V);
    if (!(ndv > R(0))) return false;
    if (ul < ps) {
        V3<R> H = ggx_half(N, alpha, u1, u2);
        V3<R> wi = reflect(d, H);
        R ndl = dot(N, wi), ndh = dot(N, H), vdh = dot(V, H);
        if (!(ndl > R(0)) || !(ndh > R(0)) || !(vdh > R(0))) return false;
        R D = d_ggx(ndh, alpha);
        R k = clamp01n((rough + R(1)) * (rough + R(1)) / R(8));
        R G = (ndl / (ndl * (R(1) - k) + k)) * (ndv / (ndv * (R(1) - k) + k));
        V3<R> F = schlick_s3(vdh, f0);
        R pdf = ggx_pdf(ndh, vdh, alpha);
        if (!(pdf > R(0)) || !(ps > R(0))) return false;
        out->wi = wi;
        out->w = F * (D * G / (R(4) * ndv * pdf * ps));
        out->spec = 1;
        return true;
    }
    R den = R(1) - ps;
    if (!(den > R(0))) return false;
    V3<R> wi = cosine_dir(N, u1, u2);
    if (!(dot(N, wi) > R(0))) return false;
    V3<R> F = schlick_s3(ndv, f0);
    V3<R> kd = (mk(R(1), R(1), R(1)) - F) * (R(1) - met);
    out->wi = wi;
    out->w = vmul(kd, alb) * (R(1) / den);
    return true;
}

// PBR lobe sampling with every decision (lobe, half vector, pdf, lobe
// probability) from the frozen material f / albedo falb and the BSDF value
// from m / alb: the weight f_m / p_f is then a smooth function of m.
template <typename R>
__device__ bool sample_pbr_frozen(const DMat<R>& m, V3<R> alb, const DMat<R>& f, V3<R> falb, V3<R> N, V3<R> d,
                                  V3<R> V, R u1, R u2, R ul, Lobe<R>* out)
{
    out->spec = 0;
    out->refr = 0;
    R met = clamp01n(m.metallic), rough = clamp01n(m.roughness);
    R alpha = rough_alpha(rough);
    R fmet = clamp01n(f.metallic), frough = clamp01n(f.roughness);
    R falpha = rough_alpha(frough);
    V3<R> f0 = lerp(mk(R(0.04), R(0.04), R(0.04)), alb, met);
    V3<R> ff0 = lerp(mk(R(0.04), R(0.04), R(0.04)), falb, fmet);
    R lum = R(0.2126) * ff0.x + R(0.7152) * ff0.y + R(0.0722) * ff0.z;
    lum = lum > R(0) ? lum : R(0);
    R ps = p_specular(fmet, frough, lum);
    R ndv = dot(N, V);
    if (!(ndv > R(0))) return false;
    if (ul < ps) {
        V3<R> H = ggx_half(N, falpha, u1, u2);
        V3<R> wi = reflect(d, H);
        R ndl = dot(N, wi), ndh = dot(N, H), vdh = dot(V, H);
        if (!(ndl > R(0)) || !(ndh > R(0)) || !(vdh > R(0))) return false;
        R D = d_ggx(ndh, alpha);
        R k = clamp01n((rough + R(1)) * (rough + R(1)) / R(8));
        R G = (ndl / (ndl * (R(1) - k) + k)) * (ndv / (ndv * (R(1) - k) + k));
        V3<R> F = schlick_s3(vdh, f0);
        R pdf = ggx_pdf(ndh, vdh, falpha);
        if (!(pdf > R(0)) || !(ps > R(0))) return false;
        out->wi = wi;
        out->w = F * (D * G / (R(4) * ndv * pdf * ps));
        out->spec = 1;
        return true;
    }
    R den = R(1) - ps;
    if (!(den > R(0))) return false;
    V3<R> wi = cosine_dir(N, u1, u2);
    if (!(dot(N, wi) > R(0))) return false;
    V3<R> F = schlick_s3(ndv, f0);
    V3<R> kd = (mk(R(1), R(1), R(1)) - F) * (R(1) - met);
    out->wi = wi;
    out->w = vmul(kd, alb) * (R(1) / den);
    return true;
}

template <typename R>
__device__ bool sample_legacy(const DMat<R>& m, V3<R> alb, V3<R> N, V3<R> d, V3<R> V, int front, R u1, R u2, R ul,
                              R ug, Lobe<R>* out)
{
    out->refr = 0;
    out->spec = 0;
    R rm = clamp01n(m.reflectivity), rt = clamp01n(m.transparency);
    R rd = R(1) - rm - rt;
    if (rd < R(0)) rd = R(0);
    R pm = rm, pt = rt, pd = R(1) - pm - pt, norm = R(1);
    if (pd < R(0)) {
        R s = pm + pt;
        if (s > R(0)) {
            pm /= s;
            pt /= s;
            if (s > R(1)) norm = s;
        }
        pd = R(0);
    }
    if (ul < pm) {
        if (!(pm > R(0))) return false;
        R w = rm / (pm * norm);
        out->wi = reflect(d, N);
        out->w = mk(w, w, w);
        out->spec = 1;
        return true;
    }
    if (ul < pm + pt) {
        if (!(pt > R(0))) return false;
        out->spec = 1;
        R ior = m.ior > R(0) ? m.ior : R(1);
        R eta = front ? R(1) / ior : ior;
        R F = fresnel_diel(clamp01n(dot(N, V)), eta);
        R lobe = rt / (pt * norm);
        if (F >= R(1)) {
            out->wi = reflect(d, N);
            out->w = mk(lobe, lobe, lobe);
            return true;
        }
        R pr = clamp_n(F, R(0.05), R(0.95));
        if (ug < pr) {
            R w = lobe * (F / pr);
            out->wi = reflect(d, N);
            out->w = mk(w, w, w);
            return true;
        }
        V3<R> wi;
        if (!refract_n(d, N, eta, &wi)) {
            R w = lobe * (F / pr);
            out->wi = reflect(d, N);
            out->w = mk(w, w, w);
            return true;
        }
        R w = lobe * ((R(1) - F) / (R(1) - pr)) * eta * eta;
        out->wi = wi;
        out->w = mk(w, w, w);
        out->refr = 1;
        return true;
    }
    if (!(pd > R(0))) return false;
    V3<R> wi = cosine_dir(N, u1, u2);
    if (!(dot(N, wi) > R(0))) return false;
    out->wi = wi;
    out->w = alb * (rd / pd);
    return true;
}

// Shadow connection through up to 6 closest hits; water and near-clear
// (transparency > 0.8) surfaces let light through with Beer-Lambert decay.
template <typename R, typename TS>
__device__ bool occluded_segment(const SceneView<R>& sc, TS ts, V3<R> o, V3<R> d, R tmin, R tmax, V3<R>* att,
                                 int* nq)
{
    *att = mk(R(1), R(1), R(1));
    if (!sc.has_transmissive) {
        // every first hit would be opaque: blocked iff any hit in [tmin, tmax]
        ++*nq;
        return any_hit(sc, ts, o, unit_dir(d), tmin, tmax);
    }
    R rem = tmax;
    // The queries run on the unit direction, so hit distances (and the
    // absorption over them) are in scene units and a sphere's unit-direction
    // quadratic holds even for an unnormalised sun direction; the step past
    // each transmissive hit goes along d as given.
    const V3<R> u = unit_dir(d);
    for (int step = 0; step < 6; ++step) {
        DHit<R> h;
        ++*nq;
        if (!intersect(sc, ts, o, u, tmin, rem, &h)) return false;
        if (h.mat < 0 || h.mat >= sc.n_mats) return true;
        const DMat<R>& m = sc.mats[h.mat];
        if (m.is_water || m.transparency > R(0.8)) {
            if (m.beer_lambert || m.is_water)
                *att = vmul(*att, mk(mexp(-m.absorption.x * h.t), mexp(-m.absorption.y * h.t),
                                     mexp(-m.absorption.z * h.t)));
            o = h.p + d * R(1e-3);
            rem -= h.t;
            if (rem <= R(1e-3)) return false;
            tmin = R(1e-3);
            continue;
        }
        return true;
    }
    return true;
}

// Returns the radiance it added to L. beta2 (optional) is attenuated alike.
template <typename R>
__device__ V3<R> apply_medium(const DMat<R>& m, R dist, V3<R>* beta, V3<R>* L, V3<R>* beta2 = nullptr)
{
    V3<R> T, deep;
    medium_terms(m, dist, &T, &deep);
    V3<R> white = mk(R(1) * T.x + deep.x, R(1) * T.y + deep.y, R(1) * T.z + deep.z);
    V3<R> Tm = white - deep;
    V3<R> add = vmul(*beta, deep);
    *L = *L + add;
    *beta = vmul(*beta, Tm);
    if (beta2) *beta2 = vmul(*beta2, Tm);
    return add;
}

// Water surface height at (x, z): the scene's Gerstner ocean when it has one
// (reference src/scene.c:1109-1153), else the constant water level. Decides
// whether a camera ray starts under water: once per camera ray, so it is kept
// out of line (inlining its four fp64 cosines into the fp64 megakernel gave
// wrong fp64 images, also in scenes that never call it).
template <typename R> __device__ __noinline__ R water_height(const SceneView<R>& sc, R x, R z)
{
    if (!sc.ocean.on) return sc.water_level;
    R y = sc.ocean.y0;
    for (int h = 0; h < 4; ++h)
        y += sc.ocean.A[h] * mcos(sc.ocean.k[h] * (sc.ocean.dx[h] * x + sc.ocean.dz[h] * z) + sc.ocean.phi[h]);
    return y;
}

// ---- fog (reference src/material.c:963-1039, src/pathtrace.c:614-794) --------

// Henyey-Greenstein lobe (1 - g^2) / (1 + g^2 - 2 g cos)^1.5, denominator
// floored at 1e-4.
template <typename R> __device__ __forceinline__ R hg_lobe(R g, R c)
{
    R den = R(1) + g * g - R(2) * g * c;
    if (den < R(1e-4)) den = R(1e-4);
    return (R(1) - g * g) / (den * msqrt(den));
}

// Density multiplier of the fog turbulence at p: 1 - a + 2 a clamp01((fbm3 + 1) / 2).
template <typename R> __device__ __forceinline__ R fog_turb(const DFog<R>& f, V3<R> p)
{
    if (!f.noise) return R(1);
    R n = fbm3<R>(p.x * f.noise_scale, p.y * f.noise_scale, p.z * f.noise_scale, 1337u);
    R t = R(0.5) * (n + R(1));
    if (t < R(0)) t = R(0);
    if (t > R(1)) t = R(1);
    return (R(1) - f.noise_amount) + f.noise_amount * t * R(2);
}

// Closed-form segment [o, o + dist d] (dist >= 1e20: to infinity) through
// exponential height fog: transmittance T and in-scatter
// (colour + sun HG strength) (1 - T), the optical depth scaled by the
// turbulence at one point. dsun (optional): d(in-scatter)/d(sun colour).
template <typename R>
__device__ void fog_segment(const DFog<R>& f, const DSky<R>& sky, V3<R> o, V3<R> d, R dist, R* T, V3<R>* ins,
                            V3<R>* dsun)
{
    *T = R(1);
    *ins = zero3<R>();
    if (dsun) *dsun = zero3<R>();
    if (f.density <= R(0) || dist <= R(1e-7)) return;
    R lam = f.falloff > R(0) ? f.falloff : R(0);
    R y0 = o.y - f.height, dy = d.y, tau;
    bool to_sky = dist >= R(1e20);
    if (lam <= R(1e-6)) {
        tau = to_sky ? R(1e6) : f.density * dist;
    } else {
        R base = f.density * mexp(-lam * y0);
        if (to_sky)
            tau = dy > R(1e-5) ? base / (lam * dy) : R(1e6);
        else if (mfabs(dy) < R(1e-5))
            tau = base * dist;
        else
            tau = base * ((R(1) - mexp(-lam * dy * dist)) / (lam * dy));
    }
    if (tau < R(0)) tau = R(0);
    R ev = to_sky ? R(20) : (dist > R(50) ? R(50) : dist * R(0.5));
    tau *= fog_turb(f, o + d * ev);
    R tr = mexp(-tau);
    if (tr < R(0)) tr = R(0);
    if (tr > R(1)) tr = R(1);
    V3<R> col = f.color;
    R ph = R(0);
    if (f.inscatter > R(0)) {
        ph = hg_lobe(f.g, dot(d, sky.sun_dir));
        col = col + sky.sun_color * (ph * f.inscatter);
    }
    *T = tr;
    *ins = col * (R(1) - tr);
    if (dsun) {
        R k = ph * f.inscatter * (R(1) - tr);
        *dsun = mk(k, k, k);
    }
}

// Camera-ray segment: shadow_steps (clamped to [4, 128]) jittered steps, each
// lit by the sun unless its shadow ray (tmin 0.05, tmax 100) is blocked; HG
// clamped at 6, mixed 0.7 / 0.3 with isotropic. The march covers at most 60
// units (to infinity: up to the fog layer's top, at most 40); the rest is the
// closed form.
template <typename R, typename TS>
__device__ void pt_fog_segment(const SceneView<R>& sc, TS ts, V3<R> o, V3<R> d, R dist, uint32_t seed, int b,
                               R* T, V3<R>* ins, V3<R>* dsun, int* nq)
{
    const DFog<R>& f = sc.fog;
    if (f.steps <= 0) {
        fog_segment(f, sc.sky, o, d, dist, T, ins, dsun);
        return;
    }
    int steps = f.steps < 4 ? 4 : (f.steps > 128 ? 128 : f.steps);
    bool to_sky = dist >= R(1e20);
    R lam = f.falloff, max_t;
    if (to_sky) {
        R top = lam > R(1e-4) ? f.height + R(3) / lam : R(25);
        if (o.y >= top && d.y >= R(0)) {
            *T = R(1);
            *ins = zero3<R>();
            if (dsun) *dsun = zero3<R>();
            return;
        }
        if (d.y > R(0.02)) {
            max_t = (top - o.y) / d.y;
            if (max_t > R(40)) max_t = R(40);
        } else {
            max_t = R(40);
        }
    } else {
        max_t = dist > R(60) ? R(60) : dist;
    }
    R dt = max_t / (R)steps;
    R ph = hg_lobe(f.g, dot(d, sc.sky.sun_dir));
    if (ph > R(6)) ph = R(6);
    R sk = (R(0.70) * ph + R(0.30)) * f.inscatter;
    V3<R> glow = sc.sky.sun_color * sk;
    R Ta = R(1);
    V3<R> acc = zero3<R>(), dacc = zero3<R>();
    R jit = rand01<R>(seed, (uint32_t)b, 0xFE01u);
    for (int i = 0; i < steps; ++i) {
        V3<R> p = o + d * (((R)i + jit) * dt);
        R rho = f.density;
        if (lam > R(1e-6)) rho *= mexp(-lam * (p.y - f.height));
        rho *= fog_turb(f, p);
        if (rho < R(0)) rho = R(0);
        R sT = mexp(-(rho * dt));
        V3<R> att;
        bool shadowed = occluded_segment(sc, ts, p, sc.sky.sun_dir, R(0.05), R(100), &att, nq);
        V3<R> col = f.color;
        if (!shadowed) col = col + vmul(glow, att);
        acc = acc + (col * (R(1) - sT)) * Ta;
        if (dsun && !shadowed) dacc = dacc + (att * (sk * (R(1) - sT))) * Ta;
        Ta *= sT;
        if (Ta < R(1e-4)) break;
    }
    if (dist > max_t) {
        R rT;
        V3<R> rins, rd;
        fog_segment(f, sc.sky, o + d * max_t, d, to_sky ? R(1e30) : dist - max_t, &rT, &rins, dsun ? &rd : nullptr);
        acc = acc + rins * Ta;
        if (dsun) dacc = dacc + rd * Ta;
        Ta *= rT;
    }
    *T = Ta;
    *ins = acc;
    if (dsun) *dsun = dacc;
}

// Segment inside medium m. Water with fog steps > 0 is marched in steps
// (clamped to [4, 16]) jittered steps: Beer-Lambert exp(-absorption dt) per
// step and in-scatter (0.4 deep colour + sun HG(0.72, clamped at 4.5) 0.18
// where the shadow ray (tmin 0.02, tmax 100) is clear) (1 - T); other media
// take the closed-form blend. Returns the radiance added to L; beta2 takes
// the same transmittance; dsun (optional) accumulates d(added)/d(sun colour).
template <typename R, typename TS>
__device__ V3<R> pt_water_segment(const SceneView<R>& sc, TS ts, const DMat<R>& m, V3<R> o, V3<R> d, R dist,
                                  uint32_t seed, int b, V3<R>* beta, V3<R>* L, V3<R>* beta2, V3<R>* dsun, int* nq)
{
    if (!m.is_water || sc.fog.steps <= 0 || dist <= R(1e-4)) return apply_medium(m, dist, beta, L, beta2);
    int steps = sc.fog.steps < 4 ? 4 : (sc.fog.steps > 16 ? 16 : sc.fog.steps);
    R dt = dist / (R)steps;
    R ph = hg_lobe(R(0.72), dot(d, sc.sky.sun_dir));
    if (ph > R(4.5)) ph = R(4.5);
    R jit = rand01<R>(seed, (uint32_t)b, 0xAC31u);
    V3<R> sT = mk(mexp(-m.absorption.x * dt), mexp(-m.absorption.y * dt), mexp(-m.absorption.z * dt));
    V3<R> om = mk(R(1) - sT.x, R(1) - sT.y, R(1) - sT.z);
    V3<R> amb = m.deep_color * R(0.40), added = zero3<R>();
    for (int i = 0; i < steps; ++i) {
        V3<R> p = o + d * (((R)i + jit) * dt);
        V3<R> att;
        bool shadowed = occluded_segment(sc, ts, p, sc.sky.sun_dir, R(0.02), R(100), &att, nq);
        V3<R> col = amb;
        if (!shadowed) col = col + vmul(sc.sky.sun_color, att) * (ph * R(0.18));
        V3<R> t = vmul(*beta, vmul(col, om));
        *L = *L + t;
        added = added + t;
        if (dsun && !shadowed) *dsun = *dsun + vmul(*beta, vmul(att * (ph * R(0.18)), om));
        *beta = vmul(*beta, sT);
        if (beta2) *beta2 = vmul(*beta2, sT);
    }
    return added;
}

// Observer of one path's radiance terms, for the path-replay adjoint. The
// forward renderers use NoHook, which compiles away. Every term added to L
// is reported once, in order, through exactly one of these calls.
struct NoHook {
    // kGrad: the hook wants sun-colour derivatives of fog in-scatter (sun_lin)
    static constexpr bool kGrad = false;
    template <typename R> __device__ void add(V3<R>) {}
    // a term of L that is linear in the sun colour: d(term)/d(sun colour)
    template <typename R> __device__ void sun_lin(V3<R>) {}
    template <typename R> __device__ void sky(V3<R>, V3<R>, V3<R>) {}
    template <typename R> __device__ void emission(V3<R>, int, V3<R>) {}
    template <typename R>
    __device__ void nee_light(const DMat<R>&, V3<R>, V3<R>, V3<R>, V3<R>, V3<R>, int, V3<R>, R, V3<R>)
    {
    }
    template <typename R>
    __device__ void nee_sun(const DMat<R>&, V3<R>, V3<R>, V3<R>, V3<R>, V3<R>, V3<R>, R, V3<R>)
    {
    }
    template <typename R> __device__ void bsdf(int, const DMat<R>&, V3<R>, V3<R>, V3<R>, const Lobe<R>&) {}
    // Frozen-sampling renders return the material whose parameters take the
    // sampling decisions at this vertex; nullptr = the vertex's own.
    template <typename R> __device__ const DMat<R>* frozen(int) const { return nullptr; }
};

// FOG: compile the fog and marched-water code (scenes with fog_kernel set).
template <bool FOG, typename R, typename TS, typename HK>
__device__ V3<R> pt_radiance_h(const SceneView<R>& sc, TS ts, V3<R> ro, V3<R> rd0, int max_depth, uint32_t seed,
                               int* nq, HK& hk)
{
    V3<R> L = zero3<R>();
    V3<R> o = ro, d = normalize(rd0);
    if (len2(d) <= R(0)) return L;
    V3<R> beta = mk(R(1), R(1), R(1)), beta_f = beta;
    int medium = -1;
    if (sc.water_material >= 0 && sc.water_material < sc.n_mats && ro.y < water_height(sc, ro.x, ro.z))
        medium = sc.water_material;
    int last_spec = 1;
    for (int b = 0; b <= max_depth; ++b) {
        DHit<R> h;
        ++*nq;
        if (!intersect(sc, ts, o, d, R(1e-4), R(1e30), &h)) {
            if (medium >= 0) {
                if (FOG) {
                    V3<R> dw = zero3<R>();
                    hk.add(pt_water_segment(sc, ts, sc.mats[medium], o, d, R(5), seed, b, &beta, &L, &beta_f,
                                            HK::kGrad ? &dw : nullptr, nq));
                    if (HK::kGrad) hk.sun_lin(dw);
                } else {
                    hk.add(apply_medium(sc.mats[medium], R(5), &beta, &L, &beta_f));
                }
                medium = -1;
            }
            V3<R> skyc = sky_sample(d, sc.sky), bs = beta;
            if (FOG && sc.fog.density > R(0)) {
                R T;
                V3<R> ins, dfs;
                if (b == 0)
                    pt_fog_segment(sc, ts, o, d, R(1e30), seed, b, &T, &ins, HK::kGrad ? &dfs : nullptr, nq);
                else
                    fog_segment(sc.fog, sc.sky, o, d, R(1e30), &T, &ins, HK::kGrad ? &dfs : nullptr);
                skyc = skyc * T + ins;
                bs = beta * T;
                if (HK::kGrad) hk.sun_lin(vmul(beta, dfs));
            }
            V3<R> t = vmul(beta, skyc);
            L = L + t;
            hk.sky(bs, d, t);
            break;
        }
        if (medium >= 0) {
            if (FOG) {
                V3<R> dw = zero3<R>();
                hk.add(pt_water_segment(sc, ts, sc.mats[medium], o, d, h.t, seed, b, &beta, &L, &beta_f,
                                        HK::kGrad ? &dw : nullptr, nq));
                if (HK::kGrad) hk.sun_lin(dw);
            } else {
                hk.add(apply_medium(sc.mats[medium], h.t, &beta, &L, &beta_f));
            }
        } else if (FOG && sc.fog.density > R(0)) {
            R T;
            V3<R> ins, dfs;
            if (b == 0)
                pt_fog_segment(sc, ts, o, d, h.t, seed, b, &T, &ins, HK::kGrad ? &dfs : nullptr, nq);
            else
                fog_segment(sc.fog, sc.sky, o, d, h.t, &T, &ins, HK::kGrad ? &dfs : nullptr);
            V3<R> t = vmul(beta, ins);
            L = L + t;
            hk.add(t);
            if (HK::kGrad) hk.sun_lin(vmul(beta, dfs));
            beta = beta * T;
            beta_f = beta_f * T;
        }
        if (h.mat < 0 || h.mat >= sc.n_mats) {
            V3<R> t = vmul(beta, sky_sample(d, sc.sky));
            L = L + t;
            hk.sky(beta, d, t);
            break;
        }
        const DMat<R>& m = sc.mats[h.mat];
        V3<R> P = h.p, N = h.n;
        if (m.is_water && mfabs(h.n.y) > R(0.999) && mfabs(h.n.x) < R(1e-4) && mfabs(h.n.z) < R(1e-4))
            N = water_normal<R>(P.x, P.z);
        if (dot(N, d) > R(0)) N = -N;
        N = normalize(N);
        if (len2(N) <= R(0)) break;
        V3<R> V = -d;
        V3<R> alb = texture_albedo(m, P, h.u, h.v);
        if (b == 0 || last_spec) {
            V3<R> t = m.has_emission ? vmul(beta, m.emissive) : zero3<R>();
            if (m.has_emission) L = L + t;
            hk.emission(beta, h.mat, t);
        }
        if (b == 0 && m.has_glow) {
            V3<R> t = vmul(beta, limb_glow(m, N, V, sc.sky.sun_dir));
            L = L + t;
            hk.add(t);
        }

        // emissive-sphere NEE, one cone sample per light
        for (int li = 0; li < sc.n_lights; ++li) {
            const DLight<R>& lt = sc.lights[li];
            if (lt.prim == h.prim) continue;
            V3<R> toc = lt.center - P;
            R dc2 = len2(toc), dc = msqrt(dc2);
            if (dc <= lt.radius) continue;
            R cmx = msqrt(R(1) - (lt.radius * lt.radius) / dc2);
            R omega = R(2) * Consts<R>::pi * (R(1) - cmx);
            if (!(omega > R(0))) continue;
            V3<R> w = toc * (R(1) / dc);
            R e1 = rand01<R>(seed ^ (uint32_t)li, (uint32_t)b, 0xA201u);
            R e2 = rand01<R>(seed ^ (uint32_t)li, (uint32_t)b, 0xA202u);
            V3<R> wi = cone_dir(w, cmx, e1, e2);
            if (!finite3(wi) || dot(N, wi) <= R(0)) continue;
            V3<R> so = P + N * R(1e-3);
            PD<R> lp;
            lp.kind = K_SPHERE;
            lp.p0 = lt.center;
            lp.r0 = lt.radius;
            R lt_t;
            if (!hit_sphere<R, false>(lp, so, unit_dir(wi), R(1e-3), R(1e30), &lt_t, nullptr)) continue;
            V3<R> att;
            if (occluded_segment(sc, ts, so, wi, R(1e-3), lt_t - R(1e-3), &att, nq)) continue;
            V3<R> lit = shade(m, alb, N, wi, V, vmul(lt.emissive, att));
            V3<R> t = vmul(beta, lit * omega);
            L = L + t;
            hk.nee_light(m, alb, N, wi, V, beta, li, att, omega, t);
        }
        // sun NEE at the primary hit (or the first hit inside a medium after refraction)
        if (b == 0 || (b == 1 && medium >= 0 && last_spec)) {
            R s1 = rand01<R>(seed, (uint32_t)b, 0xA101u);
            R s2 = rand01<R>(seed, (uint32_t)b, 0xA102u);
            V3<R> Ls = sun_disk_dir(sc.sky.sun_dir, sc.sky.sun_radius, s1, s2);
            if (finite3(Ls) && dot(N, Ls) > R(0)) {
                V3<R> att;
                if (!occluded_segment(sc, ts, P + N * R(1e-3), Ls, R(1e-3), R(1e30), &att, nq)) {
                    if (!(att.x < R(1e-4) && att.y < R(1e-4) && att.z < R(1e-4))) {
                        V3<R> lit = shade(m, alb, N, Ls, V, vmul(sc.sky.sun_color, att));
                        V3<R> t = vmul(beta, lit * sc.sky.sun_omega);
                        L = L + t;
                        hk.nee_sun(m, alb, N, Ls, V, beta, att, sc.sky.sun_omega, t);
                    }
                }
            }
        }

        uint32_t kb = hash3(seed ^ 0x5054u, (uint32_t)b, 0u);
        R u1 = rand01<R>(kb, (uint32_t)b, 0xB501u);
        R u2 = rand01<R>(kb, (uint32_t)b, 0xB502u);
        R ul = rand01<R>(kb, (uint32_t)b, 0x1001u);
        R ug = rand01<R>(kb, (uint32_t)b, 0x1002u);
        Lobe<R> lb;
        bool ok;
        const DMat<R>* fm = hk.template frozen<R>(h.mat);
        V3<R> w0 = zero3<R>();
        if (fm) {
            // decisions (and the roulette's throughput) from the frozen material
            V3<R> falb = texture_albedo(*fm, P, h.u, h.v);
            Lobe<R> l0;
            if (m.pbr) {
                ok = sample_pbr_frozen(m, alb, *fm, falb, N, d, V, u1, u2, ul, &lb);
                if (ok) ok = sample_pbr(*fm, falb, N, d, V, u1, u2, ul, &l0);
            } else {
                ok = sample_legacy(m, alb, N, d, V, h.front, u1, u2, ul, ug, &lb);
                if (ok) ok = sample_legacy(*fm, falb, N, d, V, h.front, u1, u2, ul, ug, &l0);
            }
            w0 = l0.w;
        } else {
            ok = m.pbr ? sample_pbr(m, alb, N, d, V, u1, u2, ul, &lb)
                       : sample_legacy(m, alb, N, d, V, h.front, u1, u2, ul, ug, &lb);
        }
        if (!ok) break;
        hk.bsdf(h.mat, m, alb, N, V, lb);
        last_spec = lb.spec;
        beta = vmul(beta, lb.w);
        if (fm) beta_f = vmul(beta_f, w0);
        else beta_f = beta;
        R mc = beta_f.x;
        if (beta_f.y > mc) mc = beta_f.y;
        if (beta_f.z > mc) mc = beta_f.z;
        if (!(mc > R(0))) mc = R(0);
        if (!finite3(beta_f) || !(mc > R(0))) break;
        if (lb.refr && (m.is_water || m.beer_lambert)) medium = h.front ? h.mat : -1;
        if (b >= 2) {
            R q = clamp01n(mc);
            if (!(q > R(0))) break;
            if (rand01<R>(kb, (uint32_t)b, 0x5201u) >= q) break;
            beta = beta * (R(1) / q);
            if (fm) beta_f = beta_f * (R(1) / q);
        }
        V3<R> off = lb.refr ? lb.wi * R(1e-4) : N * (dot(lb.wi, N) >= R(0) ? R(1e-4) : R(-1e-4));
        o = P + off;
        d = lb.wi;
    }
    if (!finite3(L)) return zero3<R>();
    if (L.x < R(0)) L.x = R(0);
    if (L.y < R(0)) L.y = R(0);
    if (L.z < R(0)) L.z = R(0);
    return L;
}

template <bool FOG, typename R, typename TS>
__device__ V3<R> pt_radiance(const SceneView<R>& sc, TS ts, V3<R> ro, V3<R> rd0, int max_depth, uint32_t seed,
                             int* nq)
{
    NoHook hk;
    return pt_radiance_h<FOG>(sc, ts, ro, rd0, max_depth, seed, nq, hk);
}

template <typename R>
__device__ __forceinline__ void camera_ray(const CamView<R>& c, R u, R v, R r1, R r2, V3<R>* o, V3<R>* d)
{
    V3<R> pt = c.ll + (c.hor * u + c.ver * v);
    V3<R> dir = normalize(pt - c.pos);
    if (c.aperture <= R(1e-12)) {
        *o = c.pos;
        *d = dir;
        return;
    }
    R fd = c.focus > R(0) ? c.focus : R(1);
    V3<R> focal = c.pos + dir * fd;
    R rad = c.aperture * msqrt(r1);
    R th = R(2) * Consts<R>::pi * r2;
    R s, cs;
    msincos(th, &s, &cs);
    R sq = c.squeeze;
    V3<R> off;
    if (sq <= R(0) || mfabs(sq - R(1)) < R(1e-12)) off = c.right_n * (rad * cs) + c.up_n * (rad * s);
    else off = c.right_n * ((rad / sq) * cs) + c.up_n * (rad * s);
    *o = c.pos + off;
    *d = normalize(focal - *o);
}

template <typename R> __device__ __forceinline__ unsigned char to_byte(R c)
{
    if (!(c >= R(0))) c = R(0);
    else if (c > R(1)) c = R(1);
    int iv = (int)(mpow(c, R(1.0 / 2.2)) * R(255) + R(0.5));
    iv = iv < 0 ? 0 : (iv > 255 ? 255 : iv);
    return (unsigned char)iv;
}

// ---- Whitted (explicit work stack) ------------------------------------------------
//
// The reference recursion trace()/trace_hit() is linear in the colours its
// children return, so a pixel sample is the sum over a tree of rays of
// (weight x local term). Each stack item carries its ray, its weight and,
// for rays refracted into an absorbing medium, the material whose
// Beer-Lambert blend applies once the segment length is known.

template <typename R> struct WItem {
    V3<R> o, d, w;
    int depth;
    int medium; // material index of the absorbing medium, or -1
};

template <typename R, typename TS>
__device__ V3<R> emissive_direct(const SceneView<R>& sc, TS ts, const DMat<R>& m, V3<R> alb, V3<R> P, V3<R> N,
                                 V3<R> V,
                                 int self, uint32_t seed, int* nq)
{
    V3<R> sum = zero3<R>();
    V3<R> so = P + N * R(1e-3);
    for (int li = 0; li < sc.n_lights; ++li) {
        const DLight<R>& lt = sc.lights[li];
        if (lt.prim == self) continue;
        V3<R> toc = lt.center - P;
        R dc2 = len2(toc), dc = msqrt(dc2);
        if (dc <= lt.radius) continue;
        R cmx = msqrt(R(1) - (lt.radius * lt.radius) / dc2);
        R solid = R(2) * Consts<R>::pi * (R(1) - cmx);
        if (!(solid > R(0))) continue;
        V3<R> w = toc * (R(1) / dc);
        for (int si = 0; si < 16; ++si) {
            uint32_t k = hash3(seed ^ 0xE311u, (uint32_t)li, (uint32_t)si);
            R e1 = rand01<R>(k, (uint32_t)si, 0x4c11u);
            R e2 = rand01<R>(k, (uint32_t)si, 0x9d27u);
            V3<R> wi = cone_dir(w, cmx, e1, e2);
            if (dot(N, wi) <= R(0)) continue;
            PD<R> lp;
            lp.kind = K_SPHERE;
            lp.p0 = lt.center;
            lp.r0 = lt.radius;
            R lt_t;
            if (!hit_sphere<R, false>(lp, so, unit_dir(wi), R(1e-3), R(1e30), &lt_t, nullptr)) continue;
            ++*nq;
            if (any_hit(sc, ts, so, unit_dir(wi), R(1e-3), lt_t - R(1e-3))) continue;
            sum = sum + shade(m, alb, N, wi, V, lt.emissive * solid);
        }
    }
    return sum * (R(1) / R(16));
}

// Shade one Whitted hit: add weight x local terms to acc, push children.
// Returns false on stack overflow.
template <typename R, typename TS>
__device__ bool whitted_hit(const SceneView<R>& sc, TS ts, const WItem<R>& it, const DHit<R>& h, int max_depth,
                            uint32_t seed,
                            V3<R>* acc, WItem<R>* stk, int* sp, int cap, int* nq)
{
    if (it.depth > max_depth) return true;
    V3<R> d = it.d;
    if (h.mat < 0 || h.mat >= sc.n_mats) {
        *acc = *acc + vmul(it.w, sky_sample(d, sc.sky));
        return true;
    }
    const DMat<R>& m = sc.mats[h.mat];
    V3<R> P = h.p, N = h.n;
    if (m.is_water) N = water_normal<R>(P.x, P.z);
    if (dot(N, d) > R(0)) N = -N;
    V3<R> V = -d;
    V3<R> alb = texture_albedo(m, P, h.u, h.v);
    V3<R> color = ambient(alb, N, sc.sky);
    V3<R> Ls = sc.sky.sun_dir;
    V3<R> so = P + N * R(1e-3);
    if (sc.sky.sun_radius > R(1e-9)) {
        R lit = R(0);
        if constexpr (sizeof(R) == 8 && MH_WHITTED_BUNDLE) {
            // the 16 sun-disk rays leave so within the sun's angular radius of
            // its direction: one bundle walk, the same per-ray occlusion
            V3<R> dirs[16];
            for (int i = 0; i < 16; ++i) {
                uint32_t k = seed * 0x9e3779b9u + (uint32_t)i * 2u;
                R r1 = rand01<R>(k, (uint32_t)i, 0x5a17u);
                R r2 = rand01<R>(k, (uint32_t)i, 0x7c3du);
                dirs[i] = unit_dir(sun_disk_dir(Ls, sc.sky.sun_radius, r1, r2));
                ++*nq;
            }
            const uint32_t occ = any_hit_bundle(sc, ts, so, dirs, 16, normalize(Ls),
                                                sc.sky.sun_radius * (Consts<R>::pi / R(180)), R(1e-3), R(1e30));
            lit = R(16 - __builtin_popcount(occ));
        } else {
            for (int i = 0; i < 16; ++i) {
                uint32_t k = seed * 0x9e3779b9u + (uint32_t)i * 2u;
                R r1 = rand01<R>(k, (uint32_t)i, 0x5a17u);
                R r2 = rand01<R>(k, (uint32_t)i, 0x7c3du);
                V3<R> Ld = sun_disk_dir(Ls, sc.sky.sun_radius, r1, r2);
                ++*nq;
                if (!any_hit(sc, ts, so, unit_dir(Ld), R(1e-3), R(1e30))) lit += R(1);
            }
        }
        if (lit > R(0)) color = color + shade(m, alb, N, Ls, V, sc.sky.sun_color) * (lit / R(16));
    } else {
        ++*nq;
        if (!any_hit(sc, ts, so, unit_dir(Ls), R(1e-3), R(1e30)))
            color = color + shade(m, alb, N, Ls, V, sc.sky.sun_color);
    }
    if (sc.n_lights > 0) color = color + emissive_direct(sc, ts, m, alb, P, N, V, h.prim, seed, nq);

    // Fresnel at the refracted cosine (Schlick), TIR detection
    R f0 = (R(1) - m.ior) / (R(1) + m.ior);
    f0 = f0 * f0;
    R ci = dot(N, V);
    ci = ci < R(0) ? R(0) : (ci > R(1) ? R(1) : ci);
    R eta = h.front == 1 ? R(1) / m.ior : m.ior;
    R s2t = eta * eta * (R(1) - ci * ci);
    R F;
    bool tir = false;
    if (s2t > R(1)) {
        F = R(1);
        tir = true;
    } else {
        R c = clamp01m(msqrt(R(1) - s2t)), q = R(1) - c, q2 = q * q;
        F = f0 + (R(1) - f0) * (q2 * q2 * q);
    }
    R wrefl, wrefr = R(0);
    V3<R> local;
    if (m.transparency > R(0)) {
        R tr = m.transparency;
        local = color * (R(1) - tr);
        wrefl = tir ? tr : tr * F;
        if (!tir) wrefr = tr * (R(1) - F);
    } else {
        local = color * (R(1) - m.reflectivity);
        wrefl = m.reflectivity;
        if (it.depth == 0 && m.has_glow) local = local + limb_glow(m, N, V, sc.sky.sun_dir);
    }
    if (m.pbr && m.has_emission) local = local + m.emissive;
    *acc = *acc + vmul(it.w, local);

    // refraction child (its intersection starts at tmin 1e-4 from P + rd*1e-3)
    if (wrefr != R(0)) {
        R cosi = -dot(d, N);
        R s2 = eta * eta * (R(1) - cosi * cosi);
        V3<R> rd = zero3<R>();
        if (!(s2 > R(1))) {
            R ct = msqrt(R(1) - s2), k = eta * cosi - ct;
            rd = mk(eta * d.x + k * N.x, eta * d.y + k * N.y, eta * d.z + k * N.z);
        }
        if (*sp >= cap) return false;
        WItem<R>& c = stk[(*sp)++];
        c.o = P + rd * R(1e-3);
        c.d = rd;
        c.w = it.w * wrefr;
        c.depth = it.depth + 1;
        c.medium = (m.is_water || m.beer_lambert) ? h.mat : -2; // -2: refracted, no medium
    }
    if (it.depth < max_depth && wrefl != R(0)) {
        if (m.pbr != 0 && m.reflectivity > R(0) && m.roughness > R(1e-6)) {
            for (int i = 0; i < 16; ++i) {
                uint32_t k = seed * 0x9e3779b9u + (uint32_t)it.depth * 0x85ebca6bu + (uint32_t)i * 2u;
                R r1 = rand01<R>(k, (uint32_t)i, 0x9e37u);
                R r2 = rand01<R>(k, (uint32_t)i, 0x85ebu);
                if (*sp >= cap) return false;
                WItem<R>& c = stk[(*sp)++];
                c.o = P + N * R(1e-3);
                c.d = glossy_dir(m, N, d, r1, r2);
                c.w = it.w * (wrefl * (R(1) / R(16)));
                c.depth = it.depth + 1;
                c.medium = -1;
            }
        } else {
            if (*sp >= cap) return false;
            WItem<R>& c = stk[(*sp)++];
            c.o = P + N * R(1e-3);
            c.d = reflect(d, N);
            c.w = it.w * wrefl;
            c.depth = it.depth + 1;
            c.medium = -1;
        }
    }
    return true;
}

template <typename R, typename TS>
__device__ V3<R> whitted_sample(const SceneView<R>& sc, TS ts, V3<R> o, V3<R> d, int max_depth, uint32_t seed,
                                WItem<R>* stk, int cap, int* overflow, int* nq)
{
    V3<R> acc = zero3<R>();
    int sp = 0;
    WItem<R>& root = stk[sp++];
    root.o = o;
    root.d = d;
    root.w = mk(R(1), R(1), R(1));
    root.depth = 0;
    root.medium = -1;
    while (sp > 0) {
        WItem<R> it = stk[--sp];
        if (it.depth > max_depth && it.medium == -1) continue;
        DHit<R> h;
        ++*nq;
        bool hit = intersect(sc, ts, it.o, unit_dir(it.d), R(1e-4), R(1e30), &h);
        if (it.medium >= 0) {
            // Beer-Lambert blend over the in-medium segment (5 when it escapes)
            const DMat<R>& mm = sc.mats[it.medium];
            V3<R> T, deep;
            medium_terms(mm, hit ? h.t : R(5), &T, &deep);
            acc = acc + vmul(it.w, deep);
            it.w = vmul(it.w, T);
        }
        if (it.medium == -1 && sc.fog.density > R(0)) {
            // the root and reflected rays carry fog over their segment
            // (refracted rays are shaded without it)
            R T;
            V3<R> ins;
            fog_segment(sc.fog, sc.sky, it.o, it.d, hit ? h.t : R(1e30), &T, &ins, (V3<R>*)nullptr);
            acc = acc + vmul(it.w, ins);
            it.w = it.w * T;
        }
        if (!hit) {
            acc = acc + vmul(it.w, sky_sample(it.d, sc.sky));
            continue;
        }
        if (it.depth > max_depth) continue;
        if (!whitted_hit(sc, ts, it, h, max_depth, seed, &acc, stk, &sp, cap, nq)) {
            *overflow = 1;
            break;
        }
    }
    return acc;
}

// ---- per-pixel estimates (the megakernels' bodies) ----------------------------

// Path-traced pixel (x, y) of a w x h image: samples s0 .. s0 + spp - 1 in
// order, uniform jitter (channels 2s, 2s+1), thin-lens channels 2s+100 /
// 2s+101, the dome's start-under-water test, then the mean (reference
// src/pathtrace.c:1047-1096).
template <bool FOG, typename R, typename TS>
__device__ V3<R> pt_pixel(const SceneView<R>& sc, TS ts, const CamView<R>& cam, int x, int y, int w, int h, int spp,
                          int s0, int depth, int* nq)
{
    V3<R> acc = zero3<R>();
    for (int si = 0; si < spp; ++si) {
        uint32_t s = (uint32_t)(s0 + si);
        R jx = rand01<R>((uint32_t)x, (uint32_t)y, s * 2u + 0u);
        R jy = rand01<R>((uint32_t)x, (uint32_t)y, s * 2u + 1u);
        R l1 = rand01<R>((uint32_t)x, (uint32_t)y, s * 2u + 100u);
        R l2 = rand01<R>((uint32_t)x, (uint32_t)y, s * 2u + 101u);
        R uu = ((R)x + jx) / (R)w;
        R vv = R(1) - ((R)y + jy) / (R)h;
        V3<R> o, d;
        camera_ray(cam, uu, vv, l1, l2, &o, &d);
        if (cam.dome > R(1e-4) && sc.water_material >= 0) {
            V3<R> pd = o + d * cam.dome;
            if (pd.y < water_height(sc, pd.x, pd.z)) o = pd - mk(R(0), R(1e-3), R(0));
        }
        acc = acc + pt_radiance<FOG>(sc, ts, o, d, depth, hash3((uint32_t)x, (uint32_t)y, s), nq);
    }
    return acc * (R(1) / (R)spp);
}

// Whitted pixel (x, y): stratified g x g offsets when spp = g^2 (strat), else
// uniform; lens channels 2s+2 / 2s+3 (reference src/render.c:874-889).
template <typename R, typename TS>
__device__ V3<R> whitted_pixel(const SceneView<R>& sc, TS ts, const CamView<R>& cam, int x, int y, int w, int h,
                               int spp, int s0, int depth, int g, bool strat, WItem<R>* stk, int cap, int* ovf,
                               int* nq)
{
    V3<R> acc = zero3<R>();
    for (int si = 0; si < spp; ++si) {
        int s = s0 + si;
        R du, dv;
        R jx = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 0));
        R jy = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 1));
        if (strat) {
            du = ((R)(s % g) + jx) / (R)g;
            dv = ((R)(s / g) + jy) / (R)g;
        } else {
            du = jx;
            dv = jy;
        }
        R uu = ((R)x + du) / (R)w;
        R vv = R(1) - ((R)y + dv) / (R)h;
        R l1 = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 2));
        R l2 = rand01<R>((uint32_t)x, (uint32_t)y, (uint32_t)(s * 2 + 3));
        V3<R> o, d;
        camera_ray(cam, uu, vv, l1, l2, &o, &d);
        acc = acc + whitted_sample(sc, ts, o, d, depth, hash3((uint32_t)x, (uint32_t)y, (uint32_t)s), stk, cap, ovf, nq);
    }
    return acc * (R(1) / (R)spp);
}

} // namespace mh
