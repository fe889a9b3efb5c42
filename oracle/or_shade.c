/*
 * or_shade.c - oracle: local shading (Blinn-Phong, Cook-Torrance), glossy
 * direction, ambient, procedural sky (gradient, sun glow, 2D fBm clouds),
 * sun-disk and sphere-cap sampling, water normal / attenuation and the albedo
 * textures of the GPU path, restated in FP64. TEST INFRASTRUCTURE ONLY.
 *
 * Follows reference src/material.c:25-50 (helpers), 157-188
 * (shade_local), 218-290 (roughness_to_alpha, shade_pbr), 303-349 (glossy
 * dir), 351-374 (ambient), 473-601 + 950-957 (sky_sample with 2D clouds),
 * 1062-1146 (sky_basis, sun disk, sphere cap), 1161-1267 (water), and
 * reference src/texture.c:269-327 (albedo: none, checker, uv_checker).
 */
#include "or_internal.h"

/* material.c clamp: NaN passes through */
double or_clamp01(double v)
{
    if (v < 0.0) return 0.0;
    if (v > 1.0) return 1.0;
    return v;
}

static double smooth3(double e0, double e1, double x)
{
    double den = e1 - e0;
    if (den == 0.0) return x < e0 ? 0.0 : 1.0;
    double t = or_clamp01((x - e0) / den);
    return t * t * (3.0 - 2.0 * t);
}

static double powpos(double b, double e) { return b <= 0.0 ? 0.0 : pow(b, e); }

/* Blinn-Phong: albedo*L*NdotL + specular*L*(N.H)^shininess (no 1/pi). */
dv3 or_shade_local(const mh_material *m, dv3 N, dv3 L, dv3 V, dv3 light)
{
    double ndl = ddot(N, L);
    if (ndl <= 0.0) return d3(0.0, 0.0, 0.0);
    dv3 diff = dscale(dmul(fromm(m->albedo), light), ndl);
    dv3 hv = dadd(L, V);
    dv3 spec = d3(0.0, 0.0, 0.0);
    if (dlen2(hv) > 1e-12) {
        double ndh = ddot(N, dnorm(hv));
        if (ndh > 0.0) spec = dscale(dmul(fromm(m->specular), light), powpos(ndh, m->shininess));
    }
    return dadd(diff, spec);
}

double or_roughness_to_alpha(double r)
{
    double c = or_clamp01(r);
    double a = c * c;
    return a < 1e-4 ? 1e-4 : a;
}

static double smith_k(double ndl, double ndv, double rough)
{
    double k = (rough + 1.0) * (rough + 1.0) / 8.0;
    if (k < 0.0) k = 0.0;
    else if (k > 1.0) k = 1.0;
    return (ndl / (ndl * (1.0 - k) + k)) * (ndv / (ndv * (1.0 - k) + k));
}

/* Cook-Torrance: L*NdotL*(kD*albedo/pi + D*G*F/(4 NdotL NdotV)). */
dv3 or_shade_pbr(const mh_material *m, dv3 N, dv3 L, dv3 V, dv3 light)
{
    double ndl = ddot(N, L);
    if (ndl <= 0.0) return d3(0.0, 0.0, 0.0);
    double ndv = ddot(N, V);
    if (ndv <= 0.0) return d3(0.0, 0.0, 0.0);
    double met = or_clamp01(m->metallic);
    double alpha = or_roughness_to_alpha(m->roughness);
    dv3 hv = dadd(L, V);
    double ndh = 0.0;
    if (dlen2(hv) > 1e-12) {
        ndh = ddot(N, dnorm(hv));
        if (ndh < 0.0) ndh = 0.0;
    }
    dv3 alb = fromm(m->albedo);
    dv3 f0 = dlerp(d3(0.04, 0.04, 0.04), alb, met);
    /* material.c Schlick: clamp cos only */
    double c = or_clamp01(ndv), mm = 1.0 - c, m2 = mm * mm, k5 = m2 * m2 * mm;
    dv3 F = d3(f0.x + (1.0 - f0.x) * k5, f0.y + (1.0 - f0.y) * k5, f0.z + (1.0 - f0.z) * k5);
    double D = or_d_ggx(ndh, alpha);
    double G = smith_k(ndl, ndv, or_clamp01(m->roughness));
    double den = 4.0 * ndl * ndv;
    double ss = den > 0.0 ? D * G / den : 0.0;
    dv3 spec = dscale(F, ss);
    dv3 kd = dscale(dsub(d3(1.0, 1.0, 1.0), F), 1.0 - met);
    dv3 fd = dscale(dmul(kd, alb), 1.0 / OR_PI);
    return dscale(dmul(light, dadd(fd, spec)), ndl);
}

static void sky_frame(dv3 n, dv3 *t, dv3 *b)
{
    dv3 axis;
    if (fabs(n.x) <= fabs(n.y) && fabs(n.x) <= fabs(n.z)) axis = d3(1.0, 0.0, 0.0);
    else if (fabs(n.y) <= fabs(n.z)) axis = d3(0.0, 1.0, 0.0);
    else axis = d3(0.0, 0.0, 1.0);
    *t = dnorm(dcross(axis, n));
    *b = dcross(n, *t);
}

#define U_MAX 0.9999999999999999

dv3 or_sample_glossy_dir(const mh_material *m, dv3 N, dv3 incident, double r1, double r2)
{
    dv3 mirror = dreflect(incident, N);
    if (m->roughness <= 1e-6) return mirror;
    double a = or_roughness_to_alpha(m->roughness);
    double a2 = a * a;
    if (!(r1 > 0.0)) r1 = 0.0;
    else if (r1 > U_MAX) r1 = U_MAX;
    if (!(r2 > 0.0)) r2 = 0.0;
    else if (r2 > U_MAX) r2 = U_MAX;
    double ct = sqrt((1.0 - r1) / (1.0 + (a2 - 1.0) * r1));
    double st = sqrt(1.0 - ct * ct);
    double phi = 2.0 * OR_PI * r2;
    dv3 t, b;
    sky_frame(N, &t, &b);
    dv3 H = dnorm(dadd(dadd(dscale(t, st * cos(phi)), dscale(b, st * sin(phi))), dscale(N, ct)));
    dv3 R = dreflect(incident, H);
    if (!(ddot(R, N) > 0.0)) return mirror;
    return dnorm(R);
}

dv3 or_ambient(const mh_material *m, dv3 N, const mh_sky *sky)
{
    dv3 ground = dscale(fromm(sky->horizon_color), 0.45);
    dv3 skyc = fromm(sky->zenith_color);
    double w = or_clamp01(0.5 + 0.5 * N.y);
    dv3 hemi = dlerp(ground, skyc, w);
    return dscale(dmul(fromm(m->albedo), hemi), 0.15);
}

/* Cloud-layer opacity of the planar fBm layer along dir (0 below 0.03). */
double or_sky_cloud_alpha(dv3 dir, const mh_sky *sky)
{
    if (!(dir.y > 0.03)) return 0.0;
    double tt = sky->cloud_height / dir.y;
    double px = dir.x * tt, pz = dir.z * tt;
    int oct = sky->cloud_octaves < 1 ? 1 : sky->cloud_octaves;
    double raw = or_fbm2(px * sky->cloud_scale, pz * sky->cloud_scale, oct, 2.0, 0.5, sky->seed);
    double dens = or_clamp01(0.5 * (raw + 1.0));
    double a = sky->cloud_coverage - sky->cloud_softness;
    double b = sky->cloud_coverage + sky->cloud_softness;
    double al = powpos(smooth3(a, b, dens), sky->cloud_sharpness);
    al *= smooth3(0.0, 0.15, dir.y);
    return al;
}

dv3 or_sky_sample(dv3 dir, const mh_sky *sky)
{
    double h = or_clamp01(dir.y);
    double t = powpos(h, sky->gradient_gamma);
    dv3 base = dlerp(fromm(sky->horizon_color), fromm(sky->zenith_color), t);
    dv3 sun_dir = fromm(sky->sun_dir), sun_col = fromm(sky->sun_color);
    double cs = ddot(dir, sun_dir);
    if (cs > 0.0) {
        double glow = powpos(cs, sky->sun_glow_exponent) * sky->sun_glow_strength;
        base = dadd(base, dscale(sun_col, glow));
    }
    /* planar cloud layer at cloud_height (volumetric clouds are off) */
    if (dir.y > 0.03) {
        double al = or_sky_cloud_alpha(dir, sky);
        if (al > 0.0) {
            double shade = 0.6 + 0.4 * al;
            dv3 cc = dscale(d3(0.95, 0.95, 0.98), shade);
            double lit = cs > 0.0 ? cs : 0.0;
            dv3 lc = dadd(cc, dscale(dmul(cc, sun_col), 0.35 * lit));
            base = dlerp(base, lc, al);
        }
    }
    if (!(base.x >= 0.0)) base.x = 0.0;
    if (!(base.y >= 0.0)) base.y = 0.0;
    if (!(base.z >= 0.0)) base.z = 0.0;
    return base;
}

/* Uniform disk of angular radius radius_deg around sun, small-angle tan offset. */
dv3 or_sun_disk_dir(dv3 sun, double radius_deg, double r1, double r2)
{
    if (radius_deg <= 0.0) return sun;
    dv3 n = dnorm(sun);
    if (dlen2(n) == 0.0) return d3(0.0, 1.0, 0.0);
    if (r1 < 0.0) r1 = 0.0; else if (r1 > 1.0) r1 = 1.0;
    if (r2 < 0.0) r2 = 0.0; else if (r2 > 1.0) r2 = 1.0;
    double rr = radius_deg * (OR_PI / 180.0);
    double off = sqrt(r1) * rr;
    double th = 2.0 * OR_PI * r2;
    dv3 t, b;
    sky_frame(n, &t, &b);
    dv3 radial = dadd(dscale(t, cos(th)), dscale(b, sin(th)));
    return dnorm(dadd(n, dscale(radial, tan(off))));
}

/* Uniform-in-solid-angle direction inside the cone cos >= cos_max around w. */
dv3 or_light_cone_dir(dv3 w, double cos_max, double u1, double u2)
{
    w = dnorm(w);
    if (dlen2(w) == 0.0) return d3(0.0, 1.0, 0.0);
    if (!(cos_max >= 0.0)) cos_max = 0.0;
    else if (cos_max > 1.0) cos_max = 1.0;
    if (!(u1 > 0.0)) u1 = 0.0;
    else if (u1 > U_MAX) u1 = U_MAX;
    if (!(u2 > 0.0)) u2 = 0.0;
    else if (u2 > U_MAX) u2 = U_MAX;
    double ca = 1.0 - u1 * (1.0 - cos_max);
    double sa = sqrt(1.0 - ca * ca);
    double phi = 2.0 * OR_PI * u2;
    dv3 t, b;
    sky_frame(w, &t, &b);
    return dnorm(dadd(dadd(dscale(t, sa * cos(phi)), dscale(b, sa * sin(phi))), dscale(w, ca)));
}

/* Height-field normal of 4 directional sine waves + a 4-octave fBm ripple. */
dv3 or_water_normal(double x, double z, double time)
{
    static const double W[4][6] = {
        { 0.9584, 0.2855, 2.0 * OR_PI / 24.0, 0.1800, 1.0, 0.0 },
        { 0.7071, -0.7071, 2.0 * OR_PI / 13.0, 0.1000, 1.4, 1.7 },
        { -0.4061, 0.9138, 2.0 * OR_PI / 7.0, 0.0500, 2.0, 3.1 },
        { 0.2005, 0.9797, 2.0 * OR_PI / 4.0, 0.0250, 2.6, 0.6 },
    };
    double gx = 0.0, gz = 0.0;
    for (int i = 0; i < 4; ++i) {
        double ph = (W[i][0] * x + W[i][1] * z) * W[i][2] + W[i][5] + time * W[i][4] * W[i][2];
        double c = cos(ph) * W[i][3] * W[i][2];
        gx += c * W[i][0];
        gz += c * W[i][1];
    }
    const double f = 0.35, amp = 0.06, e = 0.02;
    double hl = or_fbm2((x - e) * f, z * f, 4, 2.0, 0.5, 1337u);
    double hr = or_fbm2((x + e) * f, z * f, 4, 2.0, 0.5, 1337u);
    double hd = or_fbm2(x * f, (z - e) * f, 4, 2.0, 0.5, 1337u);
    double hu = or_fbm2(x * f, (z + e) * f, 4, 2.0, 0.5, 1337u);
    gx += amp * (hr - hl) / (2.0 * e);
    gz += amp * (hu - hd) / (2.0 * e);
    dv3 n = dnorm(d3(-gx, 1.0, -gz));
    if (dlen2(n) < 1e-12) return d3(0.0, 1.0, 0.0);
    if (n.y < 0.0) n = dneg(n);
    return n;
}

static double transmit(double a, double d)
{
    if (a == 0.0) return 1.0;
    double t = exp(-a * d);
    if (!(t > 0.0)) return 0.0;
    if (t > 1.0) return 1.0;
    return t;
}

/* inner*T + deep*(1-T), T = exp(-absorption*depth) per channel. */
dv3 or_water_attenuate(const mh_material *m, dv3 inner, double depth)
{
    double d = depth > 0.0 ? depth : 0.0;
    dv3 tr = d3(transmit(m->absorption.x, d), transmit(m->absorption.y, d), transmit(m->absorption.z, d));
    dv3 o = d3(inner.x * tr.x + m->deep_color.x * (1.0 - tr.x),
               inner.y * tr.y + m->deep_color.y * (1.0 - tr.y),
               inner.z * tr.z + m->deep_color.z * (1.0 - tr.z));
    if (!(o.x >= 0.0)) o.x = 0.0;
    if (!(o.y >= 0.0)) o.y = 0.0;
    if (!(o.z >= 0.0)) o.z = 0.0;
    return o;
}

dv3 or_texture_albedo(const mh_material *m, dv3 p, double u, double v)
{
    double s = m->texture_scale > 0.0 ? m->texture_scale : 1.0;
    if (m->texture_kind == MH_TEXTURE_CHECKER) {
        double is = 1.0 / s;
        int px = (int)(((long long)floor(p.x * is)) & 1);
        int py = (int)(((long long)floor(p.y * is)) & 1);
        int pz = (int)(((long long)floor(p.z * is)) & 1);
        int par = (px + py + pz) & 1;
        return dmul(fromm(m->albedo), fromm(par ? m->texture_color_a : m->texture_color_b));
    }
    if (m->texture_kind == MH_TEXTURE_UV_CHECKER) {
        int pu = (int)(((long long)floor(u * s)) & 1);
        int pv = (int)(((long long)floor(v * s)) & 1);
        int par = (pu + pv) & 1;
        return dmul(fromm(m->albedo), fromm(par ? m->texture_color_a : m->texture_color_b));
    }
    return fromm(m->albedo);
}
