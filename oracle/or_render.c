/*
 * or_render.c - oracle: the unidirectional path tracer, the fixed-spp
 * Whitted renderer, camera rays, tone map, the pthread tile pool and the
 * public liboracle API. FP64 restatement; TEST INFRASTRUCTURE ONLY.
 *
 * Follows:
 *   camera                 reference src/camera.c:24-134
 *   PT lobe sampling       reference src/pathtrace.c:141-200, 220-233, 251-453
 *   PT NEE + shadow        reference src/pathtrace.c:504-606
 *   PT radiance            reference src/pathtrace.c:800-1028
 *   water height           reference src/scene.c:1109-1153 (Gerstner ocean)
 *   fog                    reference src/material.c:963-1039 (closed-form
 *                          segment), src/pathtrace.c:614-729 (shadow-marched
 *                          camera segment), 737-794 (marched underwater segment)
 *   PT pixel + pool        reference src/pathtrace.c:1047-1096, 1134-1234
 *   tone map               reference src/pathtrace.c:464-480, src/render.c:892-907
 *   Whitted                reference src/render.c:507-863 (trace/trace_hit,
 *                          emissive_direct), 874-889 (sample_offset),
 *                          924-970 (render_region), 842-863 (fog)
 *   adaptive Whitted       reference src/render.c:1128-1496 (per-pixel relative
 *                          standard error of the mean luminance, refinement
 *                          passes of n0 samples), 1627-1665 (defaults)
 */
#define _POSIX_C_SOURCE 200809L

#include "oracle.h"
#include "or_internal.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

/* ---- helpers --------------------------------------------------------- */

static double cl01n(double v) /* NaN -> 0 */
{
    if (!(v > 0.0)) return 0.0;
    if (v > 1.0) return 1.0;
    return v;
}

static double clamp_n(double v, double lo, double hi)
{
    if (!(v > lo)) return lo;
    if (v > hi) return hi;
    return v;
}

static double max_comp(dv3 v)
{
    double m = v.x;
    if (v.y > m) m = v.y;
    if (v.z > m) m = v.z;
    if (!(m > 0.0)) m = 0.0;
    return m;
}

static double luma(dv3 c)
{
    double y = 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z;
    return y > 0.0 ? y : 0.0;
}

static const mh_material *mat_of(const or_scene *s, int i)
{
    if (i < 0 || i >= s->n_mats) return NULL;
    return &s->mats[i];
}

unsigned char or_to_byte(double c)
{
    if (!(c >= 0.0)) c = 0.0;
    else if (c > 1.0) c = 1.0;
    int iv = (int)(pow(c, 1.0 / 2.2) * 255.0 + 0.5);
    if (iv < 0) iv = 0;
    else if (iv > 255) iv = 255;
    return (unsigned char)iv;
}

/* ---- camera ---------------------------------------------------------- */

void or_camera_create(const double eye[3], const double at[3], const double up[3],
                      double vfov, double aspect, mh_camera *out)
{
    dv3 from = d3(eye[0], eye[1], eye[2]), to = d3(at[0], at[1], at[2]);
    dv3 upv = d3(up[0], up[1], up[2]);
    dv3 fwd = dnorm(dsub(to, from));
    if (dlen2(fwd) <= 1e-12) fwd = d3(0.0, 0.0, -1.0);
    dv3 right = dcross(fwd, upv);
    if (dlen2(right) <= 1e-12) {
        dv3 alt = fabs(fwd.y) < 0.99 ? d3(0.0, 1.0, 0.0) : d3(1.0, 0.0, 0.0);
        right = dcross(fwd, alt);
    }
    right = dnorm(right);
    dv3 tup = dcross(right, fwd);
    double hh = tan(vfov * (OR_PI / 180.0) * 0.5);
    double hw = aspect * hh;
    dv3 hor = dscale(right, 2.0 * hw), ver = dscale(tup, 2.0 * hh);
    dv3 ll = dsub(dadd(from, fwd), dadd(dscale(hor, 0.5), dscale(ver, 0.5)));
    out->position = (mh_vec3){ from.x, from.y, from.z };
    out->horizontal = (mh_vec3){ hor.x, hor.y, hor.z };
    out->vertical = (mh_vec3){ ver.x, ver.y, ver.z };
    out->lower_left = (mh_vec3){ ll.x, ll.y, ll.z };
    out->aperture = 0.0;
    out->focus_distance = dlen(dsub(to, from));
    out->dome_radius = 0.0;
    out->anamorphic_squeeze = 1.0;
}

static oray cam_ray(const mh_camera *c, double u, double v, double r1, double r2)
{
    dv3 pos = fromm(c->position), hor = fromm(c->horizontal), ver = fromm(c->vertical);
    dv3 pt = dadd(fromm(c->lower_left), dadd(dscale(hor, u), dscale(ver, v)));
    oray pin = { pos, dnorm(dsub(pt, pos)) };
    if (c->aperture <= 1e-12) return pin;
    double fd = c->focus_distance;
    if (fd <= 0.0) fd = 1.0;
    dv3 focal = dadd(pos, dscale(pin.d, fd));
    dv3 rt = dnorm(hor), upv = dnorm(ver);
    double rad = c->aperture * sqrt(r1);
    double th = 2.0 * OR_PI * r2;
    double sq = c->anamorphic_squeeze;
    dv3 off;
    if (sq <= 0.0 || fabs(sq - 1.0) < 1e-12)
        off = dadd(dscale(rt, rad * cos(th)), dscale(upv, rad * sin(th)));
    else
        off = dadd(dscale(rt, (rad / sq) * cos(th)), dscale(upv, rad * sin(th)));
    oray r;
    r.o = dadd(pos, off);
    r.d = dnorm(dsub(focal, r.o));
    return r;
}

void or_camera_ray(const mh_camera *cam, double u, double v, double r1, double r2, double out[6])
{
    oray r = cam_ray(cam, u, v, r1, r2);
    out[0] = r.o.x; out[1] = r.o.y; out[2] = r.o.z;
    out[3] = r.d.x; out[4] = r.d.y; out[5] = r.d.z;
}

/* ---- path tracer ----------------------------------------------------- */

#define K_BOUNCE 0x5054u
#define CH_U1 0xB501u
#define CH_U2 0xB502u
#define CH_LOBE 0x1001u
#define CH_GLASS 0x1002u
#define CH_RR 0x5201u
#define CH_SUN_A 0xA101u
#define CH_SUN_B 0xA102u
#define CH_EM_A 0xA201u
#define CH_EM_B 0xA202u

static void apply_medium(const mh_material *m, double dist, dv3 *beta, dv3 *rad, dv3 *beta2)
{
    dv3 white = or_water_attenuate(m, d3(1.0, 1.0, 1.0), dist);
    dv3 deep = or_water_attenuate(m, d3(0.0, 0.0, 0.0), dist);
    dv3 T = dsub(white, deep);
    *rad = dadd(*rad, dmul(*beta, deep));
    *beta = dmul(*beta, T);
    if (beta2) *beta2 = dmul(*beta2, T); /* frozen-sampling throughput */
}

static double pt_smith(double ndl, double ndv, double rough)
{
    double k = cl01n((rough + 1.0) * (rough + 1.0) / 8.0);
    return (ndl / (ndl * (1.0 - k) + k)) * (ndv / (ndv * (1.0 - k) + k));
}

static int sample_pbr(const mh_material *m, dv3 alb, dv3 N, dv3 d, dv3 V, double u1, double u2,
                      double ul, dv3 *wi_o, dv3 *w_o, int *spec)
{
    *spec = 0;
    double met = cl01n(m->metallic), rough = cl01n(m->roughness);
    double alpha = or_roughness_to_alpha(rough);
    dv3 f0 = dlerp(d3(0.04, 0.04, 0.04), alb, met);
    double ps = or_specular_probability(met, rough, luma(f0));
    double ndv = ddot(N, V);
    if (!(ndv > 0.0)) return 0;
    if (ul < ps) {
        dv3 H = or_ggx_half(N, alpha, u1, u2);
        dv3 wi = dreflect(d, H);
        double ndl = ddot(N, wi), ndh = ddot(N, H), vdh = ddot(V, H);
        if (!(ndl > 0.0) || !(ndh > 0.0) || !(vdh > 0.0)) return 0;
        double D = or_d_ggx(ndh, alpha);
        double G = pt_smith(ndl, ndv, rough);
        dv3 F = or_fresnel_schlick_rgb(vdh, f0);
        double pdf = or_ggx_pdf(ndh, vdh, alpha);
        if (!(pdf > 0.0) || !(ps > 0.0)) return 0;
        *wi_o = wi;
        *w_o = dscale(F, D * G / (4.0 * ndv * pdf * ps));
        *spec = 1;
        return 1;
    }
    double den = 1.0 - ps;
    if (!(den > 0.0)) return 0;
    dv3 wi = or_cosine_hemisphere(N, u1, u2);
    if (!(ddot(N, wi) > 0.0)) return 0;
    dv3 F = or_fresnel_schlick_rgb(ndv, f0);
    dv3 kd = dscale(dsub(d3(1.0, 1.0, 1.0), F), 1.0 - met);
    *wi_o = wi;
    *w_o = dscale(dmul(kd, alb), 1.0 / den);
    return 1;
}

static int sample_legacy(const mh_material *m, dv3 alb, dv3 N, dv3 d, dv3 V, int front,
                         double u1, double u2, double ul, double ug,
                         dv3 *wi_o, dv3 *w_o, int *refr, int *spec)
{
    *refr = 0;
    *spec = 0;
    double rm = cl01n(m->reflectivity), rt = cl01n(m->transparency);
    double rd = 1.0 - rm - rt;
    if (rd < 0.0) rd = 0.0;
    double pm = rm, pt = rt, pd = 1.0 - pm - pt, norm = 1.0;
    if (pd < 0.0) {
        double s = pm + pt;
        if (s > 0.0) {
            pm /= s;
            pt /= s;
            if (s > 1.0) norm = s;
        }
        pd = 0.0;
    }
    if (ul < pm) {
        if (!(pm > 0.0)) return 0;
        double w = rm / (pm * norm);
        *wi_o = dreflect(d, N);
        *w_o = d3(w, w, w);
        *spec = 1;
        return 1;
    }
    if (ul < pm + pt) {
        if (!(pt > 0.0)) return 0;
        *spec = 1;
        double ior = m->ior > 0.0 ? m->ior : 1.0;
        double eta = front ? 1.0 / ior : ior;
        double ci = cl01n(ddot(N, V));
        double F = or_fresnel_dielectric(ci, eta);
        double lobe = rt / (pt * norm);
        if (F >= 1.0) {
            *wi_o = dreflect(d, N);
            *w_o = d3(lobe, lobe, lobe);
            return 1;
        }
        double pr = clamp_n(F, 0.05, 0.95);
        if (ug < pr) {
            double w = lobe * (F / pr);
            *wi_o = dreflect(d, N);
            *w_o = d3(w, w, w);
            return 1;
        }
        dv3 wi;
        if (!or_refract(d, N, eta, &wi)) {
            double w = lobe * (F / pr);
            *wi_o = dreflect(d, N);
            *w_o = d3(w, w, w);
            return 1;
        }
        double w = lobe * ((1.0 - F) / (1.0 - pr)) * eta * eta;
        *wi_o = wi;
        *w_o = d3(w, w, w);
        *refr = 1;
        return 1;
    }
    if (!(pd > 0.0)) return 0;
    dv3 wi = or_cosine_hemisphere(N, u1, u2);
    if (!(ddot(N, wi) > 0.0)) return 0;
    *wi_o = wi;
    *w_o = dscale(alb, rd / pd);
    return 1;
}

/* Shadow connection through up to 6 closest hits; water and near-clear
 * (transparency > 0.8) surfaces pass light with Beer-Lambert attenuation. */
int or_anyhit_shadow = 0;

static int occluded_segment(const or_scene *s, oray r, double tmin, double tmax, dv3 *att)
{
    *att = d3(1.0, 1.0, 1.0);
    if (or_anyhit_shadow && !s->has_transmissive) return or_occluded1(s, r, tmin, tmax);
    oray cur = r;
    double rem = tmax;
    for (int step = 0; step < 6; ++step) {
        ohit h;
        if (!or_intersect1(s, cur, tmin, rem, &h)) return 0;
        const mh_material *m = mat_of(s, h.material);
        if (!m) return 1;
        if (m->is_water || m->transparency > 0.8) {
            if (m->beer_lambert || m->is_water) {
                dv3 seg = d3(exp(-m->absorption.x * h.t), exp(-m->absorption.y * h.t),
                             exp(-m->absorption.z * h.t));
                *att = dmul(*att, seg);
            }
            cur.o = dadd(h.point, dscale(cur.d, 1e-3));
            rem -= h.t;
            if (rem <= 1e-3) return 0;
            tmin = 1e-3;
            continue;
        }
        return 1;
    }
    return 1;
}

/* ---- water height ----------------------------------------------------- */

/* Water surface height at (x, z): the base level plus four cosine harmonics
 * of the scene's Gerstner ocean when it has one, else the water level. */
static double water_height(const or_scene *s, double x, double z)
{
    const mh_ocean *o = &s->ocean;
    if (!o->enabled) return s->water_level;
    double y = o->center.y;
    double dir_x = o->direction.x, dir_z = o->direction.z;
    double dlen = sqrt(dir_x * dir_x + dir_z * dir_z);
    if (dlen > 1e-12) {
        dir_x /= dlen;
        dir_z /= dlen;
    } else {
        dir_x = 1.0;
        dir_z = 0.0;
    }
    double wl = o->wavelength <= 0.0 ? 12.0 : o->wavelength;
    double amp = o->amplitude;
    double chop_wl = o->chop_wavelength > 0.0 ? o->chop_wavelength : wl * 0.25;
    double chop_amp = o->chop > 0.0 ? o->chop : amp * 0.25;
    const double lambda[4] = { wl, wl * 0.52, chop_wl, chop_wl * 0.40 };
    const double A[4] = { amp, amp * 0.38, chop_amp, chop_amp * 0.32 };
    const double angle[4] = { 0.0, 0.628, -0.488, 1.134 };
    const double phi[4] = { 0.0, 1.7 + (double)(o->seed % 100) * 0.03, 3.1 + (double)((o->seed * 7) % 100) * 0.03,
                            0.8 + (double)((o->seed * 13) % 100) * 0.03 };
    for (int h = 0; h < 4; ++h) {
        double ca = cos(angle[h]), sa = sin(angle[h]);
        double dx = ca * dir_x - sa * dir_z, dz = sa * dir_x + ca * dir_z;
        double k = (2.0 * OR_PI) / lambda[h];
        y += A[h] * cos(k * (dx * x + dz * z) + phi[h]);
    }
    return y;
}

/* ---- fog ------------------------------------------------------------- */

#define CH_FOG 0xFE01u
#define CH_WATER 0xAC31u

/* Henyey-Greenstein lobe (1 - g^2) / (1 + g^2 - 2 g cos)^1.5 with the
 * denominator floored at 1e-4. */
static double hg_lobe(double g, double cosv)
{
    double den = 1.0 + g * g - 2.0 * g * cosv;
    if (den < 1e-4) den = 1e-4;
    return (1.0 - g * g) / (den * sqrt(den));
}

static double clamp_g(double g) { return g < -0.95 ? -0.95 : (g > 0.95 ? 0.95 : g); }

/* Turbulence multiplier of the density at p: 1 - a + 2 a clamp01((fbm3 + 1)/2)
 * (3 octaves, lacunarity 2, gain 0.5, seed 1337); 1 when noise is off. */
static double fog_turb(const mh_fog *f, dv3 p)
{
    if (!(f->noise_amount > 1e-4 && f->noise_scale > 1e-4)) return 1.0;
    double n = or_fbm3(p.x * f->noise_scale, p.y * f->noise_scale, p.z * f->noise_scale,
                       3, 2.0, 0.5, 1337u);
    double t = 0.5 * (n + 1.0);
    if (t < 0.0) t = 0.0;
    if (t > 1.0) t = 1.0;
    return (1.0 - f->noise_amount) + f->noise_amount * t * 2.0;
}

/* Closed-form segment of length dist (>= 1e20: to infinity) through
 * exponential height fog: transmittance exp(-tau) and in-scatter
 * (colour + sun * HG * strength) (1 - T), tau scaled by the turbulence at one
 * point of the segment. dsun (optional): d(in-scatter)/d(sun colour), per
 * channel. */
static void fog_segment(const mh_fog *f, const mh_sky *sky, oray r, double dist, double *T,
                        dv3 *ins, dv3 *dsun)
{
    *T = 1.0;
    *ins = d3(0.0, 0.0, 0.0);
    if (dsun) *dsun = d3(0.0, 0.0, 0.0);
    if (f->density <= 0.0 || dist <= 1e-7) return;
    double lam = f->height_falloff > 0.0 ? f->height_falloff : 0.0;
    double y0 = r.o.y - f->height, dy = r.d.y, tau;
    int to_sky = dist >= 1e20;
    if (lam <= 1e-6) {
        tau = to_sky ? 1e6 : f->density * dist;
    } else {
        double base = f->density * exp(-lam * y0);
        if (to_sky)
            tau = dy > 1e-5 ? base / (lam * dy) : 1e6;
        else if (fabs(dy) < 1e-5)
            tau = base * dist;
        else
            tau = base * ((1.0 - exp(-lam * dy * dist)) / (lam * dy));
    }
    if (tau < 0.0) tau = 0.0;
    double ev = to_sky ? 20.0 : (dist > 50.0 ? 50.0 : dist * 0.5);
    tau *= fog_turb(f, dat(r.o, r.d, ev));
    double tr = exp(-tau);
    if (tr < 0.0) tr = 0.0;
    if (tr > 1.0) tr = 1.0;
    dv3 col = fromm(f->color);
    double ph = 0.0;
    if (f->inscatter_strength > 0.0) {
        ph = hg_lobe(clamp_g(f->sun_anisotropy), ddot(r.d, fromm(sky->sun_dir)));
        col = dadd(col, dscale(fromm(sky->sun_color), ph * f->inscatter_strength));
    }
    *T = tr;
    *ins = dscale(col, 1.0 - tr);
    if (dsun) {
        double k = ph * f->inscatter_strength * (1.0 - tr);
        *dsun = d3(k, k, k);
    }
}

/* Camera-ray segment marched in shadow_steps (clamped to [4, 128]) jittered
 * steps, each lit by the sun unless a shadow ray (tmin 0.05, tmax 100)
 * is blocked; HG phase clamped at 6 and mixed 0.7 / 0.3 with isotropic. The
 * march covers at most 60 units (to infinity: up to the top of the fog
 * layer, at most 40); the rest of the segment is closed form. */
static void pt_fog_segment(const or_scene *s, oray r, double dist, unsigned seed, int b,
                           double *T, dv3 *ins, dv3 *dsun)
{
    const mh_fog *f = &s->fog;
    if (f->shadow_steps <= 0) {
        fog_segment(f, &s->sky, r, dist, T, ins, dsun);
        return;
    }
    int steps = f->shadow_steps < 4 ? 4 : (f->shadow_steps > 128 ? 128 : f->shadow_steps);
    int to_sky = dist >= 1e20;
    double lam = f->height_falloff, max_t;
    if (to_sky) {
        double top = lam > 1e-4 ? f->height + 3.0 / lam : 25.0;
        if (r.o.y >= top && r.d.y >= 0.0) {
            *T = 1.0;
            *ins = d3(0.0, 0.0, 0.0);
            if (dsun) *dsun = d3(0.0, 0.0, 0.0);
            return;
        }
        if (r.d.y > 0.02) {
            max_t = (top - r.o.y) / r.d.y;
            if (max_t > 40.0) max_t = 40.0;
        } else {
            max_t = 40.0;
        }
    } else {
        max_t = dist > 60.0 ? 60.0 : dist;
    }
    double dt = max_t / (double)steps;
    double ph = hg_lobe(clamp_g(f->sun_anisotropy), ddot(r.d, fromm(s->sky.sun_dir)));
    if (ph > 6.0) ph = 6.0;
    double sk = (0.70 * ph + 0.30) * f->inscatter_strength;
    dv3 glow = dscale(fromm(s->sky.sun_color), sk);
    dv3 sun = fromm(s->sky.sun_dir);
    double Ta = 1.0;
    dv3 acc = d3(0.0, 0.0, 0.0), dacc = d3(0.0, 0.0, 0.0);
    double jit = or_rand01(seed, (unsigned)b, CH_FOG);
    for (int i = 0; i < steps; ++i) {
        dv3 p = dat(r.o, r.d, ((double)i + jit) * dt);
        double rho = f->density;
        if (lam > 1e-6) rho *= exp(-lam * (p.y - f->height));
        rho *= fog_turb(f, p);
        if (rho < 0.0) rho = 0.0;
        double sT = exp(-(rho * dt));
        oray sr = { p, sun };
        dv3 att;
        int shadowed = occluded_segment(s, sr, 0.05, 100.0, &att);
        dv3 col = fromm(f->color);
        if (!shadowed) col = dadd(col, dmul(glow, att));
        acc = dadd(acc, dscale(dscale(col, 1.0 - sT), Ta));
        if (!shadowed) dacc = dadd(dacc, dscale(dscale(att, sk * (1.0 - sT)), Ta));
        Ta *= sT;
        if (Ta < 1e-4) break;
    }
    if (dist > max_t) {
        oray rr = { dat(r.o, r.d, max_t), r.d };
        double rT;
        dv3 rins, rd;
        fog_segment(f, &s->sky, rr, to_sky ? 1e30 : dist - max_t, &rT, &rins, &rd);
        acc = dadd(acc, dscale(rins, Ta));
        dacc = dadd(dacc, dscale(rd, Ta));
        Ta *= rT;
    }
    *T = Ta;
    *ins = acc;
    if (dsun) *dsun = dacc;
}

/* Segment inside a medium. Water with fog shadow_steps > 0 is marched in
 * shadow_steps (clamped to [4, 16]) jittered steps: per-step Beer-Lambert
 * transmittance exp(-absorption dt) and in-scatter (0.4 deep colour + sun
 * * HG(0.72, clamped at 4.5) * 0.18 where a shadow ray (tmin 0.02, tmax
 * 100) is clear) (1 - T); other media use apply_medium. beta2 (optional)
 * takes the same transmittance; dsun (optional) accumulates d(added
 * radiance)/d(sun colour). */
static void pt_water_segment(const or_scene *s, const mh_material *m, oray r, double dist,
                             unsigned seed, int b, dv3 *beta, dv3 *rad, dv3 *beta2, dv3 *dsun)
{
    if (!m->is_water || s->fog.shadow_steps <= 0 || dist <= 1e-4) {
        apply_medium(m, dist, beta, rad, beta2);
        return;
    }
    int steps = s->fog.shadow_steps < 4 ? 4 : (s->fog.shadow_steps > 16 ? 16 : s->fog.shadow_steps);
    double dt = dist / (double)steps;
    double ph = hg_lobe(0.72, ddot(r.d, fromm(s->sky.sun_dir)));
    if (ph > 4.5) ph = 4.5;
    double jit = or_rand01(seed, (unsigned)b, CH_WATER);
    dv3 sT = d3(exp(-m->absorption.x * dt), exp(-m->absorption.y * dt), exp(-m->absorption.z * dt));
    dv3 sun = fromm(s->sky.sun_dir), amb = dscale(fromm(m->deep_color), 0.40);
    for (int i = 0; i < steps; ++i) {
        dv3 p = dat(r.o, r.d, ((double)i + jit) * dt);
        oray sr = { p, sun };
        dv3 att;
        int shadowed = occluded_segment(s, sr, 0.02, 100.0, &att);
        dv3 sc = d3(0.0, 0.0, 0.0);
        if (!shadowed) sc = dscale(dmul(fromm(s->sky.sun_color), att), ph * 0.18);
        dv3 col = dadd(amb, sc);
        dv3 one_m = d3(1.0 - sT.x, 1.0 - sT.y, 1.0 - sT.z);
        *rad = dadd(*rad, dmul(*beta, dmul(col, one_m)));
        if (dsun && !shadowed) *dsun = dadd(*dsun, dmul(*beta, dmul(dscale(att, ph * 0.18), one_m)));
        *beta = dmul(*beta, sT);
        if (beta2) *beta2 = dmul(*beta2, sT);
    }
}

static dv3 shade(const mh_material *m, dv3 N, dv3 L, dv3 V, dv3 c)
{
    return m->pbr ? or_shade_pbr(m, N, L, V, c) : or_shade_local(m, N, L, V, c);
}

/* ---- path-replay adjoint (FP64) ---------------------------------------
 * Observer of one path's radiance terms: with adj == NULL the radiance code
 * below is the plain estimator. Gradient slots: per material {albedo rgb,
 * roughness, metallic, emissive rgb}, then sun colour rgb. */
typedef struct or_adj {
    dv3 Lt, delta, Lacc;  /* path radiance, dLoss/dL per channel, terms so far */
    double *g;
    int n_mats;
    int mat;              /* material of the vertex being shaded */
    dv3 texf;             /* d(shaded albedo)/d(material albedo) there */
} or_adj;

static double in01(double v) { return v > 0.0 && v < 1.0 ? 1.0 : 0.0; }

static double powpos_r(double b, double e) { return b <= 0.0 ? 0.0 : pow(b, e); }

static void gput(or_adj *a, int mat, int k, double v)
{
    if (v != 0.0 && v == v) a->g[(size_t)mat * 8 + k] += v;
}

static void gsun(or_adj *a, int c, double v)
{
    if (v != 0.0 && v == v) a->g[(size_t)a->n_mats * 8 + c] += v;
}

/* d ln D / d alpha and d ln G1 / d k (GGX, Smith-Schlick). */
static double dlnD_da(double ndh, double a)
{
    double d = ndh * ndh * (a * a - 1.0) + 1.0;
    return 2.0 / a - 4.0 * a * ndh * ndh / d;
}

static double dlnG1_dk(double x, double k) { return -(1.0 - x) / (x * (1.0 - k) + k); }

static double da_dr(double r) /* alpha = max(clamp(r)^2, 1e-4) */
{
    if (!(r > 0.0) || r >= 1.0) return 0.0;
    return r * r > 1e-4 ? 2.0 * r : 0.0;
}

/* d shade(m, N, L, V, light)/d(shaded albedo, roughness, metallic). */
static void shade_grad(const mh_material *m, dv3 N, dv3 L, dv3 V, dv3 light, dv3 *ga, dv3 *gr, dv3 *gm)
{
    *ga = *gr = *gm = d3(0.0, 0.0, 0.0);
    double ndl = ddot(N, L);
    if (ndl <= 0.0) return;
    if (!m->pbr) {
        *ga = dscale(light, ndl);
        return;
    }
    double ndv = ddot(N, V);
    if (ndv <= 0.0) return;
    double met = or_clamp01(m->metallic), rr = or_clamp01(m->roughness);
    double alpha = or_roughness_to_alpha(m->roughness);
    dv3 hv = dadd(L, V);
    double ndh = 0.0;
    if (dlen2(hv) > 1e-12) {
        ndh = ddot(N, dnorm(hv));
        if (ndh < 0.0) ndh = 0.0;
    }
    double c = or_clamp01(ndv), q = 1.0 - c, k5 = q * q * q * q * q;
    double D = or_d_ggx(ndh, alpha);
    double k = (rr + 1.0) * (rr + 1.0) / 8.0;
    double G = (ndl / (ndl * (1.0 - k) + k)) * (ndv / (ndv * (1.0 - k) + k));
    double ss = D * G / (4.0 * ndl * ndv);
    double dls = dlnD_da(ndh, alpha) * da_dr(rr) + (dlnG1_dk(ndl, k) + dlnG1_dk(ndv, k)) * (rr + 1.0) / 4.0 *
                 in01(m->roughness);
    double dmt = in01(m->metallic);
    double a[3] = { m->albedo.x, m->albedo.y, m->albedo.z }, lc[3] = { light.x, light.y, light.z };
    double oa[3], orr[3], om[3];
    for (int i = 0; i < 3; ++i) {
        double f0 = 0.04 + (a[i] - 0.04) * met;
        double F = f0 + (1.0 - f0) * k5;
        double dFa = met * (1.0 - k5), dFm = (a[i] - 0.04) * (1.0 - k5) * dmt;
        oa[i] = lc[i] * ndl * (dFa * ss + (1.0 - met) / OR_PI * ((1.0 - F) - a[i] * dFa));
        om[i] = lc[i] * ndl * (dFm * ss + a[i] / OR_PI * (-dFm * (1.0 - met) - (1.0 - F) * dmt));
        orr[i] = lc[i] * ndl * F * ss * dls;
    }
    *ga = d3(oa[0], oa[1], oa[2]);
    *gr = d3(orr[0], orr[1], orr[2]);
    *gm = d3(om[0], om[1], om[2]);
}

/* A NEE term beta * shade(mm, light_col * att) * omega: its light source
 * (emissive of light material lmat, or the sun when lmat < 0) and the shaded
 * surface's parameters. */
static void adj_nee(or_adj *a, const mh_material *mm, dv3 N, dv3 L, dv3 V, dv3 beta, dv3 light_col, dv3 att,
                    double omega, int lmat, dv3 t)
{
    dv3 K = shade(mm, N, L, V, att);
    dv3 w = dscale(dmul(a->delta, beta), omega);
    double kv[3] = { w.x * K.x, w.y * K.y, w.z * K.z };
    for (int c = 0; c < 3; ++c) {
        if (lmat >= 0) gput(a, lmat, 5 + c, kv[c]);
        else gsun(a, c, kv[c]);
    }
    dv3 ga, gr, gm;
    shade_grad(mm, N, L, V, dmul(light_col, att), &ga, &gr, &gm);
    ga = dmul(ga, a->texf);
    gput(a, a->mat, 0, w.x * ga.x);
    gput(a, a->mat, 1, w.y * ga.y);
    gput(a, a->mat, 2, w.z * ga.z);
    gput(a, a->mat, 3, w.x * gr.x + w.y * gr.y + w.z * gr.z);
    gput(a, a->mat, 4, w.x * gm.x + w.y * gm.y + w.z * gm.z);
    a->Lacc = dadd(a->Lacc, t);
}

/* d(sky)/d(sun colour) per channel. */
static dv3 sky_dsun(dv3 dir, const mh_sky *sky)
{
    double h = or_clamp01(dir.y);
    double t = powpos_r(h, sky->gradient_gamma);
    dv3 base = dlerp(fromm(sky->horizon_color), fromm(sky->zenith_color), t);
    double cs = ddot(dir, fromm(sky->sun_dir));
    double gl = 0.0;
    if (cs > 0.0) {
        gl = powpos_r(cs, sky->sun_glow_exponent) * sky->sun_glow_strength;
        base = dadd(base, dscale(fromm(sky->sun_color), gl));
    }
    dv3 dv = d3(gl, gl, gl);
    double al = or_sky_cloud_alpha(dir, sky);
    if (al > 0.0) {
        dv3 cc = dscale(d3(0.95, 0.95, 0.98), 0.6 + 0.4 * al);
        double lit = cs > 0.0 ? cs : 0.0;
        dv3 lc = dadd(cc, dscale(dmul(cc, fromm(sky->sun_color)), 0.35 * lit));
        base = dlerp(base, lc, al);
        dv = dadd(dscale(dv, 1.0 - al), dscale(cc, al * 0.35 * lit));
    }
    if (!(base.x >= 0.0)) dv.x = 0.0;
    if (!(base.y >= 0.0)) dv.y = 0.0;
    if (!(base.z >= 0.0)) dv.z = 0.0;
    return dv;
}

static dv3 nee_sun(const or_scene *s, const mh_material *mm, dv3 P, dv3 N, dv3 V, dv3 beta,
                   unsigned seed, int b, or_adj *adj)
{
    double r1 = or_rand01(seed, (unsigned)b, CH_SUN_A);
    double r2 = or_rand01(seed, (unsigned)b, CH_SUN_B);
    dv3 L = or_sun_disk_dir(fromm(s->sky.sun_dir), s->sky.sun_radius, r1, r2);
    if (!dfinite(L) || ddot(N, L) <= 0.0) return d3(0.0, 0.0, 0.0);
    oray sr = { dadd(P, dscale(N, 1e-3)), L };
    dv3 att;
    if (occluded_segment(s, sr, 1e-3, 1e30, &att)) return d3(0.0, 0.0, 0.0);
    if (att.x < 1e-4 && att.y < 1e-4 && att.z < 1e-4) return d3(0.0, 0.0, 0.0);
    double rad = s->sky.sun_radius * OR_PI / 180.0;
    double omega = s->sky.sun_radius > 0.0 ? 2.0 * OR_PI * (1.0 - cos(rad)) : 1.0;
    dv3 lit = shade(mm, N, L, V, dmul(fromm(s->sky.sun_color), att));
    dv3 t = dmul(beta, dscale(lit, omega));
    if (adj) adj_nee(adj, mm, N, L, V, beta, fromm(s->sky.sun_color), att, omega, -1, t);
    return t;
}

static dv3 nee_emissive(const or_scene *s, const mh_material *mm, dv3 P, dv3 N, dv3 V,
                        dv3 beta, int self, unsigned seed, int b, or_adj *adj)
{
    dv3 sum = d3(0.0, 0.0, 0.0);
    for (int li = 0; li < s->n_lights; ++li) {
        const olight *lt = &s->lights[li];
        if (lt->prim == self) continue;
        dv3 toc = dsub(lt->center, P);
        double dc2 = dlen2(toc), dc = sqrt(dc2);
        if (dc <= lt->radius) continue;
        double cmx = sqrt(1.0 - (lt->radius * lt->radius) / dc2);
        double omega = 2.0 * OR_PI * (1.0 - cmx);
        if (!(omega > 0.0)) continue;
        dv3 w = dscale(toc, 1.0 / dc);
        double u1 = or_rand01(seed ^ (unsigned)li, (unsigned)b, CH_EM_A);
        double u2 = or_rand01(seed ^ (unsigned)li, (unsigned)b, CH_EM_B);
        dv3 wi = or_light_cone_dir(w, cmx, u1, u2);
        if (!dfinite(wi) || ddot(N, wi) <= 0.0) continue;
        oray sr = { dadd(P, dscale(N, 1e-3)), wi };
        ohit lh;
        oray srn = sr;
        double l2 = dlen2(srn.d);
        if (l2 > 0.0 && fabs(l2 - 1.0) > 1e-12) srn.d = dnorm(srn.d);
        if (!or_prim_intersect(&s->prims[lt->prim], srn, 1e-3, 1e30, &lh)) continue;
        dv3 att;
        if (occluded_segment(s, sr, 1e-3, lh.t - 1e-3, &att)) continue;
        dv3 lit = shade(mm, N, wi, V, dmul(lt->emissive, att));
        dv3 t = dmul(beta, dscale(lit, omega));
        if (adj) adj_nee(adj, mm, N, wi, V, beta, lt->emissive, att, omega, lt->mat, t);
        sum = dadd(sum, t);
    }
    return sum;
}

static dv3 limb_glow(const mh_material *mm, dv3 N, dv3 V, dv3 sun)
{
    double cv = fmax(0.0, ddot(N, V));
    double limb = pow(1.0 - cv, 3.5);
    double sd = ddot(N, sun);
    double day = 0.0;
    if (sd > -0.15) {
        day = (sd + 0.15) / 0.35;
        if (day > 1.0) day = 1.0;
    }
    dv3 g = fromm(mm->atmosphere_glow);
    if (sd > -0.10 && sd < 0.25) {
        double st = 1.0 - fabs(sd - 0.05) / 0.20;
        if (st > 0.0) g = dlerp(g, d3(1.0, 0.45, 0.15), st * 0.6);
    }
    return dscale(g, limb * day * 2.5);
}

/* PBR lobe sampling with every decision (lobe, half vector, pdf, lobe
 * probability) taken from the frozen material f / albedo falb, and the BSDF
 * value from m / alb: the weight is f_m / (p_f), a smooth function of m. */
static int sample_pbr_frozen(const mh_material *m, dv3 alb, const mh_material *f, dv3 falb, dv3 N, dv3 d, dv3 V,
                             double u1, double u2, double ul, dv3 *wi_o, dv3 *w_o, int *spec)
{
    *spec = 0;
    double met = cl01n(m->metallic), rough = cl01n(m->roughness), alpha = or_roughness_to_alpha(rough);
    double fmet = cl01n(f->metallic), frough = cl01n(f->roughness), falpha = or_roughness_to_alpha(frough);
    dv3 f0 = dlerp(d3(0.04, 0.04, 0.04), alb, met);
    dv3 ff0 = dlerp(d3(0.04, 0.04, 0.04), falb, fmet);
    double ps = or_specular_probability(fmet, frough, luma(ff0));
    double ndv = ddot(N, V);
    if (!(ndv > 0.0)) return 0;
    if (ul < ps) {
        dv3 H = or_ggx_half(N, falpha, u1, u2);
        dv3 wi = dreflect(d, H);
        double ndl = ddot(N, wi), ndh = ddot(N, H), vdh = ddot(V, H);
        if (!(ndl > 0.0) || !(ndh > 0.0) || !(vdh > 0.0)) return 0;
        double D = or_d_ggx(ndh, alpha);
        double G = pt_smith(ndl, ndv, rough);
        dv3 F = or_fresnel_schlick_rgb(vdh, f0);
        double pdf = or_ggx_pdf(ndh, vdh, falpha);
        if (!(pdf > 0.0) || !(ps > 0.0)) return 0;
        *wi_o = wi;
        *w_o = dscale(F, D * G / (4.0 * ndv * pdf * ps));
        *spec = 1;
        return 1;
    }
    double den = 1.0 - ps;
    if (!(den > 0.0)) return 0;
    dv3 wi = or_cosine_hemisphere(N, u1, u2);
    if (!(ddot(N, wi) > 0.0)) return 0;
    dv3 F = or_fresnel_schlick_rgb(ndv, f0);
    dv3 kd = dscale(dsub(d3(1.0, 1.0, 1.0), F), 1.0 - met);
    *wi_o = wi;
    *w_o = dscale(dmul(kd, alb), 1.0 / den);
    return 1;
}

/* The sampled lobe's weight w multiplies all later terms: gradient
 * delta . L_after * (dw/dtheta) / w per channel (sampling held fixed). */
static void adj_bsdf(or_adj *a, const mh_material *mm, dv3 alb, dv3 N, dv3 V, dv3 wi, dv3 w, int spec, int refr)
{
    dv3 la = dsub(a->Lt, a->Lacc);
    double sv[3] = { a->delta.x * la.x, a->delta.y * la.y, a->delta.z * la.z };
    double wv[3] = { w.x, w.y, w.z }, av[3] = { alb.x, alb.y, alb.z };
    double tf[3] = { a->texf.x, a->texf.y, a->texf.z };
    if (!mm->pbr) {
        if (spec || refr) return;
        for (int i = 0; i < 3; ++i)
            if (wv[i] != 0.0 && av[i] != 0.0) gput(a, a->mat, i, sv[i] / av[i] * tf[i]);
        return;
    }
    double met = cl01n(mm->metallic), rough = cl01n(mm->roughness), dmt = in01(mm->metallic);
    double ndv = ddot(N, V);
    double gmet = 0.0, grough = 0.0;
    if (spec) {
        dv3 H = dnorm(dadd(V, wi));
        double vdh = ddot(V, H), ndh = ddot(N, H), ndl = ddot(N, wi);
        double c = cl01n(vdh), q = 1.0 - c, k5 = q * q * q * q * q;
        double alpha = or_roughness_to_alpha(rough);
        double k = cl01n((rough + 1.0) * (rough + 1.0) / 8.0);
        double dr = dlnD_da(ndh, alpha) * da_dr(rough) + (dlnG1_dk(ndl, k) + dlnG1_dk(ndv, k)) * (rough + 1.0) / 4.0;
        for (int i = 0; i < 3; ++i) {
            if (wv[i] == 0.0) continue;
            double f0 = 0.04 + (av[i] - 0.04) * met, fin = in01(f0);
            double f = cl01n(f0), F = f + (1.0 - f) * k5;
            gput(a, a->mat, i, sv[i] * met * (1.0 - k5) * fin / F * tf[i]);
            gmet += sv[i] * (av[i] - 0.04) * (1.0 - k5) * fin * dmt / F;
            grough += sv[i] * dr;
        }
        gput(a, a->mat, 4, gmet);
        gput(a, a->mat, 3, grough * in01(mm->roughness));
    } else {
        double c = cl01n(ndv), q = 1.0 - c, k5 = q * q * q * q * q;
        for (int i = 0; i < 3; ++i) {
            if (wv[i] == 0.0) continue;
            double f0 = 0.04 + (av[i] - 0.04) * met, fin = in01(f0);
            double f = cl01n(f0), F = f + (1.0 - f) * k5;
            if (av[i] != 0.0) gput(a, a->mat, i, sv[i] * (1.0 / av[i] - met * (1.0 - k5) * fin / (1.0 - F)) * tf[i]);
            gmet += sv[i] * (-(av[i] - 0.04) * (1.0 - k5) * fin * dmt / (1.0 - F) - dmt / (1.0 - met));
        }
        gput(a, a->mat, 4, gmet);
    }
}

static dv3 pt_radiance(const or_scene *s, oray primary, int max_depth, unsigned seed, or_adj *adj)
{
    dv3 L = d3(0.0, 0.0, 0.0);
    if (max_depth < 0) max_depth = 0;
    oray r = { primary.o, dnorm(primary.d) };
    if (dlen2(r.d) <= 0.0) return L;
    dv3 beta = d3(1.0, 1.0, 1.0), beta_f = beta;
    const mh_material *medium = NULL;
    if (s->water_material >= 0 && primary.o.y < water_height(s, primary.o.x, primary.o.z))
        medium = mat_of(s, s->water_material);
    int last_spec = 1;
    for (int b = 0; b <= max_depth; ++b) {
        ohit h;
        or_tls_stats.ext_queries++;
        if (!or_intersect1(s, r, 1e-4, 1e30, &h)) {
            if (medium) {
                dv3 L0 = L, dw = d3(0.0, 0.0, 0.0);
                pt_water_segment(s, medium, r, 5.0, seed, b, &beta, &L, &beta_f, adj ? &dw : NULL);
                if (adj) {
                    adj->Lacc = dadd(adj->Lacc, dsub(L, L0));
                    dv3 ds = dmul(adj->delta, dw);
                    gsun(adj, 0, ds.x);
                    gsun(adj, 1, ds.y);
                    gsun(adj, 2, ds.z);
                }
                medium = NULL;
            }
            dv3 sky = or_sky_sample(r.d, &s->sky);
            dv3 dsk = adj ? sky_dsun(r.d, &s->sky) : d3(0.0, 0.0, 0.0);
            if (s->fog.density > 0.0) {
                double T;
                dv3 ins, dfs;
                if (b == 0) pt_fog_segment(s, r, 1e30, seed, b, &T, &ins, &dfs);
                else fog_segment(&s->fog, &s->sky, r, 1e30, &T, &ins, &dfs);
                sky = dadd(dscale(sky, T), ins);
                dsk = dadd(dscale(dsk, T), dfs);
            }
            dv3 t = dmul(beta, sky);
            if (adj) {
                dv3 ds = dmul(dmul(adj->delta, beta), dsk);
                gsun(adj, 0, ds.x);
                gsun(adj, 1, ds.y);
                gsun(adj, 2, ds.z);
                adj->Lacc = dadd(adj->Lacc, t);
            }
            L = dadd(L, t);
            break;
        }
        if (medium) {
            dv3 L0 = L, dw = d3(0.0, 0.0, 0.0);
            pt_water_segment(s, medium, r, h.t, seed, b, &beta, &L, &beta_f, adj ? &dw : NULL);
            if (adj) {
                adj->Lacc = dadd(adj->Lacc, dsub(L, L0));
                dv3 ds = dmul(adj->delta, dw);
                gsun(adj, 0, ds.x);
                gsun(adj, 1, ds.y);
                gsun(adj, 2, ds.z);
            }
        } else if (s->fog.density > 0.0) {
            double T;
            dv3 ins, dfs;
            if (b == 0) pt_fog_segment(s, r, h.t, seed, b, &T, &ins, &dfs);
            else fog_segment(&s->fog, &s->sky, r, h.t, &T, &ins, &dfs);
            dv3 t = dmul(beta, ins);
            L = dadd(L, t);
            if (adj) {
                dv3 ds = dmul(dmul(adj->delta, beta), dfs);
                gsun(adj, 0, ds.x);
                gsun(adj, 1, ds.y);
                gsun(adj, 2, ds.z);
                adj->Lacc = dadd(adj->Lacc, t);
            }
            beta = dscale(beta, T);
            beta_f = dscale(beta_f, T);
        }
        const mh_material *m = mat_of(s, h.material);
        if (!m) {
            dv3 t = dmul(beta, or_sky_sample(r.d, &s->sky));
            if (adj) {
                dv3 ds = dmul(dmul(adj->delta, beta), sky_dsun(r.d, &s->sky));
                gsun(adj, 0, ds.x);
                gsun(adj, 1, ds.y);
                gsun(adj, 2, ds.z);
                adj->Lacc = dadd(adj->Lacc, t);
            }
            L = dadd(L, t);
            break;
        }
        dv3 P = h.point, d = r.d, N = h.normal;
        if (m->is_water && fabs(h.normal.y) > 0.999 && fabs(h.normal.x) < 1e-4 &&
            fabs(h.normal.z) < 1e-4)
            N = or_water_normal(P.x, P.z, 0.0);
        if (ddot(N, d) > 0.0) N = dneg(N);
        N = dnorm(N);
        if (dlen2(N) <= 0.0) break;
        dv3 V = dneg(d);
        mh_material ml = *m;
        dv3 alb = or_texture_albedo(m, P, h.u, h.v);
        ml.albedo = (mh_vec3){ alb.x, alb.y, alb.z };
        if (adj) {
            adj->mat = h.material;
            adj->texf = d3(m->texture_kind == 0 ? 1.0 : (m->albedo.x != 0.0 ? alb.x / m->albedo.x : 0.0),
                           m->texture_kind == 0 ? 1.0 : (m->albedo.y != 0.0 ? alb.y / m->albedo.y : 0.0),
                           m->texture_kind == 0 ? 1.0 : (m->albedo.z != 0.0 ? alb.z / m->albedo.z : 0.0));
        }
        if (b == 0 || last_spec) {
            dv3 t = dmul(beta, fromm(ml.emissive));
            L = dadd(L, t);
            if (adj) {
                gput(adj, h.material, 5, adj->delta.x * beta.x);
                gput(adj, h.material, 6, adj->delta.y * beta.y);
                gput(adj, h.material, 7, adj->delta.z * beta.z);
                adj->Lacc = dadd(adj->Lacc, t);
            }
        }
        if (b == 0 && dlen2(fromm(ml.atmosphere_glow)) > 1e-6) {
            dv3 t = dmul(beta, limb_glow(&ml, N, V, fromm(s->sky.sun_dir)));
            L = dadd(L, t);
            if (adj) adj->Lacc = dadd(adj->Lacc, t);
        }
        L = dadd(L, nee_emissive(s, &ml, P, N, V, beta, h.prim, seed, b, adj));
        if (b == 0 || (b == 1 && medium != NULL && last_spec))
            L = dadd(L, nee_sun(s, &ml, P, N, V, beta, seed, b, adj));
        unsigned kb = or_hash3(seed ^ K_BOUNCE, (unsigned)b, 0u);
        double u1 = or_rand01(kb, (unsigned)b, CH_U1);
        double u2 = or_rand01(kb, (unsigned)b, CH_U2);
        double ul = or_rand01(kb, (unsigned)b, CH_LOBE);
        double ug = or_rand01(kb, (unsigned)b, CH_GLASS);
        dv3 wi, w;
        int ok, refr = 0, spec = 0;
        if (s->freeze) {
            /* frozen sampling: decisions and the roulette's throughput from
             * the frozen materials, BSDF values from the current ones */
            const mh_material *fm = &s->freeze[h.material];
            dv3 falb = or_texture_albedo(fm, P, h.u, h.v);
            mh_material fl = *fm;
            fl.albedo = (mh_vec3){ falb.x, falb.y, falb.z };
            dv3 w0;
            if (ml.pbr) {
                ok = sample_pbr_frozen(&ml, alb, &fl, falb, N, d, V, u1, u2, ul, &wi, &w, &spec);
                if (ok) ok = sample_pbr(&fl, falb, N, d, V, u1, u2, ul, &wi, &w0, &spec);
            } else {
                ok = sample_legacy(&ml, alb, N, d, V, h.front, u1, u2, ul, ug, &wi, &w, &refr, &spec);
                if (ok) ok = sample_legacy(&fl, falb, N, d, V, h.front, u1, u2, ul, ug, &wi, &w0, &refr, &spec);
            }
            if (!ok) break;
            last_spec = spec;
            beta = dmul(beta, w);
            beta_f = dmul(beta_f, w0);
            if (!dfinite(beta_f) || !(max_comp(beta_f) > 0.0)) break;
            if (refr && (m->is_water || m->beer_lambert)) medium = h.front ? m : NULL;
            if (b >= 2) {
                double q = cl01n(max_comp(beta_f));
                if (!(q > 0.0)) break;
                if (or_rand01(kb, (unsigned)b, CH_RR) >= q) break;
                beta = dscale(beta, 1.0 / q);
                beta_f = dscale(beta_f, 1.0 / q);
            }
            dv3 off = refr ? dscale(wi, 1e-4) : dscale(N, ddot(wi, N) >= 0.0 ? 1e-4 : -1e-4);
            r.o = dadd(P, off);
            r.d = wi;
            continue;
        }
        if (ml.pbr) ok = sample_pbr(&ml, alb, N, d, V, u1, u2, ul, &wi, &w, &spec);
        else ok = sample_legacy(&ml, alb, N, d, V, h.front, u1, u2, ul, ug, &wi, &w, &refr, &spec);
        if (!ok) break;
        if (adj) adj_bsdf(adj, &ml, alb, N, V, wi, w, spec, refr);
        last_spec = spec;
        beta = dmul(beta, w);
        if (!dfinite(beta) || !(max_comp(beta) > 0.0)) break;
        if (refr && (m->is_water || m->beer_lambert)) medium = h.front ? m : NULL;
        if (b >= 2) {
            double q = cl01n(max_comp(beta));
            if (!(q > 0.0)) break;
            if (or_rand01(kb, (unsigned)b, CH_RR) >= q) break;
            beta = dscale(beta, 1.0 / q);
        }
        dv3 off = refr ? dscale(wi, 1e-4) : dscale(N, ddot(wi, N) >= 0.0 ? 1e-4 : -1e-4);
        r.o = dadd(P, off);
        r.d = wi;
    }
    if (!dfinite(L)) return d3(0.0, 0.0, 0.0);
    if (L.x < 0.0) L.x = 0.0;
    if (L.y < 0.0) L.y = 0.0;
    if (L.z < 0.0) L.z = 0.0;
    return L;
}

/* ---- Whitted --------------------------------------------------------- */

static dv3 w_trace(const or_scene *s, oray r, int depth, int maxd, unsigned seed);

static dv3 emissive_direct(const or_scene *s, const mh_material *mm, dv3 P, dv3 N, dv3 V,
                           int self, unsigned seed)
{
    dv3 sum = d3(0.0, 0.0, 0.0);
    dv3 so = dadd(P, dscale(N, 1e-3));
    for (int li = 0; li < s->n_lights; ++li) {
        const olight *lt = &s->lights[li];
        if (lt->prim == self) continue;
        dv3 toc = dsub(lt->center, P);
        double dc2 = dlen2(toc), dc = sqrt(dc2);
        if (dc <= lt->radius) continue;
        double cmx = sqrt(1.0 - (lt->radius * lt->radius) / dc2);
        double solid = 2.0 * OR_PI * (1.0 - cmx);
        if (!(solid > 0.0)) continue;
        dv3 w = dscale(toc, 1.0 / dc);
        for (int si = 0; si < 16; ++si) {
            unsigned k = or_hash3(seed ^ 0xE311u, (unsigned)li, (unsigned)si);
            double u1 = or_rand01(k, (unsigned)si, 0x4c11u);
            double u2 = or_rand01(k, (unsigned)si, 0x9d27u);
            dv3 wi = or_light_cone_dir(w, cmx, u1, u2);
            if (ddot(N, wi) <= 0.0) continue;
            oray sr = { so, wi };
            oray srn = sr;
            double l2 = dlen2(srn.d);
            if (l2 > 0.0 && fabs(l2 - 1.0) > 1e-12) srn.d = dnorm(srn.d);
            ohit lh;
            int vis = 0;
            if (or_prim_intersect(&s->prims[lt->prim], srn, 1e-3, 1e30, &lh))
                vis = !or_occluded1(s, sr, 1e-3, lh.t - 1e-3);
            if (!vis) continue;
            sum = dadd(sum, shade(mm, N, wi, V, dscale(lt->emissive, solid)));
        }
    }
    return dscale(sum, 1.0 / 16.0);
}

static dv3 w_trace_hit(const or_scene *s, oray r, int depth, int maxd, const ohit *h, unsigned seed)
{
    if (depth > maxd) return d3(0.0, 0.0, 0.0);
    const mh_material *m = mat_of(s, h->material);
    if (!m) return or_sky_sample(r.d, &s->sky);
    dv3 P = h->point, N = h->normal;
    if (m->is_water) N = or_water_normal(P.x, P.z, 0.0);
    if (ddot(N, r.d) > 0) N = dneg(N);
    dv3 V = dneg(r.d);
    mh_material ml = *m;
    dv3 alb = or_texture_albedo(m, P, h->u, h->v);
    ml.albedo = (mh_vec3){ alb.x, alb.y, alb.z };
    const mh_material *mm = &ml;

    dv3 color = or_ambient(mm, N, &s->sky);
    dv3 L = fromm(s->sky.sun_dir), sunc = fromm(s->sky.sun_color);
    dv3 so = dadd(P, dscale(N, 1e-3));
    double srad = s->sky.sun_radius;
    if (srad > 1e-9) {
        double lit = 0.0;
        for (int i = 0; i < 16; ++i) {
            unsigned k = seed * 0x9e3779b9u + (unsigned)i * 2u;
            double r1 = or_rand01(k, (unsigned)i, 0x5a17u);
            double r2 = or_rand01(k, (unsigned)i, 0x7c3du);
            dv3 Ld = or_sun_disk_dir(L, srad, r1, r2);
            oray sr = { so, Ld };
            if (!or_occluded1(s, sr, 1e-3, 1e30)) lit += 1.0;
        }
        if (lit > 0.0) color = dadd(color, dscale(shade(mm, N, L, V, sunc), lit / 16.0));
    } else {
        oray sr = { so, L };
        if (!or_occluded1(s, sr, 1e-3, 1e30)) color = dadd(color, shade(mm, N, L, V, sunc));
    }
    if (s->n_lights > 0) color = dadd(color, emissive_direct(s, mm, P, N, V, h->prim, seed));

    dv3 refl = d3(0.0, 0.0, 0.0);
    if (depth < maxd) {
        if (mm->pbr != 0 && m->reflectivity > 0.0 && m->roughness > 1e-6) {
            const double inv = 1.0 / 16.0;
            for (int i = 0; i < 16; ++i) {
                unsigned k = seed * 0x9e3779b9u + (unsigned)depth * 0x85ebca6bu + (unsigned)i * 2u;
                double r1 = or_rand01(k, (unsigned)i, 0x9e37u);
                double r2 = or_rand01(k, (unsigned)i, 0x85ebu);
                dv3 rd = or_sample_glossy_dir(mm, N, r.d, r1, r2);
                oray rr = { dadd(P, dscale(N, 1e-3)), rd };
                refl = dadd(refl, dscale(w_trace(s, rr, depth + 1, maxd, seed), inv));
            }
        } else {
            oray rr = { dadd(P, dscale(N, 1e-3)), dreflect(r.d, N) };
            refl = w_trace(s, rr, depth + 1, maxd, seed);
        }
    }
    double f0 = (1.0 - m->ior) / (1.0 + m->ior);
    f0 = f0 * f0;
    double ci = ddot(N, V);
    if (ci < 0.0) ci = 0.0;
    else if (ci > 1.0) ci = 1.0;
    double eta = h->front == 1 ? 1.0 / m->ior : m->ior;
    double s2t = eta * eta * (1.0 - ci * ci);
    double F;
    int tir = 0;
    if (s2t > 1.0) {
        F = 1.0;
        tir = 1;
    } else {
        /* material.c Schlick: cos clamped to [0,1], f0 unclamped */
        double c = or_clamp01(sqrt(1.0 - s2t)), mm5 = 1.0 - c, m2 = mm5 * mm5;
        F = f0 + (1.0 - f0) * (m2 * m2 * mm5);
    }
    if (m->transparency > 0.0) {
        dv3 refr;
        if (tir) {
            refr = refl;
        } else {
            double cosi = -ddot(r.d, N);
            double s2 = eta * eta * (1.0 - cosi * cosi);
            dv3 rd;
            if (s2 > 1.0) {
                rd = d3(0.0, 0.0, 0.0);
            } else {
                double ct = sqrt(1.0 - s2), k = eta * cosi - ct;
                rd = d3(eta * r.d.x + k * N.x, eta * r.d.y + k * N.y, eta * r.d.z + k * N.z);
            }
            oray rr = { dadd(P, dscale(rd, 1e-3)), rd };
            ohit h2;
            if (or_intersect1(s, rr, 1e-4, 1e30, &h2)) {
                dv3 inner = w_trace_hit(s, rr, depth + 1, maxd, &h2, seed);
                refr = (m->is_water || m->beer_lambert) ? or_water_attenuate(m, inner, h2.t) : inner;
            } else {
                dv3 inner = or_sky_sample(rd, &s->sky);
                refr = (m->is_water || m->beer_lambert) ? or_water_attenuate(m, inner, 5.0) : inner;
            }
        }
        dv3 surf = dadd(dscale(refl, F), dscale(refr, 1.0 - F));
        dv3 out = dlerp(color, surf, m->transparency);
        if (mm->pbr && (m->emissive.x != 0.0 || m->emissive.y != 0.0 || m->emissive.z != 0.0))
            out = dadd(out, fromm(m->emissive));
        return out;
    }
    dv3 out = dlerp(color, refl, m->reflectivity);
    if (mm->pbr && (m->emissive.x != 0.0 || m->emissive.y != 0.0 || m->emissive.z != 0.0))
        out = dadd(out, fromm(m->emissive));
    if (depth == 0 && dlen2(fromm(mm->atmosphere_glow)) > 1e-6)
        out = dadd(out, limb_glow(mm, N, V, fromm(s->sky.sun_dir)));
    return out;
}

static dv3 w_trace(const or_scene *s, oray r, int depth, int maxd, unsigned seed)
{
    if (depth > maxd) return d3(0.0, 0.0, 0.0);
    ohit h;
    double T;
    dv3 ins;
    if (!or_intersect1(s, r, 1e-4, 1e30, &h)) {
        dv3 sky = or_sky_sample(r.d, &s->sky);
        if (s->fog.density <= 0.0) return sky;
        fog_segment(&s->fog, &s->sky, r, 1e30, &T, &ins, NULL);
        return dadd(dscale(sky, T), ins);
    }
    dv3 c = w_trace_hit(s, r, depth, maxd, &h, seed);
    if (s->fog.density <= 0.0) return c;
    fog_segment(&s->fog, &s->sky, r, h.t, &T, &ins, NULL);
    return dadd(dscale(c, T), ins);
}

/* ---- pixel kernels ---------------------------------------------------- */

typedef struct {
    const or_scene *s;
    const mh_camera *cam;
    int w, h, spp, depth, mode, s0, y0, y1;
    unsigned char *rgb;
    double *mean, *var;
    int tiles_x, tiles_y;
    atomic_int next;
    pthread_mutex_t lock;
    long long stats[OR_NSTATS];
} job_t;

static void render_pixel(job_t *j, int x, int y)
{
    const or_scene *s = j->s;
    /* acc: the reference's running sum; wm / wq: Welford mean and squared
     * deviations (numerically stable variance) */
    dv3 acc = d3(0.0, 0.0, 0.0), wm = d3(0.0, 0.0, 0.0), wq = d3(0.0, 0.0, 0.0);
    int g = (int)(sqrt((double)j->spp) + 0.5);
    for (int si = 0; si < j->spp; ++si) {
        int sidx = j->s0 + si;
        dv3 c;
        if (j->mode == 0) {
            double jx = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 0));
            double jy = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 1));
            double l1 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 100));
            double l2 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 101));
            double uu = ((double)x + jx) / (double)j->w;
            double vv = 1.0 - ((double)y + jy) / (double)j->h;
            oray ray = cam_ray(j->cam, uu, vv, l1, l2);
            if (j->cam->dome_radius > 1e-4 && s->water_material >= 0) {
                dv3 pd = dadd(ray.o, dscale(ray.d, j->cam->dome_radius));
                if (pd.y < water_height(s, pd.x, pd.z)) ray.o = dsub(pd, d3(0.0, 1e-3, 0.0));
            }
            c = pt_radiance(s, ray, j->depth, or_hash3((unsigned)x, (unsigned)y, (unsigned)sidx), NULL);
        } else {
            double du, dv;
            if (g > 0 && g * g == j->spp) {
                int gi = sidx % g, gj = sidx / g;
                double jx = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 0));
                double jy = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 1));
                du = ((double)gi + jx) / (double)g;
                dv = ((double)gj + jy) / (double)g;
            } else {
                du = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 0));
                dv = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 1));
            }
            double uu = ((double)x + du) / (double)j->w;
            double vv = 1.0 - ((double)y + dv) / (double)j->h;
            double l1 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 2));
            double l2 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 3));
            oray ray = cam_ray(j->cam, uu, vv, l1, l2);
            c = w_trace(s, ray, 0, j->depth, or_hash3((unsigned)x, (unsigned)y, (unsigned)sidx));
        }
        acc = dadd(acc, c);
        double k = 1.0 / (double)(si + 1);
        dv3 dl = dsub(c, wm);
        wm = dadd(wm, dscale(dl, k));
        wq = dadd(wq, dmul(dl, dsub(c, wm)));
    }
    dv3 m = dscale(acc, 1.0 / (double)j->spp);
    size_t o = ((size_t)y * (size_t)j->w + (size_t)x) * 3;
    if (j->rgb) {
        j->rgb[o] = or_to_byte(m.x);
        j->rgb[o + 1] = or_to_byte(m.y);
        j->rgb[o + 2] = or_to_byte(m.z);
    }
    if (j->mean) {
        j->mean[o] = m.x;
        j->mean[o + 1] = m.y;
        j->mean[o + 2] = m.z;
    }
    if (j->var) {
        double n = (double)j->spp;
        double k = n > 1.0 ? 1.0 / (n - 1.0) : 0.0;
        j->var[o] = wq.x * k;
        j->var[o + 1] = wq.y * k;
        j->var[o + 2] = wq.z * k;
    }
}

#define TILE 16

static void *worker(void *arg)
{
    job_t *j = (job_t *)arg;
    int nt = j->tiles_x * j->tiles_y;
    or_stats_reset();
    for (;;) {
        int t = atomic_fetch_add(&j->next, 1);
        if (t >= nt) break;
        int ty = t / j->tiles_x, tx = t - ty * j->tiles_x;
        int x0 = tx * TILE, yy0 = j->y0 + ty * TILE;
        int x1 = x0 + TILE < j->w ? x0 + TILE : j->w;
        int yy1 = yy0 + TILE < j->y1 ? yy0 + TILE : j->y1;
        for (int y = yy0; y < yy1; ++y)
            for (int x = x0; x < x1; ++x) render_pixel(j, x, y);
    }
    long long st[OR_NSTATS];
    or_stats_get(st);
    pthread_mutex_lock(&j->lock);
    for (int i = 0; i < OR_NSTATS; ++i) j->stats[i] += st[i];
    pthread_mutex_unlock(&j->lock);
    return NULL;
}

static long long g_render_stats[OR_NSTATS];

int or_render(const or_scene *s, const mh_camera *cam, int width, int height, int spp,
              int depth, int mode, int sample_offset, int y0, int y1, int threads,
              unsigned char *rgb, double *mean, double *var, double *seconds)
{
    if (!s || !cam) return 1;
    if (width <= 0 || height <= 0) return 2;
    if (spp < 1 || depth < 0) return 3;
    if (y1 <= 0 || y1 > height) y1 = height;
    if (y0 < 0) y0 = 0;
    if (y0 >= y1) return 2;
    job_t *j = (job_t *)calloc(1, sizeof(job_t));
    if (!j) return 4;
    j->s = s; j->cam = cam; j->w = width; j->h = height; j->spp = spp; j->depth = depth;
    j->mode = mode; j->s0 = sample_offset; j->y0 = y0; j->y1 = y1;
    j->rgb = rgb; j->mean = mean; j->var = var;
    j->tiles_x = (width + TILE - 1) / TILE;
    j->tiles_y = (y1 - y0 + TILE - 1) / TILE;
    atomic_init(&j->next, 0);
    pthread_mutex_init(&j->lock, NULL);
    int nt = threads;
    if (nt <= 0) {
        long nc = sysconf(_SC_NPROCESSORS_ONLN);
        nt = nc >= 1 ? (int)nc : 4;
    }
    if (nt > 64) nt = 64;
    if (nt > j->tiles_x * j->tiles_y) nt = j->tiles_x * j->tiles_y;
    if (nt < 1) nt = 1;
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    pthread_t tid[64];
    int started = 0;
    for (int i = 1; i < nt; ++i)
        if (pthread_create(&tid[started], NULL, worker, j) == 0) ++started;
    worker(j);
    for (int i = 0; i < started; ++i) pthread_join(tid[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    if (seconds) *seconds = (double)(b.tv_sec - a.tv_sec) + (double)(b.tv_nsec - a.tv_nsec) * 1e-9;
    memcpy(g_render_stats, j->stats, sizeof g_render_stats);
    pthread_mutex_destroy(&j->lock);
    free(j);
    return 0;
}

void or_render_stats_get(long long out[OR_NSTATS]) { memcpy(out, g_render_stats, sizeof g_render_stats); }

/* ---- adaptive Whitted ----------------------------------------------------- */

typedef struct {
    const or_scene *s;
    const mh_camera *cam;
    int w, h, n0, nmax, depth;
    double *sum;            /* R, G, B, Y, Y^2 per pixel */
    int *ns;
    const unsigned char *active;
    atomic_int next_row;
} adapt_job;

/* One refinement pass of one pixel: up to n0 more samples, continuing the
 * pixel's global sample index; the fixed render's jitter for s < n0, uniform
 * jitter after. The pass's sums are added to the running sums at the end. */
static void adapt_pixel(adapt_job *j, int x, int y)
{
    size_t p = (size_t)y * (size_t)j->w + (size_t)x;
    int n = j->ns[p], b = j->n0;
    if (b > j->nmax - n) b = j->nmax - n;
    if (b <= 0) return;
    int g = (int)(sqrt((double)j->n0) + 0.5);
    double sr = 0.0, sg = 0.0, sb = 0.0, sy = 0.0, sy2 = 0.0;
    for (int k = 0; k < b; ++k) {
        int s = n + k;
        double du, dv;
        if (s < j->n0 && g > 0 && g * g == j->n0) {
            double jx = or_rand01((unsigned)x, (unsigned)y, (unsigned)(s * 2 + 0));
            double jy = or_rand01((unsigned)x, (unsigned)y, (unsigned)(s * 2 + 1));
            du = ((double)(s % g) + jx) / (double)g;
            dv = ((double)(s / g) + jy) / (double)g;
        } else {
            du = or_rand01((unsigned)x, (unsigned)y, (unsigned)(s * 2 + 0));
            dv = or_rand01((unsigned)x, (unsigned)y, (unsigned)(s * 2 + 1));
        }
        double uu = ((double)x + du) / (double)j->w;
        double vv = 1.0 - ((double)y + dv) / (double)j->h;
        double l1 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(s * 2 + 2));
        double l2 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(s * 2 + 3));
        oray ray = cam_ray(j->cam, uu, vv, l1, l2);
        dv3 c = w_trace(j->s, ray, 0, j->depth, or_hash3((unsigned)x, (unsigned)y, (unsigned)s));
        sr += c.x;
        sg += c.y;
        sb += c.z;
        double Y = 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z;
        sy += Y;
        sy2 += Y * Y;
    }
    double *q = &j->sum[5 * p];
    q[0] += sr;
    q[1] += sg;
    q[2] += sb;
    q[3] += sy;
    q[4] += sy2;
    j->ns[p] = n + b;
}

static void *adapt_worker(void *arg)
{
    adapt_job *j = (adapt_job *)arg;
    for (;;) {
        int y = atomic_fetch_add(&j->next_row, 1);
        if (y >= j->h) break;
        for (int x = 0; x < j->w; ++x)
            if (j->active[(size_t)y * (size_t)j->w + (size_t)x]) adapt_pixel(j, x, y);
    }
    return NULL;
}

static void adapt_pass(adapt_job *j, int threads)
{
    atomic_store(&j->next_row, 0);
    pthread_t tid[64];
    int started = 0;
    for (int i = 1; i < threads; ++i)
        if (pthread_create(&tid[started], NULL, adapt_worker, j) == 0) ++started;
    adapt_worker(j);
    for (int i = 0; i < started; ++i) pthread_join(tid[i], NULL);
}

int or_render_adaptive(const or_scene *s, const mh_camera *cam, int width, int height, int n0, int depth,
                       int nmax, double tau, int threads, unsigned char *rgb, double *mean, int *nsamp,
                       long long *total, int *max_n)
{
    if (!s || !cam) return 1;
    if (width <= 0 || height <= 0) return 2;
    if (n0 < 1 || depth < 0) return 3;
    if (nmax <= 0) nmax = 4 * n0;
    if (!(tau > 0.0)) tau = 0.02;
    if (nmax < n0) nmax = n0;
    size_t np = (size_t)width * (size_t)height;
    adapt_job j;
    j.s = s; j.cam = cam; j.w = width; j.h = height; j.n0 = n0; j.nmax = nmax; j.depth = depth;
    j.sum = (double *)calloc(5 * np, sizeof(double));
    j.ns = (int *)calloc(np, sizeof(int));
    unsigned char *act = (unsigned char *)malloc(np);
    if (!j.sum || !j.ns || !act) {
        free(j.sum); free(j.ns); free(act);
        return 4;
    }
    j.active = act;
    int nt = threads;
    if (nt <= 0) {
        long nc = sysconf(_SC_NPROCESSORS_ONLN);
        nt = nc >= 1 ? (int)nc : 4;
    }
    if (nt > 64) nt = 64;
    int max_passes = 1;
    if (nmax > n0) max_passes += (nmax - n0 + n0 - 1) / n0;
    memset(act, 1, np);
    adapt_pass(&j, nt);
    for (int pass = 1;; ++pass) {
        int count = 0;
        for (size_t p = 0; p < np; ++p) {
            int n = j.ns[p], flag = 0;
            if (n < nmax) {
                if (n < 2) {
                    flag = 1;
                } else {
                    double nf = (double)n, *q = &j.sum[5 * p];
                    double ybar = q[3] / nf;
                    double var = (q[4] - q[3] * ybar) / (nf - 1.0);
                    if (var < 0.0) var = 0.0;
                    double e = sqrt(var / nf) / (ybar + 1e-4);
                    flag = e > tau;
                }
            }
            act[p] = (unsigned char)flag;
            count += flag;
        }
        if (count == 0 || pass >= max_passes) break;
        adapt_pass(&j, nt);
    }
    long long tot = 0;
    int mx = 0;
    for (size_t p = 0; p < np; ++p) {
        double inv = 1.0 / (double)j.ns[p];
        double *q = &j.sum[5 * p];
        if (rgb) {
            rgb[3 * p] = or_to_byte(q[0] * inv);
            rgb[3 * p + 1] = or_to_byte(q[1] * inv);
            rgb[3 * p + 2] = or_to_byte(q[2] * inv);
        }
        if (mean) {
            mean[3 * p] = q[0] * inv;
            mean[3 * p + 1] = q[1] * inv;
            mean[3 * p + 2] = q[2] * inv;
        }
        if (nsamp) nsamp[p] = j.ns[p];
        tot += j.ns[p];
        if (j.ns[p] > mx) mx = j.ns[p];
    }
    if (total) *total = tot;
    if (max_n) *max_n = mx;
    free(j.sum); free(j.ns); free(act);
    return 0;
}

/* ---- adjoint driver ----------------------------------------------------- */

typedef struct {
    const or_scene *s;
    const mh_camera *cam;
    int w, h, spp, depth, s0;
    const double *dimg;
    double *grad;        /* shared result, n_mats*8 + 3 */
    int ng;
    int tiles_x, tiles_y;
    atomic_int next;
    pthread_mutex_t lock;
} adj_job_t;

static void adj_pixel(adj_job_t *j, int x, int y, or_adj *a)
{
    const or_scene *s = j->s;
    size_t o = ((size_t)y * (size_t)j->w + (size_t)x) * 3;
    double inv = 1.0 / (double)j->spp;
    dv3 dp = d3(j->dimg[o] * inv, j->dimg[o + 1] * inv, j->dimg[o + 2] * inv);
    for (int si = 0; si < j->spp; ++si) {
        int sidx = j->s0 + si;
        double jx = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 0));
        double jy = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 1));
        double l1 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 100));
        double l2 = or_rand01((unsigned)x, (unsigned)y, (unsigned)(sidx * 2 + 101));
        double uu = ((double)x + jx) / (double)j->w;
        double vv = 1.0 - ((double)y + jy) / (double)j->h;
        oray ray = cam_ray(j->cam, uu, vv, l1, l2);
        if (j->cam->dome_radius > 1e-4 && s->water_material >= 0) {
            dv3 pd = dadd(ray.o, dscale(ray.d, j->cam->dome_radius));
            if (pd.y < water_height(s, pd.x, pd.z)) ray.o = dsub(pd, d3(0.0, 1e-3, 0.0));
        }
        unsigned seed = or_hash3((unsigned)x, (unsigned)y, (unsigned)sidx);
        /* pass 1: the raw (unclamped) sum of the path's terms; the final
         * clamp of a negative channel to 0 has zero derivative */
        a->delta = d3(0.0, 0.0, 0.0);
        a->Lacc = d3(0.0, 0.0, 0.0);
        (void)pt_radiance(s, ray, j->depth, seed, a);
        dv3 L = a->Lacc;
        if (!dfinite(L)) continue;
        a->Lt = L;
        a->Lacc = d3(0.0, 0.0, 0.0);
        a->delta = d3(L.x >= 0.0 ? dp.x : 0.0, L.y >= 0.0 ? dp.y : 0.0, L.z >= 0.0 ? dp.z : 0.0);
        (void)pt_radiance(s, ray, j->depth, seed, a);
    }
}

static void *adj_worker(void *arg)
{
    adj_job_t *j = (adj_job_t *)arg;
    or_adj a;
    memset(&a, 0, sizeof a);
    a.n_mats = j->s->n_mats;
    a.g = (double *)calloc((size_t)j->ng, sizeof(double));
    if (!a.g) return NULL;
    int nt = j->tiles_x * j->tiles_y;
    for (;;) {
        int t = atomic_fetch_add(&j->next, 1);
        if (t >= nt) break;
        int ty = t / j->tiles_x, tx = t - ty * j->tiles_x;
        int x0 = tx * TILE, y0 = ty * TILE;
        int x1 = x0 + TILE < j->w ? x0 + TILE : j->w;
        int y1 = y0 + TILE < j->h ? y0 + TILE : j->h;
        for (int y = y0; y < y1; ++y)
            for (int x = x0; x < x1; ++x) adj_pixel(j, x, y, &a);
    }
    pthread_mutex_lock(&j->lock);
    for (int i = 0; i < j->ng; ++i) j->grad[i] += a.g[i];
    pthread_mutex_unlock(&j->lock);
    free(a.g);
    return NULL;
}

int or_adjoint(const or_scene *s, const mh_camera *cam, int width, int height, int spp, int depth,
               int sample_offset, const double *dimg, double *grad, int threads)
{
    if (!s || !cam || !dimg || !grad) return 1;
    if (width <= 0 || height <= 0) return 2;
    if (spp < 1 || depth < 0 || sample_offset < 0) return 3;
    if (s->freeze) return 4;
    adj_job_t *j = (adj_job_t *)calloc(1, sizeof(adj_job_t));
    if (!j) return 4;
    j->s = s; j->cam = cam; j->w = width; j->h = height; j->spp = spp; j->depth = depth;
    j->s0 = sample_offset; j->dimg = dimg; j->grad = grad;
    j->ng = s->n_mats * 8 + 3;
    memset(grad, 0, sizeof(double) * (size_t)j->ng);
    j->tiles_x = (width + TILE - 1) / TILE;
    j->tiles_y = (height + TILE - 1) / TILE;
    atomic_init(&j->next, 0);
    pthread_mutex_init(&j->lock, NULL);
    int nt = threads;
    if (nt <= 0) {
        long nc = sysconf(_SC_NPROCESSORS_ONLN);
        nt = nc >= 1 ? (int)nc : 4;
    }
    if (nt > 64) nt = 64;
    if (nt > j->tiles_x * j->tiles_y) nt = j->tiles_x * j->tiles_y;
    if (nt < 1) nt = 1;
    pthread_t tid[64];
    int started = 0;
    for (int i = 1; i < nt; ++i)
        if (pthread_create(&tid[started], NULL, adj_worker, j) == 0) ++started;
    adj_worker(j);
    for (int i = 0; i < started; ++i) pthread_join(tid[i], NULL);
    pthread_mutex_destroy(&j->lock);
    free(j);
    return 0;
}

/* Replace the materials (same count): the emissive-light SET stays the one
 * built at creation, each light taking its material's new emission. */
int or_scene_set_materials(or_scene *s, const mh_material *mats, int n)
{
    if (!s || !mats) return 1;
    if (n != s->n_mats) return 5;
    memcpy(s->mats, mats, sizeof(mh_material) * (size_t)n);
    for (int i = 0; i < s->n_lights; ++i) s->lights[i].emissive = fromm(mats[s->lights[i].mat].emissive);
    return 0;
}

int or_scene_set_sky(or_scene *s, const mh_sky *sky)
{
    if (!s || !sky) return 1;
    s->sky = *sky;
    return 0;
}

/* Freeze sampling decisions at these materials (NULL: unfreeze). */
int or_set_freeze(or_scene *s, const mh_material *mats, int n)
{
    if (!s) return 1;
    free(s->freeze);
    s->freeze = NULL;
    if (!mats) return 0;
    if (n != s->n_mats) return 5;
    s->freeze = (mh_material *)malloc(sizeof(mh_material) * (size_t)(n > 0 ? n : 1));
    if (!s->freeze) return 4;
    memcpy(s->freeze, mats, sizeof(mh_material) * (size_t)n);
    return 0;
}

void or_set_anyhit_shadow(int on) { or_anyhit_shadow = on ? 1 : 0; }

void or_stats_reset(void) { memset(&or_tls_stats, 0, sizeof or_tls_stats); }

void or_stats_get(long long out[OR_NSTATS])
{
    out[0] = or_tls_stats.closest_queries;
    out[1] = or_tls_stats.any_queries;
    out[2] = or_tls_stats.c_inner;
    out[3] = or_tls_stats.c_prims;
    out[4] = or_tls_stats.c_tris;
    out[5] = or_tls_stats.a_inner;
    out[6] = or_tls_stats.a_prims;
    out[7] = or_tls_stats.a_tris;
    out[8] = or_tls_stats.plane_tests;
    out[9] = or_tls_stats.ext_queries;
}

/* ---- scene ------------------------------------------------------------ */

or_scene *or_scene_create(const mh_scene_soa *soa)
{
    if (!soa || soa->n_prims < 0 || soa->n_materials < 0) return NULL;
    or_scene *s = (or_scene *)calloc(1, sizeof(or_scene));
    if (!s) return NULL;
    s->n_prims = soa->n_prims;
    s->prims = (oprim *)calloc((size_t)(soa->n_prims > 0 ? soa->n_prims : 1), sizeof(oprim));
    s->n_mats = soa->n_materials;
    s->mats = (mh_material *)calloc((size_t)(soa->n_materials > 0 ? soa->n_materials : 1), sizeof(mh_material));
    if (!s->prims || !s->mats) {
        or_scene_destroy(s);
        return NULL;
    }
    if (soa->n_materials > 0) memcpy(s->mats, soa->materials, sizeof(mh_material) * (size_t)soa->n_materials);
    s->sky = soa->sky;
    s->fog = soa->fog;
    s->water_level = soa->water_level;
    s->water_material = soa->water_material;
    s->ocean = soa->ocean;
    for (int i = 0; i < s->n_mats; ++i)
        if (s->mats[i].is_water || s->mats[i].transparency > 0.8) s->has_transmissive = 1;
    for (int i = 0; i < soa->n_prims; ++i) {
        const mh_prim *q = &soa->prims[i];
        oprim *p = &s->prims[i];
        p->kind = q->kind;
        p->material = q->material;
        p->flags = soa->attrs ? q->flags : 0;
        p->p0 = fromm(q->p0);
        p->p1 = fromm(q->p1);
        p->p2 = fromm(q->p2);
        p->r0 = q->r0;
        p->r1 = q->r1;
        p->e1 = dsub(p->p1, p->p0);
        p->e2 = dsub(p->p2, p->p0);
        if (soa->attrs) {
            const mh_prim_attr *a = &soa->attrs[i];
            p->n0 = fromm(a->n0); p->n1 = fromm(a->n1); p->n2 = fromm(a->n2);
            p->uv0 = fromm(a->uv0); p->uv1 = fromm(a->uv1); p->uv2 = fromm(a->uv2);
        }
    }
    /* emissive area lights: first 8 PBR emitting spheres, file order */
    for (int i = 0; i < s->n_prims && s->n_lights < 8; ++i) {
        const oprim *p = &s->prims[i];
        if (p->kind != MH_PRIM_SPHERE) continue;
        const mh_material *m = mat_of(s, p->material);
        if (!m || !m->pbr) continue;
        if (m->emissive.x == 0.0 && m->emissive.y == 0.0 && m->emissive.z == 0.0) continue;
        if (!(p->r0 > 0.0)) continue;
        olight *l = &s->lights[s->n_lights++];
        l->prim = i;
        l->mat = p->material;
        l->center = p->p0;
        l->radius = p->r0;
        l->emissive = fromm(m->emissive);
    }
    if (or_bvh_build(s) != 0) {
        or_scene_destroy(s);
        return NULL;
    }
    return s;
}

void or_scene_destroy(or_scene *s)
{
    if (!s) return;
    free(s->prims);
    free(s->mats);
    free(s->nodes);
    free(s->order);
    free(s->leaf_prims);
    free(s->planes);
    free(s->freeze);
    free(s);
}

int or_scene_info(const or_scene *s, int *n_prims, int *n_nodes, int *n_lights)
{
    if (!s) return 1;
    if (n_prims) *n_prims = s->n_prims;
    if (n_nodes) *n_nodes = s->n_nodes;
    if (n_lights) *n_lights = s->n_lights;
    return 0;
}

/* ---- batch queries & vectors ----------------------------------------- */

static oray from_mh(const mh_ray *r)
{
    oray o = { d3(r->ox, r->oy, r->oz), d3(r->dx, r->dy, r->dz) };
    return o;
}

int or_intersect(const or_scene *s, const mh_ray *rays, int n, double tmin, double tmax, mh_hit *out)
{
    if (!s || (!rays && n > 0) || (!out && n > 0)) return 1;
    for (int i = 0; i < n; ++i) {
        ohit h;
        memset(&out[i], 0, sizeof out[i]);
        if (or_intersect1(s, from_mh(&rays[i]), tmin, tmax, &h)) {
            out[i].t = h.t;
            out[i].prim_index = h.prim;
            out[i].material = h.material;
            out[i].front_face = h.front;
            out[i].nx = h.normal.x; out[i].ny = h.normal.y; out[i].nz = h.normal.z;
            out[i].u = h.u; out[i].v = h.v;
        } else {
            out[i].prim_index = -1;
        }
    }
    return 0;
}

int or_occluded(const or_scene *s, const mh_ray *rays, int n, double tmin, double tmax, unsigned char *out)
{
    if (!s || (!rays && n > 0) || (!out && n > 0)) return 1;
    for (int i = 0; i < n; ++i) out[i] = (unsigned char)or_occluded1(s, from_mh(&rays[i]), tmin, tmax);
    return 0;
}

int or_radiance(const or_scene *s, const mh_ray *ray, int depth, unsigned seed, double out[3])
{
    if (!s || !ray) return 1;
    dv3 c = pt_radiance(s, from_mh(ray), depth, seed, NULL);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
    return 0;
}

void or_sky_samples(const or_scene *s, const double *dirs, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        dv3 c = or_sky_sample(d3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), &s->sky);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

void or_water_normals(const double *xz, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        dv3 c = or_water_normal(xz[2 * i], xz[2 * i + 1], 0.0);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

void or_noise(const double *xy, int n, int octaves, unsigned seed, double *perlin, double *fbm)
{
    for (int i = 0; i < n; ++i) {
        perlin[i] = or_perlin2(xy[2 * i], xy[2 * i + 1], seed);
        fbm[i] = or_fbm2(xy[2 * i], xy[2 * i + 1], octaves, 2.0, 0.5, seed);
    }
}

void or_noise3(const double *xyz, int n, unsigned seed, double *perlin, double *fbm)
{
    for (int i = 0; i < n; ++i) {
        perlin[i] = or_perlin3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], seed);
        fbm[i] = or_fbm3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 3, 2.0, 0.5, seed);
    }
}

static void put3(double *o, dv3 v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

void or_sun_disk_dirs(const double sun[3], double radius_deg, const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i)
        put3(out + 3 * i, or_sun_disk_dir(d3(sun[0], sun[1], sun[2]), radius_deg, u[2 * i], u[2 * i + 1]));
}

void or_light_cone_dirs(const double w[3], double cos_max, const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i)
        put3(out + 3 * i, or_light_cone_dir(d3(w[0], w[1], w[2]), cos_max, u[2 * i], u[2 * i + 1]));
}

void or_cosine_dirs(const double nrm[3], const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i)
        put3(out + 3 * i, or_cosine_hemisphere(d3(nrm[0], nrm[1], nrm[2]), u[2 * i], u[2 * i + 1]));
}

void or_ggx_halfs(const double nrm[3], double alpha, const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i)
        put3(out + 3 * i, or_ggx_half(d3(nrm[0], nrm[1], nrm[2]), alpha, u[2 * i], u[2 * i + 1]));
}

double or_fresnel_dielectric_x(double cos_i, double eta) { return or_fresnel_dielectric(cos_i, eta); }

double or_specular_probability_x(double m, double r, double f0) { return or_specular_probability(m, r, f0); }

void or_texture_albedos(const mh_material *m, const double *p, const double *uv, int n, double *out)
{
    for (int i = 0; i < n; ++i)
        put3(out + 3 * i, or_texture_albedo(m, d3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), uv[2 * i], uv[2 * i + 1]));
}

void or_shade(const or_scene *s, int material, const double N[3], const double L[3],
              const double V[3], const double light[3], double out[3])
{
    const mh_material *m = mat_of(s, material);
    if (!m) { out[0] = out[1] = out[2] = 0.0; return; }
    put3(out, shade(m, d3(N[0], N[1], N[2]), d3(L[0], L[1], L[2]), d3(V[0], V[1], V[2]),
                    d3(light[0], light[1], light[2])));
}
