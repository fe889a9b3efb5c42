/*
 * or_math.c - oracle: integer-hash PRNG, Perlin/fBm noise and Monte-Carlo
 * sampling primitives, restated in FP64. TEST INFRASTRUCTURE ONLY.
 *
 * Follows:
 *   PRNG      reference src/pathtrace.c:92-117 (== src/render.c:371-396)
 *   noise     reference src/noise.c:31-61 (hash), 68-76 (fade/lerp),
 *             158-187 (perlin2), 272-294 (fbm2), 189-253 (perlin3),
 *             296-317 (fbm3)
 *   sampling  reference src/sampling.c:52-281
 */
#include "or_internal.h"

#include <stdint.h>

__thread or_stats or_tls_stats;

/* ---- PRNG (a5) ----------------------------------------------------- */

/* 32-bit avalanche: xorshift-multiply finaliser, constants 0x7feb352d /
 * 0x846ca68b (reference src/pathtrace.c:92-100). */
unsigned or_hash_u32(unsigned x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* Three-word mix (reference src/pathtrace.c:103-110). */
unsigned or_hash3(unsigned a, unsigned b, unsigned c)
{
    unsigned h = 0x9e3779b9u ^ or_hash_u32(a + 0x85ebca6bu);
    h = or_hash_u32(h + b);
    h ^= or_hash_u32(c + 0xc2b2ae35u);
    return or_hash_u32(h);
}

/* Top 24 hash bits scaled by 2^-24: uniform on [0,1) (src/pathtrace.c:113-117). */
double or_rand01(unsigned a, unsigned b, unsigned c)
{
    return (double)(or_hash3(a, b, c) >> 8) * (1.0 / 16777216.0);
}

/* ---- noise --------------------------------------------------------- */

static unsigned lattice_hash(int x, int y, int z, unsigned seed)
{
    unsigned h = seed + 0x9e3779b9u;
    h ^= (unsigned)x * 0x85ebca6bu;
    h ^= (unsigned)y * 0xc2b2ae35u;
    h ^= (unsigned)z * 0x27d4eb2fu;
    return or_hash_u32(h);
}

static double lattice01(int x, int y, unsigned seed)
{
    return (double)lattice_hash(x, y, 0, seed) / (double)UINT32_MAX;
}

static double fade5(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
static double mixd(double a, double b, double t) { return a + t * (b - a); }
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* 2D gradient noise: a random angle per lattice corner (hash * 2 pi), dot with
 * the corner offset, quintic blend, scale by sqrt(2), clamp to [-1,1]. */
double or_perlin2(double x, double y, unsigned seed)
{
    double fx = floor(x), fy = floor(y);
    int ix = (int)fx, iy = (int)fy;
    double rx = x - fx, ry = y - fy;
    double u = fade5(rx), v = fade5(ry);
    double g00 = lattice01(ix, iy, seed) * 2.0 * OR_PI;
    double g10 = lattice01(ix + 1, iy, seed) * 2.0 * OR_PI;
    double g01 = lattice01(ix, iy + 1, seed) * 2.0 * OR_PI;
    double g11 = lattice01(ix + 1, iy + 1, seed) * 2.0 * OR_PI;
    double d00 = cos(g00) * rx + sin(g00) * ry;
    double d10 = cos(g10) * (rx - 1) + sin(g10) * ry;
    double d01 = cos(g01) * rx + sin(g01) * (ry - 1);
    double d11 = cos(g11) * (rx - 1) + sin(g11) * (ry - 1);
    double val = mixd(mixd(d00, d10, u), mixd(d01, d11, u), v);
    return clampd(val * 1.4142135623730951, -1.0, 1.0);
}

double or_fbm2(double x, double y, int octaves, double lac, double gain, unsigned seed)
{
    int n = octaves <= 0 ? 1 : octaves;
    double l = lac > 0.0 ? lac : 2.0;
    double g = gain > 0.0 ? gain : 0.5;
    double amp = 1.0, freq = 1.0, sum = 0.0, norm = 0.0;
    for (int i = 0; i < n; ++i) {
        sum += amp * or_perlin2(x * freq, y * freq, seed);
        norm += amp;
        amp *= g;
        freq *= l;
    }
    if (norm <= 0.0) return 0.0;
    return clampd(sum / norm, -1.0, 1.0);
}

/* 3D gradient noise: per corner a gradient of three signed lattice values
 * (the seed, and the seed xor 0x68bc21eb / 0x02e5be93 for y / z),
 * normalised (the x axis when degenerate); dot with the corner offset,
 * trilinear quintic blend x then y then z, scale by 2/sqrt(3), clamp. */
static double lattice_signed3(int x, int y, int z, unsigned seed)
{
    return (double)lattice_hash(x, y, z, seed) / (double)UINT32_MAX * 2.0 - 1.0;
}

double or_perlin3(double x, double y, double z, unsigned seed)
{
    double fx = floor(x), fy = floor(y), fz = floor(z);
    int ix = (int)fx, iy = (int)fy, iz = (int)fz;
    double r[3] = {x - fx, y - fy, z - fz};
    double d[8];
    for (int c = 0; c < 8; ++c) {
        int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        double gx = lattice_signed3(ix + ox, iy + oy, iz + oz, seed);
        double gy = lattice_signed3(ix + ox, iy + oy, iz + oz, seed ^ 0x68bc21ebu);
        double gz = lattice_signed3(ix + ox, iy + oy, iz + oz, seed ^ 0x02e5be93u);
        double len = sqrt(gx * gx + gy * gy + gz * gz);
        if (len < 1e-12) {
            gx = 1.0; gy = 0.0; gz = 0.0;
        } else {
            gx /= len; gy /= len; gz /= len;
        }
        d[c] = gx * (r[0] - ox) + gy * (r[1] - oy) + gz * (r[2] - oz);
    }
    double u = fade5(r[0]), v = fade5(r[1]), w = fade5(r[2]);
    double a = mixd(mixd(d[0], d[1], u), mixd(d[2], d[3], u), v);
    double b = mixd(mixd(d[4], d[5], u), mixd(d[6], d[7], u), v);
    return clampd(mixd(a, b, w) * 1.1547005383792515, -1.0, 1.0);
}

double or_fbm3(double x, double y, double z, int octaves, double lac, double gain, unsigned seed)
{
    int n = octaves <= 0 ? 1 : octaves;
    double l = lac > 0.0 ? lac : 2.0;
    double g = gain > 0.0 ? gain : 0.5;
    double amp = 1.0, freq = 1.0, sum = 0.0, norm = 0.0;
    for (int i = 0; i < n; ++i) {
        sum += amp * or_perlin3(x * freq, y * freq, z * freq, seed);
        norm += amp;
        amp *= g;
        freq *= l;
    }
    if (norm <= 0.0) return 0.0;
    return clampd(sum / norm, -1.0, 1.0);
}

/* ---- sampling ------------------------------------------------------ */

#define OR_U_MAX 0.9999999999999999

static double clamp_u(double u)
{
    if (!(u > 0.0)) return 0.0;
    if (u > OR_U_MAX) return OR_U_MAX;
    return u;
}

static double clamp_lo_hi(double v, double lo, double hi)
{
    if (!(v > lo)) return lo;
    if (v > hi) return hi;
    return v;
}

static double cl01(double v)
{
    if (!(v > 0.0)) return 0.0;
    if (v > 1.0) return 1.0;
    return v;
}

/* Tangent frame: cross with the world axis least aligned with n. */
void or_basis(dv3 n, dv3 *t, dv3 *b)
{
    dv3 nn = dnorm(n);
    if (dlen2(nn) <= 0.0) nn = d3(0.0, 1.0, 0.0);
    dv3 axis;
    if (fabs(nn.x) <= fabs(nn.y) && fabs(nn.x) <= fabs(nn.z)) axis = d3(1.0, 0.0, 0.0);
    else if (fabs(nn.y) <= fabs(nn.z)) axis = d3(0.0, 1.0, 0.0);
    else axis = d3(0.0, 0.0, 1.0);
    *t = dnorm(dcross(axis, nn));
    *b = dcross(nn, *t);
}

/* Malley: r = sqrt(u1), phi = 2 pi u2, z = sqrt(1 - u1). */
dv3 or_cosine_hemisphere(dv3 n, double u1, double u2)
{
    dv3 nn = dnorm(n);
    if (dlen2(nn) <= 0.0) nn = d3(0.0, 1.0, 0.0);
    double r1 = clamp_u(u1), r2 = clamp_u(u2);
    double r = sqrt(r1);
    double phi = 2.0 * OR_PI * r2;
    double x = r * cos(phi), y = r * sin(phi);
    double z = sqrt(fmax(0.0, 1.0 - r1));
    dv3 t, b;
    or_basis(nn, &t, &b);
    return dnorm(dadd(dadd(dscale(t, x), dscale(b, y)), dscale(nn, z)));
}

double or_fresnel_schlick(double c, double f0)
{
    double cc = cl01(c), f = cl01(f0);
    double m = 1.0 - cc, m2 = m * m;
    return f + (1.0 - f) * (m2 * m2 * m);
}

dv3 or_fresnel_schlick_rgb(double c, dv3 f0)
{
    return d3(or_fresnel_schlick(c, f0.x), or_fresnel_schlick(c, f0.y), or_fresnel_schlick(c, f0.z));
}

/* Unpolarised dielectric Fresnel; 1 under total internal reflection. */
double or_fresnel_dielectric(double cos_i, double eta)
{
    double c = cl01(cos_i);
    double e = eta > 0.0 ? eta : 0.0;
    double s2t = e * e * (1.0 - c * c);
    if (s2t >= 1.0) return 1.0;
    double ct = sqrt(1.0 - s2t);
    double rs = (e * c - ct) / (e * c + ct);
    double rp = (e * ct - c) / (e * ct + c);
    return cl01(0.5 * (rs * rs + rp * rp));
}

int or_refract(dv3 incident, dv3 n, double eta, dv3 *out)
{
    dv3 i = dnorm(incident), nn = dnorm(n);
    if (dlen2(i) <= 0.0 || dlen2(nn) <= 0.0) return 0;
    double e = eta > 0.0 ? eta : 0.0;
    double ci = -ddot(i, nn);
    double s2t = e * e * (1.0 - ci * ci);
    if (s2t >= 1.0) return 0;
    double ct = sqrt(1.0 - s2t);
    double k = e * ci - ct;
    *out = dnorm(d3(e * i.x + k * nn.x, e * i.y + k * nn.y, e * i.z + k * nn.z));
    return 1;
}

double or_specular_probability(double metallic, double roughness, double f0)
{
    double m = cl01(metallic), r = cl01(roughness), f = cl01(f0);
    double p = f + (1.0 - f) * m;
    p *= 1.0 - 0.5 * r * (1.0 - m);
    return clamp_lo_hi(p, 1e-3, 1.0);
}

double or_d_ggx(double ndh, double a)
{
    double a2 = a * a;
    double d = ndh * ndh * (a2 - 1.0) + 1.0;
    return a2 / (OR_PI * d * d);
}

/* GGX half vector by NDF inversion: cos = sqrt((1-u1)/(1+(a^2-1)u1)). */
dv3 or_ggx_half(dv3 n, double alpha, double u1, double u2)
{
    dv3 nn = dnorm(n);
    if (dlen2(nn) <= 0.0) nn = d3(0.0, 1.0, 0.0);
    double a = alpha > 1e-4 ? alpha : 1e-4;
    double a2 = a * a;
    double r1 = clamp_u(u1), r2 = clamp_u(u2);
    double ct = sqrt((1.0 - r1) / (1.0 + (a2 - 1.0) * r1));
    double st = sqrt(fmax(0.0, 1.0 - ct * ct));
    double phi = 2.0 * OR_PI * r2;
    dv3 t, b;
    or_basis(nn, &t, &b);
    dv3 h = dadd(dadd(dscale(t, st * cos(phi)), dscale(b, st * sin(phi))), dscale(nn, ct));
    return dnorm(h);
}

double or_ggx_pdf(double ndh, double hdv, double alpha)
{
    if (!(hdv > 0.0) || !(ndh > 0.0)) return 0.0;
    double a = alpha > 1e-4 ? alpha : 1e-4;
    double c = ndh > 1.0 ? 1.0 : ndh;
    return or_d_ggx(c, a) * c / (4.0 * hdv);
}
