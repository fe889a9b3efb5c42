/*
 * mt_gen.c - seeded benchmark scenes for the marmel GPU path tracer.
 *
 *   MT_GEN_SPHERES    ground plane + 64 spheres, centres U[-10,10]^2 x U[0,4],
 *                     radii U[0.2,1.5], each with its own material drawn from
 *                     Lambert / mirror / glass (IOR 1.5) / GGX-PBR
 *                     (roughness U[0.05,0.9], metallic U[0,1]).
 *   MT_GEN_TERRAIN    grid x grid fBm heightfield (2 (grid-1)^2 triangles,
 *                     vertex coordinates representable in fp32) + 200
 *                     spheres, the first 8 emissive (radiance U[2,20]), GGX
 *                     PBR materials.
 *   MT_GEN_ICOSPHERE  a randomly displaced icosphere + spheres on a plane
 *                     (the scenes of a multi-scene batch).
 *   MT_GEN_FOREST     the stress scene: `grid` seeded procedural trees (branch
 *                     cylinders tessellated to 16-segment triangle tubes,
 *                     `subdiv` branching levels, leaf quads at the tips) on a
 *                     ground plane around a pond (a water box whose flat top
 *                     takes the wave normal; refraction, Beer-Lambert), sun
 *                     radius 0.5 degrees.
 *   MT_GEN_MATERIALS  the differentiable-rendering scene: a PBR ground plane
 *                     and a 9 x 7 grid of spheres, one GGX PBR material each
 *                     (64 materials: albedo U[0.1,0.9], roughness
 *                     U[0.1,0.9], metallic U[0,1]); the first n_emissive
 *                     spheres (default 6) also emit U[1,6] per channel and
 *                     are the scene's area lights.
 *
 * Every draw comes from a splitmix64 stream seeded by mt_gen_params.seed, so a
 * scene is a pure function of its parameters.
 */
#include "marmel_host.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint64_t s;
} rng_t;

static uint64_t next64(rng_t *r)
{
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uni(rng_t *r, double lo, double hi)
{
    return lo + (hi - lo) * ((double)(next64(r) >> 11) * (1.0 / 9007199254740992.0));
}

static mh_vec3 V(double x, double y, double z)
{
    mh_vec3 v = { x, y, z };
    return v;
}

static double f32(double x) { return (double)(float)x; }

static mh_prim sphere(double x, double y, double z, double r, int mat)
{
    mh_prim p;
    memset(&p, 0, sizeof p);
    p.kind = MH_PRIM_SPHERE;
    p.material = mat;
    p.p0 = V(f32(x), f32(y), f32(z));
    p.r0 = f32(r);
    return p;
}

static mh_material random_material(rng_t *r, int type)
{
    mh_material m;
    mt_material_default(&m);
    switch (type) {
    case 0: /* Lambert */
        m.albedo = V(uni(r, 0.15, 0.9), uni(r, 0.15, 0.9), uni(r, 0.15, 0.9));
        m.specular = V(0.04, 0.04, 0.04);
        m.shininess = 16.0;
        break;
    case 1: /* mirror */
        m.albedo = V(0.9, 0.9, 0.9);
        m.specular = V(1.0, 1.0, 1.0);
        m.shininess = 256.0;
        m.reflectivity = 1.0;
        break;
    case 2: /* glass, IOR 1.5 */
        mt_material_preset("glass", &m);
        break;
    default: /* GGX PBR */
        m.pbr = 1;
        m.albedo = V(uni(r, 0.1, 0.95), uni(r, 0.1, 0.95), uni(r, 0.1, 0.95));
        m.metallic = uni(r, 0.0, 1.0);
        m.roughness = uni(r, 0.05, 0.9);
        m.specular = V(0.04, 0.04, 0.04);
        m.shininess = 32.0;
        break;
    }
    return m;
}

/* ---- heightfield noise (value-gradient fBm of the generator's own) ------ */

static uint32_t hmix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x21f0aaadu;
    x ^= x >> 15;
    x *= 0x735a2d97u;
    x ^= x >> 15;
    return x;
}

static double grad2(int ix, int iz, uint32_t seed, double dx, double dz)
{
    uint32_t h = hmix((uint32_t)ix * 0x8da6b343u ^ (uint32_t)iz * 0xd8163841u ^ seed);
    double a = (double)(h >> 8) * (6.283185307179586 / 16777216.0);
    return cos(a) * dx + sin(a) * dz;
}

static double gnoise(double x, double z, uint32_t seed)
{
    double fx = floor(x), fz = floor(z);
    int ix = (int)fx, iz = (int)fz;
    double rx = x - fx, rz = z - fz;
    double u = rx * rx * rx * (rx * (rx * 6.0 - 15.0) + 10.0);
    double v = rz * rz * rz * (rz * (rz * 6.0 - 15.0) + 10.0);
    double a = grad2(ix, iz, seed, rx, rz), b = grad2(ix + 1, iz, seed, rx - 1.0, rz);
    double c = grad2(ix, iz + 1, seed, rx, rz - 1.0), d = grad2(ix + 1, iz + 1, seed, rx - 1.0, rz - 1.0);
    double x0 = a + u * (b - a), x1 = c + u * (d - c);
    return x0 + v * (x1 - x0);
}

static double terrain_h(double x, double z, uint32_t seed)
{
    double amp = 1.0, f = 1.0 / 28.0, sum = 0.0;
    for (int o = 0; o < 6; ++o) {
        sum += amp * gnoise(x * f, z * f, seed + (uint32_t)o * 101u);
        amp *= 0.5;
        f *= 2.0;
    }
    /* gentle bowl so the far edge rises into view */
    double r2 = (x * x + z * z) / (80.0 * 80.0);
    return 9.0 * sum + 6.0 * r2 - 1.0;
}

static int gen_spheres(mt_scene *s, const mt_gen_params *p, rng_t *r, mt_gen_info *info)
{
    int nsph = p->n_spheres > 0 ? p->n_spheres : 64;
    mh_material g;
    mt_material_default(&g);
    g.albedo = V(0.5, 0.5, 0.45);
    g.specular = V(0.02, 0.02, 0.02);
    g.shininess = 8.0;
    int gm = mt_scene_add_material(s, &g);
    mh_prim plane;
    memset(&plane, 0, sizeof plane);
    plane.kind = MH_PRIM_PLANE;
    plane.material = gm;
    plane.p0 = V(0.0, 0.0, 0.0);
    plane.p1 = V(0.0, 1.0, 0.0);
    if (mt_scene_add_prims(s, &plane, NULL, 1) < 0) return -1;
    for (int i = 0; i < nsph; ++i) {
        int type = (int)(next64(r) % 4u);
        mh_material m = random_material(r, type);
        int mi = mt_scene_add_material(s, &m);
        double x = uni(r, -10.0, 10.0), z = uni(r, -10.0, 10.0), y = uni(r, 0.0, 4.0);
        double rad = uni(r, 0.2, 1.5);
        mh_prim sp = sphere(x, y, z, rad, mi);
        if (mi < 0 || mt_scene_add_prims(s, &sp, NULL, 1) < 0) return -1;
    }
    if (info) {
        info->eye[0] = 0.0; info->eye[1] = 5.5; info->eye[2] = 23.0;
        info->at[0] = 0.0; info->at[1] = 1.0; info->at[2] = 0.0;
        info->vfov = 45.0;
        info->n_spheres = nsph;
        info->n_prims = nsph + 1;
    }
    return 0;
}

static int gen_terrain(mt_scene *s, const mt_gen_params *p, rng_t *r, mt_gen_info *info)
{
    int n = p->grid > 1 ? p->grid : 1024;
    int nsph = p->n_spheres > 0 ? p->n_spheres : 200;
    int nem = p->n_emissive > 0 ? p->n_emissive : (p->n_emissive < 0 ? 0 : 8);
    if (nem > nsph) nem = nsph;
    const double S = 80.0;
    uint32_t tseed = (uint32_t)next64(r);

    /* spheres first: the emitters lead the file so the reference's light
     * collection (first 8 emissive PBR spheres) takes exactly these */
    for (int i = 0; i < nsph; ++i) {
        mh_material m;
        mt_material_default(&m);
        m.pbr = 1;
        double x = uni(r, -0.8 * S, 0.8 * S), z = uni(r, -0.8 * S, 0.8 * S);
        double rad, y;
        if (i < nem) {
            double radiance = uni(r, 2.0, 20.0);
            double cr = uni(r, 0.6, 1.0), cg = uni(r, 0.6, 1.0), cb = uni(r, 0.6, 1.0);
            m.albedo = V(cr, cg, cb);
            m.emissive = V(radiance * cr, radiance * cg, radiance * cb);
            m.roughness = 1.0;
            rad = uni(r, 0.5, 1.5);
            y = terrain_h(x, z, tseed) + uni(r, 2.5, 7.0);
        } else {
            m.albedo = V(uni(r, 0.1, 0.95), uni(r, 0.1, 0.95), uni(r, 0.1, 0.95));
            m.metallic = uni(r, 0.0, 1.0);
            m.roughness = uni(r, 0.05, 0.9);
            rad = uni(r, 0.3, 2.2);
            y = terrain_h(x, z, tseed) + 0.7 * rad;
        }
        int mi = mt_scene_add_material(s, &m);
        mh_prim sp = sphere(x, y, z, rad, mi);
        if (mi < 0 || mt_scene_add_prims(s, &sp, NULL, 1) < 0) return -1;
    }
    mh_material tm;
    mt_material_default(&tm);
    tm.pbr = 1;
    tm.albedo = V(0.34, 0.30, 0.24);
    tm.roughness = 0.85;
    tm.metallic = 0.0;
    int tmi = mt_scene_add_material(s, &tm);
    if (tmi < 0) return -1;

    double *hx = (double *)malloc(sizeof(double) * (size_t)n * (size_t)n);
    if (!hx) return -1;
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            double x = -S + 2.0 * S * (double)i / (double)(n - 1);
            double z = -S + 2.0 * S * (double)j / (double)(n - 1);
            hx[(size_t)j * n + i] = terrain_h(x, z, tseed);
        }
    int row = 2 * (n - 1);
    mh_prim *tri = (mh_prim *)calloc((size_t)row, sizeof(mh_prim));
    if (!tri) {
        free(hx);
        return -1;
    }
    for (int j = 0; j < n - 1; ++j) {
        for (int i = 0; i < n - 1; ++i) {
            double x0 = f32(-S + 2.0 * S * (double)i / (double)(n - 1));
            double x1 = f32(-S + 2.0 * S * (double)(i + 1) / (double)(n - 1));
            double z0 = f32(-S + 2.0 * S * (double)j / (double)(n - 1));
            double z1 = f32(-S + 2.0 * S * (double)(j + 1) / (double)(n - 1));
            double h00 = f32(hx[(size_t)j * n + i]), h10 = f32(hx[(size_t)j * n + i + 1]);
            double h01 = f32(hx[(size_t)(j + 1) * n + i]), h11 = f32(hx[(size_t)(j + 1) * n + i + 1]);
            mh_prim *a = &tri[2 * i], *b = &tri[2 * i + 1];
            a->kind = b->kind = MH_PRIM_TRIANGLE;
            a->material = b->material = tmi;
            a->p0 = V(x0, h00, z0); a->p1 = V(x0, h01, z1); a->p2 = V(x1, h10, z0);
            b->p0 = V(x1, h10, z0); b->p1 = V(x0, h01, z1); b->p2 = V(x1, h11, z1);
        }
        if (mt_scene_add_prims(s, tri, NULL, row) < 0) {
            free(tri);
            free(hx);
            return -1;
        }
    }
    free(tri);
    free(hx);
    if (info) {
        info->eye[0] = 0.0; info->eye[1] = 30.0; info->eye[2] = 92.0;
        info->at[0] = 0.0; info->at[1] = 2.0; info->at[2] = 0.0;
        info->vfov = 48.0;
        info->n_spheres = nsph;
        info->n_emissive = nem;
        info->n_triangles = row * (n - 1);
        info->n_prims = nsph + row * (n - 1);
    }
    return 0;
}

/* icosphere with (seeded) radial displacement, subdivided `level` times */
static int gen_icosphere(mt_scene *s, const mt_gen_params *p, rng_t *r, mt_gen_info *info)
{
    int level = p->subdiv > 0 ? p->subdiv : 3;
    if (level > 8) level = 8;
    int nsph = p->n_spheres > 0 ? p->n_spheres : 20;
    const double t = (1.0 + sqrt(5.0)) / 2.0;
    double base[12][3] = { { -1, t, 0 }, { 1, t, 0 }, { -1, -t, 0 }, { 1, -t, 0 }, { 0, -1, t }, { 0, 1, t },
                           { 0, -1, -t }, { 0, 1, -t }, { t, 0, -1 }, { t, 0, 1 }, { -t, 0, -1 }, { -t, 0, 1 } };
    int faces[20][3] = { { 0, 11, 5 }, { 0, 5, 1 }, { 0, 1, 7 }, { 0, 7, 10 }, { 0, 10, 11 }, { 1, 5, 9 }, { 5, 11, 4 },
                         { 11, 10, 2 }, { 10, 7, 6 }, { 7, 1, 8 }, { 3, 9, 4 }, { 3, 4, 2 }, { 3, 2, 6 }, { 3, 6, 8 },
                         { 3, 8, 9 }, { 4, 9, 5 }, { 2, 4, 11 }, { 6, 2, 10 }, { 8, 6, 7 }, { 9, 8, 1 } };
    long ntri = 20L << (2 * level);
    double *tri = (double *)malloc(sizeof(double) * 9 * (size_t)ntri);
    double *tmp = (double *)malloc(sizeof(double) * 9 * (size_t)ntri);
    if (!tri || !tmp) {
        free(tri);
        free(tmp);
        return -1;
    }
    for (int f = 0; f < 20; ++f)
        for (int v = 0; v < 3; ++v) {
            double *q = base[faces[f][v]];
            double l = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            for (int k = 0; k < 3; ++k) tri[9 * f + 3 * v + k] = q[k] / l;
        }
    long cur = 20;
    for (int lv = 0; lv < level; ++lv) {
        for (long f = 0; f < cur; ++f) {
            double *a = &tri[9 * f], *b = a + 3, *c = a + 6;
            double m[3][3];
            for (int k = 0; k < 3; ++k) {
                m[0][k] = 0.5 * (a[k] + b[k]);
                m[1][k] = 0.5 * (b[k] + c[k]);
                m[2][k] = 0.5 * (c[k] + a[k]);
            }
            for (int e = 0; e < 3; ++e) {
                double l = sqrt(m[e][0] * m[e][0] + m[e][1] * m[e][1] + m[e][2] * m[e][2]);
                for (int k = 0; k < 3; ++k) m[e][k] /= l;
            }
            double *o = &tmp[36 * f];
            const double *ts[4][3] = { { a, m[0], m[2] }, { m[0], b, m[1] }, { m[2], m[1], c }, { m[0], m[1], m[2] } };
            for (int q = 0; q < 4; ++q)
                for (int v = 0; v < 3; ++v)
                    for (int k = 0; k < 3; ++k) o[9 * q + 3 * v + k] = ts[q][v][k];
        }
        cur *= 4;
        double *sw = tri;
        tri = tmp;
        tmp = sw;
    }
    free(tmp);
    uint32_t dseed = (uint32_t)next64(r);
    double R = uni(r, 2.0, 3.5), amp = uni(r, 0.1, 0.35);
    double cx = uni(r, -2.0, 2.0), cz = uni(r, -2.0, 2.0);
    double cy = R + 0.5;

    mh_material g;
    mt_material_default(&g);
    g.albedo = V(0.5, 0.5, 0.5);
    g.specular = V(0.02, 0.02, 0.02);
    g.shininess = 8.0;
    int gm = mt_scene_add_material(s, &g);
    mh_prim plane;
    memset(&plane, 0, sizeof plane);
    plane.kind = MH_PRIM_PLANE;
    plane.material = gm;
    plane.p1 = V(0.0, 1.0, 0.0);
    if (mt_scene_add_prims(s, &plane, NULL, 1) < 0) {
        free(tri);
        return -1;
    }
    for (int i = 0; i < nsph; ++i) {
        mh_material m = random_material(r, (int)(next64(r) % 4u));
        int mi = mt_scene_add_material(s, &m);
        double ang = uni(r, 0.0, 6.283185307179586), dist = uni(r, R + 1.5, R + 9.0);
        double rad = uni(r, 0.2, 1.0);
        mh_prim sp = sphere(cx + dist * cos(ang), rad, cz + dist * sin(ang), rad, mi);
        if (mi < 0 || mt_scene_add_prims(s, &sp, NULL, 1) < 0) {
            free(tri);
            return -1;
        }
    }
    mh_material im = random_material(r, 3);
    int imi = mt_scene_add_material(s, &im);
    mh_prim *pp = (mh_prim *)calloc((size_t)cur, sizeof(mh_prim));
    if (!pp || imi < 0) {
        free(tri);
        free(pp);
        return -1;
    }
    for (long f = 0; f < cur; ++f) {
        double v[3][3];
        for (int q = 0; q < 3; ++q) {
            double *u = &tri[9 * f + 3 * q];
            double disp = 1.0 + amp * gnoise(u[0] * 3.0 + u[2] * 1.7, u[1] * 3.0 - u[2] * 2.3, dseed);
            for (int k = 0; k < 3; ++k) v[q][k] = u[k] * R * disp;
        }
        pp[f].kind = MH_PRIM_TRIANGLE;
        pp[f].material = imi;
        pp[f].p0 = V(f32(cx + v[0][0]), f32(cy + v[0][1]), f32(cz + v[0][2]));
        pp[f].p1 = V(f32(cx + v[1][0]), f32(cy + v[1][1]), f32(cz + v[1][2]));
        pp[f].p2 = V(f32(cx + v[2][0]), f32(cy + v[2][1]), f32(cz + v[2][2]));
    }
    int rc = mt_scene_add_prims(s, pp, NULL, (int)cur) < 0 ? -1 : 0;
    /* further icospheres (count > 1): rows of four behind the first, each
     * with its own radius, displacement and material; drawn after everything
     * above, so a scene of one icosphere is unchanged */
    int count = p->count > 1 ? p->count : 1;
    for (int k = 1; k < count && rc == 0; ++k) {
        uint32_t ks = (uint32_t)next64(r);
        double Rk = uni(r, 1.0, 2.5), ak = uni(r, 0.1, 0.35);
        int col = (k - 1) % 4, rowk = (k - 1) / 4;
        double kx = cx + (col - 1.5) * 6.0 + uni(r, -0.5, 0.5);
        double kz = cz - R - 4.0 - rowk * 6.0 + uni(r, -0.5, 0.5);
        double ky = Rk + 0.3;
        mh_material km = random_material(r, (int)(next64(r) % 4u));
        int kmi = mt_scene_add_material(s, &km);
        if (kmi < 0) {
            rc = -1;
            break;
        }
        for (long f = 0; f < cur; ++f) {
            double v[3][3];
            for (int q = 0; q < 3; ++q) {
                double *u = &tri[9 * f + 3 * q];
                double disp = 1.0 + ak * gnoise(u[0] * 3.0 + u[2] * 1.7, u[1] * 3.0 - u[2] * 2.3, ks);
                for (int c = 0; c < 3; ++c) v[q][c] = u[c] * Rk * disp;
            }
            pp[f].kind = MH_PRIM_TRIANGLE;
            pp[f].material = kmi;
            pp[f].p0 = V(f32(kx + v[0][0]), f32(ky + v[0][1]), f32(kz + v[0][2]));
            pp[f].p1 = V(f32(kx + v[1][0]), f32(ky + v[1][1]), f32(kz + v[1][2]));
            pp[f].p2 = V(f32(kx + v[2][0]), f32(ky + v[2][1]), f32(kz + v[2][2]));
        }
        rc = mt_scene_add_prims(s, pp, NULL, (int)cur) < 0 ? -1 : 0;
    }
    free(pp);
    free(tri);
    if (info) {
        info->eye[0] = cx + 2.0; info->eye[1] = cy + 4.0; info->eye[2] = cz + R + 14.0;
        info->at[0] = cx; info->at[1] = cy * 0.6; info->at[2] = cz;
        info->vfov = 45.0;
        info->n_spheres = nsph;
        info->n_triangles = (int)(cur * count);
        info->n_prims = nsph + 1 + (int)(cur * count);
    }
    return rc;
}

/* ---- procedural trees ------------------------------------------------------ */

typedef struct {
    mh_prim *p;
    long n, cap;
} tribuf;

static int tb_push(tribuf *b, mh_vec3 a, mh_vec3 c, mh_vec3 d, int mat)
{
    if (b->n == b->cap) {
        long cap = b->cap ? 2 * b->cap : 4096;
        mh_prim *q = (mh_prim *)realloc(b->p, sizeof(mh_prim) * (size_t)cap);
        if (!q) return -1;
        b->p = q;
        b->cap = cap;
    }
    mh_prim *t = &b->p[b->n++];
    memset(t, 0, sizeof *t);
    t->kind = MH_PRIM_TRIANGLE;
    t->material = mat;
    t->p0 = V(f32(a.x), f32(a.y), f32(a.z));
    t->p1 = V(f32(c.x), f32(c.y), f32(c.z));
    t->p2 = V(f32(d.x), f32(d.y), f32(d.z));
    return 0;
}

static mh_vec3 vadd(mh_vec3 a, mh_vec3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static mh_vec3 vsc(mh_vec3 a, double k) { return V(a.x * k, a.y * k, a.z * k); }
static mh_vec3 vcross(mh_vec3 a, mh_vec3 b)
{
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static mh_vec3 vnorm(mh_vec3 a)
{
    double l = sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return l > 0.0 ? vsc(a, 1.0 / l) : V(0.0, 1.0, 0.0);
}

/* Frame (u, v) perpendicular to unit axis w. */
static void perp(mh_vec3 w, mh_vec3 *u, mh_vec3 *v)
{
    mh_vec3 a = fabs(w.y) < 0.9 ? V(0.0, 1.0, 0.0) : V(1.0, 0.0, 0.0);
    *u = vnorm(vcross(a, w));
    *v = vcross(w, *u);
}

/* Open tapered tube from a (radius ra) to b (radius rb), 16 segments. */
static int tube(tribuf *tb, mh_vec3 a, mh_vec3 b, double ra, double rb, int mat)
{
    mh_vec3 w = vnorm(vadd(b, vsc(a, -1.0))), u, v;
    perp(w, &u, &v);
    for (int k = 0; k < 16; ++k) {
        double t0 = 6.283185307179586 * k / 16.0, t1 = 6.283185307179586 * (k + 1) / 16.0;
        mh_vec3 d0 = vadd(vsc(u, cos(t0)), vsc(v, sin(t0))), d1 = vadd(vsc(u, cos(t1)), vsc(v, sin(t1)));
        mh_vec3 a0 = vadd(a, vsc(d0, ra)), a1 = vadd(a, vsc(d1, ra));
        mh_vec3 b0 = vadd(b, vsc(d0, rb)), b1 = vadd(b, vsc(d1, rb));
        if (tb_push(tb, a0, a1, b1, mat) || tb_push(tb, a0, b1, b0, mat)) return -1;
    }
    return 0;
}

static int branch(tribuf *tb, rng_t *r, mh_vec3 base, mh_vec3 dir, double len, double rad, int level, int bark,
                  int leaf)
{
    mh_vec3 tip = vadd(base, vsc(dir, len));
    if (tube(tb, base, tip, rad, rad * 0.7, bark)) return -1;
    if (level <= 0) {
        /* leaf quads around the tip */
        for (int i = 0; i < 10; ++i) {
            mh_vec3 c = vadd(tip, V(uni(r, -0.5, 0.5) * len, uni(r, -0.2, 0.5) * len, uni(r, -0.5, 0.5) * len));
            mh_vec3 n = vnorm(V(uni(r, -1.0, 1.0), uni(r, -1.0, 1.0), uni(r, -1.0, 1.0))), u, v;
            perp(n, &u, &v);
            double s = 0.12 + 0.1 * uni(r, 0.0, 1.0);
            mh_vec3 p0 = vadd(c, vadd(vsc(u, -s), vsc(v, -s))), p1 = vadd(c, vadd(vsc(u, s), vsc(v, -s)));
            mh_vec3 p2 = vadd(c, vadd(vsc(u, s), vsc(v, s))), p3 = vadd(c, vadd(vsc(u, -s), vsc(v, s)));
            if (tb_push(tb, p0, p1, p2, leaf) || tb_push(tb, p0, p2, p3, leaf)) return -1;
        }
        return 0;
    }
    int kids = 3;
    mh_vec3 u, v;
    perp(dir, &u, &v);
    double phase = uni(r, 0.0, 6.283185307179586);
    for (int k = 0; k < kids; ++k) {
        double az = phase + 6.283185307179586 * k / kids + uni(r, -0.3, 0.3);
        double tilt = uni(r, 0.45, 0.9);
        mh_vec3 side = vadd(vsc(u, cos(az)), vsc(v, sin(az)));
        mh_vec3 nd = vnorm(vadd(vsc(dir, cos(tilt)), vsc(side, sin(tilt))));
        nd = vnorm(vadd(nd, V(0.0, 0.25, 0.0)));
        if (branch(tb, r, tip, nd, len * uni(r, 0.62, 0.78), rad * 0.62, level - 1, bark, leaf)) return -1;
    }
    return 0;
}

static int gen_forest(mt_scene *s, const mt_gen_params *p, rng_t *r, mt_gen_info *info)
{
    int ntrees = p->grid > 0 ? p->grid : 8;
    int levels = p->subdiv > 0 ? p->subdiv : 4;
    if (levels > 9) levels = 9;
    double radius = 6.0 + 2.0 * sqrt((double)ntrees); /* forest extent */
    mh_material ground;
    mt_material_default(&ground);
    ground.albedo = V(0.32, 0.30, 0.22);
    ground.specular = V(0.02, 0.02, 0.02);
    ground.shininess = 8.0;
    mh_material bark;
    mt_material_default(&bark);
    bark.albedo = V(0.28, 0.18, 0.10);
    bark.specular = V(0.02, 0.02, 0.02);
    bark.shininess = 8.0;
    mh_material leaf;
    mt_material_default(&leaf);
    leaf.pbr = 1;
    leaf.albedo = V(0.18, 0.42, 0.10);
    leaf.roughness = 0.55;
    leaf.metallic = 0.0;
    mh_material water;
    mt_material_preset("water", &water);
    int gm = mt_scene_add_material(s, &ground), bm = mt_scene_add_material(s, &bark);
    int lm = mt_scene_add_material(s, &leaf), wm = mt_scene_add_material(s, &water);
    if (gm < 0 || bm < 0 || lm < 0 || wm < 0) return -1;
    mh_prim plane;
    memset(&plane, 0, sizeof plane);
    plane.kind = MH_PRIM_PLANE;
    plane.material = gm;
    plane.p0 = V(0.0, -1.0, 0.0); /* pond bed and ground level */
    plane.p1 = V(0.0, 1.0, 0.0);
    /* the pond: a water box of depth 0.98 whose top (y = 0) is the water
     * level. Its bottom stays clear of the bed: a face coincident with the
     * plane would tie with it on every ray through the water, and which of
     * the two is nearer would then turn on the last bit of the ray. */
    mh_prim pond;
    memset(&pond, 0, sizeof pond);
    pond.kind = MH_PRIM_BOX;
    pond.material = wm;
    pond.p0 = V(0.0, -0.49, 0.0);
    pond.p1 = V(4.0, 0.49, 3.0);
    if (mt_scene_add_prims(s, &plane, NULL, 1) < 0 || mt_scene_add_prims(s, &pond, NULL, 1) < 0) return -1;
    mt_scene_set_water(s, 0.0, wm);
    /* banks around the pond: a ring of ground tiles at y = 0 */
    tribuf tb = { NULL, 0, 0 };
    double R = radius + 4.0;
    mh_vec3 q[8][2] = { { V(-R, 0, -R), V(R, 0, -3) }, { V(-R, 0, 3), V(R, 0, R) },
                        { V(-R, 0, -3), V(-4, 0, 3) }, { V(4, 0, -3), V(R, 0, 3) } };
    for (int i = 0; i < 4; ++i) {
        mh_vec3 a = q[i][0], b = q[i][1];
        if (tb_push(&tb, V(a.x, 0.0, a.z), V(b.x, 0.0, b.z), V(b.x, 0.0, a.z), gm) ||
            tb_push(&tb, V(a.x, 0.0, a.z), V(a.x, 0.0, b.z), V(b.x, 0.0, b.z), gm)) {
            free(tb.p);
            return -1;
        }
    }
    for (int t = 0; t < ntrees; ++t) {
        double ang = uni(r, 0.0, 6.283185307179586), dist = uni(r, 5.5, radius);
        mh_vec3 base = V(dist * cos(ang), 0.0, dist * sin(ang));
        double h = uni(r, 1.6, 2.6);
        mh_vec3 dir = vnorm(V(uni(r, -0.1, 0.1), 1.0, uni(r, -0.1, 0.1)));
        if (branch(&tb, r, base, dir, h, 0.16 + 0.06 * uni(r, 0.0, 1.0), levels, bm, lm)) {
            free(tb.p);
            return -1;
        }
    }
    int rc = mt_scene_add_prims(s, tb.p, NULL, (int)tb.n) < 0 ? -1 : 0;
    mh_sky sky;
    mt_sky_default(&sky);
    sky.sun_radius = 0.5;
    mt_scene_set_sky(s, &sky);
    if (info) {
        info->eye[0] = 0.0; info->eye[1] = 4.5; info->eye[2] = radius + 6.0;
        info->at[0] = 0.0; info->at[1] = 1.0; info->at[2] = 0.0;
        info->vfov = 55.0;
        info->n_triangles = (int)tb.n;
        info->n_prims = (int)tb.n + 2;
    }
    free(tb.p);
    return rc;
}

static mh_material pbr_material(rng_t *r)
{
    mh_material m;
    mt_material_default(&m);
    m.pbr = 1;
    m.albedo = V(uni(r, 0.1, 0.9), uni(r, 0.1, 0.9), uni(r, 0.1, 0.9));
    m.roughness = uni(r, 0.1, 0.9);
    m.metallic = uni(r, 0.0, 1.0);
    m.specular = V(0.04, 0.04, 0.04);
    m.shininess = 32.0;
    return m;
}

static int gen_materials(mt_scene *s, const mt_gen_params *p, rng_t *r, mt_gen_info *info)
{
    int nem = p->n_emissive > 0 ? p->n_emissive : 6;
    if (nem > 8) nem = 8;
    mh_material g = pbr_material(r);
    g.metallic = 0.1 * g.metallic; /* a mostly dielectric ground, inside (0, 1) */
    int gm = mt_scene_add_material(s, &g);
    mh_prim plane;
    memset(&plane, 0, sizeof plane);
    plane.kind = MH_PRIM_PLANE;
    plane.material = gm;
    plane.p1 = V(0.0, 1.0, 0.0);
    if (gm < 0 || mt_scene_add_prims(s, &plane, NULL, 1) < 0) return -1;
    int k = 0;
    for (int row = 0; row < 7; ++row)
        for (int col = 0; col < 9; ++col, ++k) {
            mh_material m = pbr_material(r);
            if (k < nem) m.emissive = V(uni(r, 1.0, 6.0), uni(r, 1.0, 6.0), uni(r, 1.0, 6.0));
            int mi = mt_scene_add_material(s, &m);
            double rad = uni(r, 0.45, 0.8);
            double x = (col - 4) * 2.0 + uni(r, -0.3, 0.3), z = (row - 3) * 2.0 + uni(r, -0.3, 0.3);
            mh_prim sp = sphere(x, rad, z, rad, mi);
            if (mi < 0 || mt_scene_add_prims(s, &sp, NULL, 1) < 0) return -1;
        }
    if (info) {
        info->eye[0] = 0.0; info->eye[1] = 9.0; info->eye[2] = 14.0;
        info->at[0] = 0.0; info->at[1] = 0.0; info->at[2] = 0.5;
        info->vfov = 50.0;
        info->n_spheres = 63;
        info->n_emissive = nem;
        info->n_prims = 64;
    }
    return 0;
}

int mt_scene_generate(mt_scene *s, const mt_gen_params *p, mt_gen_info *info)
{
    if (!s || !p) return 1;
    rng_t r = { 0x243F6A8885A308D3ull ^ ((uint64_t)p->seed * 0x9E3779B97F4A7C15ull) };
    if (info) {
        memset(info, 0, sizeof *info);
        info->up[1] = 1.0;
    }
    mh_sky sky;
    mt_sky_default(&sky);
    mt_scene_set_sky(s, &sky);
    switch (p->kind) {
    case MT_GEN_SPHERES: return gen_spheres(s, p, &r, info);
    case MT_GEN_TERRAIN: return gen_terrain(s, p, &r, info);
    case MT_GEN_ICOSPHERE: return gen_icosphere(s, p, &r, info);
    case MT_GEN_MATERIALS: return gen_materials(s, p, &r, info);
    case MT_GEN_FOREST: return gen_forest(s, p, &r, info);
    default: return 2;
    }
}
