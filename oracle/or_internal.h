/*
 * or_internal.h - internal types of the CPU oracle (FP64 restatement of the
 * marmel-trace path-tracing hot path). TEST INFRASTRUCTURE ONLY: the product
 * path never includes, links or calls anything under oracle/.
 *
 * The arithmetic below is written to perform the same IEEE-754 double
 * operations, in the same order, as the reference (vec3.h/vec3.c), so the
 * oracle reproduces the reference images byte for byte.
 */
#ifndef OR_INTERNAL_H
#define OR_INTERNAL_H

#include "marmel_hip.h"

#include <math.h>
#include <stdint.h>

typedef struct { double x, y, z; } dv3;

static inline dv3 d3(double x, double y, double z) { dv3 r = { x, y, z }; return r; }
static inline dv3 dadd(dv3 a, dv3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline dv3 dsub(dv3 a, dv3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline dv3 dneg(dv3 a) { return d3(-a.x, -a.y, -a.z); }
static inline dv3 dscale(dv3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
static inline dv3 dmul(dv3 a, dv3 b) { return d3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline double ddot(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline dv3 dcross(dv3 a, dv3 b)
{
    return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double dlen2(dv3 a) { return ddot(a, a); }
static inline double dlen(dv3 a) { return sqrt(dlen2(a)); }
/* lerp as a + (b - a) * t (reference src/vec3.h vec3_lerp) */
static inline dv3 dlerp(dv3 a, dv3 b, double t)
{
    return d3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t);
}
/* i - (2 d) n (reference src/vec3.h vec3_reflect) */
static inline dv3 dreflect(dv3 i, dv3 n)
{
    double d = ddot(i, n);
    return d3(i.x - 2.0 * d * n.x, i.y - 2.0 * d * n.y, i.z - 2.0 * d * n.z);
}
static inline dv3 dat(dv3 o, dv3 d, double t) { return d3(o.x + t * d.x, o.y + t * d.y, o.z + t * d.z); }
static inline dv3 dmin3(dv3 a, dv3 b) { return d3(a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z); }
static inline dv3 dmax3(dv3 a, dv3 b) { return d3(a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z); }
/* zero vector when |a|^2 <= 1e-12, else a * (1/|a|) (reference src/vec3.c:26-34) */
static inline dv3 dnorm(dv3 a)
{
    double l2 = dlen2(a);
    if (l2 <= 1e-12) return d3(0.0, 0.0, 0.0);
    return dscale(a, 1.0 / sqrt(l2));
}
static inline int dfinite(dv3 v) { return isfinite(v.x) && isfinite(v.y) && isfinite(v.z); }
static inline dv3 fromm(mh_vec3 v) { return d3(v.x, v.y, v.z); }

#define OR_PI 3.14159265358979323846

typedef struct { dv3 o, d; } oray;

typedef struct {
    double t;
    dv3 point, normal;
    int material, prim, front;
    double u, v;
} ohit;

typedef struct {
    int kind, material, flags;
    dv3 p0, p1, p2;          /* see mh_prim */
    double r0, r1;
    dv3 e1, e2;              /* triangle edges, precomputed like prim_triangle */
    dv3 n0, n1, n2;          /* smooth normals */
    dv3 uv0, uv1, uv2;
} oprim;

typedef struct {
    dv3 bmin, bmax;
    int left, right, first, count;
} onode;

typedef struct {
    int prim;
    int mat;
    dv3 center;
    double radius;
    dv3 emissive;
} olight;

typedef struct or_scene {
    int n_prims;
    oprim *prims;
    int n_mats;
    mh_material *mats;
    mh_sky sky;
    mh_fog fog;
    mh_ocean ocean;
    double water_level;
    int water_material;
    int has_transmissive;   /* some material is water or transparency > 0.8 */
    mh_material *freeze;    /* non-NULL: sampling decisions use these materials */
    int n_lights;
    olight lights[8];
    /* binary SAH BVH restated from the reference build */
    onode *nodes;
    int n_nodes, cap_nodes;
    int *order;          /* tree primitive indices in leaf order */
    int n_order;
    oprim *leaf_prims;   /* primitive copies in leaf order */
    int *planes;
    int n_planes;
} or_scene;

/* traversal statistics (thread-local), used to price algorithmic bytes:
 * per query kind, interior nodes whose two child boxes were tested and
 * primitives tested (all / triangles), plus plane tests and path bounces */
typedef struct {
    long long closest_queries, any_queries;
    long long c_inner, c_prims, c_tris;
    long long a_inner, a_prims, a_tris;
    long long plane_tests, ext_queries;
} or_stats;
#define OR_NSTATS 10
extern __thread or_stats or_tls_stats;
/* 1: the path tracer's shadow segment runs as one any-hit query when the
 * scene has no transmissive material (same answer: the first hit would be
 * opaque), so the counts price the GPU's traversal; 0: the reference's
 * closest-hit walk. */
extern int or_anyhit_shadow;

/* or_math.c */
unsigned or_hash_u32(unsigned x);
unsigned or_hash3(unsigned a, unsigned b, unsigned c);
double or_rand01(unsigned a, unsigned b, unsigned c);
double or_perlin2(double x, double y, unsigned seed);
double or_fbm2(double x, double y, int octaves, double lac, double gain, unsigned seed);
double or_perlin3(double x, double y, double z, unsigned seed);
double or_fbm3(double x, double y, double z, int octaves, double lac, double gain, unsigned seed);
void or_basis(dv3 n, dv3 *t, dv3 *b);
dv3 or_cosine_hemisphere(dv3 n, double u1, double u2);
double or_fresnel_schlick(double c, double f0);
dv3 or_fresnel_schlick_rgb(double c, dv3 f0);
double or_fresnel_dielectric(double cos_i, double eta);
int or_refract(dv3 incident, dv3 n, double eta, dv3 *out);
double or_specular_probability(double metallic, double roughness, double f0);
dv3 or_ggx_half(dv3 n, double alpha, double u1, double u2);
double or_ggx_pdf(double ndh, double hdv, double alpha);
double or_d_ggx(double ndh, double a);

/* or_geom.c */
int or_prim_intersect(const oprim *p, oray r, double tmin, double tmax, ohit *h);
int or_prim_occluded(const oprim *p, oray r, double tmin, double tmax);
void or_prim_bounds(const oprim *p, dv3 *mn, dv3 *mx);
int or_bvh_build(or_scene *s);
int or_intersect1(const or_scene *s, oray r, double tmin, double tmax, ohit *h);
int or_occluded1(const or_scene *s, oray r, double tmin, double tmax);

/* or_shade.c */
double or_clamp01(double v);
dv3 or_shade_local(const mh_material *m, dv3 N, dv3 L, dv3 V, dv3 light);
dv3 or_shade_pbr(const mh_material *m, dv3 N, dv3 L, dv3 V, dv3 light);
double or_sky_cloud_alpha(dv3 dir, const mh_sky *sky);
double or_roughness_to_alpha(double r);
dv3 or_sample_glossy_dir(const mh_material *m, dv3 N, dv3 incident, double r1, double r2);
dv3 or_ambient(const mh_material *m, dv3 N, const mh_sky *sky);
dv3 or_sky_sample(dv3 dir, const mh_sky *sky);
dv3 or_sun_disk_dir(dv3 sun, double radius_deg, double r1, double r2);
dv3 or_light_cone_dir(dv3 w, double cos_max, double u1, double u2);
dv3 or_water_normal(double x, double z, double time);
dv3 or_water_attenuate(const mh_material *m, dv3 inner, double depth);
dv3 or_texture_albedo(const mh_material *m, dv3 p, double u, double v);

#endif /* OR_INTERNAL_H */
