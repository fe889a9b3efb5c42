/*
 * oracle.h - C API of liboracle.so, the CPU FP64 restatement of the
 * marmel-trace path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the GPU library
 * (libmarmel_hip) and the CPU baseline timed beside it. The product path never
 * includes, links or calls it.
 *
 * Parity status of each function is pinned by tests/test_oracle.py against the
 * reference itself (oracle/_ref) and against golden fixtures generated from
 * the reference (tests/golden/).
 */
#ifndef MARMEL_ORACLE_H
#define MARMEL_ORACLE_H

#include "marmel_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_scene or_scene;

/* Build from the flattened ABI scene. Uses prims/attrs/materials/sky/water;
 * ignores the host BVH and builds the reference's own binned-SAH tree. */
or_scene *or_scene_create(const mh_scene_soa *soa);
void or_scene_destroy(or_scene *s);
int or_scene_info(const or_scene *s, int *n_prims, int *n_nodes, int *n_lights);

/* Batch queries (reference scene_intersect / scene_occluded). */
int or_intersect(const or_scene *s, const mh_ray *rays, int n, double tmin,
                 double tmax, mh_hit *out);
int or_occluded(const or_scene *s, const mh_ray *rays, int n, double tmin,
                double tmax, unsigned char *out);

/* One path (reference pathtrace_radiance). */
int or_radiance(const or_scene *s, const mh_ray *ray, int depth, unsigned seed,
                double out[3]);

/*
 * Render rows [y0, y1) of a width x height image (y1 <= 0 means height).
 * mode 0 = path tracer (pathtrace_render), 1 = fixed-spp Whitted
 * (render_image). Samples sample_offset .. sample_offset+spp-1 of every pixel.
 * rgb (w*h*3) receives the tone-mapped mean; mean/var (w*h*3 doubles, may be
 * NULL) receive the per-pixel sample mean and unbiased sample variance.
 * threads <= 0 uses every online CPU (capped at 64). seconds receives the
 * wall time of the render loop only. Returns the reference codes 0/1/2/3.
 */
int or_render(const or_scene *s, const mh_camera *cam, int width, int height,
              int spp, int depth, int mode, int sample_offset, int y0, int y1,
              int threads, unsigned char *rgb, double *mean, double *var,
              double *seconds);

/* Adaptive fixed-batch Whitted (reference render_image_ex with adaptive on):
 * per-pixel linear means, 8-bit image and sample counts; total / max samples.
 * nmax <= 0: 4 n0; tau <= 0: 0.02. */
int or_render_adaptive(const or_scene *s, const mh_camera *cam, int width, int height, int n0, int depth,
                       int nmax, double tau, int threads, unsigned char *rgb, double *mean, int *nsamp,
                       long long *total, int *max_n);

/* Traversal counts accumulated by the calling thread's queries since the
 * last reset (used to price algorithmic bytes per sample). */
void or_stats_reset(void);
/* out[10]: closest queries, any-hit queries, then for closest and for any-hit
 * queries in turn: interior nodes tested, primitives tested, triangles
 * tested; then plane tests and path bounces (extension queries). */
void or_stats_get(long long out[10]);
/* Counts of a full render (all worker threads), filled by or_render. */
void or_render_stats_get(long long out[10]);
/* 1: price shadow segments as the GPU runs them (one any-hit query when no
 * material is transmissive; identical result). 0 (default): the reference's
 * closest-hit walk. */
void or_set_anyhit_shadow(int on);

/*
 * Path-replay adjoint of the path tracer in FP64 (the algorithm of the GPU
 * adjoint): grad[n_mats*8 + 3] = dLoss/dtheta for dimg = dLoss/d(pixel mean)
 * (w*h*3): per material albedo rgb, roughness, metallic, emissive rgb; then
 * sun colour rgb. Sampling decisions held fixed. Returns 4 while frozen.
 */
int or_adjoint(const or_scene *s, const mh_camera *cam, int width, int height, int spp,
               int depth, int sample_offset, const double *dimg, double *grad, int threads);

/* Parameter updates: materials (same count; the emissive-light set stays the
 * one built at creation) and sky. */
int or_scene_set_materials(or_scene *s, const mh_material *mats, int n);
int or_scene_set_sky(or_scene *s, const mh_sky *sky);

/*
 * Frozen sampling: every sampling decision (lobe choice and probability,
 * GGX half vector and pdf, Russian roulette on a throughput computed from
 * these materials) uses `mats`, the BSDF values the scene's current
 * materials. The path-traced image is then a smooth function of the
 * current material parameters whose derivative at mats is what the adjoint
 * computes (finite differences of it pin the adjoint exactly). NULL unfreezes.
 */
int or_set_freeze(or_scene *s, const mh_material *mats, int n);

/* Pure-function vectors. */
unsigned or_hash3(unsigned a, unsigned b, unsigned c);
double or_rand01(unsigned a, unsigned b, unsigned c);
void or_camera_create(const double eye[3], const double at[3], const double up[3],
                      double vfov, double aspect, mh_camera *out);
void or_camera_ray(const mh_camera *cam, double u, double v, double r1, double r2,
                   double out[6]);
void or_sky_samples(const or_scene *s, const double *dirs, int n, double *out);
void or_water_normals(const double *xz, int n, double *out);
void or_noise(const double *xy, int n, int octaves, unsigned seed, double *perlin,
              double *fbm);
/* 3D gradient noise and its 3-octave fBm (the fog turbulence) at n points. */
void or_noise3(const double *xyz, int n, unsigned seed, double *perlin, double *fbm);

/* Anamorphic streak flare on a top-down RGB8 image, in place (reference
 * flare_apply). 1: invalid arguments, 2: out of memory. */
int or_flare_apply(unsigned char *rgb, int width, int height, const mh_flare_params *p);
void or_sun_disk_dirs(const double sun[3], double radius_deg, const double *u,
                      int n, double *out);
void or_light_cone_dirs(const double w[3], double cos_max, const double *u, int n,
                        double *out);
void or_cosine_dirs(const double nrm[3], const double *u, int n, double *out);
void or_ggx_halfs(const double nrm[3], double alpha, const double *u, int n,
                  double *out);
double or_fresnel_dielectric_x(double cos_i, double eta);
double or_specular_probability_x(double metallic, double roughness, double f0);
void or_shade(const or_scene *s, int material, const double N[3], const double L[3],
              const double V[3], const double light[3], double out[3]);
unsigned char or_to_byte(double c);
/* Albedo of material m at n points p (3 doubles each) with surface uvs uv
 * (2 doubles each): textures NONE, CHECKER, UV_CHECKER (reference
 * texture_albedo_uv, src/texture.c:269-327). out: 3 doubles per point. */
void or_texture_albedos(const mh_material *m, const double *p, const double *uv, int n,
                        double *out);

#ifdef __cplusplus
}
#endif

#endif /* MARMEL_ORACLE_H */
