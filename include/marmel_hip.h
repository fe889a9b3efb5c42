/*
 * marmel_hip.h - C ABI of libmarmel_hip, the MI355X (gfx950) path-tracing
 * library that replaces the CPU renderers of marmel-trace.
 *
 * Everything here is plain C: fixed-width integers, doubles, floats and
 * pointers. No HIP or torch type crosses the boundary. The caller owns every
 * host buffer it passes in; the library owns all device memory.
 *
 * Scene input (mh_scene_soa) is the FLATTENED scene of the C11 host:
 *   - primitives in file order (the index into `prims` is the prim_index that
 *     breaks nearest-hit ties, as Geometry.prims does in
 *     reference src/geometry.h:68-72 / src/bvh.c:641-646),
 *   - the material table (reference src/material.h Material),
 *   - the sky (reference src/material.h SkyParams),
 *   - water metadata (reference src/scene.h:81-106),
 *   - the host-built binary BVH over the non-plane primitives (node array +
 *     leaf-order primitive index array) and the plane side list, as
 *     reference src/bvh.c:66-76 keeps them.
 * mh_scene_create() converts that into the device layout (fp32 child-box
 * nodes, leaf-ordered fp32 and fp64 primitive records) and uploads it once.
 *
 * Return codes keep the reference meanings (reference src/pathtrace.h:69-71):
 *   0 OK, 1 NULL pointer, 2 non-positive image dimensions,
 *   3 spp < 1 or depth < 0.
 * Library-specific failures use:
 *   4  scene uses a feature outside the GPU path (text in mh_last_error),
 *   5  invalid scene data (bad index, malformed BVH),
 *   100+ HIP runtime error (text in mh_last_error).
 *
 * Threading: calls on one mh_scene are synchronous on that scene's own HIP
 * stream and must not be issued from two threads at once; distinct scenes
 * may be used concurrently. mh_last_error / mh_last_seconds are per thread.
 */
#ifndef MARMEL_HIP_H
#define MARMEL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MH_ABI_VERSION 2

/* Opaque device-resident scene handle. */
struct mh_scene;
typedef struct mh_scene mh_scene;

/* ------------------------------------------------------------------ */
/* Scene input types                                                   */
/* ------------------------------------------------------------------ */

typedef struct mh_vec3 {
    double x, y, z;
} mh_vec3;

/* Primitive kinds (reference src/geometry.h:32-40; CSG and SDF are not on
 * the GPU path and are rejected with code 4). */
enum {
    MH_PRIM_SPHERE   = 0,
    MH_PRIM_PLANE    = 1,
    MH_PRIM_BOX      = 2,
    MH_PRIM_TRIANGLE = 3,
    MH_PRIM_CYLINDER = 4
};

/* mh_prim.flags */
enum {
    MH_PRIM_SMOOTH = 1, /* triangle: interpolate attrs[i].n0..n2 (Primitive.radius > 0.5) */
    MH_PRIM_UV     = 2  /* triangle: interpolate attrs[i].uv0..uv2 (Primitive.has_uv)     */
};

/*
 * One primitive (reference Primitive, src/geometry.h:42-66):
 *   SPHERE    p0 = center, r0 = radius
 *   PLANE     p0 = point on the plane, p1 = unit normal
 *   BOX       p0 = center, p1 = half extents
 *   TRIANGLE  p0, p1, p2 = vertices a, b, c
 *   CYLINDER  p0 = base centre, p1 = top centre, r0 = bottom radius, r1 = top radius
 */
typedef struct mh_prim {
    int32_t kind;
    int32_t material;
    int32_t flags;
    int32_t reserved;
    mh_vec3 p0, p1, p2;
    double  r0, r1;
} mh_prim;

/* Optional per-primitive vertex attributes (smooth normals, uvs). */
typedef struct mh_prim_attr {
    mh_vec3 n0, n1, n2;
    mh_vec3 uv0, uv1, uv2;
} mh_prim_attr;

/* Texture kinds on the GPU path (reference src/material.h TextureKind). */
enum {
    MH_TEXTURE_NONE       = 0,
    MH_TEXTURE_CHECKER    = 1,
    MH_TEXTURE_UV_CHECKER = 7
};

/* Surface material (reference Material, src/material.h). */
typedef struct mh_material {
    mh_vec3 albedo;
    mh_vec3 specular;
    double  shininess;
    double  reflectivity;
    double  transparency;
    double  ior;
    int32_t is_water;
    int32_t beer_lambert;
    int32_t pbr;
    int32_t texture_kind;
    mh_vec3 absorption;
    mh_vec3 deep_color;
    double  metallic;
    double  roughness;
    mh_vec3 emissive;
    double  texture_scale;
    mh_vec3 texture_color_a;
    mh_vec3 texture_color_b;
    mh_vec3 atmosphere_glow;
    double  bump_strength;   /* must be <= 1e-6 on the GPU path */
} mh_material;

/* Sky and sun (reference SkyParams, src/material.h). The opt-in volumetric
 * clouds, stars, nebula, galaxy and fog must be off (0) on the GPU path. */
typedef struct mh_sky {
    mh_vec3  sun_dir;
    mh_vec3  sun_color;
    mh_vec3  horizon_color;
    mh_vec3  zenith_color;
    double   gradient_gamma;
    double   sun_glow_exponent;
    double   sun_glow_strength;
    double   cloud_height;
    double   cloud_scale;
    double   cloud_coverage;
    double   cloud_softness;
    double   cloud_sharpness;
    int32_t  cloud_octaves;
    uint32_t seed;
    double   sun_radius;        /* degrees; 0 = point sun */
    double   cloud_thickness;   /* must be 0 */
    double   star_intensity;    /* must be 0 */
    double   nebula_intensity;  /* must be 0 */
    double   galaxy_intensity;  /* must be 0 */
    double   fog_density;       /* must be 0: fog is mh_scene_soa.fog */
} mh_sky;

/* Atmospheric fog (reference FogParams, src/material.h:199-209). density 0
 * disables it. shadow_steps > 0 marches camera-ray segments with a shadow
 * ray to the sun per step (god rays) and marches underwater segments; 0 uses
 * the closed form. */
typedef struct mh_fog {
    double  density;             /* extinction coefficient >= 0            */
    mh_vec3 color;               /* ambient in-scatter colour (linear)     */
    double  height;              /* reference height of the falloff        */
    double  height_falloff;      /* exponential decay with height, >= 0    */
    double  inscatter_strength;  /* sun in-scatter intensity >= 0          */
    double  sun_anisotropy;      /* Henyey-Greenstein g, clamped to 0.95   */
    double  noise_scale;         /* 3D fBm turbulence frequency (0 = none) */
    double  noise_amount;        /* turbulence strength in [0, 1]          */
    int32_t shadow_steps;        /* march steps (0 = closed form)          */
    int32_t reserved;
} mh_fog;

/* Procedural Gerstner ocean of the scene (reference SceneOceanParams,
 * src/scene.h; the first ocean block of a scene). When enabled, the water
 * height at (x, z) -- which decides whether a camera ray starts under water
 * -- is the 4-harmonic surface of src/scene.c:1109-1153 instead of
 * water_level. The ocean's triangles are ordinary prims (mt_scene_add_ocean). */
typedef struct mh_ocean {
    int32_t  enabled;
    uint32_t seed;               /* harmonic phases                        */
    mh_vec3  center;             /* center.y: base water level             */
    mh_vec3  direction;          /* swell direction in the XZ plane        */
    double   amplitude, wavelength, steepness, chop, chop_wavelength;
} mh_ocean;

/* Pinhole / thin-lens camera (reference Camera, src/camera.h:17-48). */
typedef struct mh_camera {
    mh_vec3 position;
    mh_vec3 lower_left;
    mh_vec3 horizontal;
    mh_vec3 vertical;
    double  aperture;
    double  focus_distance;
    double  dome_radius;
    double  anamorphic_squeeze;
} mh_camera;

/* Host BVH node (binary; reference BvhNode src/bvh.c:50-57 with fp32 boxes
 * rounded OUTWARD from the double bounds). Leaf: count > 0, primitives
 * leaf_prims[first .. first+count). Interior: count == 0, children left/right. */
typedef struct mh_bvh_node {
    float   bmin[3];
    float   bmax[3];
    int32_t left;
    int32_t right;
    int32_t first;
    int32_t count;
} mh_bvh_node;

typedef struct mh_scene_soa {
    int32_t             abi_version;     /* MH_ABI_VERSION */
    int32_t             n_prims;
    const mh_prim      *prims;           /* file order                       */
    const mh_prim_attr *attrs;           /* NULL, or n_prims entries         */
    int32_t             n_materials;
    const mh_material  *materials;
    mh_sky              sky;
    double              water_level;
    int32_t             water_material;  /* -1 = none                        */
    int32_t             n_nodes;         /* 0 when every primitive is a plane */
    const mh_bvh_node  *nodes;           /* node 0 is the root               */
    int32_t             n_leaf_prims;
    const int32_t      *leaf_prims;      /* prim indices in leaf order       */
    int32_t             n_planes;
    const int32_t      *planes;          /* plane prim indices, file order   */
    mh_fog              fog;             /* reference Scene.fog              */
    mh_ocean            ocean;           /* reference Scene.ocean_params     */
} mh_scene_soa;

/* ------------------------------------------------------------------ */
/* Ray query types                                                     */
/* ------------------------------------------------------------------ */

typedef struct mh_ray {
    double ox, oy, oz;
    double dx, dy, dz;
} mh_ray;

/* Nearest hit (reference Hit, src/geometry.h:74-83). prim_index -1 = miss. */
typedef struct mh_hit {
    double  t;
    int32_t prim_index;
    int32_t material;
    int32_t front_face;
    int32_t reserved;
    double  nx, ny, nz;
    double  u, v;
} mh_hit;

/* ------------------------------------------------------------------ */
/* Render parameters                                                   */
/* ------------------------------------------------------------------ */

enum {
    MH_MODE_PATHTRACE = 0,  /* reference pathtrace_render      */
    MH_MODE_WHITTED   = 1   /* reference render_image (fixed)  */
};

enum {
    MH_PREC_FP32 = 0,       /* fp32 geometry and shading (throughput path) */
    MH_PREC_FP64 = 1        /* fp64 throughout (parity path)               */
};

enum {
    MH_ENGINE_MEGAKERNEL = 0, /* one thread per pixel, whole path in one kernel */
    MH_ENGINE_WAVEFRONT  = 1  /* queue-based extend / shade / shadow kernels (fp32) */
};

typedef struct mh_render_params {
    int32_t width;
    int32_t height;
    int32_t spp;
    int32_t max_depth;
    int32_t mode;           /* MH_MODE_*  */
    int32_t precision;      /* MH_PREC_*  */
    int32_t sample_offset;  /* first sample index (progressive chunks) */
    int32_t engine;         /* MH_ENGINE_* (path tracing only)          */
    int32_t pool;           /* wavefront: paths per batch (0 = min(2^24, 8*w*h)); at most
                               w*h*spp and 2^26; from 2*w*h up, rounded down to whole
                               samples of every pixel, else capped at w*h            */
    int32_t grid_blocks;    /* wavefront: persistent blocks of 256 (0 = default)       */
    int32_t use_graph;      /* wavefront: replay the bounce sequence as a hipGraph     */
    int32_t kernel_clock;   /* wavefront: 1 = record a HIP event pair around every
                               kernel launch (launched directly, not from a graph);
                               per-kernel totals via mh_last_kernel_times          */
    int32_t bvh_width;      /* 0 = default (4 when the scene's 4-wide tree fits the
                               traversal stack, else 2), 2 = binary, 4 = 4-wide;
                               both give identical hits                           */
} mh_render_params;

/* Per-call timing of the last render on the calling thread, measured with
 * HIP events recorded on the scene's stream around each kernel. */
typedef struct mh_timing {
    double total_ms;        /* first event to last event               */
    double kernel_ms;       /* the render kernel(s) only               */
    double tonemap_ms;      /* mean -> 8-bit conversion                */
    int64_t samples;        /* width * height * spp                    */
    int64_t rays;           /* extension + shadow queries (counted)    */
} mh_timing;

/* Per-kernel totals of the last wavefront render made with kernel_clock = 1:
 * summed HIP-event durations and launch counts, indexed by MH_K_*. */
enum {
    MH_K_GEN = 0,       /* camera rays of a batch                  */
    MH_K_EXTEND = 1,    /* nearest hit of the extension queue      */
    MH_K_SHADE = 2,     /* shading, NEE emission, BSDF sampling    */
    MH_K_SHADOW = 3,    /* any-hit / transmissive shadow queries   */
    MH_K_FINALIZE = 4,  /* ended paths -> pixel accumulators       */
    MH_K_OTHER = 5,     /* queue resets, batch advance, resolve    */
    MH_K_COUNT = 6
};

typedef struct mh_kernel_times {
    double  ms[8];
    int64_t launches[8];
} mh_kernel_times;

/* ------------------------------------------------------------------ */
/* Entry points                                                        */
/* ------------------------------------------------------------------ */

/* Library / device info. Returns 0 and fills `name` (NUL-terminated). */
int mh_device_info(char *name, size_t name_len, int *cu_count, size_t *hbm_bytes);

/* Upload a flattened scene (one host->device copy of every array).
 * Replaces the scene handed to reference pathtrace_render (src/pathtrace.h:78). */
int mh_scene_create(const mh_scene_soa *soa, struct mh_scene **out);
void mh_scene_destroy(struct mh_scene *s);

/* Tree statistics of an uploaded scene. */
typedef struct mh_bvh_info {
    int32_t n_nodes;       /* binary nodes                                  */
    int32_t n_leaf;        /* leaf-order primitive slots                    */
    int32_t max_depth;     /* levels (device build; 0 for a host tree)      */
    int32_t width4;        /* 1: the 4-wide tree was built too              */
    int32_t device_built;  /* 1: built by mh_scene_create_gpu_bvh           */
    int32_t reserved;
    double  build_ms;      /* device build time (tree + device layout)      */
} mh_bvh_info;

/* Upload a scene WITHOUT a host tree and build its BVH on the device:
 * binned SAH over `bins` bins (0: 16) on all three axes with leaves of at
 * most max_leaf prims (0: 4) -- the split decisions mt_scene_build_bvh makes
 * for the same options, so the tree has the host tree's nodes (prim order
 * inside a leaf may differ). soa->nodes / leaf_prims / planes are ignored.
 * Replaces the reference's host build before a render (bvh_build,
 * src/bvh.c:180-422, called from scene setup); the reference has no device
 * build. Returns 5 for a tree deeper than 60 levels. */
int mh_scene_create_gpu_bvh(const mh_scene_soa *soa, int max_leaf, int bins,
                            struct mh_scene **out);
int mh_scene_bvh_info(struct mh_scene *s, mh_bvh_info *info);
/* The device-built binary tree (mh_bvh_node, node 0 = root; leaf ranges
 * index leaf_prims) into host arrays of info.n_nodes / info.n_leaf entries.
 * Returns 4 for a scene created from a host tree. */
int mh_scene_bvh_export(struct mh_scene *s, mh_bvh_node *nodes, int32_t *leaf_prims);

/* Device bytes held by the scene (nodes + primitive records + tables). */
size_t mh_scene_device_bytes(const struct mh_scene *s);

/* Nearest-hit queries for a ray batch in [tmin, tmax]
 * (reference scene_intersect, src/scene.h:162 / src/scene.c:767-775). */
int mh_trace_closest(struct mh_scene *s, const mh_ray *rays, int n,
                     double tmin, double tmax, int precision, mh_hit *out);

/* Any-hit occlusion for a ray batch in [tmin, tmax]
 * (reference scene_occluded, src/scene.h:169 / src/bvh.c:714-813). */
int mh_trace_any(struct mh_scene *s, const mh_ray *rays, int n,
                 double tmin, double tmax, int precision, uint8_t *occluded);

/* Drop-in for reference pathtrace_render (src/pathtrace.h:78-80): path-trace
 * a width x height image with `samples_per_pixel` and `max_depth` into the
 * TOP-DOWN RGB8 buffer rgb_out[(y*width+x)*3+c]. fp32 precision. */
int mh_pathtrace_render(struct mh_scene *s, const mh_camera *cam, int width,
                        int height, int samples_per_pixel, int max_depth,
                        unsigned char *rgb_out);

/* Drop-in for reference render_image (src/render.h:29-31; render_image_ex
 * with adaptive off, src/render.c:1627-1651): fixed-spp Whitted render. */
int mh_render_image(struct mh_scene *s, const mh_camera *cam, int width,
                    int height, int samples_per_pixel, int max_depth,
                    unsigned char *rgb_out);

/* Parameters of render_image_ex, field for field the reference's
 * RenderParams (src/render.h:49-55): base samples n0 (>= 1), depth (>= 0),
 * adaptive (0 = the fixed render), the per-pixel cap (<= 0: 4 * n0; below
 * n0: n0) and the relative-error tolerance (<= 0: 0.02). */
typedef struct mh_render_ex_params {
    int32_t samples_per_pixel;
    int32_t max_depth;
    int32_t adaptive;
    int32_t adaptive_max_spp;
    double adaptive_tau;
} mh_render_ex_params;

/* Drop-in for reference render_image_ex (src/render.h:65-66,
 * src/render.c:1627-1680): with adaptive on, every pixel takes n0 Whitted
 * samples, then passes of up to n0 more refine the pixels whose relative
 * standard error of mean luminance exceeds tau, until none does or the cap
 * is met (at most 1 + ceil((cap - n0) / n0) passes); with adaptive off, the
 * fixed render (mh_render_image). fp64, the same return codes. */
int mh_render_image_ex(struct mh_scene *s, const mh_camera *cam, int width,
                       int height, const mh_render_ex_params *params,
                       unsigned char *rgb_out);

/* The same with a choice of precision (MH_PREC_*) and optional outputs:
 * linear_out (w*h*3 floats, per-pixel means) and nsamp_out (w*h samples
 * taken per pixel). Any output pointer may be NULL. */
int mh_render_adaptive(struct mh_scene *s, const mh_camera *cam, int width,
                       int height, const mh_render_ex_params *params,
                       int precision, unsigned char *rgb_out,
                       float *linear_out, int32_t *nsamp_out);

/* Primary samples of the last render on the calling thread: the total and
 * the largest per pixel (reference render_last_total_samples and
 * render_last_max_samples, src/render.h:79-85; w*h*spp and spp for the
 * fixed renders). */
int mh_last_sample_stats(unsigned long long *total, int *max_per_pixel);

/* Anamorphic streak flare, field for field the reference's FlareParams
 * (src/flare.h:30-36; the C layout puts 4 bytes of padding after `enabled`):
 * intensity (<= 0: no flare), highlight threshold (clamped to [0, 0.999]),
 * streak half-width as a fraction of the image width (clamped to
 * [0.005, 1]), tint (|tint|^2 < 1e-6: (0.15, 0.45, 1.0)). */
typedef struct mh_flare_params {
    int32_t enabled;
    int32_t reserved;
    double intensity;
    double threshold;
    double streak_length;
    mh_vec3 tint;
} mh_flare_params;

/* Drop-in for reference flare_apply (src/flare.h:55, src/flare.c:60-173):
 * horizontal streaks from the highlights of each scanline of the top-down
 * RGB8 image, composited in place. Runs on the GPU (host buffer in and
 * out); bytes equal the reference's. Returns 1 on invalid arguments, 0 when
 * disabled (nothing changes). */
int mh_flare_apply(unsigned char *rgb, int width, int height,
                   const mh_flare_params *params);

/* The same on the scene's device framebuffer (the 8-bit image of the last
 * render on `s`), on the scene's stream: mh_framebuffer_download then
 * returns the composited bytes. No reference counterpart (the reference
 * flares its host image); saves the host round trip between render and
 * post-pass. */
int mh_scene_flare(struct mh_scene *s, const mh_flare_params *params);

/* General render. rgb_out (w*h*3 bytes) and linear_out (w*h*3 floats, the
 * per-pixel mean radiance) are host buffers and may each be NULL: the result
 * then stays in the scene's device framebuffer (fetch it with
 * mh_framebuffer_download). */
int mh_render(struct mh_scene *s, const mh_camera *cam,
              const mh_render_params *p, unsigned char *rgb_out,
              float *linear_out);

/* Largest image (width * height) mh_render_noalloc accepts. */
#define MH_NOALLOC_MAX_PIXELS (1 << 21)

/* mh_render without any device allocation. No reference counterpart: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It runs the same
 * path-tracing megakernel as mh_render (p->mode MH_MODE_PATHTRACE, p->engine
 * MH_ENGINE_MEGAKERNEL, fp32 or fp64) and writes the image into the caller's
 * DEVICE buffers: d_rgb_out (w*h*3 bytes, required) and d_linear_out (w*h*3
 * floats, may be NULL). Its scratch -- the ray counter and, when
 * d_linear_out is NULL, the linear framebuffer -- is one fixed array per GPU
 * (24 MiB + 256 B, held until the process ends), so w*h must not exceed
 * MH_NOALLOC_MAX_PIXELS (else code 2). Calls that overlap on one GPU share
 * that array and race: issue them one at a time per GPU. Synchronous on the
 * scene's stream; the scene's own framebuffer is left untouched. Returns 4
 * for Whitted or wavefront params, other codes as mh_render. */
int mh_render_noalloc(struct mh_scene *s, const mh_camera *cam,
                      const mh_render_params *p, unsigned char *d_rgb_out,
                      float *d_linear_out);

/* Copy the device framebuffer of the last mh_render on `s` to the host. */
int mh_framebuffer_download(struct mh_scene *s, unsigned char *rgb_out,
                            float *linear_out);

/* Replace every material (same count as at creation) / the sky / the fog
 * of an uploaded scene, keeping geometry and BVH. The emissive-light list
 * and the transmissive flag are recomputed as mh_scene_create would. No
 * reference counterpart (the reference rebuilds its Scene); used by the
 * differentiable renderer and parameter sweeps. */
int mh_scene_set_materials(struct mh_scene *s, const mh_material *mats, int n);
int mh_scene_set_sky(struct mh_scene *s, const mh_sky *sky);
int mh_scene_set_fog(struct mh_scene *s, const mh_fog *fog);

/* The fp32 path tracer as the adjoint computes it (same code and rounding),
 * also recording every sample's radiance sum on the scene (16 bytes per
 * sample, at most 16 GiB). A following mh_adjoint with the same camera and
 * params (size, spp, sample_offset, max_depth) then replays each path once
 * instead of twice. Any material / sky / fog update drops the record; the
 * next recording render replaces it. No reference counterpart. */
int mh_render_record(struct mh_scene *s, const mh_camera *cam,
                     const mh_render_params *p, unsigned char *rgb_out,
                     float *linear_out);

/* Path-replay adjoint of the fp32 path tracer (no reference counterpart).
 * d_image: dLoss/d(linear pixel mean), w*h*3 floats, top-down like the
 * render. grad receives n_materials*8 + 3 floats: for each material
 * {albedo r, g, b, roughness, metallic, emissive r, g, b}, then
 * dLoss/d(sky sun_color) r, g, b. Same sample streams as mh_render's path
 * tracer with the same params (mode PATHTRACE, precision FP32). Sampling
 * decisions are held fixed (gradient of the estimator's expectation);
 * visibility is not differentiated. Returns 4 for Whitted or fp64 params. */
int mh_adjoint(struct mh_scene *s, const mh_camera *cam,
               const mh_render_params *p, const float *d_image, float *grad);

/* fp32 path tracer whose sampling decisions (lobe choice and probability,
 * GGX half vector and pdf, Russian roulette on its own throughput) are taken
 * with the materials `frozen` (same count as the scene's) and whose BSDF
 * values use the scene's current materials. The image is then a smooth
 * function of the current parameters, with derivative at `frozen` equal to
 * what mh_adjoint computes: finite differences of it check the adjoint
 * without lobe-switching noise. No reference counterpart. */
int mh_render_frozen(struct mh_scene *s, const mh_camera *cam,
                     const mh_render_params *p, const mh_material *frozen, int n,
                     unsigned char *rgb_out, float *linear_out);

/* ------------------------------------------------------------------ */
/* Batched rendering (no reference counterpart: the reference renders  */
/* one image per pathtrace_render call, src/pathtrace.c:1256-1363)      */
/* ------------------------------------------------------------------ */

/* One image of a batch: a camera view of scenes[scene] with its own size,
 * spp, depth and first sample index. Each job's pixels are the same
 * estimator as mh_render's path tracer (same sample streams). */
typedef struct mh_job {
    int32_t   scene;
    int32_t   width;
    int32_t   height;
    int32_t   spp;
    int32_t   max_depth;
    int32_t   sample_offset;
    mh_camera cam;
} mh_job;

/* Launch settings of the batched engine. Zeros pick the defaults. On
 * return from mh_render_batch, mh_batch_last_tune gives what was used. */
typedef struct mh_tune {
    int32_t pool;          /* paths in flight per batch (queue capacity); 0 = 2^21;
                              at most 2^26                                           */
    int32_t grid_blocks;   /* workgroups of each persistent kernel; 0 = 4096             */
    int32_t block;         /* lanes per workgroup: 64, 128 or 256; 0 = 256               */
    int32_t use_graph;     /* replay each batch's bounce sequence as a hipGraph          */
    int32_t autotune;      /* 1 = choose pool / grid_blocks / block / fixed_batches by
                              timing trial batches of this workload (cached per
                              workload on the handle)                                */
    int32_t kernel_clock;  /* 1 = per-kernel HIP-event timing (direct launches)          */
    int32_t bvh_width;     /* BVH node width: 0 = default (4 when built, else 2), 2, 4;
                              autotune also tries both widths                        */
    int32_t fixed_batches; /* 0 (default): path regeneration, every bounce step refills
                              the slots of ended paths with new ones; 1: batches of
                              `pool` paths run to completion one after another     */
    int32_t engine;        /* MH_BATCH_WAVEFRONT (0, default) or MH_BATCH_MEGAKERNEL:
                              one launch of (job, 16x16 tile, 16 samples) units, each
                              lane tracing whole paths; autotune times both       */
} mh_tune;

enum { MH_BATCH_WAVEFRONT = 0, MH_BATCH_MEGAKERNEL = 1 };

struct mh_batch;
typedef struct mh_batch mh_batch;

/* A batch renderer owns its stream, path state and graph. One handle must
 * not be used from two threads at once. */
int  mh_batch_create(mh_batch **out);
void mh_batch_destroy(mh_batch *b);

/* Samples per pixel of one batched job (pixel sums are 64-bit fixed point
 * with each sample capped at 2^26). */
#define MH_BATCH_MAX_SPP 16383

/* Render n_jobs images of n_scenes uploaded scenes in one persistent
 * wavefront launch sequence (fp32). Outputs are concatenated in job order,
 * each job top-down RGB: rgb_out holds sum(w*h*3) bytes, linear_out the same
 * count of floats; either may be NULL. Return codes as mh_render; 4 when a
 * job's max_depth exceeds 254 or its spp exceeds MH_BATCH_MAX_SPP. */
int mh_render_batch(mh_batch *b, struct mh_scene *const *scenes, int n_scenes,
                    const mh_job *jobs, int n_jobs, const mh_tune *tune,
                    unsigned char *rgb_out, float *linear_out);

/* Settings used by the last mh_render_batch on this handle. */
int mh_batch_last_tune(const mh_batch *b, mh_tune *out);

/* Traversal work of the batched engine's extension (closest-hit) and any-hit
 * shadow queries in the last mh_render_batch made with counting on: queries,
 * interior BVH nodes fetched, triangles and other primitives (spheres,
 * boxes, cylinders, planes) tested. Counted on the binary-tree wavefront
 * kernels only (zeros for 4-wide launches, transmissive shadow walks and
 * the batched megakernel). No reference counterpart: it prices the
 * algorithmic bytes of the tree the GPU actually traverses. */
typedef struct mh_traversal_counts {
    int64_t ext_queries, ext_nodes, ext_triangles, ext_other;
    int64_t shadow_queries, shadow_nodes, shadow_triangles, shadow_other;
} mh_traversal_counts;

/* Counting on (1) or off (0) for the following renders on `b`; counting
 * launches run separate instantiations of the two kernels. */
int mh_batch_set_counting(mh_batch *b, int on);
int mh_batch_last_counts(const mh_batch *b, mh_traversal_counts *out);

/* Timing and error state of the last call on the calling thread. */
int mh_last_timing(mh_timing *out);
int mh_last_kernel_times(mh_kernel_times *out);  /* no reference counterpart */
double mh_last_seconds(void);   /* like reference render_last_seconds() */
const char *mh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MARMEL_HIP_H */
