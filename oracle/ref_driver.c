/*
 * ref_driver.c - thin C entry points over the REFERENCE renderer, compiled
 * together with the reference sources (where they lie under the reference
 * tree) into oracle/_ref/libmtref.so by oracle/Makefile.
 *
 * This file is test infrastructure only. It calls the reference's public API
 * (scene_desc_load_string, scene_build_from_desc, camera_create,
 * pathtrace_render, render_image, scene_intersect, scene_occluded,
 * pathtrace_radiance, sky_sample, water_normal, noise_*) so that tests can
 * (a) generate golden fixtures and (b) check the C restatement in oracle/ and
 * the GPU library against the reference on identical scenes. It never ships.
 */
#define _POSIX_C_SOURCE 200809L

#include "marmel_hip.h"

#include "bvh.h"
#include "camera.h"
#include "flare.h"
#include "geometry.h"
#include "material.h"
#include "noise.h"
#include "obj.h"
#include "pathtrace.h"
#include "render.h"
#include "sampling.h"
#include "scene.h"
#include "scene_desc.h"
#include "texture.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    Scene scene;
    int has_camera;
    CameraDesc camera;
} RefScene;

static Vec3 v3(mh_vec3 a) { return vec3(a.x, a.y, a.z); }
static mh_vec3 m3(Vec3 a) { mh_vec3 r = { a.x, a.y, a.z }; return r; }

static void quiet_env(int threads)
{
    char buf[32];
    setenv("RAYTRACER_NO_PROGRESS", "1", 1);
    if (threads > 0) {
        snprintf(buf, sizeof buf, "%d", threads);
        setenv("RAYTRACER_THREADS", buf, 1);
    } else {
        unsetenv("RAYTRACER_THREADS");
    }
}

/* Build a reference Scene from scene text (the path main.c takes). */
void *mtref_load_text(const char *text, char *err, size_t errlen)
{
    SceneDesc desc;
    RefScene *rs = (RefScene *)calloc(1, sizeof(RefScene));
    if (!rs) return NULL;
    scene_desc_init(&desc);
    if (scene_desc_load_string(&desc, text, "<text>", err, errlen) != 0) {
        scene_desc_free(&desc);
        free(rs);
        return NULL;
    }
    if (scene_build_from_desc(&rs->scene, &desc) != 0) {
        if (err && errlen) snprintf(err, errlen, "scene_build_from_desc failed");
        scene_desc_free(&desc);
        free(rs);
        return NULL;
    }
    rs->has_camera = desc.camera.present;
    rs->camera = desc.camera;
    scene_desc_free(&desc);
    return rs;
}

/* Emissive-light list with the semantics of the reference scene builder:
 * first 8 PBR spheres with non-zero emission and positive radius, file order. */
static void collect_lights(Scene *s)
{
    s->emissive_light_count = 0;
    for (int i = 0; i < s->geo.count; ++i) {
        if (s->emissive_light_count >= SCENE_MAX_EMISSIVE_LIGHTS) break;
        const Primitive *p = &s->geo.prims[i];
        if (p->kind != PRIM_SPHERE) continue;
        if (p->material_index < 0 || p->material_index >= s->material_count) continue;
        const Material *m = &s->materials[p->material_index];
        if (!m->pbr) continue;
        if (m->emissive.x == 0.0 && m->emissive.y == 0.0 && m->emissive.z == 0.0) continue;
        if (!(p->radius > 0.0)) continue;
        EmissiveLight *l = &s->emissive_lights[s->emissive_light_count++];
        l->prim_index = i;
        l->center = p->center;
        l->radius = p->radius;
        l->emissive = m->emissive;
    }
}

/* Build a reference Scene from flattened arrays (same data the GPU gets). */
/* One flattened material as the reference's Material. */
static void mat_from(const mh_material *a, Material *m)
{
    material_texture_defaults(m);
    m->albedo = v3(a->albedo);
    m->specular = v3(a->specular);
    m->shininess = a->shininess;
    m->reflectivity = a->reflectivity;
    m->transparency = a->transparency;
    m->ior = a->ior;
    m->is_water = a->is_water;
    m->beer_lambert = a->beer_lambert;
    m->pbr = a->pbr;
    m->texture_kind = a->texture_kind;
    m->absorption = v3(a->absorption);
    m->deep_color = v3(a->deep_color);
    m->metallic = a->metallic;
    m->roughness = a->roughness;
    m->emissive = v3(a->emissive);
    m->texture_scale = a->texture_scale;
    m->texture_color_a = v3(a->texture_color_a);
    m->texture_color_b = v3(a->texture_color_b);
    m->atmosphere_glow = v3(a->atmosphere_glow);
    m->bump_strength = a->bump_strength;
    m->bump_scale = 1.0;
}

void *mtref_from_soa(const mh_scene_soa *soa)
{
    RefScene *rs = (RefScene *)calloc(1, sizeof(RefScene));
    if (!rs || !soa) return NULL;
    Scene *s = &rs->scene;
    geometry_init(&s->geo);
    s->water_level = soa->water_level;
    s->water_material = soa->water_material;
    sky_default_params(&s->sky);
    fog_default_params(&s->fog);
    const mh_sky *k = &soa->sky;
    s->sky.sun_dir = v3(k->sun_dir);
    s->sky.sun_color = v3(k->sun_color);
    s->sky.horizon_color = v3(k->horizon_color);
    s->sky.zenith_color = v3(k->zenith_color);
    s->sky.gradient_gamma = k->gradient_gamma;
    s->sky.sun_glow_exponent = k->sun_glow_exponent;
    s->sky.sun_glow_strength = k->sun_glow_strength;
    s->sky.cloud_height = k->cloud_height;
    s->sky.cloud_scale = k->cloud_scale;
    s->sky.cloud_coverage = k->cloud_coverage;
    s->sky.cloud_softness = k->cloud_softness;
    s->sky.cloud_sharpness = k->cloud_sharpness;
    s->sky.cloud_octaves = k->cloud_octaves;
    s->sky.seed = k->seed;
    s->sky.sun_radius = k->sun_radius;
    s->sky.cloud_thickness = k->cloud_thickness;
    s->sky.star_intensity = k->star_intensity;
    s->sky.nebula_intensity = k->nebula_intensity;
    s->sky.galaxy_intensity = k->galaxy_intensity;
    s->fog.density = soa->fog.density;
    s->fog.color = v3(soa->fog.color);
    s->fog.height = soa->fog.height;
    s->fog.height_falloff = soa->fog.height_falloff;
    s->fog.inscatter_strength = soa->fog.inscatter_strength;
    s->fog.sun_anisotropy = soa->fog.sun_anisotropy;
    s->fog.noise_scale = soa->fog.noise_scale;
    s->fog.noise_amount = soa->fog.noise_amount;
    s->fog.shadow_steps = soa->fog.shadow_steps;
    if (soa->ocean.enabled) {
        s->has_ocean = 1;
        s->ocean_params.enabled = 1;
        s->ocean_params.center = v3(soa->ocean.center);
        s->ocean_params.direction = v3(soa->ocean.direction);
        s->ocean_params.amplitude = soa->ocean.amplitude;
        s->ocean_params.wavelength = soa->ocean.wavelength;
        s->ocean_params.steepness = soa->ocean.steepness;
        s->ocean_params.chop = soa->ocean.chop;
        s->ocean_params.chop_wavelength = soa->ocean.chop_wavelength;
        s->ocean_params.seed = soa->ocean.seed;
    }

    s->material_count = soa->n_materials;
    s->materials = (Material *)calloc((size_t)(soa->n_materials > 0 ? soa->n_materials : 1), sizeof(Material));
    for (int i = 0; i < soa->n_materials; ++i) {
        const mh_material *a = &soa->materials[i];
        Material *m = &s->materials[i];
        mat_from(a, m);
    }

    for (int i = 0; i < soa->n_prims; ++i) {
        const mh_prim *q = &soa->prims[i];
        Primitive p;
        switch (q->kind) {
        case MH_PRIM_SPHERE:
            p = prim_sphere(v3(q->p0), q->r0, q->material);
            break;
        case MH_PRIM_PLANE:
            p = prim_plane(v3(q->p0), v3(q->p1), q->material);
            p.axis = v3(q->p1); /* stored exactly as given */
            break;
        case MH_PRIM_BOX:
            p = prim_box(v3(q->p0), v3(q->p1), q->material);
            break;
        case MH_PRIM_TRIANGLE:
            if ((q->flags & MH_PRIM_SMOOTH) && soa->attrs) {
                const mh_prim_attr *at = &soa->attrs[i];
                p = prim_triangle_smooth(v3(q->p0), v3(q->p1), v3(q->p2),
                                         v3(at->n0), v3(at->n1), v3(at->n2), q->material);
            } else {
                p = prim_triangle(v3(q->p0), v3(q->p1), v3(q->p2), q->material);
            }
            if ((q->flags & MH_PRIM_UV) && soa->attrs) {
                const mh_prim_attr *at = &soa->attrs[i];
                p.has_uv = 1;
                p.uva = v3(at->uv0);
                p.uvb = v3(at->uv1);
                p.uvc = v3(at->uv2);
            }
            break;
        case MH_PRIM_CYLINDER:
            p = prim_cylinder(v3(q->p0), v3(q->p1), q->r0, q->r1, q->material);
            break;
        default:
            scene_free(s);
            free(rs);
            return NULL;
        }
        if (geometry_add(&s->geo, p) < 0) {
            scene_free(s);
            free(rs);
            return NULL;
        }
    }
    collect_lights(s);
    s->bvh = bvh_build(&s->geo);
    return rs;
}

void mtref_free(void *h)
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return;
    scene_free(&rs->scene);
    free(rs);
}

int mtref_counts(void *h, int *n_prims, int *n_materials, int *n_lights)
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return 1;
    if (n_prims) *n_prims = rs->scene.geo.count;
    if (n_materials) *n_materials = rs->scene.material_count;
    if (n_lights) *n_lights = rs->scene.emissive_light_count;
    return 0;
}

/* One reference Primitive in the ABI record; returns 4 for kinds outside
 * the GPU path. */
static int flatten_prim(const Primitive *p, mh_prim *q, mh_prim_attr *at)
{
    memset(q, 0, sizeof *q);
    memset(at, 0, sizeof *at);
    q->material = p->material_index;
    switch (p->kind) {
    case PRIM_SPHERE:
        q->kind = MH_PRIM_SPHERE; q->p0 = m3(p->center); q->r0 = p->radius; return 0;
    case PRIM_PLANE:
        q->kind = MH_PRIM_PLANE; q->p0 = m3(p->center); q->p1 = m3(p->axis); return 0;
    case PRIM_BOX:
        q->kind = MH_PRIM_BOX; q->p0 = m3(p->center); q->p1 = m3(p->half); return 0;
    case PRIM_TRIANGLE:
        q->kind = MH_PRIM_TRIANGLE;
        q->p0 = m3(p->a); q->p1 = m3(p->b); q->p2 = m3(p->c);
        if (p->radius > 0.5) {
            q->flags |= MH_PRIM_SMOOTH;
            at->n0 = m3(p->center); at->n1 = m3(p->axis); at->n2 = m3(p->half);
        }
        if (p->has_uv) {
            q->flags |= MH_PRIM_UV;
            at->uv0 = m3(p->uva); at->uv1 = m3(p->uvb); at->uv2 = m3(p->uvc);
        }
        return 0;
    case PRIM_CYLINDER:
        q->kind = MH_PRIM_CYLINDER;
        q->p0 = m3(p->a); q->p1 = m3(p->b); q->r0 = p->radius; q->r1 = p->radius2; return 0;
    default:
        q->kind = -1;
        return 4;
    }
}

/* Export the reference scene in the flattened ABI layout. Returns 0, or 4
 * when the scene holds a primitive kind outside the GPU path. */
int mtref_export(void *h, mh_prim *prims, mh_prim_attr *attrs, mh_material *mats,
                 mh_sky *sky, double *water_level, int *water_material, mh_fog *fog)
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return 1;
    const Scene *s = &rs->scene;
    int rc = 0;
    for (int i = 0; i < s->geo.count; ++i)
        if (flatten_prim(&s->geo.prims[i], &prims[i], &attrs[i])) rc = 4;
    for (int i = 0; i < s->material_count; ++i) {
        const Material *m = &s->materials[i];
        mh_material *a = &mats[i];
        memset(a, 0, sizeof *a);
        a->albedo = m3(m->albedo); a->specular = m3(m->specular);
        a->shininess = m->shininess; a->reflectivity = m->reflectivity;
        a->transparency = m->transparency; a->ior = m->ior;
        a->is_water = m->is_water; a->beer_lambert = m->beer_lambert;
        a->pbr = m->pbr; a->texture_kind = m->texture_kind;
        a->absorption = m3(m->absorption); a->deep_color = m3(m->deep_color);
        a->metallic = m->metallic; a->roughness = m->roughness;
        a->emissive = m3(m->emissive); a->texture_scale = m->texture_scale;
        a->texture_color_a = m3(m->texture_color_a); a->texture_color_b = m3(m->texture_color_b);
        a->atmosphere_glow = m3(m->atmosphere_glow); a->bump_strength = m->bump_strength;
    }
    const SkyParams *k = &s->sky;
    memset(sky, 0, sizeof *sky);
    sky->sun_dir = m3(k->sun_dir); sky->sun_color = m3(k->sun_color);
    sky->horizon_color = m3(k->horizon_color); sky->zenith_color = m3(k->zenith_color);
    sky->gradient_gamma = k->gradient_gamma; sky->sun_glow_exponent = k->sun_glow_exponent;
    sky->sun_glow_strength = k->sun_glow_strength; sky->cloud_height = k->cloud_height;
    sky->cloud_scale = k->cloud_scale; sky->cloud_coverage = k->cloud_coverage;
    sky->cloud_softness = k->cloud_softness; sky->cloud_sharpness = k->cloud_sharpness;
    sky->cloud_octaves = k->cloud_octaves; sky->seed = k->seed; sky->sun_radius = k->sun_radius;
    sky->cloud_thickness = k->cloud_thickness; sky->star_intensity = k->star_intensity;
    sky->nebula_intensity = k->nebula_intensity; sky->galaxy_intensity = k->galaxy_intensity;
    sky->fog_density = 0.0;
    if (fog) {
        const FogParams *f = &s->fog;
        memset(fog, 0, sizeof *fog);
        fog->density = f->density; fog->color = m3(f->color);
        fog->height = f->height; fog->height_falloff = f->height_falloff;
        fog->inscatter_strength = f->inscatter_strength;
        fog->sun_anisotropy = f->sun_anisotropy;
        fog->noise_scale = f->noise_scale; fog->noise_amount = f->noise_amount;
        fog->shadow_steps = f->shadow_steps;
    }
    *water_level = s->water_level;
    *water_material = s->water_material;
    return rc;
}

/* The scene's ocean (the water-height surface); returns has_ocean. */
int mtref_ocean(void *h, mh_ocean *o)
{
    RefScene *rs = (RefScene *)h;
    memset(o, 0, sizeof *o);
    if (!rs || !rs->scene.has_ocean) return 0;
    const SceneOceanParams *p = &rs->scene.ocean_params;
    o->enabled = 1;
    o->seed = p->seed;
    o->center = m3(p->center);
    o->direction = m3(p->direction);
    o->amplitude = p->amplitude;
    o->wavelength = p->wavelength;
    o->steepness = p->steepness;
    o->chop = p->chop;
    o->chop_wavelength = p->chop_wavelength;
    return 1;
}

/* The reference OBJ loader (obj_load_mem) on text: up to cap triangles in
 * the ABI record; returns the loader's count (or < 0). */
int mtref_obj_prims(const char *data, size_t len, int material, const double pos[3], const double scale[3],
                    const double rot[3], int smooth, int auto_center, double auto_scale, mh_prim *out,
                    mh_prim_attr *attrs, int cap)
{
    Geometry g;
    geometry_init(&g);
    ObjTransform xf = obj_transform_default();
    xf.position = vec3(pos[0], pos[1], pos[2]);
    xf.scale = vec3(scale[0], scale[1], scale[2]);
    xf.rotation_deg = vec3(rot[0], rot[1], rot[2]);
    xf.smooth_normals = smooth;
    xf.auto_center = auto_center;
    xf.auto_scale = auto_scale;
    int n = obj_load_mem(data, len, &g, material, &xf);
    for (int i = 0; i < g.count && i < cap; ++i) flatten_prim(&g.prims[i], &out[i], &attrs[i]);
    geometry_free(&g);
    return n;
}

static Camera cam_from(const mh_camera *c)
{
    Camera cam;
    memset(&cam, 0, sizeof cam);
    cam.position = v3(c->position);
    cam.lower_left = v3(c->lower_left);
    cam.horizontal = v3(c->horizontal);
    cam.vertical = v3(c->vertical);
    cam.aperture = c->aperture;
    cam.focus_distance = c->focus_distance;
    cam.dome_radius = c->dome_radius;
    cam.anamorphic_squeeze = c->anamorphic_squeeze;
    flare_default_params(&cam.flare);
    return cam;
}

static mh_camera cam_to(const Camera *c)
{
    mh_camera o;
    o.position = m3(c->position);
    o.lower_left = m3(c->lower_left);
    o.horizontal = m3(c->horizontal);
    o.vertical = m3(c->vertical);
    o.aperture = c->aperture;
    o.focus_distance = c->focus_distance;
    o.dome_radius = c->dome_radius;
    o.anamorphic_squeeze = c->anamorphic_squeeze;
    return o;
}

/* camera_create (reference src/camera.c:24-66). */
void mtref_camera_create(const double eye[3], const double at[3], const double up[3],
                         double vfov, double aspect, mh_camera *out)
{
    Camera c = camera_create(vec3(eye[0], eye[1], eye[2]), vec3(at[0], at[1], at[2]),
                             vec3(up[0], up[1], up[2]), vfov, aspect);
    *out = cam_to(&c);
}

/* Camera of a text-loaded scene, built the way main.c does. */
int mtref_scene_camera(void *h, double aspect, mh_camera *out)
{
    RefScene *rs = (RefScene *)h;
    Vec3 eye, target, up;
    double vfov;
    if (!rs) return 1;
    if (rs->has_camera) {
        eye = rs->camera.eye; target = rs->camera.target; up = rs->camera.up;
        vfov = rs->camera.vfov_deg;
    } else {
        scene_default_view(&eye, &target, &up, &vfov);
    }
    Camera cam = camera_create(eye, target, up, vfov, aspect);
    cam.aperture = rs->camera.aperture;
    if (rs->camera.focus_distance > 0.0) cam.focus_distance = rs->camera.focus_distance;
    if (rs->camera.dome_radius > 0.0) cam.dome_radius = rs->camera.dome_radius;
    if (rs->camera.anamorphic_squeeze > 0.0) cam.anamorphic_squeeze = rs->camera.anamorphic_squeeze;
    *out = cam_to(&cam);
    return 0;
}

/* mode 0 = pathtrace_render, 1 = render_image (fixed-spp Whitted). */
int mtref_render(void *h, const mh_camera *c, int w, int hh, int spp, int depth,
                 int mode, unsigned char *rgb, int threads, double *seconds)
{
    RefScene *rs = (RefScene *)h;
    if (!rs || !c) return 1;
    quiet_env(threads);
    Camera cam = cam_from(c);
    int rc = (mode == 0) ? pathtrace_render(&rs->scene, &cam, w, hh, spp, depth, rgb)
                         : render_image(&rs->scene, &cam, w, hh, spp, depth, rgb);
    if (seconds) *seconds = render_last_seconds();
    return rc;
}

int mtref_intersect(void *h, const mh_ray *rays, int n, double tmin, double tmax, mh_hit *out)
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return 1;
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.origin = vec3(rays[i].ox, rays[i].oy, rays[i].oz);
        r.dir = vec3(rays[i].dx, rays[i].dy, rays[i].dz);
        Hit hit;
        memset(&out[i], 0, sizeof out[i]);
        if (scene_intersect(&rs->scene, r, tmin, tmax, &hit)) {
            out[i].t = hit.t;
            out[i].prim_index = hit.prim_index;
            out[i].material = hit.material_index;
            out[i].front_face = hit.front_face;
            out[i].nx = hit.normal.x; out[i].ny = hit.normal.y; out[i].nz = hit.normal.z;
            out[i].u = hit.u; out[i].v = hit.v;
        } else {
            out[i].prim_index = -1;
        }
    }
    return 0;
}

int mtref_occluded(void *h, const mh_ray *rays, int n, double tmin, double tmax, uint8_t *out)
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return 1;
    for (int i = 0; i < n; ++i) {
        Ray r;
        r.origin = vec3(rays[i].ox, rays[i].oy, rays[i].oz);
        r.dir = vec3(rays[i].dx, rays[i].dy, rays[i].dz);
        out[i] = (uint8_t)scene_occluded_cached(&rs->scene, r, tmin, tmax, NULL);
    }
    return 0;
}

int mtref_radiance(void *h, const mh_ray *ray, int depth, unsigned seed, double out[3])
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return 1;
    Ray r;
    r.origin = vec3(ray->ox, ray->oy, ray->oz);
    r.dir = vec3(ray->dx, ray->dy, ray->dz);
    Vec3 L = pathtrace_radiance(&rs->scene, r, depth, seed);
    out[0] = L.x; out[1] = L.y; out[2] = L.z;
    return 0;
}

/* ---- pure-function vectors ---------------------------------------- */

void mtref_sky_sample(void *h, const double *dirs, int n, double *out)
{
    RefScene *rs = (RefScene *)h;
    for (int i = 0; i < n; ++i) {
        Vec3 c = sky_sample(vec3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]), &rs->scene.sky);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

void mtref_water_normal(const double *xz, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 c = water_normal(xz[2 * i], xz[2 * i + 1], 0.0);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

void mtref_noise(const double *xy, int n, int octaves, unsigned seed, double *perlin, double *fbm)
{
    for (int i = 0; i < n; ++i) {
        perlin[i] = noise_perlin2(xy[2 * i], xy[2 * i + 1], seed);
        fbm[i] = noise_fbm2(xy[2 * i], xy[2 * i + 1], octaves, 2.0, 0.5, seed);
    }
}

void mtref_noise3(const double *xyz, int n, unsigned seed, double *perlin, double *fbm)
{
    for (int i = 0; i < n; ++i) {
        perlin[i] = noise_perlin3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], seed);
        fbm[i] = noise_fbm3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], 3, 2.0, 0.5, seed);
    }
}

void mtref_sun_disk_dir(const double sun[3], double radius_deg, const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 d = sky_sun_disk_dir(vec3(sun[0], sun[1], sun[2]), radius_deg, u[2 * i], u[2 * i + 1]);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}

void mtref_light_cone_dir(const double w[3], double cos_max, const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 d = light_sphere_sample_dir(vec3(w[0], w[1], w[2]), cos_max, u[2 * i], u[2 * i + 1]);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}

void mtref_cosine_dir(const double nrm[3], const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 d = sampling_cosine_hemisphere(vec3(nrm[0], nrm[1], nrm[2]), u[2 * i], u[2 * i + 1]);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}

void mtref_ggx_half(const double nrm[3], double alpha, const double *u, int n, double *out)
{
    for (int i = 0; i < n; ++i) {
        Vec3 d = sampling_ggx_half_vector(vec3(nrm[0], nrm[1], nrm[2]), alpha, u[2 * i], u[2 * i + 1]);
        out[3 * i] = d.x; out[3 * i + 1] = d.y; out[3 * i + 2] = d.z;
    }
}

double mtref_fresnel_dielectric(double cos_i, double eta)
{
    return sampling_fresnel_dielectric(cos_i, eta);
}

double mtref_specular_probability(double metallic, double roughness, double f0)
{
    return sampling_specular_probability(metallic, roughness, f0);
}

/* material_shade_local / material_shade_pbr on material `mi` of scene h. */
void mtref_shade(void *h, int mi, const double N[3], const double L[3], const double V[3],
                 const double light[3], double out[3])
{
    RefScene *rs = (RefScene *)h;
    const Material *m = scene_material(&rs->scene, mi);
    Vec3 n = vec3(N[0], N[1], N[2]), l = vec3(L[0], L[1], L[2]), v = vec3(V[0], V[1], V[2]);
    Vec3 c = vec3(light[0], light[1], light[2]);
    Vec3 r = (m && m->pbr) ? material_shade_pbr(m, n, l, v, c) : material_shade_local(m, n, l, v, c);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

unsigned char mtref_to_byte(double c)
{
    return pathtrace_to_byte(c);
}

/* camera_ray_dof (reference src/camera.c:83-134) for n (u, v, r1, r2)
 * quadruples: out holds origin and direction, 6 doubles per ray. */
void mtref_camera_ray_dof(const mh_camera *c, const double *uvr, int n, double *out)
{
    Camera cam = cam_from(c);
    for (int i = 0; i < n; ++i) {
        Ray r = camera_ray_dof(&cam, uvr[4 * i], uvr[4 * i + 1], uvr[4 * i + 2], uvr[4 * i + 3]);
        out[6 * i] = r.origin.x; out[6 * i + 1] = r.origin.y; out[6 * i + 2] = r.origin.z;
        out[6 * i + 3] = r.dir.x; out[6 * i + 4] = r.dir.y; out[6 * i + 5] = r.dir.z;
    }
}

/* texture_albedo_uv (reference src/texture.c:269-327) of one material at n
 * points (3 doubles each) with uvs (2 doubles each). */
void mtref_texture_albedo(const mh_material *a, const double *p, const double *uv, int n, double *out)
{
    Material m;
    memset(&m, 0, sizeof m);
    mat_from(a, &m);
    for (int i = 0; i < n; ++i) {
        Vec3 c = texture_albedo_uv(&m, vec3(p[3 * i], p[3 * i + 1], p[3 * i + 2]), uv[2 * i], uv[2 * i + 1]);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

/* render_image_ex (adaptive or fixed Whitted) with the last render's sample
 * statistics. */
int mtref_render_ex(void *h, const mh_camera *c, int w, int hgt, int n0, int depth, int adaptive, int nmax,
                    double tau, unsigned char *rgb, unsigned long long *total, int *max_n)
{
    RefScene *rs = (RefScene *)h;
    if (!rs) return 1;
    Camera cam = cam_from(c);
    RenderParams p;
    memset(&p, 0, sizeof p);
    p.samples_per_pixel = n0;
    p.max_depth = depth;
    p.adaptive = adaptive;
    p.adaptive_max_spp = nmax;
    p.adaptive_tau = tau;
    int rc = render_image_ex(&rs->scene, &cam, w, hgt, &p, rgb);
    if (total) *total = render_last_total_samples();
    if (max_n) *max_n = render_last_max_samples();
    return rc;
}

/* flare_apply with FlareParams built field by field from mh_flare_params. */
int mtref_flare(unsigned char *rgb, int w, int hgt, const mh_flare_params *p)
{
    if (!p) return flare_apply(rgb, w, hgt, NULL);
    FlareParams f;
    f.enabled = p->enabled;
    f.intensity = p->intensity;
    f.threshold = p->threshold;
    f.streak_length = p->streak_length;
    f.tint = vec3(p->tint.x, p->tint.y, p->tint.z);
    return flare_apply(rgb, w, hgt, &f);
}
