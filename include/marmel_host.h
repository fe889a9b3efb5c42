/*
 * marmel_host.h - C ABI of libmarmel_host, the C11 host side of the
 * marmel MI355X path tracer.
 *
 * The host owns what the reference keeps on the CPU around the renderer:
 * the scene container (reference src/scene.h Scene), material / sky presets
 * (reference src/scene_desc.c:655-700, src/material.c:59-96), the look-at
 * camera (reference src/camera.c:24-66), the binned-SAH BVH build (reference
 * src/bvh.c:180-422), flattened here into the mh_scene_soa arrays that
 * libmarmel_hip uploads once, the seeded benchmark scene generators, and the
 * BMP / PPM writers (reference src/bmp.c:38-230).
 *
 * All arrays returned through mh_scene_soa stay owned by the mt_scene and
 * remain valid until the next mutating call or mt_scene_free.
 */
#ifndef MARMEL_HOST_H
#define MARMEL_HOST_H

#include "marmel_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mt_scene mt_scene;

mt_scene *mt_scene_new(void);
void mt_scene_free(mt_scene *s);

/* Defaults: reference sky_default_params (src/material.c:59-96) and a zeroed
 * opaque material with ior 1 (reference sd_opaque_preset). */
void mt_sky_default(mh_sky *sky);
void mt_material_default(mh_material *m);
/* Named presets "water", "glass", "opaque" (reference src/scene_desc.c:655-713).
 * Returns 0, or -1 for an unknown name. */
int mt_material_preset(const char *name, mh_material *m);

/* Append a material; returns its index or -1. */
int mt_scene_add_material(mt_scene *s, const mh_material *m);
/* Append n primitives in file order (attrs may be NULL); returns the index of
 * the first one or -1. Plane normals are normalised as the reference's
 * prim_plane does. */
int mt_scene_add_prims(mt_scene *s, const mh_prim *prims, const mh_prim_attr *attrs, int n);
void mt_scene_set_sky(mt_scene *s, const mh_sky *sky);
void mt_scene_set_water(mt_scene *s, double level, int material);
/* Fog (default: reference fog_default_params, disabled). */
void mt_fog_default(mh_fog *fog);
void mt_scene_set_fog(mt_scene *s, const mh_fog *fog);
int mt_scene_counts(const mt_scene *s, int *n_prims, int *n_materials);

/* BVH build options. */
typedef struct mt_bvh_options {
    int max_leaf;      /* 1..8, default 4 */
    int bins;          /* SAH bins per axis, default 16 */
    int threads;       /* reserved; 0 = default */
} mt_bvh_options;

typedef struct mt_bvh_stats {
    int n_nodes, n_leaves, max_depth;
    double sah_cost;    /* expected cost, interior 1 + leaf prims */
    double build_seconds;
} mt_bvh_stats;

/* Build the binary SAH BVH over the non-plane primitives (planes go to the
 * side list, as reference src/bvh.c:364-422). opts may be NULL. */
int mt_scene_build_bvh(mt_scene *s, const mt_bvh_options *opts, mt_bvh_stats *stats);

/* Flattened view for mh_scene_create (requires a built BVH). */
int mt_scene_soa(const mt_scene *s, mh_scene_soa *out);
/* The same without the tree (n_nodes = n_leaf_prims = n_planes = 0), for
 * mh_scene_create_gpu_bvh, which builds it on the device. */
int mt_scene_soa_no_bvh(const mt_scene *s, mh_scene_soa *out);

/* ---- geometry ingest at scale ------------------------------------------ */

/* Placement of an OBJ mesh (reference ObjTransform, src/obj.h; the scene's
 * `mesh` block, src/scene_desc.h:163-174). A vertex v becomes
 * rotate((v - c) * k * scale) + position, with c the bounding-box centre
 * when auto_center (else 0), k = auto_scale / largest extent when
 * auto_scale > 0 (else 1), rotation about X, then Y, then Z (degrees). */
typedef struct mt_obj_xform {
    double  position[3];
    double  scale[3];
    double  rotate_deg[3];
    int32_t smooth;        /* 1: the file's vertex normals where a face has them */
    int32_t auto_center;
    double  auto_scale;
} mt_obj_xform;
void mt_obj_xform_default(mt_obj_xform *x);

/* Add the triangles of Wavefront OBJ text (v, vn, vt, f; polygons
 * fan-triangulated; negative indices count back from the file's totals)
 * with material `material`: the triangles, normals and uvs the reference
 * loader (src/obj.c:123-300) makes of the same text, parsed on `threads`
 * threads (0: all cores). Returns the number of triangles added, or < 0. */
long mt_scene_add_obj(mt_scene *s, const char *data, size_t len, int material,
                      const mt_obj_xform *xf, int threads);
long mt_scene_add_obj_file(mt_scene *s, const char *path, int material,
                           const mt_obj_xform *xf, int threads);

/* A procedural ocean (reference SceneOceanDesc, src/scene_desc.h:177-192;
 * scene_grow_ocean, src/scene.c:1155-1331): a res_x x res_z Gerstner grid of
 * smooth triangles over size[0] x size[1] around center, plus side skirts
 * and a floor when depth > 0. */
typedef struct mt_ocean_desc {
    double   center[3];
    double   size[2];
    int32_t  res_x, res_z;
    double   amplitude, wavelength;
    double   direction[3];
    double   steepness, chop, chop_wavelength, depth;
    uint32_t seed;
    int32_t  reserved;
} mt_ocean_desc;
void mt_ocean_default(mt_ocean_desc *o);   /* the reference block's defaults */
/* Adds the ocean's triangles; the scene's first ocean also sets the water
 * level, the water material and the water-height surface (mh_ocean). */
long mt_scene_add_ocean(mt_scene *s, const mt_ocean_desc *o, int material, int threads);

/* Look-at camera with vertical fov in degrees (reference camera_create). */
void mt_camera_create(const double eye[3], const double at[3], const double up[3],
                      double vfov_deg, double aspect, mh_camera *out);

/* Seeded benchmark scenes. */
enum {
    MT_GEN_SPHERES = 1,   /* plane + 64 spheres (Lambert / mirror / glass / GGX) */
    MT_GEN_TERRAIN = 2,   /* n x n fBm heightfield + 200 spheres, 8 emissive   */
    MT_GEN_ICOSPHERE = 3, /* displaced icosphere + spheres (batched scenes)     */
    MT_GEN_MATERIALS = 4, /* plane + 63 spheres, 64 PBR materials (adjoint)    */
    MT_GEN_FOREST = 5     /* `grid` trees of `subdiv` levels + pond (stress)   */
};

typedef struct mt_gen_params {
    int kind;             /* MT_GEN_* */
    unsigned seed;
    int grid;             /* terrain: vertices per side (default 1024)     */
    int n_spheres;        /* sphere count (defaults per kind)              */
    int n_emissive;       /* emissive spheres (terrain default 8)          */
    int subdiv;           /* icosphere subdivision level                   */
    int count;            /* icospheres per scene (0 or 1: one)            */
} mt_gen_params;

typedef struct mt_gen_info {
    double eye[3], at[3], up[3];
    double vfov;
    int n_prims, n_triangles, n_spheres, n_emissive;
} mt_gen_info;

int mt_scene_generate(mt_scene *s, const mt_gen_params *p, mt_gen_info *info);

/* Image output (reference bmp_write / ppm_write byte layout). 0 on success. */
int mt_write_bmp(const char *path, const unsigned char *rgb, int width, int height);
int mt_write_ppm(const char *path, const unsigned char *rgb, int width, int height);
/* BMP file bytes into buf (size >= 54 + padded rows); returns the size. */
long mt_encode_bmp(const unsigned char *rgb, int width, int height, unsigned char *buf, long cap);

const char *mt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MARMEL_HOST_H */
