/* hd_host_asan.c - the host scene library (marmel-trace_amd/host) and the
 * FP64 oracle compiled into one executable under AddressSanitizer +
 * UndefinedBehaviorSanitizer (test infrastructure only). It drives every
 * entry point the tests use: the five seeded generators, the BVH build and
 * flattening, OBJ ingest (including malformed and out-of-range faces),
 * oceans, BMP encoding; then the oracle's scene build, ray queries, path
 * tracer, Whitted, adaptive Whitted, adjoint and flare on those scenes.
 * Prints one line per stage and "ok" at the end; any sanitizer report aborts
 * with a non-zero status. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "marmel_host.h"
#include "oracle.h"

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__);       \
            exit(2);                                                              \
        }                                                                         \
    } while (0)

static const char *OBJ_TEXT =
    "# quad, triangle, negative indices, normals and uvs, malformed lines\n"
    "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nv 0.5 0.5 1\n"
    "vn 0 0 1\nvn 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\n"
    "f 1/1/1 2/2/1 3/3/1 4/1/1\n"
    "f -1 -2 -3\n"
    "f 1//2 3//2 5//2\n"
    "f 1 2\n"
    "f 1 99 2\n"
    "f 0 1 2\n"
    "v nan 0 0\n"
    "f garbage\n"
    "vt 0.5\n";

static void render_checks(const mh_scene_soa *soa, const mh_camera *cam, const char *tag)
{
    or_scene *o = or_scene_create(soa);
    CHECK(o);
    enum { W = 24, H = 18 };
    unsigned char rgb[W * H * 3];
    double mean[W * H * 3], var[W * H * 3], sec;
    CHECK(or_render(o, cam, W, H, 4, 6, 0, 0, 0, 0, 2, rgb, mean, var, &sec) == 0);
    CHECK(or_render(o, cam, W, H, 4, 6, 1, 3, 2, 9, 2, rgb, mean, NULL, &sec) == 0);
    int ns[W * H], mx;
    long long tot;
    CHECK(or_render_adaptive(o, cam, W, H, 2, 4, 0, 0.0, 2, rgb, mean, ns, &tot, &mx) == 0);
    mh_ray rays[64];
    mh_hit hits[64];
    unsigned char occ[64];
    for (int i = 0; i < 64; ++i) {
        double u = (i % 8 + 0.5) / 8, v = (i / 8 + 0.5) / 8, r[6];
        or_camera_ray(cam, u, v, 0.3, 0.7, r);
        rays[i].ox = r[0]; rays[i].oy = r[1]; rays[i].oz = r[2];
        rays[i].dx = r[3]; rays[i].dy = r[4]; rays[i].dz = r[5];
    }
    CHECK(or_intersect(o, rays, 64, 1e-4, 1e30, hits) == 0);
    CHECK(or_occluded(o, rays, 64, 1e-4, 1e30, occ) == 0);
    double c[3];
    CHECK(or_radiance(o, &rays[5], 6, 77u, c) == 0);
    int nm = soa->n_materials;
    double *grad = calloc((size_t)nm * 8 + 3, sizeof(double));
    double *dimg = malloc(sizeof(double) * W * H * 3);
    for (int i = 0; i < W * H * 3; ++i) dimg[i] = (i % 7) * 0.1 - 0.3;
    CHECK(or_adjoint(o, cam, W, H, 2, 4, 0, dimg, grad, 2) == 0);
    free(grad);
    free(dimg);
    mh_flare_params fp;
    memset(&fp, 0, sizeof fp);
    fp.enabled = 1;
    fp.intensity = 1.0;
    fp.threshold = 0.5;
    fp.streak_length = 0.5;
    fp.tint.x = 0.2; fp.tint.y = 0.5; fp.tint.z = 1.0;
    CHECK(or_flare_apply(rgb, W, H, &fp) == 0);
    or_scene_destroy(o);
    printf("%s: oracle renders, queries, adjoint, flare\n", tag);
}

int main(void)
{
    for (int kind = MT_GEN_SPHERES; kind <= MT_GEN_FOREST; ++kind) {
        for (unsigned seed = 1; seed <= 2; ++seed) {
            mt_scene *s = mt_scene_new();
            CHECK(s);
            mt_gen_params p;
            memset(&p, 0, sizeof p);
            p.kind = kind;
            p.seed = seed;
            p.grid = kind == MT_GEN_TERRAIN ? 24 : (kind == MT_GEN_FOREST ? 2 : 0);
            p.subdiv = kind == MT_GEN_ICOSPHERE || kind == MT_GEN_FOREST ? 2 : 0;
            p.count = kind == MT_GEN_ICOSPHERE ? 3 : 0;
            mt_gen_info info;
            CHECK(mt_scene_generate(s, &p, &info) == 0);
            mt_bvh_stats st;
            CHECK(mt_scene_build_bvh(s, NULL, &st) == 0);
            mh_scene_soa soa;
            CHECK(mt_scene_soa(s, &soa) == 0);
            mh_scene_soa nob;
            CHECK(mt_scene_soa_no_bvh(s, &nob) == 0);
            mh_camera cam;
            mt_camera_create(info.eye, info.at, info.up, info.vfov, 4.0 / 3.0, &cam);
            char tag[64];
            snprintf(tag, sizeof tag, "generator %d seed %u (%d prims, %d nodes)", kind, seed, soa.n_prims,
                     soa.n_nodes);
            render_checks(&soa, &cam, tag);
            mt_scene_free(s);
        }
    }

    /* OBJ ingest (with a default material) and an ocean */
    mt_scene *s = mt_scene_new();
    mh_material m;
    mt_material_default(&m);
    CHECK(mt_scene_add_material(s, &m) == 0);
    m.is_water = 1;
    m.transparency = 1.0;
    m.ior = 1.333;
    CHECK(mt_scene_add_material(s, &m) == 1);
    mt_obj_xform xf;
    mt_obj_xform_default(&xf);
    xf.smooth = 1;
    long nt = mt_scene_add_obj(s, OBJ_TEXT, strlen(OBJ_TEXT), 0, &xf, 3);
    printf("obj: %ld triangles\n", nt);
    printf("obj: empty text -> %ld, faces without vertices -> %ld\n", mt_scene_add_obj(s, "", 0, 0, &xf, 1),
           mt_scene_add_obj(s, "f 1 2 3\n", 8, 0, &xf, 1));
    mt_ocean_desc od;
    mt_ocean_default(&od);
    od.res_x = 12;
    od.res_z = 10;
    od.depth = 3.0;
    long no = mt_scene_add_ocean(s, &od, 1, 2);
    CHECK(no > 0);
    printf("ocean: %ld triangles\n", no);
    mh_sky sky;
    mt_sky_default(&sky);
    mt_scene_set_sky(s, &sky);
    mh_fog fog;
    mt_fog_default(&fog);
    fog.density = 0.05;
    mt_scene_set_fog(s, &fog);
    mt_bvh_options bo = {2, 8, 0};
    CHECK(mt_scene_build_bvh(s, &bo, NULL) == 0);
    mh_scene_soa soa;
    CHECK(mt_scene_soa(s, &soa) == 0);
    double eye[3] = {0, 1.5, -8}, at[3] = {0, 0, 0}, up[3] = {0, 1, 0};
    mh_camera cam;
    mt_camera_create(eye, at, up, 50.0, 1.5, &cam);
    render_checks(&soa, &cam, "obj + ocean + fog");
    mt_scene_free(s);

    unsigned char img[5 * 3 * 3], buf[256];
    for (int i = 0; i < 45; ++i) img[i] = (unsigned char)(i * 5);
    CHECK(mt_encode_bmp(img, 5, 3, buf, sizeof buf) > 54);
    CHECK(mt_encode_bmp(img, 5, 3, buf, 10) < 0);
    printf("ok\n");
    return 0;
}
