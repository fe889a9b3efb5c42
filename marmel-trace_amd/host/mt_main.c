/*
 * mt_main.c - `marmel` command-line renderer: generate a seeded scene, build
 * the BVH on the host, upload it and render on the GPU, write BMP / PPM.
 *
 * Flags follow the reference CLI where they overlap (reference
 * src/main.c:69-477): --width --height --samples --depth --out
 * --pathtrace / --no-pathtrace (Whitted) --no-adaptive (accepted; the GPU
 * Whitted path is the fixed-spp renderer). Scene selection is by generator:
 * --scene-gen spheres|terrain|icosphere --seed N [--grid N].
 * Exit codes: 0 success, 1 usage / setup / render failure.
 */
#define _POSIX_C_SOURCE 200809L

#include "marmel_hip.h"
#include "marmel_host.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void usage(void)
{
    fprintf(stderr,
            "usage: marmel [--width N] [--height N] [--samples N] [--depth N] [--out FILE.bmp|.ppm]\n"
            "              [--pathtrace | --no-pathtrace] [--no-adaptive] [--fp64]\n"
            "              [--scene-gen spheres|terrain|icosphere] [--seed N] [--grid N]\n");
}

int main(int argc, char **argv)
{
    int w = 320, h = 240, spp = 16, depth = 6, pt = 1, fp64 = 0, grid = 0;
    unsigned seed = 1;
    const char *out = "output/marmel.bmp";
    const char *gen = "spheres";
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!strcmp(a, "--width") && v) { w = atoi(v); ++i; }
        else if (!strcmp(a, "--height") && v) { h = atoi(v); ++i; }
        else if (!strcmp(a, "--samples") && v) { spp = atoi(v); ++i; }
        else if (!strcmp(a, "--depth") && v) { depth = atoi(v); ++i; }
        else if (!strcmp(a, "--out") && v) { out = v; ++i; }
        else if (!strcmp(a, "--seed") && v) { seed = (unsigned)strtoul(v, NULL, 10); ++i; }
        else if (!strcmp(a, "--grid") && v) { grid = atoi(v); ++i; }
        else if (!strcmp(a, "--scene-gen") && v) { gen = v; ++i; }
        else if (!strcmp(a, "--pathtrace")) pt = 1;
        else if (!strcmp(a, "--no-pathtrace")) pt = 0;
        else if (!strcmp(a, "--no-adaptive") || !strcmp(a, "--no-progress")) { }
        else if (!strcmp(a, "--fp64")) fp64 = 1;
        else { usage(); return 1; }
    }
    if (w <= 0 || h <= 0 || spp < 1 || depth < 0) { usage(); return 1; }
    mt_gen_params gp;
    memset(&gp, 0, sizeof gp);
    gp.seed = seed;
    gp.grid = grid;
    if (!strcmp(gen, "spheres")) gp.kind = MT_GEN_SPHERES;
    else if (!strcmp(gen, "terrain")) gp.kind = MT_GEN_TERRAIN;
    else if (!strcmp(gen, "icosphere")) gp.kind = MT_GEN_ICOSPHERE;
    else { usage(); return 1; }

    mt_scene *sc = mt_scene_new();
    mt_gen_info info;
    mt_bvh_stats bs;
    mh_scene_soa soa;
    if (!sc || mt_scene_generate(sc, &gp, &info) || mt_scene_build_bvh(sc, NULL, &bs) || mt_scene_soa(sc, &soa)) {
        fprintf(stderr, "error: scene setup failed: %s\n", mt_last_error());
        mt_scene_free(sc);
        return 1;
    }
    mh_scene *ds = NULL;
    if (mh_scene_create(&soa, &ds)) {
        fprintf(stderr, "error: scene upload failed: %s\n", mh_last_error());
        mt_scene_free(sc);
        return 1;
    }
    mh_camera cam;
    mt_camera_create(info.eye, info.at, info.up, info.vfov, (double)w / (double)h, &cam);
    unsigned char *rgb = (unsigned char *)malloc((size_t)w * (size_t)h * 3);
    mh_render_params rp;
    memset(&rp, 0, sizeof rp);
    rp.width = w;
    rp.height = h;
    rp.spp = spp;
    rp.max_depth = depth;
    rp.mode = pt ? MH_MODE_PATHTRACE : MH_MODE_WHITTED;
    rp.precision = fp64 ? MH_PREC_FP64 : MH_PREC_FP32;
    int rc = rgb ? mh_render(ds, &cam, &rp, rgb, NULL) : 1;
    if (rc) {
        fprintf(stderr, "error: rendering failed (code %d): %s\n", rc, mh_last_error());
    } else {
        size_t n = strlen(out);
        int ppm = n > 4 && !strcmp(out + n - 4, ".ppm");
        rc = ppm ? mt_write_ppm(out, rgb, w, h) : mt_write_bmp(out, rgb, w, h);
        if (rc) fprintf(stderr, "error: %s\n", mt_last_error());
        mh_timing t;
        mh_last_timing(&t);
        printf("scene: %d prims (%d triangles, %d spheres), bvh %d nodes depth %d (%.2f s)\n", info.n_prims,
               info.n_triangles, info.n_spheres, bs.n_nodes, bs.max_depth, bs.build_seconds);
        printf("mode: %s %s, %dx%d, %d spp, depth %d\n", pt ? "pathtrace" : "whitted", fp64 ? "fp64" : "fp32", w, h,
               spp, depth);
        printf("render: %.3f ms, %.2f Msamples/s, %.2f Mrays/s -> %s\n", t.kernel_ms,
               (double)t.samples / (t.kernel_ms * 1e3), (double)t.rays / (t.kernel_ms * 1e3), out);
    }
    free(rgb);
    mh_scene_destroy(ds);
    mt_scene_free(sc);
    return rc ? 1 : 0;
}
