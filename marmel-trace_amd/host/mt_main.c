/*
 * mt_main.c - `marmel` command-line renderer: generate a seeded scene, build
 * the BVH on the host, upload it and render on the GPU, write BMP / PPM.
 *
 * Flags follow the reference CLI where they overlap (reference
 * src/main.c:69-477, main 690-725): --width --height --samples --depth --out
 * --pathtrace / --no-pathtrace (Whitted, through render_image_ex's drop-in:
 * --adaptive (the default) / --no-adaptive, --adaptive-max N, --adaptive-tau
 * T), --lens-flare / --no-lens-flare with --flare-intensity X (>= 0) and
 * --flare-threshold X (in [0, 1]) for the streak post-pass. Scene selection
 * is by generator: --scene-gen spheres|terrain|icosphere|forest --seed N
 * [--grid N]. Precision: the path tracer runs fp32 (--fp64 for fp64), the
 * Whitted renderer fp64, the reference-parity path (--fp32 for fp32).
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
            "              [--pathtrace | --no-pathtrace] [--fp64 | --fp32]\n"
            "              [--adaptive | --no-adaptive] [--adaptive-max N] [--adaptive-tau T]\n"
            "              [--lens-flare | --no-lens-flare] [--flare-intensity X] [--flare-threshold X]\n"
            "              [--scene-gen spheres|terrain|icosphere|forest] [--seed N] [--grid N]\n");
}

int main(int argc, char **argv)
{
    int w = 320, h = 240, spp = 16, depth = 6, pt = 1, fp64 = 0, fp32 = 0, grid = 0;
    int adaptive = 1, adaptive_max = 0, flare = 0;
    double adaptive_tau = 0.0;
    mh_flare_params fl;
    memset(&fl, 0, sizeof fl);
    fl.intensity = 0.6;        /* the reference's defaults (src/flare.h:38-41) */
    fl.threshold = 0.85;
    fl.streak_length = 0.25;
    fl.tint.x = 0.15;
    fl.tint.y = 0.45;
    fl.tint.z = 1.0;
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
        else if (!strcmp(a, "--adaptive")) adaptive = 1;
        else if (!strcmp(a, "--no-adaptive")) adaptive = 0;
        else if (!strcmp(a, "--adaptive-max") && v) { adaptive_max = atoi(v); ++i; }
        else if (!strcmp(a, "--adaptive-tau") && v) { adaptive_tau = strtod(v, NULL); ++i; }
        else if (!strcmp(a, "--lens-flare")) flare = 1;
        else if (!strcmp(a, "--no-lens-flare")) flare = 0;
        else if (!strcmp(a, "--flare-intensity") && v) {
            fl.intensity = strtod(v, NULL);
            ++i;
            if (!(fl.intensity >= 0.0)) { usage(); return 1; }
        } else if (!strcmp(a, "--flare-threshold") && v) {
            fl.threshold = strtod(v, NULL);
            ++i;
            if (!(fl.threshold >= 0.0 && fl.threshold <= 1.0)) { usage(); return 1; }
        }
        else if (!strcmp(a, "--no-progress")) { }
        else if (!strcmp(a, "--fp64")) fp64 = 1;
        else if (!strcmp(a, "--fp32")) fp32 = 1;
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
    else if (!strcmp(gen, "forest")) gp.kind = MT_GEN_FOREST;
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
    int rc = 1;
    unsigned long long total = 0;
    int max_spp = 0;
    if (rgb && pt) {
        mh_render_params rp;
        memset(&rp, 0, sizeof rp);
        rp.width = w;
        rp.height = h;
        rp.spp = spp;
        rp.max_depth = depth;
        rp.mode = MH_MODE_PATHTRACE;
        rp.precision = fp64 ? MH_PREC_FP64 : MH_PREC_FP32;
        rc = mh_render(ds, &cam, &rp, rgb, NULL);
    } else if (rgb) {
        mh_render_ex_params ep;
        memset(&ep, 0, sizeof ep);
        ep.samples_per_pixel = spp;
        ep.max_depth = depth;
        ep.adaptive = adaptive;
        ep.adaptive_max_spp = adaptive_max;
        ep.adaptive_tau = adaptive_tau;
        rc = mh_render_adaptive(ds, &cam, w, h, &ep, fp32 ? MH_PREC_FP32 : MH_PREC_FP64, rgb, NULL, NULL);
    }
    mh_timing t;
    mh_last_timing(&t);
    mh_last_sample_stats(&total, &max_spp);
    if (rc == 0 && flare) {
        fl.enabled = 1;
        rc = mh_flare_apply(rgb, w, h, &fl);
    }
    if (rc) {
        fprintf(stderr, "error: rendering failed (code %d): %s\n", rc, mh_last_error());
    } else {
        size_t n = strlen(out);
        int ppm = n > 4 && !strcmp(out + n - 4, ".ppm");
        rc = ppm ? mt_write_ppm(out, rgb, w, h) : mt_write_bmp(out, rgb, w, h);
        if (rc) fprintf(stderr, "error: %s\n", mt_last_error());
        printf("scene: %d prims (%d triangles, %d spheres), bvh %d nodes depth %d (%.2f s)\n", info.n_prims,
               info.n_triangles, info.n_spheres, bs.n_nodes, bs.max_depth, bs.build_seconds);
        printf("mode: %s %s, %dx%d, %d spp, depth %d%s\n", pt ? "pathtrace" : (adaptive ? "whitted adaptive" : "whitted"),
               (pt ? fp64 : !fp32) ? "fp64" : "fp32", w, h, spp, depth, flare ? ", lens flare" : "");
        printf("samples: %llu total, %d max per pixel\n", total, max_spp);
        printf("render: %.3f ms, %.2f Msamples/s, %.2f Mrays/s -> %s\n", t.kernel_ms,
               (double)t.samples / (t.kernel_ms * 1e3), (double)t.rays / (t.kernel_ms * 1e3), out);
    }
    free(rgb);
    mh_scene_destroy(ds);
    mt_scene_free(sc);
    return rc ? 1 : 0;
}
