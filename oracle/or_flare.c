/* or_flare.c - CPU restatement of the reference's anamorphic streak flare
 * (reference src/flare.c:60-173, flare_filter_channel src/flare.c:28-58).
 *
 * Test infrastructure: the checker the GPU post-pass is compared with, never
 * a product path. Pinned by the live reference (tests/test_flare.py, when
 * oracle/_ref is built), by golden vectors the reference wrote
 * (tests/golden/flare.npz) and by the reference's own expectations
 * (tests/test_anamorphic.c:131-172: invalid arguments, below-threshold rows
 * unchanged, a symmetric streak around a point light, other rows untouched).
 *
 * Arithmetic follows the reference type for type: highlight extraction in
 * double, the recurrences and the composite in float, alpha = expf(-1/r).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>

/* Streak of one channel at one scale added to dst: forward and backward
 * first-order recurrences acc = src + alpha * acc, combined as
 * (fwd + bwd - src) * weight * 0.045. */
static void streak(const float *src, float *dst, float *fwd, int n, float radius, float weight)
{
    if (radius < 1.0f) radius = 1.0f;
    float alpha = expf(-1.0f / radius);
    float norm = weight * 0.045f;
    float acc = 0.0f;
    for (int x = 0; x < n; ++x) {
        acc = src[x] + alpha * acc;
        fwd[x] = acc;
    }
    acc = 0.0f;
    float *bwd = fwd + n;
    for (int x = n - 1; x >= 0; --x) {
        acc = src[x] + alpha * acc;
        bwd[x] = acc;
    }
    for (int x = 0; x < n; ++x) dst[x] += (fwd[x] + bwd[x] - src[x]) * norm;
}

int or_flare_apply(unsigned char *rgb, int width, int height, const mh_flare_params *p)
{
    if (!rgb || width <= 0 || height <= 0 || !p) return 1;
    if (!p->enabled || p->intensity <= 0.0) return 0;
    double th = p->threshold < 0.0 ? 0.0 : (p->threshold > 0.999 ? 0.999 : p->threshold);
    double len = p->streak_length < 0.005 ? 0.005 : (p->streak_length > 1.0 ? 1.0 : p->streak_length);
    double tx = p->tint.x, ty = p->tint.y, tz = p->tint.z;
    if (tx * tx + ty * ty + tz * tz < 1e-6) {
        tx = 0.15;
        ty = 0.45;
        tz = 1.0;
    }
    size_t w = (size_t)width;
    float *buf = (float *)malloc(8 * w * sizeof(float));
    if (!buf) return 2;
    float *src[3] = {buf, buf + w, buf + 2 * w};
    float *acc[3] = {buf + 3 * w, buf + 4 * w, buf + 5 * w};
    float *tmp = buf + 6 * w;
    float r1 = (float)(len * (double)width * 0.12);
    if (r1 < 2.0f) r1 = 2.0f;
    float r2 = (float)(len * (double)width);
    if (r2 < 6.0f) r2 = 6.0f;
    for (int y = 0; y < height; ++y) {
        unsigned char *row = rgb + (size_t)y * w * 3;
        int any = 0;
        for (size_t x = 0; x < w; ++x) {
            double c[3] = {(double)row[3 * x] / 255.0, (double)row[3 * x + 1] / 255.0, (double)row[3 * x + 2] / 255.0};
            double lum = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2];
            double mx = c[0] > c[1] ? (c[0] > c[2] ? c[0] : c[2]) : (c[1] > c[2] ? c[1] : c[2]);
            double m = 0.5 * (lum + mx);
            double t[3] = {tx, ty, tz};
            for (int k = 0; k < 3; ++k) {
                acc[k][x] = 0.0f;
                src[k][x] = 0.0f;
            }
            if (m > th) {
                double e = (m - th) / (1.0 - th) * p->intensity;
                for (int k = 0; k < 3; ++k) src[k][x] = (float)((0.25 * c[k] + 0.75 * t[k]) * e);
                any = 1;
            }
        }
        if (!any) continue;
        for (int k = 0; k < 3; ++k) streak(src[k], acc[k], tmp, width, r1, 0.45f);
        for (int k = 0; k < 3; ++k) streak(src[k], acc[k], tmp, width, r2, 0.55f);
        for (size_t x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) {
                int v = (int)((float)row[3 * x + k] + acc[k][x] * 255.0f + 0.5f);
                row[3 * x + k] = (unsigned char)(v > 255 ? 255 : (v < 0 ? 0 : v));
            }
    }
    free(buf);
    return 0;
}
