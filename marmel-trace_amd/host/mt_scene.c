/*
 * mt_scene.c - host scene container, presets, camera, binned-SAH BVH build
 * and flattening into mh_scene_soa, and the BMP/PPM writers.
 *
 * BVH: binary, binned SAH over all three centroid axes (16 bins by default),
 * leaves of at most 4 primitives, planes kept out of the tree in a side list
 * (as reference src/bvh.c:364-422 does). Node boxes are the union of the
 * primitive boxes in double, padded by 1e-7 like reference src/bvh.c:163-164
 * and rounded OUTWARD to fp32, so every node box contains its primitives
 * exactly as the double-precision reference box does.
 */
#define _POSIX_C_SOURCE 200809L

#include "marmel_host.h"
#include "mt_internal.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static __thread char g_err[256];

int mt_set_err(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

const char *mt_last_error(void) { return g_err; }

static void drop_bvh(mt_scene *s)
{
    free(s->nodes);
    free(s->leaf_prims);
    free(s->planes);
    s->nodes = NULL;
    s->leaf_prims = NULL;
    s->planes = NULL;
    s->n_nodes = s->n_leaf = s->n_planes = 0;
    s->bvh_valid = 0;
}

mt_scene *mt_scene_new(void)
{
    mt_scene *s = (mt_scene *)calloc(1, sizeof(mt_scene));
    if (!s) return NULL;
    mt_sky_default(&s->sky);
    mt_fog_default(&s->fog);
    s->water_material = -1;
    return s;
}

void mt_scene_free(mt_scene *s)
{
    if (!s) return;
    drop_bvh(s);
    free(s->prims);
    free(s->attrs);
    free(s->mats);
    free(s);
}

static mh_vec3 V(double x, double y, double z)
{
    mh_vec3 v = { x, y, z };
    return v;
}

void mt_sky_default(mh_sky *p)
{
    memset(p, 0, sizeof *p);
    double x = 0.45, y = 0.75, z = -0.50;
    double l2 = x * x + y * y + z * z;
    double k = 1.0 / sqrt(l2);
    p->sun_dir = V(x * k, y * k, z * k);
    p->sun_color = V(1.00, 0.95, 0.85);
    p->horizon_color = V(0.75, 0.85, 1.00);
    p->zenith_color = V(0.35, 0.55, 0.95);
    p->gradient_gamma = 0.6;
    p->sun_glow_exponent = 350.0;
    p->sun_glow_strength = 0.8;
    p->cloud_height = 120.0;
    p->cloud_scale = 0.0025;
    p->cloud_coverage = 0.5;
    p->cloud_softness = 0.12;
    p->cloud_sharpness = 1.5;
    p->cloud_octaves = 5;
    p->seed = 1337u;
}

void mt_material_default(mh_material *m)
{
    memset(m, 0, sizeof *m);
    m->ior = 1.0;
    m->texture_scale = 1.0;
    m->texture_color_a = V(1.0, 1.0, 1.0);
}

int mt_material_preset(const char *name, mh_material *m)
{
    mt_material_default(m);
    if (!name) return -1;
    if (strcmp(name, "opaque") == 0) return 0;
    if (strcmp(name, "water") == 0) {
        m->albedo = V(0.05, 0.15, 0.20);
        m->specular = V(0.90, 0.90, 0.90);
        m->shininess = 256.0;
        m->reflectivity = 1.0;
        m->transparency = 0.85;
        m->ior = 1.33;
        m->is_water = 1;
        m->absorption = V(0.45, 0.12, 0.06);
        m->deep_color = V(0.02, 0.10, 0.16);
        return 0;
    }
    if (strcmp(name, "glass") == 0) {
        m->albedo = V(0.02, 0.02, 0.02);
        m->specular = V(1.0, 1.0, 1.0);
        m->shininess = 256.0;
        m->transparency = 1.0;
        m->ior = 1.5;
        m->deep_color = V(0.5, 0.5, 0.5);
        return 0;
    }
    return -1;
}

int mt_scene_add_material(mt_scene *s, const mh_material *m)
{
    if (!s || !m) return -1;
    if (s->nm == s->capm) {
        int cap = s->capm ? s->capm * 2 : 16;
        mh_material *nm = (mh_material *)realloc(s->mats, sizeof(mh_material) * (size_t)cap);
        if (!nm) return -1;
        s->mats = nm;
        s->capm = cap;
    }
    s->mats[s->nm] = *m;
    return s->nm++;
}

long mt_reserve(mt_scene *s, long extra, int attrs)
{
    if (!s || extra < 0) return -1;
    if ((long)s->n + extra > 0x7fffffffL) return mt_set_err(-1, "too many primitives");
    if (s->n + extra > s->cap) {
        long cap = s->cap ? s->cap : 64;
        while (cap < s->n + extra) cap *= 2;
        if (cap > 0x7fffffffL) cap = 0x7fffffffL;
        mh_prim *np = (mh_prim *)realloc(s->prims, sizeof(mh_prim) * (size_t)cap);
        if (!np) return mt_set_err(-1, "out of memory");
        s->prims = np;
        if (s->attrs) {
            mh_prim_attr *na = (mh_prim_attr *)realloc(s->attrs, sizeof(mh_prim_attr) * (size_t)cap);
            if (!na) return mt_set_err(-1, "out of memory");
            memset(na + s->cap, 0, sizeof(mh_prim_attr) * (size_t)(cap - s->cap));
            s->attrs = na;
        }
        s->cap = (int)cap;
    }
    if (attrs && !s->attrs) {
        s->attrs = (mh_prim_attr *)calloc((size_t)s->cap, sizeof(mh_prim_attr));
        if (!s->attrs) return mt_set_err(-1, "out of memory");
    }
    s->bvh_valid = 0;
    return s->n;
}

int mt_scene_add_prims(mt_scene *s, const mh_prim *prims, const mh_prim_attr *attrs, int n)
{
    if (!s || (!prims && n > 0) || n < 0) return -1;
    if (s->n + n > s->cap) {
        long cap = s->cap ? s->cap : 64;
        while (cap < s->n + n) cap *= 2;
        if (cap > 0x7fffffffL) return mt_set_err(-1, "too many primitives");
        mh_prim *np = (mh_prim *)realloc(s->prims, sizeof(mh_prim) * (size_t)cap);
        if (!np) return -1;
        s->prims = np;
        if (s->attrs) {
            mh_prim_attr *na = (mh_prim_attr *)realloc(s->attrs, sizeof(mh_prim_attr) * (size_t)cap);
            if (!na) return -1;
            s->attrs = na;
        }
        s->cap = (int)cap;
    }
    if (attrs && !s->attrs) {
        s->attrs = (mh_prim_attr *)calloc((size_t)s->cap, sizeof(mh_prim_attr));
        if (!s->attrs) return -1;
    }
    int first = s->n;
    for (int i = 0; i < n; ++i) {
        mh_prim p = prims[i];
        if (p.kind == MH_PRIM_PLANE) {
            /* unit normal, as prim_plane stores it */
            double l2 = p.p1.x * p.p1.x + p.p1.y * p.p1.y + p.p1.z * p.p1.z;
            if (l2 > 1e-12) {
                double k = 1.0 / sqrt(l2);
                p.p1 = V(p.p1.x * k, p.p1.y * k, p.p1.z * k);
            } else {
                p.p1 = V(0.0, 0.0, 0.0);
            }
        }
        if (!attrs) p.flags = 0;
        s->prims[s->n] = p;
        if (s->attrs) {
            if (attrs) s->attrs[s->n] = attrs[i];
            else memset(&s->attrs[s->n], 0, sizeof(mh_prim_attr));
        }
        s->n++;
    }
    s->bvh_valid = 0;
    return first;
}

void mt_scene_set_sky(mt_scene *s, const mh_sky *sky)
{
    if (s && sky) s->sky = *sky;
}

/* Reference fog_default_params (src/material.h:189-197): disabled. */
void mt_fog_default(mh_fog *f)
{
    if (!f) return;
    memset(f, 0, sizeof *f);
    f->density = 0.0;
    f->color.x = 0.70;
    f->color.y = 0.75;
    f->color.z = 0.80;
    f->height = 0.0;
    f->height_falloff = 0.0;
    f->inscatter_strength = 0.50;
    f->sun_anisotropy = 0.70;
    f->noise_scale = 0.0;
    f->noise_amount = 0.0;
    f->shadow_steps = 0;
}

void mt_scene_set_fog(mt_scene *s, const mh_fog *fog)
{
    if (s && fog) s->fog = *fog;
}

void mt_scene_set_water(mt_scene *s, double level, int material)
{
    if (!s) return;
    s->water_level = level;
    s->water_material = material;
}

int mt_scene_counts(const mt_scene *s, int *n_prims, int *n_materials)
{
    if (!s) return 1;
    if (n_prims) *n_prims = s->n;
    if (n_materials) *n_materials = s->nm;
    return 0;
}

/* ---- camera (reference src/camera.c:24-66) ----------------------------- */

void mt_camera_create(const double eye[3], const double at[3], const double up[3], double vfov_deg,
                      double aspect, mh_camera *out)
{
    double f[3] = { at[0] - eye[0], at[1] - eye[1], at[2] - eye[2] };
    double l2 = f[0] * f[0] + f[1] * f[1] + f[2] * f[2];
    if (l2 <= 1e-12) {
        f[0] = 0.0; f[1] = 0.0; f[2] = -1.0;
    } else {
        double k = 1.0 / sqrt(l2);
        f[0] *= k; f[1] *= k; f[2] *= k;
        if (f[0] * f[0] + f[1] * f[1] + f[2] * f[2] <= 1e-12) { f[0] = 0.0; f[1] = 0.0; f[2] = -1.0; }
    }
    double r[3] = { f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0] };
    if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] <= 1e-12) {
        double alt[3] = { 0.0, 1.0, 0.0 };
        if (!(fabs(f[1]) < 0.99)) { alt[0] = 1.0; alt[1] = 0.0; }
        r[0] = f[1] * alt[2] - f[2] * alt[1];
        r[1] = f[2] * alt[0] - f[0] * alt[2];
        r[2] = f[0] * alt[1] - f[1] * alt[0];
    }
    double rl2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (rl2 <= 1e-12) {
        r[0] = r[1] = r[2] = 0.0;
    } else {
        double k = 1.0 / sqrt(rl2);
        r[0] *= k; r[1] *= k; r[2] *= k;
    }
    double u[3] = { r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0] };
    double hh = tan(vfov_deg * (3.14159265358979323846 / 180.0) * 0.5);
    double hw = aspect * hh;
    mh_vec3 hor = V(r[0] * (2.0 * hw), r[1] * (2.0 * hw), r[2] * (2.0 * hw));
    mh_vec3 ver = V(u[0] * (2.0 * hh), u[1] * (2.0 * hh), u[2] * (2.0 * hh));
    out->position = V(eye[0], eye[1], eye[2]);
    out->horizontal = hor;
    out->vertical = ver;
    double c[3] = { eye[0] + f[0], eye[1] + f[1], eye[2] + f[2] };
    double hv[3] = { hor.x * 0.5 + ver.x * 0.5, hor.y * 0.5 + ver.y * 0.5, hor.z * 0.5 + ver.z * 0.5 };
    out->lower_left = V(c[0] - hv[0], c[1] - hv[1], c[2] - hv[2]);
    out->aperture = 0.0;
    double d[3] = { at[0] - eye[0], at[1] - eye[1], at[2] - eye[2] };
    out->focus_distance = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    out->dome_radius = 0.0;
    out->anamorphic_squeeze = 1.0;
}

/* ---- BVH build -------------------------------------------------------- */

typedef struct {
    double mn[3], mx[3];
} aabb;

static void box_empty(aabb *b)
{
    for (int k = 0; k < 3; ++k) {
        b->mn[k] = DBL_MAX;
        b->mx[k] = -DBL_MAX;
    }
}

static void box_grow(aabb *b, const aabb *o)
{
    for (int k = 0; k < 3; ++k) {
        if (o->mn[k] < b->mn[k]) b->mn[k] = o->mn[k];
        if (o->mx[k] > b->mx[k]) b->mx[k] = o->mx[k];
    }
}

static double box_area(const aabb *b)
{
    double dx = b->mx[0] - b->mn[0], dy = b->mx[1] - b->mn[1], dz = b->mx[2] - b->mn[2];
    if (!(dx > 0.0)) dx = 0.0;
    if (!(dy > 0.0)) dy = 0.0;
    if (!(dz > 0.0)) dz = 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

static void prim_box(const mh_prim *p, aabb *b)
{
    const mh_vec3 *a = &p->p0, *c = &p->p1, *e = &p->p2;
    switch (p->kind) {
    case MH_PRIM_SPHERE:
        b->mn[0] = a->x - p->r0; b->mn[1] = a->y - p->r0; b->mn[2] = a->z - p->r0;
        b->mx[0] = a->x + p->r0; b->mx[1] = a->y + p->r0; b->mx[2] = a->z + p->r0;
        break;
    case MH_PRIM_BOX:
        b->mn[0] = a->x - fabs(c->x); b->mn[1] = a->y - fabs(c->y); b->mn[2] = a->z - fabs(c->z);
        b->mx[0] = a->x + fabs(c->x); b->mx[1] = a->y + fabs(c->y); b->mx[2] = a->z + fabs(c->z);
        break;
    case MH_PRIM_TRIANGLE:
        b->mn[0] = fmin(a->x, fmin(c->x, e->x)); b->mx[0] = fmax(a->x, fmax(c->x, e->x));
        b->mn[1] = fmin(a->y, fmin(c->y, e->y)); b->mx[1] = fmax(a->y, fmax(c->y, e->y));
        b->mn[2] = fmin(a->z, fmin(c->z, e->z)); b->mx[2] = fmax(a->z, fmax(c->z, e->z));
        break;
    case MH_PRIM_CYLINDER: {
        double r = p->r0 > p->r1 ? p->r0 : p->r1;
        if (r < 0.0) r = 0.0;
        b->mn[0] = fmin(a->x, c->x) - r; b->mn[1] = fmin(a->y, c->y) - r; b->mn[2] = fmin(a->z, c->z) - r;
        b->mx[0] = fmax(a->x, c->x) + r; b->mx[1] = fmax(a->y, c->y) + r; b->mx[2] = fmax(a->z, c->z) + r;
        break;
    }
    default:
        box_empty(b);
    }
}

static float round_down(double x)
{
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}

static float round_up(double x)
{
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

typedef struct {
    int node, begin, end, depth;
} task_t;

#define MAX_BINS 64

typedef struct {
    mh_bvh_node *nodes;
    int n, cap;
} nodevec;

static int push_node(nodevec *v)
{
    if (v->n == v->cap) {
        int cap = v->cap ? v->cap * 2 : 1024;
        mh_bvh_node *nn = (mh_bvh_node *)realloc(v->nodes, sizeof(mh_bvh_node) * (size_t)cap);
        if (!nn) return -1;
        v->nodes = nn;
        v->cap = cap;
    }
    memset(&v->nodes[v->n], 0, sizeof(mh_bvh_node));
    return v->n++;
}

int mt_scene_build_bvh(mt_scene *s, const mt_bvh_options *opts, mt_bvh_stats *stats)
{
    if (!s) return mt_set_err(1, "NULL scene");
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    drop_bvh(s);
    int max_leaf = opts && opts->max_leaf > 0 ? opts->max_leaf : 4;
    if (max_leaf > 8) max_leaf = 8;
    int nbins = opts && opts->bins > 1 ? opts->bins : 16;
    if (nbins > MAX_BINS) nbins = MAX_BINS;

    int nt = 0, np = 0;
    for (int i = 0; i < s->n; ++i) {
        if (s->prims[i].kind == MH_PRIM_PLANE) ++np;
        else ++nt;
    }
    s->planes = (int *)malloc(sizeof(int) * (size_t)(np ? np : 1));
    int *idx = (int *)malloc(sizeof(int) * (size_t)(nt ? nt : 1));
    aabb *pb = (aabb *)malloc(sizeof(aabb) * (size_t)(nt ? nt : 1));
    double *cen = (double *)malloc(sizeof(double) * 3 * (size_t)(nt ? nt : 1));
    if (!s->planes || !idx || !pb || !cen) {
        free(idx); free(pb); free(cen);
        return mt_set_err(2, "out of memory");
    }
    np = nt = 0;
    for (int i = 0; i < s->n; ++i) {
        if (s->prims[i].kind == MH_PRIM_PLANE) {
            s->planes[np++] = i;
            continue;
        }
        prim_box(&s->prims[i], &pb[nt]);
        for (int k = 0; k < 3; ++k) cen[3 * nt + k] = 0.5 * (pb[nt].mn[k] + pb[nt].mx[k]);
        idx[nt++] = i;
    }
    s->n_planes = np;

    nodevec nv = { NULL, 0, 0 };
    int leaves = 0, max_depth = 0;
    double sah = 0.0, root_area = 0.0;
    /* idx holds LOCAL slots 0..nt-1 during the build; map to prims at the end */
    int *slot = (int *)malloc(sizeof(int) * (size_t)(nt ? nt : 1));
    task_t *stack = (task_t *)malloc(sizeof(task_t) * 128);
    if (!slot || !stack) {
        free(idx); free(pb); free(cen); free(slot); free(stack);
        return mt_set_err(2, "out of memory");
    }
    for (int i = 0; i < nt; ++i) slot[i] = i;
    int sp = 0, scap = 128;
    if (nt > 0) {
        int root = push_node(&nv);
        if (root < 0) goto oom;
        stack[sp++] = (task_t){ root, 0, nt, 1 };
    }
    while (sp > 0) {
        task_t t = stack[--sp];
        int count = t.end - t.begin;
        aabb nb, cb;
        box_empty(&nb);
        box_empty(&cb);
        for (int i = t.begin; i < t.end; ++i) {
            int q = slot[i];
            box_grow(&nb, &pb[q]);
            aabb c = { { cen[3 * q], cen[3 * q + 1], cen[3 * q + 2] }, { cen[3 * q], cen[3 * q + 1], cen[3 * q + 2] } };
            box_grow(&cb, &c);
        }
        mh_bvh_node *nd = &nv.nodes[t.node];
        for (int k = 0; k < 3; ++k) {
            nd->bmin[k] = round_down(nb.mn[k] - 1e-7);
            nd->bmax[k] = round_up(nb.mx[k] + 1e-7);
        }
        double area = box_area(&nb);
        if (t.node == 0) root_area = area > 0.0 ? area : 1.0;
        if (t.depth > max_depth) max_depth = t.depth;

        /* best binned SAH split over the three axes */
        int best_axis = -1, best_bin = -1;
        double best_cost = DBL_MAX;
        if (count > 1) {
            for (int ax = 0; ax < 3; ++ax) {
                double lo = cb.mn[ax], ext = cb.mx[ax] - cb.mn[ax];
                if (!(ext > 0.0)) continue;
                double scale = (double)nbins / ext;
                aabb bb[MAX_BINS];
                int bc[MAX_BINS];
                for (int b = 0; b < nbins; ++b) { box_empty(&bb[b]); bc[b] = 0; }
                for (int i = t.begin; i < t.end; ++i) {
                    int q = slot[i];
                    int b = (int)((cen[3 * q + ax] - lo) * scale);
                    if (b >= nbins) b = nbins - 1;
                    if (b < 0) b = 0;
                    bc[b]++;
                    box_grow(&bb[b], &pb[q]);
                }
                double ra[MAX_BINS];
                int rc[MAX_BINS];
                aabb acc;
                box_empty(&acc);
                int n = 0;
                for (int b = nbins - 1; b > 0; --b) {
                    if (bc[b]) { box_grow(&acc, &bb[b]); n += bc[b]; }
                    ra[b] = n ? box_area(&acc) : 0.0;
                    rc[b] = n;
                }
                box_empty(&acc);
                n = 0;
                for (int b = 0; b < nbins - 1; ++b) {
                    if (bc[b]) { box_grow(&acc, &bb[b]); n += bc[b]; }
                    if (n == 0 || rc[b + 1] == 0) continue;
                    double c = box_area(&acc) * (double)n + ra[b + 1] * (double)rc[b + 1];
                    if (c < best_cost) {
                        best_cost = c;
                        best_axis = ax;
                        best_bin = b;
                    }
                }
            }
        }
        double split_cost = best_axis >= 0 ? 1.0 + best_cost / (area > 0.0 ? area : 1.0) : DBL_MAX;
        int make_leaf = count <= 1 || (count <= max_leaf && (double)count <= split_cost);
        int mid = -1;
        if (!make_leaf) {
            if (best_axis >= 0) {
                double lo = cb.mn[best_axis], scale = (double)nbins / (cb.mx[best_axis] - cb.mn[best_axis]);
                int i = t.begin, j = t.end - 1;
                while (i <= j) {
                    int q = slot[i];
                    int b = (int)((cen[3 * q + best_axis] - lo) * scale);
                    if (b >= nbins) b = nbins - 1;
                    if (b < 0) b = 0;
                    if (b <= best_bin) {
                        ++i;
                    } else {
                        int tmp = slot[i]; slot[i] = slot[j]; slot[j] = tmp; --j;
                    }
                }
                mid = i;
            }
            if (mid <= t.begin || mid >= t.end) {
                if (count <= 8 && best_axis < 0) make_leaf = 1; /* coincident centroids */
                else mid = t.begin + count / 2;                  /* object median */
            }
        }
        if (make_leaf) {
            nd->left = nd->right = -1;
            nd->first = t.begin;
            nd->count = count;
            ++leaves;
            sah += area / root_area * (double)count;
            continue;
        }
        sah += area / root_area;
        int l = push_node(&nv);
        int r = push_node(&nv);
        if (l < 0 || r < 0) goto oom;
        nd = &nv.nodes[t.node];
        nd->left = l;
        nd->right = r;
        nd->first = 0;
        nd->count = 0;
        if (sp + 2 > scap) {
            scap *= 2;
            task_t *ns = (task_t *)realloc(stack, sizeof(task_t) * (size_t)scap);
            if (!ns) goto oom;
            stack = ns;
        }
        stack[sp++] = (task_t){ r, mid, t.end, t.depth + 1 };
        stack[sp++] = (task_t){ l, t.begin, mid, t.depth + 1 };
    }
    s->leaf_prims = (int *)malloc(sizeof(int) * (size_t)(nt ? nt : 1));
    if (!s->leaf_prims) goto oom;
    for (int i = 0; i < nt; ++i) s->leaf_prims[i] = idx[slot[i]];
    s->n_leaf = nt;
    s->nodes = nv.nodes;
    s->n_nodes = nv.n;
    s->bvh_valid = 1;
    free(idx); free(pb); free(cen); free(slot); free(stack);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (stats) {
        stats->n_nodes = s->n_nodes;
        stats->n_leaves = leaves;
        stats->max_depth = max_depth;
        stats->sah_cost = sah;
        stats->build_seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    }
    if (max_depth > 60) {
        drop_bvh(s);
        return mt_set_err(3, "BVH depth %d exceeds the GPU traversal bound (60)", max_depth);
    }
    return 0;
oom:
    free(nv.nodes); free(idx); free(pb); free(cen); free(slot); free(stack);
    drop_bvh(s);
    return mt_set_err(2, "out of memory");
}

int mt_scene_soa_no_bvh(const mt_scene *s, mh_scene_soa *out)
{
    if (!s || !out) return mt_set_err(1, "NULL argument");
    memset(out, 0, sizeof *out);
    out->abi_version = MH_ABI_VERSION;
    out->n_prims = s->n;
    out->prims = s->prims;
    out->attrs = s->attrs;
    out->n_materials = s->nm;
    out->materials = s->mats;
    out->sky = s->sky;
    out->water_level = s->water_level;
    out->water_material = s->water_material;
    out->fog = s->fog;
    out->ocean = s->ocean;
    return 0;
}

int mt_scene_soa(const mt_scene *s, mh_scene_soa *out)
{
    if (!s || !out) return mt_set_err(1, "NULL argument");
    if (!s->bvh_valid) return mt_set_err(4, "BVH not built (call mt_scene_build_bvh after the last change)");
    memset(out, 0, sizeof *out);
    out->abi_version = MH_ABI_VERSION;
    out->n_prims = s->n;
    out->prims = s->prims;
    out->attrs = s->attrs;
    out->n_materials = s->nm;
    out->materials = s->mats;
    out->sky = s->sky;
    out->water_level = s->water_level;
    out->water_material = s->water_material;
    out->fog = s->fog;
    out->ocean = s->ocean;
    out->n_nodes = s->n_nodes;
    out->nodes = s->nodes;
    out->n_leaf_prims = s->n_leaf;
    out->leaf_prims = s->leaf_prims;
    out->n_planes = s->n_planes;
    out->planes = s->planes;
    return 0;
}

/* ---- image output (reference src/bmp.c:38-230 byte layout) ------------ */

static void put_u32(unsigned char *p, unsigned v)
{
    p[0] = (unsigned char)(v & 0xFFu);
    p[1] = (unsigned char)((v >> 8) & 0xFFu);
    p[2] = (unsigned char)((v >> 16) & 0xFFu);
    p[3] = (unsigned char)((v >> 24) & 0xFFu);
}

long mt_encode_bmp(const unsigned char *rgb, int w, int h, unsigned char *buf, long cap)
{
    if (!rgb || w <= 0 || h <= 0) return -1;
    long row = ((long)w * 3 + 3) / 4 * 4;
    long total = 54 + row * h;
    if (!buf) return total;
    if (cap < total) return -1;
    memset(buf, 0, 54);
    buf[0] = 'B';
    buf[1] = 'M';
    put_u32(buf + 2, (unsigned)total);
    put_u32(buf + 10, 54u);
    put_u32(buf + 14, 40u);
    put_u32(buf + 18, (unsigned)w);
    put_u32(buf + 22, (unsigned)h);
    buf[26] = 1;
    buf[28] = 24;
    put_u32(buf + 34, (unsigned)(row * h));
    put_u32(buf + 38, 2835u);
    put_u32(buf + 42, 2835u);
    for (int y = h - 1, r = 0; y >= 0; --y, ++r) {
        unsigned char *dst = buf + 54 + (long)r * row;
        const unsigned char *src = rgb + (long)y * w * 3;
        for (int x = 0; x < w; ++x) {
            dst[3 * x] = src[3 * x + 2];
            dst[3 * x + 1] = src[3 * x + 1];
            dst[3 * x + 2] = src[3 * x];
        }
        for (long k = (long)w * 3; k < row; ++k) dst[k] = 0;
    }
    return total;
}

int mt_write_bmp(const char *path, const unsigned char *rgb, int w, int h)
{
    long n = mt_encode_bmp(rgb, w, h, NULL, 0);
    if (n < 0 || !path) return mt_set_err(-2, "bmp: invalid arguments");
    unsigned char *buf = (unsigned char *)malloc((size_t)n);
    if (!buf) return mt_set_err(-1, "bmp: out of memory");
    mt_encode_bmp(rgb, w, h, buf, n);
    FILE *f = fopen(path, "wb");
    int rc = 0;
    if (!f || fwrite(buf, 1, (size_t)n, f) != (size_t)n) rc = mt_set_err(-1, "bmp: cannot write '%s'", path);
    if (f && fclose(f) != 0) rc = -1;
    free(buf);
    return rc;
}

int mt_write_ppm(const char *path, const unsigned char *rgb, int w, int h)
{
    if (!path || !rgb || w <= 0 || h <= 0) return mt_set_err(-2, "ppm: invalid arguments");
    FILE *f = fopen(path, "wb");
    if (!f) return mt_set_err(-1, "ppm: cannot open '%s'", path);
    int rc = 0;
    if (fprintf(f, "P6\n%d %d\n255\n", w, h) < 0) rc = -1;
    size_t n = (size_t)w * (size_t)h * 3;
    if (rc == 0 && fwrite(rgb, 1, n, f) != n) rc = -1;
    if (fclose(f) != 0) rc = -1;
    return rc ? mt_set_err(-1, "ppm: write failed") : 0;
}
