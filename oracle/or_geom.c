/*
 * or_geom.c - oracle: ray/primitive intersection, bounds, the binned-SAH
 * BVH build and its nearest-hit / any-hit traversal, restated in FP64.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Follows:
 *   hit storage        reference src/geometry.c:78-91
 *   sphere/plane/box   reference src/geometry.c:338-479
 *   triangle           reference src/geometry.c:482-538 (Moller-Trumbore)
 *   cylinder           reference src/geometry.c:549-699
 *   occlusion          reference src/geometry.c:1373-1461
 *   bounds             reference src/geometry.c:1517-1584
 *   BVH build          reference src/bvh.c:140-422 (12 bins, leaf <= 4)
 *   BVH traversal      reference src/bvh.c:441-813
 */
#include "or_internal.h"

#include <stdlib.h>
#include <string.h>

#define TRI_EPS 1e-9
#define BARY_EPS 1e-9

static void store_hit(ohit *h, double t, dv3 point, dv3 n, oray r, int mat, double u, double v)
{
    int front = ddot(r.d, n) < 0.0;
    h->t = t;
    h->point = point;
    h->normal = front ? n : dneg(n);
    h->material = mat;
    h->front = front ? 1 : 0;
    h->u = u;
    h->v = v;
}

static int hit_sphere(const oprim *p, oray r, double tmin, double tmax, ohit *h)
{
    if (p->r0 <= 0.0) return 0;
    dv3 l = dsub(r.o, p->p0);
    double hb = ddot(l, r.d);
    double c = dlen2(l) - p->r0 * p->r0;
    double disc = hb * hb - c;
    if (disc < 0.0) return 0;
    double sq = sqrt(disc);
    double t = -hb - sq;
    if (t < tmin) {
        t = -hb + sq;
        if (t < tmin || t > tmax) return 0;
    } else if (t > tmax) {
        return 0;
    }
    dv3 pt = dat(r.o, r.d, t);
    dv3 n = dscale(dsub(pt, p->p0), 1.0 / p->r0);
    double phi = atan2(n.z, n.x);
    double th = asin(fmax(-1.0, fmin(1.0, n.y)));
    double u = 0.5 + phi / (2.0 * OR_PI);
    double v = 0.5 + th / OR_PI;
    u = u - floor(u);
    v = fmax(0.0, fmin(1.0, v));
    store_hit(h, t, pt, n, r, p->material, u, v);
    return 1;
}

static int hit_plane(const oprim *p, oray r, double tmin, double tmax, ohit *h)
{
    dv3 n = dnorm(p->p1);
    double den = ddot(r.d, n);
    if (fabs(den) < 1e-12) return 0;
    double t = ddot(dsub(p->p0, r.o), n) / den;
    if (t < tmin || t > tmax) return 0;
    dv3 pt = dat(r.o, r.d, t);
    dv3 hint = fabs(n.y) > 0.9 ? d3(0.0, 0.0, 1.0) : d3(0.0, 1.0, 0.0);
    dv3 T = dnorm(dcross(hint, n));
    dv3 B = dcross(n, T);
    dv3 df = dsub(pt, p->p0);
    double u = ddot(df, T), v = ddot(df, B);
    u = u - floor(u);
    v = v - floor(v);
    store_hit(h, t, pt, n, r, p->material, u, v);
    return 1;
}

static int hit_box(const oprim *p, oray r, double tmin, double tmax, ohit *h)
{
    dv3 lo3 = dsub(p->p0, p->p1), hi3 = dadd(p->p0, p->p1);
    double o[3] = { r.o.x, r.o.y, r.o.z }, d[3] = { r.d.x, r.d.y, r.d.z };
    double lo[3] = { lo3.x, lo3.y, lo3.z }, hi[3] = { hi3.x, hi3.y, hi3.z };
    double tn = -INFINITY, tf = INFINITY;
    int ea = -1, es = 0, xa = -1, xs = 0;
    for (int i = 0; i < 3; ++i) {
        double inv = 1.0 / d[i];
        double t1 = (lo[i] - o[i]) * inv, t2 = (hi[i] - o[i]) * inv;
        double tnear = fmin(t1, t2), tfar = fmax(t1, t2);
        if (tnear > tn) { tn = tnear; ea = i; es = inv > 0.0 ? -1 : 1; }
        if (tfar < tf) { tf = tfar; xa = i; xs = inv > 0.0 ? 1 : -1; }
        if (tf < tn) return 0;
    }
    double th;
    int axis, sgn;
    if (tn >= tmin && tn <= tmax) { th = tn; axis = ea; sgn = es; }
    else if (tf >= tmin && tf <= tmax) { th = tf; axis = xa; sgn = xs; }
    else return 0;
    dv3 pt = dat(r.o, r.d, th);
    dv3 n;
    if (axis < 0) {
        n = dneg(r.d);
    } else {
        n = d3(0.0, 0.0, 0.0);
        if (axis == 0) n.x = (double)sgn;
        else if (axis == 1) n.y = (double)sgn;
        else n.z = (double)sgn;
    }
    double u = 0.0, v = 0.0;
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    if (axis == 0) {
        u = dz > 1e-12 ? (pt.z - lo[2]) / dz : 0.0;
        v = dy > 1e-12 ? (pt.y - lo[1]) / dy : 0.0;
    } else if (axis == 1) {
        u = dx > 1e-12 ? (pt.x - lo[0]) / dx : 0.0;
        v = dz > 1e-12 ? (pt.z - lo[2]) / dz : 0.0;
    } else {
        u = dx > 1e-12 ? (pt.x - lo[0]) / dx : 0.0;
        v = dy > 1e-12 ? (pt.y - lo[1]) / dy : 0.0;
    }
    u = fmax(0.0, fmin(1.0, u));
    v = fmax(0.0, fmin(1.0, v));
    store_hit(h, th, pt, n, r, p->material, u, v);
    return 1;
}

static int hit_triangle(const oprim *p, oray r, double tmin, double tmax, ohit *h)
{
    dv3 e1 = dsub(p->p1, p->p0), e2 = dsub(p->p2, p->p0);
    dv3 pv = dcross(r.d, e2);
    double det = ddot(e1, pv);
    if (fabs(det) < TRI_EPS) return 0;
    double idet = 1.0 / det;
    dv3 tv = dsub(r.o, p->p0);
    double u = ddot(tv, pv) * idet;
    if (u < -BARY_EPS || u > 1.0 + BARY_EPS) return 0;
    dv3 qv = dcross(tv, e1);
    double v = ddot(r.d, qv) * idet;
    if (v < -BARY_EPS || u + v > 1.0 + BARY_EPS) return 0;
    double t = ddot(e2, qv) * idet;
    if (t < tmin || t > tmax) return 0;
    dv3 pt = dat(r.o, r.d, t);
    dv3 n;
    if (p->flags & MH_PRIM_SMOOTH) {
        double w = 1.0 - u - v;
        n = dadd(dscale(p->n0, w), dadd(dscale(p->n1, u), dscale(p->n2, v)));
        double nl = dlen(n);
        if (nl > 1e-12) {
            n = dscale(n, 1.0 / nl);
        } else {
            n = dcross(e1, e2);
            double ng = dlen(n);
            if (ng < 1e-15) return 0;
            n = dscale(n, 1.0 / ng);
        }
    } else {
        n = dcross(e1, e2);
        double nl = dlen(n);
        if (nl < 1e-15) return 0;
        n = dscale(n, 1.0 / nl);
    }
    double hu = u, hv = v;
    if (p->flags & MH_PRIM_UV) {
        double w = 1.0 - u - v;
        hu = w * p->uv0.x + u * p->uv1.x + v * p->uv2.x;
        hv = w * p->uv0.y + u * p->uv1.y + v * p->uv2.y;
    }
    store_hit(h, t, pt, n, r, p->material, hu, hv);
    return 1;
}

/* Tapered capped cylinder: lateral quadratic in the base frame plus two
 * axis-perpendicular cap discs; nearest accepted root wins. */
static int hit_cylinder(const oprim *p, oray r, double tmin, double tmax, ohit *h)
{
    dv3 A = dsub(p->p1, p->p0);
    double hgt = dlen(A);
    if (hgt < 1e-12) return 0;
    dv3 ax = dscale(A, 1.0 / hgt);
    double r0 = p->r0, r1 = p->r1;
    double k = (r1 - r0) / hgt;
    dv3 d0 = dsub(r.o, p->p0);
    double od = ddot(r.d, ax), d0d = ddot(d0, ax);
    dv3 dp = dsub(r.d, dscale(ax, od));
    dv3 d0p = dsub(d0, dscale(ax, d0d));
    double ra = k * od, rb = r0 + k * d0d;
    double qa = dlen2(dp) - ra * ra;
    double qb = ddot(dp, d0p) - rb * ra;
    double qc = dlen2(d0p) - rb * rb;

    int found = 0, is_cap = 0, cap_id = 0;
    double bt = 0.0, bs = 0.0;
    dv3 bn = d3(0.0, 0.0, 0.0);

    if (fabs(qa) > 1e-15) {
        double disc = qb * qb - qa * qc;
        if (disc >= 0.0) {
            double sq = sqrt(disc);
            double ro[2] = { (-qb - sq) / qa, (-qb + sq) / qa };
            if (ro[0] > ro[1]) { double tt = ro[0]; ro[0] = ro[1]; ro[1] = tt; }
            for (int i = 0; i < 2; ++i) {
                double t = ro[i];
                if (t < tmin || t > tmax) continue;
                double s = (d0d + t * od) / hgt;
                if (s < 0.0 || s > 1.0) continue;
                dv3 pt = dat(r.o, r.d, t);
                dv3 ap = dadd(p->p0, dscale(ax, s * hgt));
                dv3 rad = dsub(pt, ap);
                double rl = dlen(rad);
                dv3 rd = rl > 1e-15 ? dscale(rad, 1.0 / rl) : d3(0.0, 0.0, 0.0);
                dv3 n = dnorm(dsub(dscale(rd, hgt), dscale(ax, r1 - r0)));
                if (dlen2(n) < 0.5) n = rd;
                if (!found || t < bt) { found = 1; bt = t; bn = n; is_cap = 0; bs = s; }
            }
        }
    } else if (fabs(qb) > 1e-15) {
        double t = -qc / (2.0 * qb);
        if (t >= tmin && t <= tmax) {
            double s = (d0d + t * od) / hgt;
            if (s >= 0.0 && s <= 1.0) {
                dv3 pt = dat(r.o, r.d, t);
                dv3 ap = dadd(p->p0, dscale(ax, s * hgt));
                dv3 rd = dnorm(dsub(pt, ap));
                dv3 n = dnorm(dsub(dscale(rd, hgt), dscale(ax, r1 - r0)));
                if (dlen2(n) < 0.5) n = rd;
                found = 1; bt = t; bn = n; is_cap = 0; bs = s;
            }
        }
    }
    if (fabs(od) > 1e-15) {
        for (int cap = 0; cap < 2; ++cap) {
            double sc = cap == 0 ? 0.0 : 1.0;
            double rc = cap == 0 ? r0 : r1;
            if (rc <= 0.0) continue;
            dv3 cc = cap == 0 ? p->p0 : p->p1;
            double t = ((sc * hgt) - d0d) / od;
            if (t < tmin || t > tmax) continue;
            dv3 pt = dat(r.o, r.d, t);
            dv3 rad = dsub(pt, cc);
            if (dlen2(rad) > rc * rc) continue;
            dv3 n = cap == 0 ? dneg(ax) : ax;
            if (!found || t < bt) { found = 1; bt = t; bn = n; is_cap = 1; cap_id = cap; }
        }
    }
    if (!found) return 0;

    dv3 pt = dat(r.o, r.d, bt);
    double u, v;
    dv3 hint = fabs(ax.y) > 0.9 ? d3(0.0, 0.0, 1.0) : d3(0.0, 1.0, 0.0);
    dv3 T = dnorm(dcross(hint, ax));
    dv3 B = dcross(ax, T);
    if (is_cap) {
        dv3 cc = cap_id == 0 ? p->p0 : p->p1;
        double rc = cap_id == 0 ? r0 : r1;
        dv3 rad = dsub(pt, cc);
        double rx = rc > 1e-12 ? ddot(rad, T) / (2.0 * rc) : 0.0;
        double ry = rc > 1e-12 ? ddot(rad, B) / (2.0 * rc) : 0.0;
        u = fmax(0.0, fmin(1.0, 0.5 + rx));
        v = fmax(0.0, fmin(1.0, 0.5 + ry));
    } else {
        dv3 ap = dadd(p->p0, dscale(ax, bs * hgt));
        dv3 rad = dsub(pt, ap);
        double phi = atan2(ddot(rad, B), ddot(rad, T));
        u = 0.5 + phi / (2.0 * OR_PI);
        u = u - floor(u);
        v = fmax(0.0, fmin(1.0, bs));
    }
    store_hit(h, bt, pt, bn, r, p->material, u, v);
    return 1;
}

int or_prim_intersect(const oprim *p, oray r, double tmin, double tmax, ohit *h)
{
    switch (p->kind) {
    case MH_PRIM_SPHERE: return hit_sphere(p, r, tmin, tmax, h);
    case MH_PRIM_PLANE: return hit_plane(p, r, tmin, tmax, h);
    case MH_PRIM_BOX: return hit_box(p, r, tmin, tmax, h);
    case MH_PRIM_TRIANGLE: return hit_triangle(p, r, tmin, tmax, h);
    case MH_PRIM_CYLINDER: return hit_cylinder(p, r, tmin, tmax, h);
    default: return 0;
    }
}

/* ---- any-hit ------------------------------------------------------- */

static int occ_sphere(const oprim *p, oray r, double tmin, double tmax)
{
    if (p->r0 <= 0.0) return 0;
    dv3 l = dsub(r.o, p->p0);
    double hb = ddot(l, r.d);
    double c = dlen2(l) - p->r0 * p->r0;
    /* nearest root of t^2 + 2 hb t + c in [tmin, tmax] (a == 1) */
    double disc = hb * hb - 1.0 * c;
    if (disc < 0.0) return 0;
    double sq = sqrt(disc);
    double t0 = (-hb - sq) / 1.0, t1 = (-hb + sq) / 1.0;
    if (t0 > t1) { double tt = t0; t0 = t1; t1 = tt; }
    if (t0 >= tmin && t0 <= tmax) return 1;
    if (t1 >= tmin && t1 <= tmax) return 1;
    return 0;
}

static int occ_box(const oprim *p, oray r, double tmin, double tmax)
{
    dv3 lo3 = dsub(p->p0, p->p1), hi3 = dadd(p->p0, p->p1);
    double o[3] = { r.o.x, r.o.y, r.o.z }, d[3] = { r.d.x, r.d.y, r.d.z };
    double lo[3] = { lo3.x, lo3.y, lo3.z }, hi[3] = { hi3.x, hi3.y, hi3.z };
    double tn = -INFINITY, tf = INFINITY;
    for (int i = 0; i < 3; ++i) {
        double inv = 1.0 / d[i];
        double t1 = (lo[i] - o[i]) * inv, t2 = (hi[i] - o[i]) * inv;
        if (t1 > t2) { double tt = t1; t1 = t2; t2 = tt; }
        if (t1 > tn) tn = t1;
        if (t2 < tf) tf = t2;
        if (tn > tf) return 0;
    }
    return (tn >= tmin && tn <= tmax) || (tf >= tmin && tf <= tmax);
}

static int occ_triangle(const oprim *p, oray r, double tmin, double tmax)
{
    dv3 pv = dcross(r.d, p->e2);
    double det = ddot(p->e1, pv);
    if (fabs(det) < TRI_EPS) return 0;
    double idet = 1.0 / det;
    dv3 tv = dsub(r.o, p->p0);
    double u = ddot(tv, pv) * idet;
    if (u < -BARY_EPS || u > 1.0 + BARY_EPS) return 0;
    dv3 qv = dcross(tv, p->e1);
    double v = ddot(r.d, qv) * idet;
    if (v < -BARY_EPS || u + v > 1.0 + BARY_EPS) return 0;
    double t = ddot(p->e2, qv) * idet;
    return t >= tmin && t <= tmax;
}

int or_prim_occluded(const oprim *p, oray r, double tmin, double tmax)
{
    ohit h;
    switch (p->kind) {
    case MH_PRIM_SPHERE: return occ_sphere(p, r, tmin, tmax);
    case MH_PRIM_PLANE: return hit_plane(p, r, tmin, tmax, &h);
    case MH_PRIM_BOX: return occ_box(p, r, tmin, tmax);
    case MH_PRIM_TRIANGLE: return occ_triangle(p, r, tmin, tmax);
    case MH_PRIM_CYLINDER: return hit_cylinder(p, r, tmin, tmax, &h);
    default: return 0;
    }
}

void or_prim_bounds(const oprim *p, dv3 *mn, dv3 *mx)
{
    switch (p->kind) {
    case MH_PRIM_SPHERE: {
        dv3 e = d3(p->r0, p->r0, p->r0);
        *mn = dsub(p->p0, e);
        *mx = dadd(p->p0, e);
        break;
    }
    case MH_PRIM_PLANE:
        *mn = d3(-1e4, -1e4, -1e4);
        *mx = d3(1e4, 1e4, 1e4);
        break;
    case MH_PRIM_BOX: {
        dv3 e = d3(fabs(p->p1.x), fabs(p->p1.y), fabs(p->p1.z));
        *mn = dsub(p->p0, e);
        *mx = dadd(p->p0, e);
        break;
    }
    case MH_PRIM_TRIANGLE:
        *mn = dmin3(p->p0, dmin3(p->p1, p->p2));
        *mx = dmax3(p->p0, dmax3(p->p1, p->p2));
        break;
    case MH_PRIM_CYLINDER: {
        double rm = p->r0 > p->r1 ? p->r0 : p->r1;
        if (rm < 0.0) rm = 0.0;
        dv3 e = d3(rm, rm, rm);
        *mn = dsub(dmin3(p->p0, p->p1), e);
        *mx = dadd(dmax3(p->p0, p->p1), e);
        break;
    }
    default:
        *mn = *mx = d3(0.0, 0.0, 0.0);
    }
}

/* ---- BVH build ----------------------------------------------------- */

#define NBINS 12
#define LEAF_MAX 4
#define PAD 1e-7

static double axis_of(dv3 v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

static double area(dv3 mn, dv3 mx)
{
    double dx = mx.x - mn.x, dy = mx.y - mn.y, dz = mx.z - mn.z;
    if (dx < 0.0) dx = 0.0;
    if (dy < 0.0) dy = 0.0;
    if (dz < 0.0) dz = 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

static int push_node(or_scene *s, onode n)
{
    if (s->n_nodes >= s->cap_nodes) {
        int cap = s->cap_nodes ? s->cap_nodes * 2 : 64;
        onode *nn = (onode *)realloc(s->nodes, (size_t)cap * sizeof(onode));
        if (!nn) return -1;
        s->nodes = nn;
        s->cap_nodes = cap;
    }
    s->nodes[s->n_nodes] = n;
    return s->n_nodes++;
}

static dv3 centroid(const oprim *p)
{
    dv3 mn, mx;
    or_prim_bounds(p, &mn, &mx);
    return dscale(dadd(mn, mx), 0.5);
}

static int bin_of(double cv, double cmin, double scale)
{
    int b = (int)((cv - cmin) * scale);
    if (b < 0) b = 0;
    if (b >= NBINS) b = NBINS - 1;
    return b;
}

static int build_rec(or_scene *s, int first, int count)
{
    const dv3 big = d3(1e300, 1e300, 1e300), nbig = d3(-1e300, -1e300, -1e300);
    dv3 mn = big, mx = nbig, cmn = big, cmx = nbig;
    for (int i = 0; i < count; ++i) {
        dv3 a, b;
        or_prim_bounds(&s->prims[s->order[first + i]], &a, &b);
        mn = dmin3(mn, a);
        mx = dmax3(mx, b);
        dv3 c = dscale(dadd(a, b), 0.5);
        cmn = dmin3(cmn, c);
        cmx = dmax3(cmx, c);
    }
    mn = d3(mn.x - PAD, mn.y - PAD, mn.z - PAD);
    mx = d3(mx.x + PAD, mx.y + PAD, mx.z + PAD);
    onode leaf = { mn, mx, -1, -1, first, count };
    if (count <= LEAF_MAX) return push_node(s, leaf);

    dv3 ext = dsub(cmx, cmn);
    int axis = 0;
    double me = ext.x;
    if (ext.y > me) { me = ext.y; axis = 1; }
    if (ext.z > me) { me = ext.z; axis = 2; }
    if (!(me > 0.0)) return push_node(s, leaf);

    double cmin = axis_of(cmn, axis), cmax = axis_of(cmx, axis);
    double scale = NBINS / me;
    dv3 bmn[NBINS], bmx[NBINS];
    int bc[NBINS];
    for (int i = 0; i < NBINS; ++i) { bmn[i] = big; bmx[i] = nbig; bc[i] = 0; }
    for (int i = 0; i < count; ++i) {
        dv3 a, b;
        or_prim_bounds(&s->prims[s->order[first + i]], &a, &b);
        dv3 c = dscale(dadd(a, b), 0.5);
        int k = bin_of(axis_of(c, axis), cmin, scale);
        bc[k]++;
        bmn[k] = dmin3(bmn[k], a);
        bmx[k] = dmax3(bmx[k], b);
    }
    dv3 rmn[NBINS], rmx[NBINS];
    int rc[NBINS];
    dv3 am = big, aM = nbig;
    int an = 0;
    for (int i = NBINS - 1; i >= 0; --i) {
        if (bc[i] > 0) { am = dmin3(am, bmn[i]); aM = dmax3(aM, bmx[i]); an += bc[i]; }
        rmn[i] = am; rmx[i] = aM; rc[i] = an;
    }
    dv3 lm = big, lM = nbig;
    int ln = 0, best = -1;
    double best_cost = 1e300;
    for (int i = 0; i < NBINS - 1; ++i) {
        if (bc[i] > 0) { lm = dmin3(lm, bmn[i]); lM = dmax3(lM, bmx[i]); ln += bc[i]; }
        int rn = rc[i + 1];
        if (ln == 0 || rn == 0) continue;
        double cost = area(lm, lM) * (double)ln + area(rmn[i + 1], rmx[i + 1]) * (double)rn;
        if (cost < best_cost) { best_cost = cost; best = i; }
    }
    double leaf_cost = area(mn, mx) * (double)count;
    int use = (best >= 0 && best_cost < leaf_cost) ? best : -1;

    int mid;
    int i = first, j = first + count - 1;
    if (use < 0) {
        /* centroid-midpoint partition when SAH does not beat a leaf */
        double half = 0.5 * (cmin + cmax);
        while (i <= j) {
            double cv = axis_of(centroid(&s->prims[s->order[i]]), axis);
            if (cv < half) {
                ++i;
            } else {
                int t = s->order[i]; s->order[i] = s->order[j]; s->order[j] = t; --j;
            }
        }
        mid = i;
        if (mid <= first || mid >= first + count) mid = first + count / 2;
    } else {
        while (i <= j) {
            double cv = axis_of(centroid(&s->prims[s->order[i]]), axis);
            if (bin_of(cv, cmin, scale) <= use) {
                ++i;
            } else {
                int t = s->order[i]; s->order[i] = s->order[j]; s->order[j] = t; --j;
            }
        }
        mid = i;
    }
    if (mid == first || mid == first + count) mid = first + count / 2;
    int nl = mid - first, nr = count - nl;
    if (nl <= 0 || nr <= 0) return push_node(s, leaf);

    onode inner = { mn, mx, -1, -1, 0, 0 };
    int me_idx = push_node(s, inner);
    if (me_idx < 0) return -1;
    int l = build_rec(s, first, nl);
    if (l < 0) return -1;
    int r = build_rec(s, mid, nr);
    if (r < 0) return -1;
    s->nodes[me_idx].left = l;
    s->nodes[me_idx].right = r;
    return me_idx;
}

int or_bvh_build(or_scene *s)
{
    int nt = 0, np = 0;
    for (int i = 0; i < s->n_prims; ++i) {
        if (s->prims[i].kind == MH_PRIM_PLANE) ++np;
        else ++nt;
    }
    s->order = (int *)malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1));
    s->planes = (int *)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    if (!s->order || !s->planes) return -1;
    nt = np = 0;
    for (int i = 0; i < s->n_prims; ++i) {
        if (s->prims[i].kind == MH_PRIM_PLANE) s->planes[np++] = i;
        else s->order[nt++] = i;
    }
    s->n_order = nt;
    s->n_planes = np;
    if (nt > 0) {
        if (build_rec(s, 0, nt) < 0) return -1;
        s->leaf_prims = (oprim *)malloc(sizeof(oprim) * (size_t)nt);
        if (!s->leaf_prims) return -1;
        for (int i = 0; i < nt; ++i) s->leaf_prims[i] = s->prims[s->order[i]];
    }
    return 0;
}

/* ---- traversal ----------------------------------------------------- */

/* Slab test of one box; NaN-safe through fmin/fmax. */
static int slab(dv3 mn, dv3 mx, dv3 o, dv3 inv, double tmin, double tmax, double *te)
{
    double ax = (mn.x - o.x) * inv.x, bx = (mx.x - o.x) * inv.x;
    double ay = (mn.y - o.y) * inv.y, by = (mx.y - o.y) * inv.y;
    double az = (mn.z - o.z) * inv.z, bz = (mx.z - o.z) * inv.z;
    double en = fmax(fmax(fmin(ax, bx), fmin(ay, by)), fmin(az, bz));
    double ex = fmin(fmin(fmax(ax, bx), fmax(ay, by)), fmax(az, bz));
    if (ex < en || ex < tmin || en > tmax) return 0;
    *te = en < tmin ? tmin : en;
    return 1;
}

/* Child box test with the acceptance of the reference's paired slab test. */
static int slab_child(const onode *n, dv3 o, dv3 inv, double tmin, double tmax, double *te)
{
    double ax = (n->bmin.x - o.x) * inv.x, bx = (n->bmax.x - o.x) * inv.x;
    double ay = (n->bmin.y - o.y) * inv.y, by = (n->bmax.y - o.y) * inv.y;
    double az = (n->bmin.z - o.z) * inv.z, bz = (n->bmax.z - o.z) * inv.z;
    double en = fmax(fmax(fmin(ax, bx), fmin(ay, by)), fmin(az, bz));
    double ex = fmin(fmin(fmax(ax, bx), fmax(ay, by)), fmax(az, bz));
    *te = en < tmin ? tmin : en;
    return ex >= en && ex >= tmin && en <= tmax;
}

static oray unit_ray(oray r)
{
    double l2 = dlen2(r.d);
    if (l2 > 0.0 && fabs(l2 - 1.0) > 1e-12) r.d = dnorm(r.d);
    return r;
}

typedef struct { int idx; double t; } sentry;

int or_intersect1(const or_scene *s, oray r, double tmin, double tmax, ohit *out)
{
    r = unit_ray(r);
    dv3 inv = d3(1.0 / r.d.x, 1.0 / r.d.y, 1.0 / r.d.z);
    int found = 0;
    double closest = tmax;
    ohit best;
    or_tls_stats.closest_queries++;

    for (int i = 0; i < s->n_planes; ++i) {
        int pi = s->planes[i];
        ohit h;
        or_tls_stats.plane_tests++;
        if (or_prim_intersect(&s->prims[pi], r, tmin, closest, &h)) {
            h.prim = pi;
            if (!found || h.t < best.t || (h.t == best.t && pi < best.prim)) {
                best = h;
                closest = h.t;
                found = 1;
            }
        }
    }
    if (s->n_nodes > 0) {
        sentry fixed[64];
        sentry *st = fixed;
        int cap = 64, sp = 0;
        double t0;
        if (slab(s->nodes[0].bmin, s->nodes[0].bmax, r.o, inv, tmin, closest, &t0)) {
            st[sp].idx = 0; st[sp].t = t0; ++sp;
        }
        while (sp > 0) {
            sentry e = st[--sp];
            if (found && e.t > closest) continue;
            const onode *n = &s->nodes[e.idx];
            if (n->count > 0) {
                for (int i = 0; i < n->count; ++i) {
                    int pi = s->order[n->first + i];
                    ohit h;
                    or_tls_stats.c_prims++;
                    if (s->leaf_prims[n->first + i].kind == MH_PRIM_TRIANGLE) or_tls_stats.c_tris++;
                    if (or_prim_intersect(&s->leaf_prims[n->first + i], r, tmin, closest, &h)) {
                        h.prim = pi;
                        if (!found || h.t < best.t || (h.t == best.t && pi < best.prim)) {
                            best = h;
                            closest = h.t;
                            found = 1;
                        }
                    }
                }
            } else {
                double tl, tr;
                or_tls_stats.c_inner++;
                int hl = slab_child(&s->nodes[n->left], r.o, inv, tmin, closest, &tl);
                int hr = slab_child(&s->nodes[n->right], r.o, inv, tmin, closest, &tr);
                if (hl || hr) {
                    if (sp + 2 > cap) {
                        int nc = cap * 2;
                        sentry *ns = (sentry *)malloc(sizeof(sentry) * (size_t)nc);
                        if (!ns) break;
                        memcpy(ns, st, sizeof(sentry) * (size_t)sp);
                        if (st != fixed) free(st);
                        st = ns;
                        cap = nc;
                    }
                    if (hl && hr) {
                        if (tl <= tr) {
                            st[sp].idx = n->right; st[sp].t = tr; ++sp;
                            st[sp].idx = n->left; st[sp].t = tl; ++sp;
                        } else {
                            st[sp].idx = n->left; st[sp].t = tl; ++sp;
                            st[sp].idx = n->right; st[sp].t = tr; ++sp;
                        }
                    } else if (hl) {
                        st[sp].idx = n->left; st[sp].t = tl; ++sp;
                    } else {
                        st[sp].idx = n->right; st[sp].t = tr; ++sp;
                    }
                }
            }
        }
        if (st != fixed) free(st);
    }
    if (!found) return 0;
    *out = best;
    return 1;
}

int or_occluded1(const or_scene *s, oray r, double tmin, double tmax)
{
    r = unit_ray(r);
    or_tls_stats.any_queries++;
    for (int i = 0; i < s->n_planes; ++i) {
        or_tls_stats.plane_tests++;
        if (or_prim_occluded(&s->prims[s->planes[i]], r, tmin, tmax)) return 1;
    }
    if (s->n_nodes == 0) return 0;
    dv3 inv = d3(1.0 / r.d.x, 1.0 / r.d.y, 1.0 / r.d.z);
    int fixed[64];
    int *st = fixed;
    int cap = 64, sp = 0, hit = 0;
    double t0;
    if (slab(s->nodes[0].bmin, s->nodes[0].bmax, r.o, inv, tmin, tmax, &t0)) st[sp++] = 0;
    while (sp > 0 && !hit) {
        const onode *n = &s->nodes[st[--sp]];
        if (n->count > 0) {
            for (int i = 0; i < n->count; ++i) {
                or_tls_stats.a_prims++;
                if (s->leaf_prims[n->first + i].kind == MH_PRIM_TRIANGLE) or_tls_stats.a_tris++;
                if (or_prim_occluded(&s->leaf_prims[n->first + i], r, tmin, tmax)) { hit = 1; break; }
            }
        } else {
            double tl, tr;
            or_tls_stats.a_inner++;
            int hl = slab_child(&s->nodes[n->left], r.o, inv, tmin, tmax, &tl);
            int hr = slab_child(&s->nodes[n->right], r.o, inv, tmin, tmax, &tr);
            if (hl || hr) {
                if (sp + 2 > cap) {
                    int nc = cap * 2;
                    int *ns = (int *)malloc(sizeof(int) * (size_t)nc);
                    if (!ns) break;
                    memcpy(ns, st, sizeof(int) * (size_t)sp);
                    if (st != fixed) free(st);
                    st = ns;
                    cap = nc;
                }
                if (hl && hr) {
                    if (tl <= tr) { st[sp++] = n->right; st[sp++] = n->left; }
                    else { st[sp++] = n->left; st[sp++] = n->right; }
                } else if (hl) {
                    st[sp++] = n->left;
                } else {
                    st[sp++] = n->right;
                }
            }
        }
    }
    if (st != fixed) free(st);
    return hit;
}
