/*
 * mt_ingest.c - geometry ingest at scale: Wavefront OBJ text and procedural
 * Gerstner oceans, added to an mt_scene on several threads.
 *
 * The triangles are the ones the reference makes of the same input, bit for
 * bit:
 *   OBJ    reference src/obj.c:39-85 (Euler rotation, normal and position
 *          transforms), 87-121 (index resolution, face-vertex syntax),
 *          123-300 (two passes: v / vn / vt with bounds, then faces, fan
 *          triangulation, smooth / uv selection)
 *   ocean  reference src/scene.c:1155-1331 (Gerstner grid, tangents and
 *          normals, surface triangles, skirts, floor)
 * The reference reads each v / vn / vt line with sscanf("%lf %lf %lf") and
 * each face from a 1023-byte copy of its line split at " \t\r\n" starting two
 * bytes after the line start; both behaviours are kept (a vertex line with
 * fewer than three numbers reads on into the next line, as sscanf does).
 * Decimal numbers take Clinger's exact fast path (at most 19 significant
 * digits, value below 2^53, power of ten within 1e22: one correctly rounded
 * operation, equal to strtod) and strtod otherwise.
 */
#define _POSIX_C_SOURCE 200809L

#include "marmel_host.h"
#include "mt_internal.h"

#include <ctype.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define MT_PI 3.14159265358979323846

typedef struct { double x, y, z; } v3;

static v3 mk3(double x, double y, double z)
{
    v3 r = { x, y, z };
    return r;
}

static int n_threads(int want, size_t work)
{
    long n = want > 0 ? want : sysconf(_SC_NPROCESSORS_ONLN);
    if (n < 1) n = 1;
    if (n > 64) n = 64;
    if ((size_t)n > work) n = (long)(work ? work : 1);
    return (int)n;
}

/* Run fn(arg_i) for i < n on n threads (the caller's thread runs the last). */
static int run_threads(int n, void *(*fn)(void *), void *args, size_t arg_size)
{
    pthread_t th[64];
    int started = 0;
    for (int i = 0; i < n - 1; ++i) {
        if (pthread_create(&th[i], NULL, fn, (char *)args + (size_t)i * arg_size) != 0) break;
        ++started;
    }
    for (int i = started; i < n; ++i) fn((char *)args + (size_t)i * arg_size);
    for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
    return 0;
}

/* ---- numbers ------------------------------------------------------------ */

static const double P10[23] = { 1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22 };

/* The byte at q, or 0 at / past the end (the reference's buffers end in NUL). */
static int at(const char *q, const char *end) { return q < end ? (unsigned char)*q : 0; }

/* One floating-point number as scanf's %lf reads it: leading white space,
 * then the longest prefix strtod accepts. Returns 1 and advances *pq, or 0. */
static int scan_double(const char **pq, const char *end, double *out)
{
    const char *q = *pq;
    while (q < end && isspace((unsigned char)*q)) ++q;
    const char *start = q;
    int neg = 0;
    if (at(q, end) == '+' || at(q, end) == '-') {
        neg = at(q, end) == '-';
        ++q;
    }
    uint64_t w = 0;
    int sig = 0, digits = 0, exp10 = 0, simple = 1;
    /* hexadecimal, infinity and nan forms: strtod */
    int c0 = at(q, end), c1 = at(q + 1, end);
    if (!(c0 == '0' && (c1 == 'x' || c1 == 'X')) && c0 != 'i' && c0 != 'I' && c0 != 'n' && c0 != 'N') {
        while (isdigit(at(q, end))) {
            int d = at(q, end) - '0';
            if (sig < 19) {
                if (w || d) {
                    w = w * 10 + (uint64_t)d;
                    ++sig;
                }
            } else {
                simple = 0;
            }
            ++digits;
            ++q;
        }
        if (at(q, end) == '.') {
            ++q;
            while (isdigit(at(q, end))) {
                int d = at(q, end) - '0';
                if (sig < 19) {
                    if (w || d) {
                        w = w * 10 + (uint64_t)d;
                        ++sig;
                    }
                    --exp10;
                } else {
                    simple = 0;
                }
                ++digits;
                ++q;
            }
        }
        if (digits > 0) {
            int e = at(q, end);
            if (e == 'e' || e == 'E') {
                const char *r = q + 1;
                int eneg = 0;
                if (at(r, end) == '+' || at(r, end) == '-') {
                    eneg = at(r, end) == '-';
                    ++r;
                }
                if (isdigit(at(r, end))) {
                    long ev = 0;
                    while (isdigit(at(r, end))) {
                        if (ev < 100000) ev = ev * 10 + (at(r, end) - '0');
                        ++r;
                    }
                    exp10 += (int)(eneg ? -ev : ev);
                    q = r;
                }
            }
            if (simple && w <= (1ull << 53) && exp10 >= -22 && exp10 <= 22) {
                double v = (double)w;
                v = exp10 < 0 ? v / P10[-exp10] : v * P10[exp10];
                *out = neg ? -v : v;
                *pq = q;
                return 1;
            }
        }
    }
    /* strtod on a NUL-terminated copy of the token */
    char buf[512];
    size_t n = 0;
    for (const char *r = start; r < end && n < sizeof buf - 1 && !isspace((unsigned char)*r); ++r) buf[n++] = *r;
    buf[n] = '\0';
    char *e = NULL;
    double v = strtod(buf, &e);
    if (e == buf) return 0;
    *out = v;
    *pq = start + (e - buf);
    return 1;
}

/* sscanf(p, "%lf %lf %lf", ...): the number of values read (0..3). */
static int scan3(const char *p, const char *end, double v[3])
{
    int k = 0;
    for (; k < 3; ++k) {
        if (!scan_double(&p, end, &v[k])) break;
        while (p < end && isspace((unsigned char)*p)) ++p;
    }
    return k;
}

/* strtol(base 10) as the reference's face parser uses it. */
static long parse_long(const char *q, const char *end, const char **qe)
{
    const char *s = q;
    int neg = 0;
    if (at(q, end) == '+' || at(q, end) == '-') {
        neg = at(q, end) == '-';
        ++q;
    }
    if (!isdigit(at(q, end))) {
        *qe = s;
        return 0;
    }
    long v = 0;
    while (isdigit(at(q, end))) {
        v = v * 10 + (at(q, end) - '0');
        ++q;
    }
    *qe = q;
    return neg ? -v : v;
}

/* "v", "v/vt", "v//vn", "v/vt/vn" -> raw indices; returns v != 0. */
static int face_vertex(const char *t, const char *te, int *v, int *vt, int *vn)
{
    const char *e;
    *v = (int)parse_long(t, te, &e);
    *vt = 0;
    *vn = 0;
    if (at(e, te) == '/') {
        ++e;
        if (at(e, te) != '/') *vt = (int)parse_long(e, te, &e);
        if (at(e, te) == '/') {
            ++e;
            *vn = (int)parse_long(e, te, &e);
        }
    }
    return *v != 0;
}

static int resolve(int idx, int count)
{
    if (idx > 0) return idx - 1;
    if (idx < 0) return count + idx;
    return -1;
}

/* ---- OBJ ------------------------------------------------------------------- */

typedef struct {
    double *a;
    long n, cap;
} dvec;

static int dvec_push(dvec *v, double x, double y, double z)
{
    if (v->n == v->cap) {
        long cap = v->cap ? v->cap * 2 : 1024;
        double *na = (double *)realloc(v->a, sizeof(double) * 3 * (size_t)cap);
        if (!na) return -1;
        v->a = na;
        v->cap = cap;
    }
    v->a[3 * v->n] = x;
    v->a[3 * v->n + 1] = y;
    v->a[3 * v->n + 2] = z;
    v->n++;
    return 0;
}

typedef struct {
    const char *data, *end, *c0, *c1; /* buffer, this chunk's lines [c0, c1) */
    dvec pos, nrm, tex;
    double bmin[3], bmax[3];
    /* pass 2 */
    int pos_n, nrm_n, tex_n;
    const double *P, *N, *T;          /* transformed positions / normals, uvs */
    int smooth, material;
    long ntri, base;
    mh_prim *prims;
    mh_prim_attr *attrs;
    int fail;
} obj_chunk;

/* Pass 1: the vertex records of the chunk's lines. */
static void *obj_pass1(void *arg)
{
    obj_chunk *c = (obj_chunk *)arg;
    for (int k = 0; k < 3; ++k) {
        c->bmin[k] = 1e30;
        c->bmax[k] = -1e30;
    }
    const char *p = c->c0;
    while (p < c->c1) {
        const char *eol = memchr(p, '\n', (size_t)(c->end - p));
        size_t ll = eol ? (size_t)(eol - p) : (size_t)(c->end - p);
        const char *line = p, *le = p + ll;
        while (line < le && isspace((unsigned char)*line)) ++line;
        if (line < le && *line != '#') {
            double v[3] = { 0.0, 0.0, 0.0 };
            if (at(line, c->end) == 'v' && isspace(at(line + 1, c->end))) {
                if (scan3(line + 2, c->end, v) == 3) {
                    if (dvec_push(&c->pos, v[0], v[1], v[2])) c->fail = 1;
                    for (int k = 0; k < 3; ++k) {
                        c->bmin[k] = fmin(c->bmin[k], v[k]);
                        c->bmax[k] = fmax(c->bmax[k], v[k]);
                    }
                }
            } else if (at(line, c->end) == 'v' && at(line + 1, c->end) == 'n' && isspace(at(line + 2, c->end))) {
                if (scan3(line + 3, c->end, v) == 3 && dvec_push(&c->nrm, v[0], v[1], v[2])) c->fail = 1;
            } else if (at(line, c->end) == 'v' && at(line + 1, c->end) == 't' && isspace(at(line + 2, c->end))) {
                int m = scan3(line + 3, c->end, v);
                if (m >= 2 && dvec_push(&c->tex, v[0], v[1], m >= 3 ? v[2] : 0.0)) c->fail = 1;
            }
        }
        p += ll + (eol ? 1 : 0);
    }
    return NULL;
}

/* The valid face vertices of a face line (at most 64), as the reference
 * tokenises its 1023-byte copy of the line from offset 2. */
static int face_of(const char *p, size_t ll, int pos_n, int tex_n, int nrm_n, int fv[64][3])
{
    size_t cl = ll < 1023 ? ll : 1023;
    if (cl <= 2) return 0;
    const char *q = p + 2, *qe = p + cl;
    int n = 0;
    while (q < qe && n < 64) {
        while (q < qe && (*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n')) ++q;
        if (q >= qe) break;
        const char *t = q;
        while (q < qe && !(*q == ' ' || *q == '\t' || *q == '\r' || *q == '\n')) ++q;
        int vi, vti, vni;
        if (face_vertex(t, q, &vi, &vti, &vni)) {
            int r = resolve(vi, pos_n);
            if (r >= 0 && r < pos_n) {
                fv[n][0] = r;
                fv[n][1] = resolve(vti, tex_n);
                fv[n][2] = resolve(vni, nrm_n);
                ++n;
            }
        }
    }
    return n;
}

static int is_face_line(const char *p, size_t ll, const char *end)
{
    const char *line = p, *le = p + ll;
    while (line < le && isspace((unsigned char)*line)) ++line;
    return line < le && *line == 'f' && isspace(at(line + 1, end));
}

/* Pass 2a: count the chunk's triangles; 2b (prims set): write them. */
static void *obj_pass2(void *arg)
{
    obj_chunk *c = (obj_chunk *)arg;
    int fv[64][3];
    long k = c->base;
    long count = 0;
    const char *p = c->c0;
    while (p < c->c1) {
        const char *eol = memchr(p, '\n', (size_t)(c->end - p));
        size_t ll = eol ? (size_t)(eol - p) : (size_t)(c->end - p);
        if (is_face_line(p, ll, c->end)) {
            int n = face_of(p, ll, c->pos_n, c->tex_n, c->nrm_n, fv);
            for (int i = 1; i < n - 1; ++i) {
                ++count;
                if (!c->prims) continue;
                int ia = fv[0][0], ib = fv[i][0], ic = fv[i + 1][0];
                int ta = fv[0][1], tb = fv[i][1], tc = fv[i + 1][1];
                int na = fv[0][2], nb = fv[i][2], nc = fv[i + 1][2];
                int uv = ta >= 0 && ta < c->tex_n && tb >= 0 && tb < c->tex_n && tc >= 0 && tc < c->tex_n;
                int sm = c->smooth && na >= 0 && na < c->nrm_n && nb >= 0 && nb < c->nrm_n && nc >= 0 &&
                         nc < c->nrm_n;
                mh_prim *q = &c->prims[k];
                mh_prim_attr *a = &c->attrs[k];
                memset(q, 0, sizeof *q);
                memset(a, 0, sizeof *a);
                q->kind = MH_PRIM_TRIANGLE;
                q->material = c->material;
                const double *A = &c->P[3 * (size_t)ia], *B = &c->P[3 * (size_t)ib], *Cc = &c->P[3 * (size_t)ic];
                q->p0.x = A[0]; q->p0.y = A[1]; q->p0.z = A[2];
                q->p1.x = B[0]; q->p1.y = B[1]; q->p1.z = B[2];
                q->p2.x = Cc[0]; q->p2.y = Cc[1]; q->p2.z = Cc[2];
                if (sm) {
                    q->flags |= MH_PRIM_SMOOTH;
                    const double *n0 = &c->N[3 * (size_t)na], *n1 = &c->N[3 * (size_t)nb], *n2 = &c->N[3 * (size_t)nc];
                    a->n0.x = n0[0]; a->n0.y = n0[1]; a->n0.z = n0[2];
                    a->n1.x = n1[0]; a->n1.y = n1[1]; a->n1.z = n1[2];
                    a->n2.x = n2[0]; a->n2.y = n2[1]; a->n2.z = n2[2];
                }
                if (uv) {
                    q->flags |= MH_PRIM_UV;
                    const double *u0 = &c->T[3 * (size_t)ta], *u1 = &c->T[3 * (size_t)tb], *u2 = &c->T[3 * (size_t)tc];
                    a->uv0.x = u0[0]; a->uv0.y = u0[1]; a->uv0.z = u0[2];
                    a->uv1.x = u1[0]; a->uv1.y = u1[1]; a->uv1.z = u1[2];
                    a->uv2.x = u2[0]; a->uv2.y = u2[1]; a->uv2.z = u2[2];
                }
                ++k;
            }
        }
        p += ll + (eol ? 1 : 0);
    }
    c->ntri = count;
    return NULL;
}

typedef struct {
    double rc[3], rs[3];  /* cos / sin of the three Euler angles */
    int rotate;           /* some angle is not ~0 */
} euler;

static euler euler_of(const double deg[3])
{
    euler e;
    e.rotate = !(fabs(deg[0]) < 1e-9 && fabs(deg[1]) < 1e-9 && fabs(deg[2]) < 1e-9);
    for (int k = 0; k < 3; ++k) {
        double r = deg[k] * (MT_PI / 180.0);
        e.rc[k] = cos(r);
        e.rs[k] = sin(r);
    }
    return e;
}

/* Rotate about X, then Y, then Z. */
static v3 rotate3(v3 v, const euler *e)
{
    if (!e->rotate) return v;
    double cx = e->rc[0], sx = e->rs[0], cy = e->rc[1], sy = e->rs[1], cz = e->rc[2], sz = e->rs[2];
    v3 v1 = mk3(v.x, v.y * cx - v.z * sx, v.y * sx + v.z * cx);
    v3 v2 = mk3(v1.x * cy + v1.z * sy, v1.y, -v1.x * sy + v1.z * cy);
    return mk3(v2.x * cz - v2.y * sz, v2.x * sz + v2.y * cz, v2.z);
}

typedef struct {
    const double *in;
    double *out;
    long i0, i1;
    int normals;
    const mt_obj_xform *xf;
    const euler *e;
    v3 off;
    double k;
} xform_job;

static void *xform_run(void *arg)
{
    xform_job *j = (xform_job *)arg;
    const mt_obj_xform *xf = j->xf;
    for (long i = j->i0; i < j->i1; ++i) {
        v3 v = mk3(j->in[3 * i], j->in[3 * i + 1], j->in[3 * i + 2]);
        v3 r;
        if (j->normals) {
            v3 sn = mk3(fabs(xf->scale[0]) > 1e-12 ? v.x / xf->scale[0] : v.x,
                        fabs(xf->scale[1]) > 1e-12 ? v.y / xf->scale[1] : v.y,
                        fabs(xf->scale[2]) > 1e-12 ? v.z / xf->scale[2] : v.z);
            v3 rn = rotate3(sn, j->e);
            double len = sqrt(rn.x * rn.x + rn.y * rn.y + rn.z * rn.z);
            if (len > 1e-12) {
                double s = 1.0 / len;
                rn = mk3(rn.x * s, rn.y * s, rn.z * s);
            }
            r = rn;
        } else {
            v3 p = mk3(v.x - j->off.x, v.y - j->off.y, v.z - j->off.z);
            p = mk3(p.x * j->k, p.y * j->k, p.z * j->k);
            p = mk3(p.x * xf->scale[0], p.y * xf->scale[1], p.z * xf->scale[2]);
            p = rotate3(p, j->e);
            r = mk3(p.x + xf->position[0], p.y + xf->position[1], p.z + xf->position[2]);
        }
        j->out[3 * i] = r.x;
        j->out[3 * i + 1] = r.y;
        j->out[3 * i + 2] = r.z;
    }
    return NULL;
}

static int xform_all(const double *in, double *out, long n, int normals, const mt_obj_xform *xf, const euler *e,
                     v3 off, double k, int threads)
{
    int nt = n_threads(threads, (size_t)(n / 65536 + 1));
    xform_job jobs[64];
    for (int t = 0; t < nt; ++t) {
        jobs[t] = (xform_job){ in, out, n * t / nt, n * (t + 1) / nt, normals, xf, e, off, k };
    }
    return run_threads(nt, xform_run, jobs, sizeof(xform_job));
}

void mt_obj_xform_default(mt_obj_xform *x)
{
    if (!x) return;
    memset(x, 0, sizeof *x);
    x->scale[0] = x->scale[1] = x->scale[2] = 1.0;
    x->smooth = 1;
}

static double *concat(obj_chunk *c, int nt, size_t field, long *total)
{
    long n = 0;
    for (int t = 0; t < nt; ++t) n += ((dvec *)((char *)&c[t] + field))->n;
    double *out = (double *)malloc(sizeof(double) * 3 * (size_t)(n ? n : 1));
    if (!out) return NULL;
    long k = 0;
    for (int t = 0; t < nt; ++t) {
        dvec *v = (dvec *)((char *)&c[t] + field);
        if (v->n) memcpy(out + 3 * k, v->a, sizeof(double) * 3 * (size_t)v->n); /* v->a is NULL when empty */
        k += v->n;
    }
    *total = n;
    return out;
}

long mt_scene_add_obj(mt_scene *s, const char *data, size_t len, int material, const mt_obj_xform *xform,
                      int threads)
{
    if (!s || (!data && len > 0)) return mt_set_err(-1, "NULL argument");
    if (len == 0) return mt_set_err(-1, "empty OBJ text");
    mt_obj_xform xf;
    if (xform) xf = *xform;
    else mt_obj_xform_default(&xf);
    const char *end = data + len;
    int nt = n_threads(threads, len / (1 << 20) + 1);
    obj_chunk *ch = (obj_chunk *)calloc((size_t)nt, sizeof(obj_chunk));
    if (!ch) return mt_set_err(-2, "out of memory");
    /* chunks start at line starts */
    const char *cut[65];
    cut[0] = data;
    for (int t = 1; t < nt; ++t) {
        const char *q = data + len * (size_t)t / (size_t)nt;
        if (q < cut[t - 1]) q = cut[t - 1];
        const char *nl = memchr(q, '\n', (size_t)(end - q));
        cut[t] = nl ? nl + 1 : end;
    }
    cut[nt] = end;
    for (int t = 0; t < nt; ++t) {
        ch[t].data = data;
        ch[t].end = end;
        ch[t].c0 = cut[t];
        ch[t].c1 = cut[t + 1];
    }
    run_threads(nt, obj_pass1, ch, sizeof(obj_chunk));
    long npos = 0, nnrm = 0, ntex = 0, ntri = 0, ret = -2;
    double *pos = NULL, *nrm = NULL, *tex = NULL, *P = NULL, *N = NULL;
    double bmin[3] = { 1e30, 1e30, 1e30 }, bmax[3] = { -1e30, -1e30, -1e30 };
    int fail = 0;
    for (int t = 0; t < nt; ++t) {
        fail |= ch[t].fail;
        for (int k = 0; k < 3; ++k) {
            bmin[k] = fmin(bmin[k], ch[t].bmin[k]);
            bmax[k] = fmax(bmax[k], ch[t].bmax[k]);
        }
    }
    pos = concat(ch, nt, offsetof(obj_chunk, pos), &npos);
    nrm = concat(ch, nt, offsetof(obj_chunk, nrm), &nnrm);
    tex = concat(ch, nt, offsetof(obj_chunk, tex), &ntex);
    if (fail || !pos || !nrm || !tex) {
        mt_set_err(-2, "out of memory");
        goto done;
    }
    if (npos == 0) {
        ret = 0;
        goto done;
    }
    if (npos > 0x7fffffffL || nnrm > 0x7fffffffL || ntex > 0x7fffffffL) {
        ret = mt_set_err(-3, "too many OBJ records");
        goto done;
    }
    {
        v3 off = mk3(0.0, 0.0, 0.0);
        if (xf.auto_center) off = mk3((bmin[0] + bmax[0]) * 0.5, (bmin[1] + bmax[1]) * 0.5, (bmin[2] + bmax[2]) * 0.5);
        double k = 1.0;
        if (xf.auto_scale > 0.0) {
            double md = fmax(bmax[0] - bmin[0], fmax(bmax[1] - bmin[1], bmax[2] - bmin[2]));
            if (md > 1e-9) k = xf.auto_scale / md;
        }
        euler e = euler_of(xf.rotate_deg);
        P = (double *)malloc(sizeof(double) * 3 * (size_t)npos);
        N = (double *)malloc(sizeof(double) * 3 * (size_t)(nnrm ? nnrm : 1));
        if (!P || !N) {
            mt_set_err(-2, "out of memory");
            goto done;
        }
        xform_all(pos, P, npos, 0, &xf, &e, off, k, threads);
        xform_all(nrm, N, nnrm, 1, &xf, &e, off, k, threads);
    }
    for (int t = 0; t < nt; ++t) {
        ch[t].pos_n = (int)npos;
        ch[t].nrm_n = (int)nnrm;
        ch[t].tex_n = (int)ntex;
        ch[t].P = P;
        ch[t].N = N;
        ch[t].T = tex;
        ch[t].smooth = xf.smooth;
        ch[t].material = material;
    }
    run_threads(nt, obj_pass2, ch, sizeof(obj_chunk));  /* counts */
    for (int t = 0; t < nt; ++t) {
        ch[t].base = ntri;
        ntri += ch[t].ntri;
    }
    {
        long first = mt_reserve(s, ntri, 1);
        if (first < 0) {
            ret = first;
            goto done;
        }
        for (int t = 0; t < nt; ++t) {
            ch[t].base += first;
            ch[t].prims = s->prims;
            ch[t].attrs = s->attrs;
        }
        run_threads(nt, obj_pass2, ch, sizeof(obj_chunk));  /* triangles */
        s->n = (int)(first + ntri);
        ret = ntri;
    }
done:
    for (int t = 0; t < nt; ++t) {
        free(ch[t].pos.a);
        free(ch[t].nrm.a);
        free(ch[t].tex.a);
    }
    free(ch);
    free(pos);
    free(nrm);
    free(tex);
    free(P);
    free(N);
    return ret;
}

long mt_scene_add_obj_file(mt_scene *s, const char *path, int material, const mt_obj_xform *xf, int threads)
{
    if (!s || !path) return mt_set_err(-1, "NULL argument");
    FILE *f = fopen(path, "rb");
    if (!f) return mt_set_err(-1, "cannot open %s", path);
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return mt_set_err(-2, "cannot read %s", path);
    }
    long size = ftell(f);
    rewind(f);
    if (size <= 0) {
        fclose(f);
        return size == 0 ? 0 : mt_set_err(-2, "cannot read %s", path);
    }
    char *buf = (char *)malloc((size_t)size + 1);
    if (!buf) {
        fclose(f);
        return mt_set_err(-2, "out of memory");
    }
    size_t got = fread(buf, 1, (size_t)size, f);
    fclose(f);
    buf[got] = '\0';
    long rc = got ? mt_scene_add_obj(s, buf, got, material, xf, threads) : 0;
    free(buf);
    return rc;
}

/* ---- ocean ------------------------------------------------------------------ */

void mt_ocean_default(mt_ocean_desc *o)
{
    if (!o) return;
    memset(o, 0, sizeof *o);
    o->size[0] = 100.0;
    o->size[1] = 100.0;
    o->res_x = 128;
    o->res_z = 128;
    o->amplitude = 0.25;
    o->wavelength = 12.0;
    o->direction[0] = 1.0;
    o->steepness = 0.5;
    o->chop = 0.08;
    o->chop_wavelength = 4.0;
    o->depth = 10.0;
    o->seed = 42;
}

typedef struct {
    double dx[4], dz[4], k[4], Q[4], A[4], phi[4];
} harmonics;

/* The four Gerstner harmonics of an ocean (reference src/scene.c:1162-1208). */
static harmonics harmonics_of(const mt_ocean_desc *o, double *amp_out)
{
    harmonics h;
    double amp = o->amplitude;
    double wl = o->wavelength <= 0.0 ? 12.0 : o->wavelength;
    double steep = o->steepness < 0.0 ? 0.0 : (o->steepness > 0.90 ? 0.90 : o->steepness);
    double chop_amp = o->chop > 0.0 ? o->chop : amp * 0.25;
    double chop_wl = o->chop_wavelength > 0.0 ? o->chop_wavelength : wl * 0.25;
    double dir_x = o->direction[0], dir_z = o->direction[2];
    double dlen = sqrt(dir_x * dir_x + dir_z * dir_z);
    if (dlen > 1e-12) {
        dir_x /= dlen;
        dir_z /= dlen;
    } else {
        dir_x = 1.0;
        dir_z = 0.0;
    }
    const double lambda[4] = { wl, wl * 0.52, chop_wl, chop_wl * 0.40 };
    const double A[4] = { amp, amp * 0.38, chop_amp, chop_amp * 0.32 };
    const double angle[4] = { 0.0, 0.628, -0.488, 1.134 };
    const double qs[4] = { 0.55, 0.25, 0.15, 0.05 };
    h.phi[0] = 0.0;
    h.phi[1] = 1.7 + (double)(o->seed % 100) * 0.03;
    h.phi[2] = 3.1 + (double)((o->seed * 7) % 100) * 0.03;
    h.phi[3] = 0.8 + (double)((o->seed * 13) % 100) * 0.03;
    for (int i = 0; i < 4; ++i) {
        double ca = cos(angle[i]), sa = sin(angle[i]);
        h.dx[i] = ca * dir_x - sa * dir_z;
        h.dz[i] = sa * dir_x + ca * dir_z;
        h.k[i] = (2.0 * MT_PI) / lambda[i];
        h.Q[i] = steep * qs[i];
        if (h.k[i] * A[i] > 1e-6) {
            double qmax = 0.85 / (h.k[i] * A[i]);
            if (h.Q[i] > qmax) h.Q[i] = qmax;
        }
        if (h.Q[i] > 0.85) h.Q[i] = 0.85;
        h.A[i] = A[i];
    }
    if (amp_out) *amp_out = amp;
    return h;
}

typedef struct {
    const mt_ocean_desc *o;
    const harmonics *h;
    int rx, j0, j1;
    double dx, dz, x_min, z_min;
    v3 *verts, *norms;
} grid_job;

static void *grid_run(void *arg)
{
    grid_job *g = (grid_job *)arg;
    const harmonics *h = g->h;
    for (int j = g->j0; j < g->j1; ++j) {
        double z0 = g->z_min + j * g->dz;
        for (int i = 0; i < g->rx; ++i) {
            double x0 = g->x_min + i * g->dx;
            double x = x0, z = z0, y = g->o->center[1];
            double dxdx0 = 1.0, dxdz0 = 0.0, dzdx0 = 0.0, dzdz0 = 1.0, dydx0 = 0.0, dydz0 = 0.0;
            for (int k = 0; k < 4; ++k) {
                double theta = h->k[k] * (h->dx[k] * x0 + h->dz[k] * z0) + h->phi[k];
                double sn = sin(theta), cs = cos(theta);
                x -= (h->Q[k] / h->k[k]) * h->dx[k] * sn;
                z -= (h->Q[k] / h->k[k]) * h->dz[k] * sn;
                y += h->A[k] * cs;
                dxdx0 -= h->Q[k] * h->dx[k] * h->dx[k] * cs;
                dxdz0 -= h->Q[k] * h->dx[k] * h->dz[k] * cs;
                dzdx0 -= h->Q[k] * h->dx[k] * h->dz[k] * cs;
                dzdz0 -= h->Q[k] * h->dz[k] * h->dz[k] * cs;
                dydx0 -= h->k[k] * h->A[k] * h->dx[k] * sn;
                dydz0 -= h->k[k] * h->A[k] * h->dz[k] * sn;
            }
            /* N = Tz x Tx with Tx = dP/dx0, Tz = dP/dz0, oriented up */
            v3 Tx = mk3(dxdx0, dydx0, dzdx0), Tz = mk3(dxdz0, dydz0, dzdz0);
            v3 N = mk3(Tz.y * Tx.z - Tz.z * Tx.y, Tz.z * Tx.x - Tz.x * Tx.z, Tz.x * Tx.y - Tz.y * Tx.x);
            double nl = sqrt(N.x * N.x + N.y * N.y + N.z * N.z);
            if (nl > 1e-12) {
                double s = 1.0 / nl;
                N = mk3(N.x * s, N.y * s, N.z * s);
            } else {
                N = mk3(0.0, 1.0, 0.0);
            }
            if (N.y < 0.0) N = mk3(-N.x, -N.y, -N.z);
            size_t idx = (size_t)j * (size_t)g->rx + (size_t)i;
            g->verts[idx] = mk3(x, y, z);
            g->norms[idx] = N;
        }
    }
    return NULL;
}

static void put_tri(mh_prim *q, mh_prim_attr *a, int mat, v3 A, v3 B, v3 C, v3 na, v3 nb, v3 nc)
{
    memset(q, 0, sizeof *q);
    memset(a, 0, sizeof *a);
    q->kind = MH_PRIM_TRIANGLE;
    q->material = mat;
    q->flags = MH_PRIM_SMOOTH;
    q->p0.x = A.x; q->p0.y = A.y; q->p0.z = A.z;
    q->p1.x = B.x; q->p1.y = B.y; q->p1.z = B.z;
    q->p2.x = C.x; q->p2.y = C.y; q->p2.z = C.z;
    a->n0.x = na.x; a->n0.y = na.y; a->n0.z = na.z;
    a->n1.x = nb.x; a->n1.y = nb.y; a->n1.z = nb.z;
    a->n2.x = nc.x; a->n2.y = nc.y; a->n2.z = nc.z;
}

typedef struct {
    int rx, j0, j1, mat;
    const v3 *verts, *norms;
    mh_prim *prims;
    mh_prim_attr *attrs;
    long base;
} surf_job;

static void *surf_run(void *arg)
{
    surf_job *g = (surf_job *)arg;
    int rx = g->rx;
    for (int j = g->j0; j < g->j1; ++j)
        for (int i = 0; i < rx - 1; ++i) {
            size_t i00 = (size_t)j * rx + i, i10 = i00 + 1, i01 = i00 + rx, i11 = i01 + 1;
            long k = g->base + 2 * ((long)j * (rx - 1) + i);
            const v3 *v = g->verts, *n = g->norms;
            put_tri(&g->prims[k], &g->attrs[k], g->mat, v[i00], v[i01], v[i11], n[i00], n[i01], n[i11]);
            put_tri(&g->prims[k + 1], &g->attrs[k + 1], g->mat, v[i00], v[i11], v[i10], n[i00], n[i11], n[i10]);
        }
    return NULL;
}

long mt_scene_add_ocean(mt_scene *s, const mt_ocean_desc *o, int material, int threads)
{
    if (!s || !o) return mt_set_err(-1, "NULL argument");
    int rx = o->res_x < 2 ? 2 : o->res_x;
    int rz = o->res_z < 2 ? 2 : o->res_z;
    if ((long)rx * (long)rz > 100000000L) return mt_set_err(-3, "ocean grid %d x %d too large", rx, rz);
    double sx = o->size[0] <= 0.0 ? 100.0 : o->size[0];
    double sz = o->size[1] <= 0.0 ? 100.0 : o->size[1];
    double depth = o->depth < 0.0 ? 0.0 : o->depth;
    harmonics h = harmonics_of(o, NULL);
    size_t nv = (size_t)rx * (size_t)rz;
    v3 *verts = (v3 *)malloc(nv * sizeof(v3));
    v3 *norms = (v3 *)malloc(nv * sizeof(v3));
    if (!verts || !norms) {
        free(verts);
        free(norms);
        return mt_set_err(-2, "out of memory");
    }
    double dx = sx / (double)(rx - 1), dz = sz / (double)(rz - 1);
    double x_min = o->center[0] - sx * 0.5, z_min = o->center[2] - sz * 0.5;
    int nt = n_threads(threads, (size_t)rz);
    grid_job gj[64];
    for (int t = 0; t < nt; ++t)
        gj[t] = (grid_job){ o, &h, rx, rz * t / nt, rz * (t + 1) / nt, dx, dz, x_min, z_min, verts, norms };
    run_threads(nt, grid_run, gj, sizeof(grid_job));

    long nsurf = 2L * (rx - 1) * (rz - 1);
    long nskirt = depth > 0.0 ? 4L * (rx - 1) + 4L * (rz - 1) + 2 : 0;
    long first = mt_reserve(s, nsurf + nskirt, 1);
    if (first < 0) {
        free(verts);
        free(norms);
        return first;
    }
    int nts = n_threads(threads, (size_t)(rz - 1));
    surf_job sj[64];
    for (int t = 0; t < nts; ++t)
        sj[t] = (surf_job){ rx, (rz - 1) * t / nts, (rz - 1) * (t + 1) / nts, material, verts, norms, s->prims,
                            s->attrs, first };
    run_threads(nts, surf_run, sj, sizeof(surf_job));
    long k = first + nsurf;
    if (depth > 0.0) {
        double yf = o->center[1] - depth;
        mh_prim *P = s->prims;
        mh_prim_attr *A = s->attrs;
        v3 ns = mk3(0.0, 0.0, -1.0), nn = mk3(0.0, 0.0, 1.0), nw = mk3(-1.0, 0.0, 0.0), ne = mk3(1.0, 0.0, 0.0);
        for (int i = 0; i < rx - 1; ++i) {   /* south, j = 0 */
            v3 t0 = verts[i], t1 = verts[i + 1], b0 = mk3(t0.x, yf, t0.z), b1 = mk3(t1.x, yf, t1.z);
            put_tri(&P[k], &A[k], material, t0, b0, b1, ns, ns, ns); ++k;
            put_tri(&P[k], &A[k], material, t0, b1, t1, ns, ns, ns); ++k;
        }
        size_t jn = (size_t)(rz - 1) * rx;
        for (int i = 0; i < rx - 1; ++i) {   /* north, j = rz - 1 */
            v3 t0 = verts[jn + i], t1 = verts[jn + i + 1], b0 = mk3(t0.x, yf, t0.z), b1 = mk3(t1.x, yf, t1.z);
            put_tri(&P[k], &A[k], material, t0, b1, b0, nn, nn, nn); ++k;
            put_tri(&P[k], &A[k], material, t0, t1, b1, nn, nn, nn); ++k;
        }
        for (int j = 0; j < rz - 1; ++j) {   /* west, i = 0 */
            v3 t0 = verts[(size_t)j * rx], t1 = verts[(size_t)(j + 1) * rx];
            v3 b0 = mk3(t0.x, yf, t0.z), b1 = mk3(t1.x, yf, t1.z);
            put_tri(&P[k], &A[k], material, t0, b1, b0, nw, nw, nw); ++k;
            put_tri(&P[k], &A[k], material, t0, t1, b1, nw, nw, nw); ++k;
        }
        for (int j = 0; j < rz - 1; ++j) {   /* east, i = rx - 1 */
            v3 t0 = verts[(size_t)j * rx + rx - 1], t1 = verts[(size_t)(j + 1) * rx + rx - 1];
            v3 b0 = mk3(t0.x, yf, t0.z), b1 = mk3(t1.x, yf, t1.z);
            put_tri(&P[k], &A[k], material, t0, b0, b1, ne, ne, ne); ++k;
            put_tri(&P[k], &A[k], material, t0, b1, t1, ne, ne, ne); ++k;
        }
        v3 nf = mk3(0.0, -1.0, 0.0);
        v3 c00 = mk3(verts[0].x, yf, verts[0].z), c10 = mk3(verts[rx - 1].x, yf, verts[rx - 1].z);
        v3 c01 = mk3(verts[jn].x, yf, verts[jn].z), c11 = mk3(verts[nv - 1].x, yf, verts[nv - 1].z);
        put_tri(&P[k], &A[k], material, c00, c10, c11, nf, nf, nf); ++k;
        put_tri(&P[k], &A[k], material, c00, c11, c01, nf, nf, nf); ++k;
    }
    s->n = (int)k;
    free(verts);
    free(norms);
    if (!s->ocean.enabled) {
        s->ocean.enabled = 1;
        s->ocean.seed = o->seed;
        s->ocean.center.x = o->center[0];
        s->ocean.center.y = o->center[1];
        s->ocean.center.z = o->center[2];
        s->ocean.direction.x = o->direction[0];
        s->ocean.direction.y = o->direction[1];
        s->ocean.direction.z = o->direction[2];
        s->ocean.amplitude = o->amplitude;
        s->ocean.wavelength = o->wavelength;
        s->ocean.steepness = o->steepness;
        s->ocean.chop = o->chop;
        s->ocean.chop_wavelength = o->chop_wavelength;
        s->water_level = o->center[1];
        s->water_material = material;
    }
    return k - first;
}
