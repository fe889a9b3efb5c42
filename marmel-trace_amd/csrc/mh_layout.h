// mh_layout.h - the device layout of a scene, built on the host: the BVH
// as 64-byte binary and 128-byte 4-wide nodes, the leaf-ordered fp32 / fp64
// primitive records and vertex attributes, the material / sky / fog / ocean
// tables and the emissive lights, and the SceneView fields that point the
// kernels at them (layout described in mh_scene_dev.h). Nothing here calls
// the HIP runtime: mh_api.cpp uploads these arrays, and a host build of the
// render code (tests/hostdev) uses them in place.
#pragma once

#include "marmel_hip.h"
#include "mh_pack.h"
#include "mh_scene_dev.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace mh {

// Error reporting of the layout checks (code + printf-style message), defined
// by the user of this header; returns `code`.
int layout_fail(int code, const char* fmt, ...);

// A sphere that may be an emissive light (file order).
struct LightSphere {
    int prim, mat;
    mh_vec3 c;
    double r;
};

inline std::vector<LightSphere> light_spheres(const mh_scene_soa* soa)
{
    std::vector<LightSphere> out;
    for (int i = 0; i < soa->n_prims; ++i) {
        const mh_prim& p = soa->prims[i];
        if (p.kind == MH_PRIM_SPHERE) out.push_back({i, p.material, p.p0, p.r0});
    }
    return out;
}

inline bool has_attrs(const mh_scene_soa* soa)
{
    if (!soa->attrs) return false;
    for (int i = 0; i < soa->n_prims; ++i)
        if (soa->prims[i].flags != 0) return true;
    return false;
}

template <typename R> DMat<R> conv_mat(const mh_material& a)
{
    DMat<R> m;
    memset(&m, 0, sizeof m);
    m.albedo = cv<R>(a.albedo);
    m.specular = cv<R>(a.specular);
    m.shininess = (R)a.shininess;
    m.reflectivity = (R)a.reflectivity;
    m.transparency = (R)a.transparency;
    m.ior = (R)a.ior;
    m.is_water = a.is_water;
    m.beer_lambert = a.beer_lambert;
    m.pbr = a.pbr;
    m.texture_kind = a.texture_kind;
    m.absorption = cv<R>(a.absorption);
    m.deep_color = cv<R>(a.deep_color);
    m.metallic = (R)a.metallic;
    m.roughness = (R)a.roughness;
    m.emissive = cv<R>(a.emissive);
    m.texture_scale = (R)a.texture_scale;
    m.tex_a = cv<R>(a.texture_color_a);
    m.tex_b = cv<R>(a.texture_color_b);
    m.glow = cv<R>(a.atmosphere_glow);
    m.has_emission = (a.emissive.x != 0.0 || a.emissive.y != 0.0 || a.emissive.z != 0.0) ? 1 : 0;
    double g2 = a.atmosphere_glow.x * a.atmosphere_glow.x + a.atmosphere_glow.y * a.atmosphere_glow.y +
                a.atmosphere_glow.z * a.atmosphere_glow.z;
    m.has_glow = g2 > 1e-6 ? 1 : 0;
    return m;
}

template <typename R> DSky<R> conv_sky(const mh_sky& k)
{
    DSky<R> s;
    memset(&s, 0, sizeof s);
    s.sun_dir = cv<R>(k.sun_dir);
    s.sun_color = cv<R>(k.sun_color);
    s.horizon = cv<R>(k.horizon_color);
    s.zenith = cv<R>(k.zenith_color);
    s.gamma = (R)k.gradient_gamma;
    s.glow_exp = (R)k.sun_glow_exponent;
    s.glow_strength = (R)k.sun_glow_strength;
    s.cloud_height = (R)k.cloud_height;
    s.cloud_scale = (R)k.cloud_scale;
    s.coverage = (R)k.cloud_coverage;
    s.softness = (R)k.cloud_softness;
    s.sharpness = (R)k.cloud_sharpness;
    s.octaves = k.cloud_octaves < 1 ? 1 : k.cloud_octaves;
    s.seed = k.seed;
    s.sun_radius = (R)k.sun_radius;
    // solid angle of the sun disk as the reference NEE uses it (1 for a point sun)
    double rad = k.sun_radius * 3.14159265358979323846 / 180.0;
    s.sun_omega = (R)(k.sun_radius > 0.0 ? 2.0 * 3.14159265358979323846 * (1.0 - cos(rad)) : 1.0);
    return s;
}

// Device fog parameters: the anisotropy clamped to [-0.95, 0.95] and the
// turbulence switch evaluated once, as the reference evaluates them per call.
template <typename R> DFog<R> conv_fog(const mh_fog& f)
{
    DFog<R> o;
    memset(&o, 0, sizeof o);
    o.density = (R)f.density;
    o.color = cv<R>(f.color);
    o.height = (R)f.height;
    o.falloff = (R)f.height_falloff;
    o.inscatter = (R)f.inscatter_strength;
    double g = f.sun_anisotropy;
    if (g < -0.95) g = -0.95;
    if (g > 0.95) g = 0.95;
    o.g = (R)g;
    o.noise_scale = (R)f.noise_scale;
    o.noise_amount = (R)f.noise_amount;
    o.steps = f.shadow_steps;
    o.noise = f.noise_amount > 1e-4 && f.noise_scale > 1e-4 ? 1 : 0;
    return o;
}

// The ocean's harmonics as the reference forms them for each height query
// (src/scene.c:1109-1153), once.
template <typename R> DOcean<R> conv_ocean(const mh_ocean& o)
{
    DOcean<R> d;
    memset(&d, 0, sizeof d);
    d.on = o.enabled ? 1 : 0;
    if (!d.on) return d;
    d.y0 = (R)o.center.y;
    double dir_x = o.direction.x, dir_z = o.direction.z;
    double dlen = sqrt(dir_x * dir_x + dir_z * dir_z);
    if (dlen > 1e-12) {
        dir_x /= dlen;
        dir_z /= dlen;
    } else {
        dir_x = 1.0;
        dir_z = 0.0;
    }
    double wl = o.wavelength <= 0.0 ? 12.0 : o.wavelength;
    double amp = o.amplitude;
    double chop_wl = o.chop_wavelength > 0.0 ? o.chop_wavelength : wl * 0.25;
    double chop_amp = o.chop > 0.0 ? o.chop : amp * 0.25;
    const double lambda[4] = {wl, wl * 0.52, chop_wl, chop_wl * 0.40};
    const double A[4] = {amp, amp * 0.38, chop_amp, chop_amp * 0.32};
    const double angle[4] = {0.0, 0.628, -0.488, 1.134};
    const double phi[4] = {0.0, 1.7 + (double)(o.seed % 100u) * 0.03, 3.1 + (double)((o.seed * 7u) % 100u) * 0.03,
                           0.8 + (double)((o.seed * 13u) % 100u) * 0.03};
    for (int h = 0; h < 4; ++h) {
        double ca = cos(angle[h]), sa = sin(angle[h]);
        d.dx[h] = (R)(ca * dir_x - sa * dir_z);
        d.dz[h] = (R)(sa * dir_x + ca * dir_z);
        d.k[h] = (R)((2.0 * 3.14159265358979323846) / lambda[h]);
        d.phi[h] = (R)phi[h];
        d.A[h] = (R)A[h];
    }
    return d;
}

inline int any_water(const mh_material* mats, int n)
{
    for (int i = 0; i < n; ++i)
        if (mats[i].is_water) return 1;
    return 0;
}

// Fog on, or marched underwater segments possible (they need fog steps > 0
// and a water material, fog density or not).
inline int needs_fog_kernel(const mh_fog& f, int has_water) { return f.density > 0.0 || (f.shadow_steps > 0 && has_water); }

// Walk the host tree from the root: checks child indices, leaf ranges, that
// every leaf slot is covered once, and the depth bound of the GPU stack.
inline int check_tree(const mh_scene_soa* soa, std::vector<int>* order)
{
    if (soa->n_nodes == 0) return 0;
    std::vector<char> seen(soa->n_nodes, 0);
    std::vector<char> slot(soa->n_leaf_prims, 0);
    std::vector<std::pair<int, int>> st;
    st.push_back({0, 1});
    int covered = 0;
    while (!st.empty()) {
        auto [n, depth] = st.back();
        st.pop_back();
        if (n < 0 || n >= soa->n_nodes) return layout_fail(5, "BVH child index %d out of range", n);
        if (seen[n]) return layout_fail(5, "BVH node %d reached twice", n);
        seen[n] = 1;
        if (depth > 60) return layout_fail(5, "BVH deeper than 60 levels");
        const mh_bvh_node& nd = soa->nodes[n];
        if (nd.count > 0) {
            if (nd.count > 8) return layout_fail(5, "BVH leaf %d holds %d > 8 primitives", n, nd.count);
            if (nd.first < 0 || nd.first + nd.count > soa->n_leaf_prims) return layout_fail(5, "BVH leaf %d range invalid", n);
            for (int i = 0; i < nd.count; ++i) {
                if (slot[nd.first + i]) return layout_fail(5, "leaf slot %d referenced twice", nd.first + i);
                slot[nd.first + i] = 1;
            }
            covered += nd.count;
        } else {
            order->push_back(n);
            st.push_back({nd.right, depth + 1});
            st.push_back({nd.left, depth + 1});
        }
    }
    if (covered != soa->n_leaf_prims) return layout_fail(5, "BVH leaves cover %d of %d primitives", covered, soa->n_leaf_prims);
    return 0;
}

// Collapse the binary tree into 4-wide nodes (DNode4): a node adopts its two
// binary children, then replaces its largest-area interior child by that
// child's own two children (in place, keeping left-to-right order) until it
// holds four or only leaves. Nodes are numbered in depth-first pre-order.
struct Bvh4Builder {
    const mh_scene_soa* soa;
    std::vector<DNode4> nodes;
    int max_depth = 0;

    static int leaf_ref(const mh_bvh_node& nd) { return ~((nd.first << 3) | (nd.count - 1)); }
    double area(int n) const
    {
        const mh_bvh_node& b = soa->nodes[n];
        double dx = (double)b.bmax[0] - b.bmin[0], dy = (double)b.bmax[1] - b.bmin[1], dz = (double)b.bmax[2] - b.bmin[2];
        return dx * dy + dy * dz + dz * dx;
    }
    int make(int n2, int depth)
    {
        if (depth > max_depth) max_depth = depth;
        const mh_bvh_node& nd = soa->nodes[n2];
        std::vector<int> c = {nd.left, nd.right};
        while (c.size() < 4) {
            int best = -1;
            double ba = -1.0;
            for (size_t i = 0; i < c.size(); ++i)
                if (soa->nodes[c[i]].count == 0 && area(c[i]) > ba) {
                    ba = area(c[i]);
                    best = (int)i;
                }
            if (best < 0) break;
            const mh_bvh_node& e = soa->nodes[c[best]];
            c[best] = e.left;
            c.insert(c.begin() + best + 1, e.right);
        }
        int idx = (int)nodes.size();
        nodes.push_back(DNode4{});
        float lo[3][4], hi[3][4];
        int ref[4];
        for (int k = 0; k < 4; ++k) {
            for (int a = 0; a < 3; ++a) {
                lo[a][k] = 0.f;
                hi[a][k] = 0.f;
            }
            ref[k] = INT_MIN;
        }
        for (size_t k = 0; k < c.size(); ++k) {
            const mh_bvh_node& ch = soa->nodes[c[k]];
            for (int a = 0; a < 3; ++a) {
                lo[a][k] = ch.bmin[a];
                hi[a][k] = ch.bmax[a];
            }
            ref[k] = ch.count > 0 ? leaf_ref(ch) : make(c[k], depth + 1);
        }
        DNode4& o = nodes[idx];
        o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
        o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
        o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
        o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
        o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
        o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
        o.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
        o.meta = make_int4((int)c.size(), 0, 0, 0);
        return idx;
    }
};

template <typename R> CamView<R> conv_cam(const mh_camera& c)
{
    CamView<R> v;
    v.pos = cv<R>(c.position);
    v.ll = cv<R>(c.lower_left);
    v.hor = cv<R>(c.horizontal);
    v.ver = cv<R>(c.vertical);
    v.aperture = (R)c.aperture;
    v.focus = (R)c.focus_distance;
    v.dome = (R)c.dome_radius;
    v.squeeze = (R)c.anamorphic_squeeze;
    // normalised viewport axes, as the thin-lens sampler forms them (in double)
    auto nrm = [](mh_vec3 a) {
        double l2 = a.x * a.x + a.y * a.y + a.z * a.z;
        if (l2 <= 1e-12) return mh_vec3{0, 0, 0};
        double k = 1.0 / sqrt(l2);
        return mh_vec3{a.x * k, a.y * k, a.z * k};
    };
    v.right_n = cv<R>(nrm(c.horizontal));
    v.up_n = cv<R>(nrm(c.vertical));
    return v;
}

// Emissive area lights: the first 8 PBR spheres with emission, file order
// (reference scene_collect_emissive_lights, src/scene.c:1384-1426).
inline int collect_lights(const std::vector<LightSphere>& spheres, const mh_material* mats, int n_mats, DLight<float>* l32,
                   DLight<double>* l64)
{
    int nl = 0;
    for (const auto& sp : spheres) {
        if (nl >= 8) break;
        if (sp.mat < 0 || sp.mat >= n_mats) continue;
        const mh_material& m = mats[sp.mat];
        if (!m.pbr) continue;
        if (m.emissive.x == 0.0 && m.emissive.y == 0.0 && m.emissive.z == 0.0) continue;
        if (!(sp.r > 0.0)) continue;
        l32[nl] = {sp.prim, sp.mat, cv<float>(sp.c), (float)sp.r, cv<float>(m.emissive)};
        l64[nl] = {sp.prim, sp.mat, cv<double>(sp.c), sp.r, cv<double>(m.emissive)};
        ++nl;
    }
    for (int i = nl; i < 8; ++i) {
        l32[i] = DLight<float>{};
        l64[i] = DLight<double>{};
    }
    return nl;
}

inline bool transmissive(const mh_material& m) { return m.is_water || m.transparency > 0.8; }

inline int any_transmissive(const mh_material* mats, int n)
{
    for (int i = 0; i < n; ++i)
        if (transmissive(mats[i])) return 1;
    return 0;
}

// Bounding box of each material's primitives: box[6 * m .. 6 * m + 6) = lo
// xyz, hi xyz (lo > hi when no primitive uses material m; infinite when a
// plane, or a primitive of unknown kind, does). Primitives whose material
// index is out of range are left out: a shadow walk treats them as opaque.
inline std::vector<double> material_boxes(const mh_scene_soa* soa)
{
    const int nm = soa->n_materials > 0 ? soa->n_materials : 0;
    const double inf = std::numeric_limits<double>::infinity();
    std::vector<double> box((size_t)nm * 6);
    for (int m = 0; m < nm; ++m)
        for (int a = 0; a < 3; ++a) {
            box[6 * m + a] = inf;
            box[6 * m + 3 + a] = -inf;
        }
    for (int i = 0; i < soa->n_prims; ++i) {
        const mh_prim& p = soa->prims[i];
        if (p.material < 0 || p.material >= nm) continue;
        double lo[3], hi[3];
        const double c0[3] = {p.p0.x, p.p0.y, p.p0.z}, c1[3] = {p.p1.x, p.p1.y, p.p1.z},
                     c2[3] = {p.p2.x, p.p2.y, p.p2.z};
        const double rmax = p.r0 > p.r1 ? p.r0 : p.r1;
        for (int a = 0; a < 3; ++a) {
            switch (p.kind) {
            case MH_PRIM_SPHERE:
                lo[a] = c0[a] - std::fabs(p.r0);
                hi[a] = c0[a] + std::fabs(p.r0);
                break;
            case MH_PRIM_BOX:
                lo[a] = c0[a] - std::fabs(c1[a]);
                hi[a] = c0[a] + std::fabs(c1[a]);
                break;
            case MH_PRIM_TRIANGLE:
                lo[a] = std::min(c0[a], std::min(c1[a], c2[a]));
                hi[a] = std::max(c0[a], std::max(c1[a], c2[a]));
                break;
            case MH_PRIM_CYLINDER:
                lo[a] = std::min(c0[a], c1[a]) - std::fabs(rmax);
                hi[a] = std::max(c0[a], c1[a]) + std::fabs(rmax);
                break;
            default:  // planes: no finite box
                lo[a] = -inf;
                hi[a] = inf;
            }
            // NaN coordinates: no finite box either
            if (!(lo[a] <= hi[a])) {
                lo[a] = -inf;
                hi[a] = inf;
            }
            double* b = &box[6 * (size_t)p.material];
            b[a] = std::min(b[a], lo[a]);
            b[3 + a] = std::max(b[3 + a], hi[a]);
        }
    }
    return box;
}

// Whether the shadow kernels should split their queue by the transmissive
// box (SceneView::trans_split): the box's surface area, the chance that a
// segment through the scene meets it, is under a quarter of the area of the
// box around every finite primitive, or no primitive is transmissive. On
// scenes with transmissive primitives all over (glass spheres scattered
// through the scene) nearly every ray would be listed for the closest-hit
// steps, and the extra any-hit pass over the queue would cost more than it
// saves.
inline int trans_split_pays(const std::vector<double>& box, const float tlo[3], const float thi[3])
{
    if (!(tlo[0] <= thi[0] && tlo[1] <= thi[1] && tlo[2] <= thi[2])) return 1;
    const float big = 1e30f;
    for (int a = 0; a < 3; ++a)
        if (!(tlo[a] > -big && thi[a] < big)) return 0;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t m = 0; 6 * m + 5 < box.size(); ++m) {
        bool finite = true;
        for (int a = 0; a < 3; ++a)
            finite = finite && std::fabs(box[6 * m + a]) < 1e30 && std::fabs(box[6 * m + 3 + a]) < 1e30 &&
                     box[6 * m + a] <= box[6 * m + 3 + a];
        if (!finite) continue;
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], box[6 * m + a]);
            hi[a] = std::max(hi[a], box[6 * m + 3 + a]);
        }
    }
    auto area = [](const double* l, const double* h) {
        const double dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
        return dx * dy + dy * dz + dz * dx;
    };
    const double tl[3] = {tlo[0], tlo[1], tlo[2]}, th[3] = {thi[0], thi[1], thi[2]};
    return area(tl, th) < 0.25 * area(lo, hi) ? 1 : 0;
}

// The box around every primitive of a transmissive material (see
// SceneView::tlo), rounded outwards to fp32.
inline void transmissive_box(const std::vector<double>& box, const mh_material* mats, int n, float lo[3],
                             float hi[3])
{
    const float inf = std::numeric_limits<float>::infinity();
    for (int a = 0; a < 3; ++a) {
        lo[a] = inf;
        hi[a] = -inf;
    }
    for (int m = 0; m < n && 6 * (size_t)m + 5 < box.size(); ++m) {
        if (!transmissive(mats[m])) continue;
        for (int a = 0; a < 3; ++a) {
            const double l = box[6 * (size_t)m + a], h = box[6 * (size_t)m + 3 + a];
            if (!(l <= h)) continue;
            float fl = (float)l, fh = (float)h;
            if ((double)fl > l) fl = std::nextafter(fl, -inf);
            if ((double)fh < h) fh = std::nextafter(fh, inf);
            lo[a] = std::min(lo[a], fl);
            hi[a] = std::max(hi[a], fh);
        }
    }
}

// Pointers of a scene's arrays (device memory, or host memory in a host
// build) and the roots the kernels start from.
struct Layout {
    const void *nodes = nullptr, *nodes4 = nullptr;
    int n_nodes = 0, root_ref = INT_MIN, root4 = INT_MIN;
    bool bvh4 = false;
    const void *p32 = nullptr, *p64 = nullptr, *pl32 = nullptr, *pl64 = nullptr, *plidx = nullptr;
    const void *a32 = nullptr, *a64 = nullptr;
    const void *m32 = nullptr, *m64 = nullptr;
    int n_planes = 0;
};

// The arrays of a scene with a host-built tree (mh_scene_soa nodes): interior
// nodes numbered in the depth-first order check_tree gives, the 4-wide
// collapse when its depth fits the traversal stack, leaf-ordered records.
struct HostLayout {
    std::vector<DNode> nodes;
    std::vector<DNode4> nodes4;
    int root_ref = INT_MIN, root4 = INT_MIN;
    bool bvh4 = false;
    std::vector<P32> p32, pl32;
    std::vector<P64> p64, pl64;
    std::vector<DAttr<float>> a32;
    std::vector<DAttr<double>> a64;
    std::vector<DMat<float>> m32;
    std::vector<DMat<double>> m64;
};

// Fills `out` from a validated scene (returns check_tree's code on a
// malformed tree).
inline int build_host_layout(const mh_scene_soa* soa, HostLayout* out)
{
    std::vector<int> inner;
    int rc = check_tree(soa, &inner);
    if (rc) return rc;
    // device node numbering: interior nodes in DFS order
    std::vector<int> dev_index(soa->n_nodes, -1);
    for (size_t i = 0; i < inner.size(); ++i) dev_index[inner[i]] = (int)i;
    auto ref_of = [&](int n) {
        const mh_bvh_node& nd = soa->nodes[n];
        if (nd.count > 0) return ~((nd.first << 3) | (nd.count - 1));
        return dev_index[n];
    };
    out->nodes.assign(inner.size(), DNode{});
    for (size_t i = 0; i < inner.size(); ++i) {
        const mh_bvh_node& nd = soa->nodes[inner[i]];
        const mh_bvh_node& L = soa->nodes[nd.left];
        const mh_bvh_node& Rn = soa->nodes[nd.right];
        DNode d;
        d.c0 = make_float4(L.bmin[0], L.bmax[0], L.bmin[1], L.bmax[1]);
        d.c1 = make_float4(L.bmin[2], L.bmax[2], Rn.bmin[0], Rn.bmax[0]);
        d.c2 = make_float4(Rn.bmin[1], Rn.bmax[1], Rn.bmin[2], Rn.bmax[2]);
        d.c3 = make_int4(ref_of(nd.left), ref_of(nd.right), 0, 0);
        out->nodes[i] = d;
    }
    out->root_ref = soa->n_nodes > 0 ? ref_of(0) : INT_MIN;
    Bvh4Builder b4{soa, {}, 0};
    out->root4 = out->root_ref;
    if (soa->n_nodes > 0 && soa->nodes[0].count == 0) out->root4 = b4.make(0, 1);
    // 4-wide traversal pushes up to 3 entries per level; the shadow walk
    // addresses 4-wide nodes with 32-bit byte offsets (below 2^25 nodes)
    bool use4 = soa->n_nodes > 0 && 3 * b4.max_depth + 1 <= MH_STACK && b4.nodes.size() < ((size_t)1 << 25);
    if (!use4) {
        b4.nodes.clear();
        out->root4 = out->root_ref;
    }
    out->bvh4 = use4 && !b4.nodes.empty();
    out->nodes4.swap(b4.nodes);

    bool any_attr = has_attrs(soa);
    out->p32.assign(soa->n_leaf_prims, P32{});
    out->p64.assign(soa->n_leaf_prims, P64{});
    out->a32.clear();
    out->a64.clear();
    if (any_attr) {
        out->a32.resize(soa->n_leaf_prims);
        out->a64.resize(soa->n_leaf_prims);
    }
    for (int i = 0; i < soa->n_leaf_prims; ++i) {
        int pi = soa->leaf_prims[i];
        mh_prim q = soa->prims[pi];
        if (!soa->attrs) q.flags = 0;
        out->p32[i] = pack32(q);
        set_prim_index32(out->p32[i], pi);
        out->p64[i] = pack64(q, pi);
        if (any_attr) {
            out->a32[i] = conv_attr<float>(soa->attrs[pi]);
            out->a64[i] = conv_attr<double>(soa->attrs[pi]);
        }
    }
    out->pl32.assign(soa->n_planes, P32{});
    out->pl64.assign(soa->n_planes, P64{});
    for (int i = 0; i < soa->n_planes; ++i) {
        int pi = soa->planes[i];
        out->pl32[i] = pack32(soa->prims[pi]);
        set_prim_index32(out->pl32[i], pi);
        out->pl64[i] = pack64(soa->prims[pi], pi);
    }
    out->m32.resize(soa->n_materials);
    out->m64.resize(soa->n_materials);
    for (int i = 0; i < soa->n_materials; ++i) {
        out->m32[i] = conv_mat<float>(soa->materials[i]);
        out->m64[i] = conv_mat<double>(soa->materials[i]);
    }
    return 0;
}

// The view of one precision: arrays from `lay`, tables from the scene.
template <typename R>
inline void fill_view(SceneView<R>& v, const Layout& lay, const mh_scene_soa* soa, int has_water,
                      const DLight<R>* lights, int nl, int has_transmissive, const float tlo[3],
                      const float thi[3], int trans_split)
{
    const bool f = sizeof(R) == 4;
    v.nodes = (const DNode*)lay.nodes;
    v.n_nodes = lay.n_nodes;
    v.root_ref = lay.root_ref;
    // no 4-wide tree in the view unless it is usable: walks that pick the
    // tree by its pointer (the shadow kernels) would pair it with root4 =
    // the binary root
    v.nodes4 = lay.bvh4 ? (const DNode4*)lay.nodes4 : nullptr;
    v.root4 = lay.root4;
    v.bvh4 = lay.bvh4 ? 1 : 0;
    v.prims = (const typename PrimT<R>::type*)(f ? lay.p32 : lay.p64);
    v.planes = (const typename PrimT<R>::type*)(f ? lay.pl32 : lay.pl64);
    v.n_planes = lay.n_planes;
    v.plane_index = (const int*)lay.plidx;
    v.attrs = (const DAttr<R>*)(f ? lay.a32 : lay.a64);
    v.mats = (const DMat<R>*)(f ? lay.m32 : lay.m64);
    v.n_mats = soa->n_materials;
    v.sky = conv_sky<R>(soa->sky);
    v.fog = conv_fog<R>(soa->fog);
    v.ocean = conv_ocean<R>(soa->ocean);
    v.fog_kernel = needs_fog_kernel(soa->fog, has_water);
    v.water_level = (R)soa->water_level;
    v.water_material = soa->water_material;
    v.has_transmissive = has_transmissive;
    v.tlo = V3<R>{(R)tlo[0], (R)tlo[1], (R)tlo[2]};
    v.thi = V3<R>{(R)thi[0], (R)thi[1], (R)thi[2]};
    v.trans_split = trans_split;
    v.n_lights = nl;
    for (int i = 0; i < 8; ++i) v.lights[i] = lights[i];
}

} // namespace mh
