// mh_scene_dev.h - device-side scene layout of libmarmel_hip.
//
// HBM layout (one upload in mh_scene_create):
//   nodes   BVH2, 64 B per node: the two CHILD boxes of the node as fp32
//           (rounded outward from the host's double bounds) + two child refs.
//           A child ref >= 0 is a node index; a ref < 0 is a leaf,
//           ~ref = (first << 3) | (count - 1) into the leaf-ordered records.
//   prims32 leaf-ordered primitive records, 48 B (fp32 throughput path)
//   prims64 leaf-ordered primitive records, 104 B (fp64 parity path)
//   planes  plane side list (not in the tree), fp32 and fp64 records
//   attrs   leaf-ordered smooth normals / uvs (only when a scene has them)
//   mats    material table (fp32 and fp64)
// Sky, lights and water metadata travel in the kernel argument block.
#pragma once

#include "mh_math.h"

namespace mh {

enum { K_SPHERE = 0, K_PLANE = 1, K_BOX = 2, K_TRIANGLE = 3, K_CYLINDER = 4 };

// Traversal stack entries per lane: the host accepts binary trees of depth
// <= 60 and uses a 4-wide tree only when 3 * its depth + 1 fits.
#ifndef MH_STACK
#define MH_STACK 64
#endif
enum { F_SMOOTH = 1, F_UV = 2 };

struct __attribute__((aligned(16))) DNode {
    float4 c0; // left  (minx, maxx, miny, maxy)
    float4 c1; // left  (minz, maxz) right (minx, maxx)
    float4 c2; // right (miny, maxy, minz, maxz)
    int4 c3;   // (left ref, right ref, -, -)
};

// BVH4 node, 128 B: the boxes of up to four children in SoA form (child k in
// component k of each float4) + their refs (same encoding as DNode) and the
// child count. Collapsed on the host from the binary tree; the children of a
// node keep the binary tree's left-to-right order.
struct __attribute__((aligned(16))) DNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;   // INT_MIN in the slots past the child count
    int4 meta;  // x = child count (2..4)
};

// The first 112 B of a DNode4 (boxes and refs), for walks that tell the
// unused slots by their INT_MIN ref and need no count.
struct __attribute__((aligned(16))) DNode4Box {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;
};

// fp32 record: a.w = kind | flags << 8, b.w = prim index, c.w = material
//   TRIANGLE a = v0, b = v1 - v0, c = v2 - v0
//   SPHERE   a = centre, b.x = radius
//   BOX      a = centre, b = half extents
//   CYLINDER a = base, b = top, c.x = r0, c.y = r1
//   PLANE    a = point, b = unit normal
struct __attribute__((aligned(16))) P32 {
    float4 a, b, c;
};

// fp64 record: the host's double data verbatim.
struct P64 {
    double p0[3], p1[3], p2[3];
    double r0, r1;
    int32_t kind, prim, material, flags;
};

template <typename R> struct DAttr {
    V3<R> n0, n1, n2, uv0, uv1, uv2;
};

template <typename R> struct DMat {
    V3<R> albedo, specular;
    R shininess, reflectivity, transparency, ior;
    int is_water, beer_lambert, pbr, texture_kind;
    V3<R> absorption, deep_color;
    R metallic, roughness;
    V3<R> emissive;
    R texture_scale;
    V3<R> tex_a, tex_b, glow;
    int has_emission;   // emissive != 0
    int has_glow;       // |glow|^2 > 1e-6
};

template <typename R> struct DSky {
    V3<R> sun_dir, sun_color, horizon, zenith;
    R gamma, glow_exp, glow_strength;
    R cloud_height, cloud_scale, coverage, softness, sharpness;
    int octaves;
    uint32_t seed;
    R sun_radius;      // degrees
    R sun_omega;       // solid angle used by sun NEE (1 for a point sun)
};

// Atmospheric fog. density > 0 turns it on; steps > 0 marches camera-ray
// segments (and underwater segments, fog or not) with sun shadow rays.
template <typename R> struct DFog {
    R density;
    V3<R> color;
    R height;
    R falloff;        // as given (the closed form uses max(falloff, 0))
    R inscatter;
    R g;              // Henyey-Greenstein anisotropy clamped to [-0.95, 0.95]
    R noise_scale, noise_amount;
    int steps;        // shadow_steps as given
    int noise;        // noise_amount > 1e-4 && noise_scale > 1e-4
};

// Gerstner ocean surface for the water-height query: base level y0 plus
// four harmonics A cos(k (dx x + dz z) + phi), precomputed on the host.
template <typename R> struct DOcean {
    int on;
    R y0;
    R dx[4], dz[4], k[4], phi[4], A[4];
};

template <typename R> struct DLight {
    int prim;
    int mat;
    V3<R> center;
    R radius;
    V3<R> emissive;
};

template <typename R> struct PrimT;
template <> struct PrimT<float> { typedef P32 type; };
template <> struct PrimT<double> { typedef P64 type; };

template <typename R> struct SceneView {
    const DNode* nodes;
    int n_nodes;
    int root_ref;                  // ref of the tree root, or -2^31 when no tree
    const DNode4* nodes4;          // the same tree collapsed to 4-wide nodes
    int root4;                     // its root ref (a leaf ref when the tree is one leaf)
    int bvh4;                      // traverse nodes4 (1) or nodes (0)
    const typename PrimT<R>::type* prims;
    const typename PrimT<R>::type* planes;
    int n_planes;
    const int* plane_index;        // prim index of each plane
    const DAttr<R>* attrs;         // leaf order, may be null
    const DMat<R>* mats;
    int n_mats;
    DSky<R> sky;
    DFog<R> fog;
    DOcean<R> ocean;
    // 1 when fog is on or marched underwater segments can occur: renders
    // launch the kernel variants that carry the fog code
    int fog_kernel;
    R water_level;
    int water_material;
    // 0 when no material is water or has transparency > 0.8: a shadow
    // segment is then blocked exactly when anything lies on it (any-hit)
    int has_transmissive;
    // box around every primitive of such a material (tlo > thi: none;
    // unbounded when a plane has one): a shadow segment that misses it meets
    // opaque surfaces only, and any-hit answers it too
    V3<R> tlo, thi;
    // 1 when the wavefront shadow kernels split their queue by that box
    // (it is small against the scene); 0: every shadow ray of a transmissive
    // scene takes the closest-hit steps
    int trans_split;
    int n_lights;
    DLight<R> lights[8];
};

template <typename R> struct CamView {
    V3<R> pos, ll, hor, ver;
    R aperture, focus, dome, squeeze;
    V3<R> right_n, up_n; // normalized horizontal / vertical
};

} // namespace mh
