// mh_adjoint.hip - path-replay adjoint of the fp32 path tracer.
//
// Gradient of a scalar image loss with respect to every material's albedo
// (RGB), roughness, metallic and emissive (RGB), and the sky's sun colour
// (RGB), given dLoss/d(pixel mean) for every pixel and channel.
//
// Path replay (no path storage): for each pixel sample the path is traced
// twice from its seed (the same pure function of pixel, sample and bounce as
// the forward render). Pass 1 gives the path radiance L. Pass 2 walks the same
// path again and, at every vertex,
//   * terms added to L at the vertex (emission, NEE to the lights and the sun,
//     sky at a miss) contribute delta . dTerm/dtheta directly,
//   * the BSDF weight w of the sampled lobe multiplies everything the path
//     adds after the vertex, L_after = L - (terms added so far), so it
//     contributes delta . L_after * (dw/dtheta) / w, channel by channel.
// Sampling decisions (lobe choice, GGX half vector, Russian roulette) are
// held fixed: their probabilities are treated as constants ("detached"),
// which differentiates the expectation of the estimator. Visibility is not
// differentiated (no parameter moves geometry). Parameters with no effect on
// the estimator's weights get no gradient: roughness / metallic of legacy
// (non-PBR) materials, and a material's emissive through NEE when it was not
// one of the scene's emissive lights. Textured albedo is the material albedo
// times a texture colour, so its gradient carries that factor.
//
// Built without FMA contraction so that both passes round identically and
// take identical decisions.
// box hits at o + t d in fp32 here (see mh_trace.h): the adjoint kernels sit at
// the register ceiling, where the snap halved the two-pass kernel's occupancy;
// for the same reason the fp32 slab setup runs without its tiny-component guard
// and the closest-hit walk loads whole 4-wide nodes (the near / far variant
// takes the recorded adjoint from 2 waves per SIMD to 1)
#define MH_SNAP_BOX 0
#define MH_SLAB_GUARD 0
#define MH_TRACE_NF 0
#include "mh_kernels.h"
#include "mh_render.h"

namespace mh {

// Gradient slots of one material: albedo rgb, roughness, metallic, emissive rgb.
enum { G_ALB = 0, G_ROUGH = 3, G_MET = 4, G_EM = 5, G_PER_MAT = 8 };

// d(sky radiance)/d(sun colour), per channel (sky_sample is affine in it).
__device__ V3<float> sky_dsun(V3<float> dir, const DSky<float>& sky)
{
    float h = clamp01m(dir.y);
    float t = powpos(h, sky.gamma);
    V3<float> base = lerp(sky.horizon, sky.zenith, t);
    float cs = dot(dir, sky.sun_dir);
    float g = 0.f;
    if (cs > 0.f) {
        g = powpos(cs, sky.glow_exp) * sky.glow_strength;
        base = base + sky.sun_color * g;
    }
    V3<float> dv = mk(g, g, g);
    if (dir.y > 0.03f) {
        float tt = sky.cloud_height / dir.y;
        float px = dir.x * tt, pz = dir.z * tt;
        float raw = fbm2<float>(px * sky.cloud_scale, pz * sky.cloud_scale, sky.octaves, sky.seed);
        float dens = clamp01m(0.5f * (raw + 1.f));
        float al = powpos(smooth3(sky.coverage - sky.softness, sky.coverage + sky.softness, dens), sky.sharpness);
        al *= smooth3(0.f, 0.15f, dir.y);
        if (al > 0.f) {
            V3<float> cc = mk(0.95f, 0.95f, 0.98f) * (0.6f + 0.4f * al);
            float lit = cs > 0.f ? cs : 0.f;
            V3<float> lc = cc + vmul(cc, sky.sun_color) * (0.35f * lit);
            base = lerp(base, lc, al);
            dv = dv * (1.f - al) + cc * (al * 0.35f * lit);
        }
    }
    // sky_sample clamps negative channels to 0
    if (!(base.x >= 0.f)) dv.x = 0.f;
    if (!(base.y >= 0.f)) dv.y = 0.f;
    if (!(base.z >= 0.f)) dv.z = 0.f;
    return dv;
}

// d ln D_GGX / d alpha at n.h.
__device__ __forceinline__ float dlnD_dalpha(float ndh, float a)
{
    float d = ndh * ndh * (a * a - 1.f) + 1.f;
    return 2.f / a - 4.f * a * ndh * ndh / d;
}

// d ln G1 / d k for G1(x) = x / (x (1 - k) + k).
__device__ __forceinline__ float dlnG1_dk(float x, float k) { return -(1.f - x) / (x * (1.f - k) + k); }

// d alpha / d roughness for alpha = max(clamp(r)^2, 1e-4).
__device__ __forceinline__ float dalpha_dr(float rough)
{
    if (!(rough > 0.f) || rough >= 1.f) return 0.f;
    return rough * rough > 1e-4f ? 2.f * rough : 0.f;
}

__device__ __forceinline__ float in01(float v) { return v > 0.f && v < 1.f ? 1.f : 0.f; }

// d shade(...)/d theta of the NEE response for a light colour `light`:
// albedo (per channel), roughness and metallic (per channel vectors).
__device__ void shade_grad(const DMat<float>& m, V3<float> alb, V3<float> N, V3<float> L, V3<float> V,
                           V3<float> light, V3<float>* d_alb, V3<float>* d_rough, V3<float>* d_met)
{
    *d_alb = *d_rough = *d_met = zero3<float>();
    float ndl = dot(N, L);
    if (ndl <= 0.f) return;
    if (!m.pbr) {
        // Blinn-Phong: albedo * light * ndl + specular term (no parameter)
        *d_alb = light * ndl;
        return;
    }
    float ndv = dot(N, V);
    if (ndv <= 0.f) return;
    float met = clamp01m(m.metallic);
    float rr = clamp01m(m.roughness);
    float alpha = rough_alpha(m.roughness);
    V3<float> hv = L + V;
    float ndh = 0.f;
    if (len2(hv) > 1e-12f) {
        ndh = dot(N, normalize(hv));
        if (ndh < 0.f) ndh = 0.f;
    }
    V3<float> f0 = lerp(mk(0.04f, 0.04f, 0.04f), alb, met);
    float c = clamp01m(ndv), q = 1.f - c, q2 = q * q, k5 = q2 * q2 * q;
    V3<float> F = mk(f0.x + (1.f - f0.x) * k5, f0.y + (1.f - f0.y) * k5, f0.z + (1.f - f0.z) * k5);
    float D = d_ggx(ndh, alpha);
    float k = (rr + 1.f) * (rr + 1.f) / 8.f;
    if (k < 0.f) k = 0.f;
    else if (k > 1.f) k = 1.f;
    float G = (ndl / (ndl * (1.f - k) + k)) * (ndv / (ndv * (1.f - k) + k));
    float den = 4.f * ndl * ndv;
    float ss = den > 0.f ? D * G / den : 0.f;
    const float ip = 1.f / Consts<float>::pi;
    float dm = in01(m.metallic);
    float dF0a = met * (1.f - k5);  // dF/dalbedo (per channel)
    float a[3] = {alb.x, alb.y, alb.z}, Fc[3] = {F.x, F.y, F.z}, f0c[3] = {f0.x, f0.y, f0.z},
          lc[3] = {light.x, light.y, light.z};
    float ga[3], gm[3], gr[3];
    float dlnss = ss > 0.f ? dlnD_dalpha(ndh, alpha) * dalpha_dr(rr) +
                                 (dlnG1_dk(ndl, k) + dlnG1_dk(ndv, k)) * (rr + 1.f) * 0.25f * in01(m.roughness)
                           : 0.f;
    (void)f0c;
    for (int i = 0; i < 3; ++i) {
        float dFa = dF0a;  // shade_pbr's Schlick term does not clamp f0
        float dFm = (a[i] - 0.04f) * (1.f - k5) * dm;
        ga[i] = lc[i] * ndl * (dFa * ss + (1.f - met) * ip * ((1.f - Fc[i]) - a[i] * dFa));
        gm[i] = lc[i] * ndl * (dFm * ss + a[i] * ip * (-dFm * (1.f - met) - (1.f - Fc[i]) * dm));
        gr[i] = lc[i] * ndl * Fc[i] * ss * dlnss;
    }
    *d_alb = mk(ga[0], ga[1], ga[2]);
    *d_met = mk(gm[0], gm[1], gm[2]);
    *d_rough = mk(gr[0], gr[1], gr[2]);
}

// d(shaded albedo)/d(material albedo): 1 untextured, the texture colour
// (shaded / material albedo) for the procedural textures.
__device__ __forceinline__ V3<float> tex_factor(const DMat<float>& m, V3<float> alb)
{
    if (m.texture_kind == 0) return mk(1.f, 1.f, 1.f);
    return mk(m.albedo.x != 0.f ? alb.x / m.albedo.x : 0.f, m.albedo.y != 0.f ? alb.y / m.albedo.y : 0.f,
              m.albedo.z != 0.f ? alb.z / m.albedo.z : 0.f);
}

// Records the raw sum of a path's radiance terms.
struct SumHook : NoHook {
    V3<float> L = zero3<float>();
    __device__ void add(V3<float> t) { L = L + t; }
    __device__ void sky(V3<float>, V3<float>, V3<float> t) { add(t); }
    __device__ void emission(V3<float>, int, V3<float> t) { add(t); }
    __device__ void nee_light(const DMat<float>&, V3<float>, V3<float>, V3<float>, V3<float>, V3<float>, int,
                              V3<float>, float, V3<float> t)
    {
        add(t);
    }
    __device__ void nee_sun(const DMat<float>&, V3<float>, V3<float>, V3<float>, V3<float>, V3<float>, V3<float>,
                            float, V3<float> t)
    {
        add(t);
    }
};

struct AdjParams {
    const float* dimg;   // dLoss/d(pixel mean), w*h*3 (device)
    double* grad;        // n_mats*8 + 3 (device, accumulated)
    int n_mats;
    int lds_mats;        // materials whose slots live in LDS (the rest go straight to grad)
    const float4* rec;   // recorded per-sample sums (k_render_record), or null
};

#define MH_ADJ_LDS_MATS 256

struct AdjHook : NoHook {
    const SceneView<float>* sc;
    V3<float> Lt, delta, Lacc;
    float* g;            // LDS slots (materials < lds_mats, then 3 sun slots at the end)
    double* gg;          // global slots
    int lds_mats, n_mats;

    __device__ void put(int slot_mat, int k, float v)
    {
        if (v == 0.f || !(v == v)) return;
        if (slot_mat < lds_mats) atomicAdd(&g[slot_mat * G_PER_MAT + k], v);
        else atomicAdd(&gg[(size_t)slot_mat * G_PER_MAT + k], (double)v);
    }
    __device__ void put_sun(int c, float v)
    {
        if (v == 0.f || !(v == v)) return;
        atomicAdd(&g[MH_ADJ_LDS_MATS * G_PER_MAT + c], v);
    }
    static constexpr bool kGrad = true;
    __device__ void add(V3<float> t) { Lacc = Lacc + t; }
    // fog / marched-water in-scatter: w = d(term)/d(sun colour), per channel
    __device__ void sun_lin(V3<float> w)
    {
        put_sun(0, delta.x * w.x);
        put_sun(1, delta.y * w.y);
        put_sun(2, delta.z * w.z);
    }
    __device__ void sky(V3<float> beta, V3<float> d, V3<float> t)
    {
        V3<float> ds = sky_dsun(d, sc->sky);
        put_sun(0, delta.x * beta.x * ds.x);
        put_sun(1, delta.y * beta.y * ds.y);
        put_sun(2, delta.z * beta.z * ds.z);
        add(t);
    }
    __device__ void emission(V3<float> beta, int mat, V3<float> t)
    {
        put(mat, G_EM + 0, delta.x * beta.x);
        put(mat, G_EM + 1, delta.y * beta.y);
        put(mat, G_EM + 2, delta.z * beta.z);
        add(t);
    }
    // direct derivative of a NEE term beta * shade(light) * omega w.r.t. the
    // shaded material's parameters
    __device__ void surface(const DMat<float>& m, int mat, V3<float> alb, V3<float> N, V3<float> wi, V3<float> V,
                            V3<float> beta, V3<float> light, float omega)
    {
        V3<float> da, dr, dm;
        shade_grad(m, alb, N, wi, V, light, &da, &dr, &dm);
        da = vmul(da, tex_factor(m, alb));
        V3<float> w = vmul(delta, beta) * omega;
        put(mat, G_ALB + 0, w.x * da.x);
        put(mat, G_ALB + 1, w.y * da.y);
        put(mat, G_ALB + 2, w.z * da.z);
        put(mat, G_ROUGH, w.x * dr.x + w.y * dr.y + w.z * dr.z);
        put(mat, G_MET, w.x * dm.x + w.y * dm.y + w.z * dm.z);
    }
    // index of a material given by reference into the scene's table
    __device__ int mat_of(const DMat<float>& m) const { return (int)(&m - sc->mats); }
    __device__ void nee_light(const DMat<float>& m, V3<float> alb, V3<float> N, V3<float> wi, V3<float> V,
                              V3<float> beta, int li, V3<float> att, float omega, V3<float> t)
    {
        const DLight<float>& lt = sc->lights[li];
        V3<float> K = shade(m, alb, N, wi, V, att);  // response per unit emission
        V3<float> w = vmul(delta, beta) * omega;
        put(lt.mat, G_EM + 0, w.x * K.x);
        put(lt.mat, G_EM + 1, w.y * K.y);
        put(lt.mat, G_EM + 2, w.z * K.z);
        surface(m, mat_of(m), alb, N, wi, V, beta, vmul(lt.emissive, att), omega);
        add(t);
    }
    __device__ void nee_sun(const DMat<float>& m, V3<float> alb, V3<float> N, V3<float> Ls, V3<float> V,
                            V3<float> beta, V3<float> att, float omega, V3<float> t)
    {
        V3<float> K = shade(m, alb, N, Ls, V, att);
        V3<float> w = vmul(delta, beta) * omega;
        put_sun(0, w.x * K.x);
        put_sun(1, w.y * K.y);
        put_sun(2, w.z * K.z);
        surface(m, mat_of(m), alb, N, Ls, V, beta, vmul(sc->sky.sun_color, att), omega);
        add(t);
    }
    // the sampled lobe's weight multiplies every term after this vertex
    __device__ void bsdf(int mat, const DMat<float>& m, V3<float> alb, V3<float> N, V3<float> V, const Lobe<float>& lb)
    {
        V3<float> La = Lt - Lacc;
        V3<float> s = vmul(delta, La);  // delta . L_after per channel
        float wv[3] = {lb.w.x, lb.w.y, lb.w.z}, sv[3] = {s.x, s.y, s.z}, a[3] = {alb.x, alb.y, alb.z};
        V3<float> tf = tex_factor(m, alb);
        float tfv[3] = {tf.x, tf.y, tf.z};
        if (!m.pbr) {
            // legacy Lambert lobe: w = albedo * const
            if (lb.spec || lb.refr) return;
            for (int i = 0; i < 3; ++i)
                if (wv[i] != 0.f && a[i] != 0.f) put(mat, G_ALB + i, sv[i] / a[i] * tfv[i]);
            return;
        }
        float met = clamp01n(m.metallic), rough = clamp01n(m.roughness);
        float dm = in01(m.metallic);
        V3<float> f0 = lerp(mk(0.04f, 0.04f, 0.04f), alb, met);
        float f0c[3] = {f0.x, f0.y, f0.z};
        float ndv = dot(N, V);
        if (lb.spec) {
            V3<float> H = normalize(V + lb.wi);
            float vdh = dot(V, H), ndh = dot(N, H), ndl = dot(N, lb.wi);
            float c = clamp01n(vdh), q = 1.f - c, q2 = q * q, k5 = q2 * q2 * q;
            float alpha = rough_alpha(rough);
            float k = clamp01n((rough + 1.f) * (rough + 1.f) / 8.f);
            float dlnw_r = dlnD_dalpha(ndh, alpha) * dalpha_dr(rough) +
                           (dlnG1_dk(ndl, k) + dlnG1_dk(ndv, k)) * (rough + 1.f) * 0.25f;
            float gr = 0.f, gm = 0.f;
            for (int i = 0; i < 3; ++i) {
                if (wv[i] == 0.f) continue;
                float f = clamp01n(f0c[i]);
                float F = f + (1.f - f) * k5;
                float fin = in01(f0c[i]);
                put(mat, G_ALB + i, sv[i] * met * (1.f - k5) * fin / F * tfv[i]);
                gm += sv[i] * (a[i] - 0.04f) * (1.f - k5) * fin * dm / F;
                gr += sv[i] * dlnw_r;
            }
            put(mat, G_MET, gm);
            put(mat, G_ROUGH, gr * in01(m.roughness));
        } else {
            float c = clamp01n(ndv), q = 1.f - c, q2 = q * q, k5 = q2 * q2 * q;
            float gm = 0.f;
            for (int i = 0; i < 3; ++i) {
                if (wv[i] == 0.f) continue;
                float f = clamp01n(f0c[i]);
                float F = f + (1.f - f) * k5;
                float fin = in01(f0c[i]);
                if (a[i] != 0.f) put(mat, G_ALB + i, sv[i] * (1.f / a[i] - met * (1.f - k5) * fin / (1.f - F)) * tfv[i]);
                gm += sv[i] * (-(a[i] - 0.04f) * (1.f - k5) * fin * dm / (1.f - F) - dm / (1.f - met));
            }
            put(mat, G_MET, gm);
        }
    }
};

// One thread per pixel (16x16 tiles), samples in order; each sample is
// traced twice (primal, then replay with gradients), or, with REC, once: the
// primal sum comes from the record of a recording render (k_render_record,
// the same code, so the same bits).
template <bool FOG, bool REC>
__global__ __launch_bounds__(256) void k_adjoint(SceneView<float> sc, CamView<float> cam, int w, int h, int spp,
                                                  int s0, int depth, int tiles_x, int ntiles, AdjParams ap)
{
    __shared__ float gl[MH_ADJ_LDS_MATS * G_PER_MAT + 3];
    MH_STACK_PRIV(ts)
    for (int i = threadIdx.x; i < MH_ADJ_LDS_MATS * G_PER_MAT + 3; i += blockDim.x) gl[i] = 0.f;
    __syncthreads();
    int tile = blockIdx.x;
    int nq = 0;
    if (tile < ntiles) {
        int tx = tile % tiles_x, ty = tile / tiles_x;
        int x = tx * 16 + (threadIdx.x & 15);
        int y = ty * 16 + (threadIdx.x >> 4);
        if (x < w && y < h) {
            size_t p = ((size_t)y * (size_t)w + (size_t)x) * 3;
            float inv = 1.f / (float)spp;
            V3<float> dpix = mk(ap.dimg[p] * inv, ap.dimg[p + 1] * inv, ap.dimg[p + 2] * inv);
            AdjHook hk;
            hk.sc = &sc;
            hk.g = gl;
            hk.gg = ap.grad;
            hk.lds_mats = ap.lds_mats;
            hk.n_mats = ap.n_mats;
            for (int si = 0; si < spp; ++si) {
                uint32_t s = (uint32_t)(s0 + si);
                float jx = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 0u);
                float jy = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 1u);
                float l1 = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 100u);
                float l2 = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 101u);
                float uu = ((float)x + jx) / (float)w;
                float vv = 1.f - ((float)y + jy) / (float)h;
                V3<float> o, d;
                camera_ray(cam, uu, vv, l1, l2, &o, &d);
                if (cam.dome > 1e-4f && sc.water_material >= 0) {
                    V3<float> pd = o + d * cam.dome;
                    if (pd.y < water_height(sc, pd.x, pd.z)) o = pd - mk(0.f, 1e-3f, 0.f);
                }
                uint32_t seed = hash3((uint32_t)x, (uint32_t)y, s);
                // pass 1: the raw (unclamped) sum of the path's terms
                V3<float> L;
                if (REC) {
                    float4 r = ap.rec[(size_t)si * (size_t)w * (size_t)h + (size_t)y * (size_t)w + (size_t)x];
                    L = mk(r.x, r.y, r.z);
                } else {
                    SumHook sum;
                    (void)pt_radiance_h<FOG>(sc, ts, o, d, depth, seed, &nq, sum);
                    L = sum.L;
                }
                // the final clamp of a negative channel to 0 (and of a
                // non-finite sum) has zero derivative
                if (!finite3(L)) continue;
                hk.Lt = L;
                hk.Lacc = zero3<float>();
                hk.delta = mk(L.x >= 0.f ? dpix.x : 0.f, L.y >= 0.f ? dpix.y : 0.f, L.z >= 0.f ? dpix.z : 0.f);
                // pass 2: replay with gradients
                (void)pt_radiance_h<FOG>(sc, ts, o, d, depth, seed, &nq, hk);
            }
        }
    }
    __syncthreads();
    int nm = ap.n_mats < ap.lds_mats ? ap.n_mats : ap.lds_mats;
    for (int i = threadIdx.x; i < nm * G_PER_MAT; i += blockDim.x)
        if (gl[i] != 0.f) atomicAdd(&ap.grad[i], (double)gl[i]);
    if (threadIdx.x < 3) {
        float v = gl[MH_ADJ_LDS_MATS * G_PER_MAT + threadIdx.x];
        if (v != 0.f) atomicAdd(&ap.grad[(size_t)ap.n_mats * G_PER_MAT + threadIdx.x], (double)v);
    }
}

// Recording render: the fp32 path tracer in the adjoint's code (same
// flags, same hooks as its first pass), writing each sample's raw term sum
// to rec (one float4 per sample and pixel, sample-major) and the pixel
// means to out.
template <bool FOG>
__global__ __launch_bounds__(256) void k_render_record(SceneView<float> sc, CamView<float> cam, int w, int h, int spp,
                                                        int s0, int depth, int tiles_x, int ntiles, float4* rec,
                                                        LaunchOut out)
{
    MH_STACK_PRIV(ts)
    int tile = blockIdx.x;
    int nq = 0;
    if (tile < ntiles) {
        int tx = tile % tiles_x, ty = tile / tiles_x;
        int x = tx * 16 + (threadIdx.x & 15);
        int y = ty * 16 + (threadIdx.x >> 4);
        if (x < w && y < h) {
            V3<float> acc = zero3<float>();
            for (int si = 0; si < spp; ++si) {
                uint32_t s = (uint32_t)(s0 + si);
                float jx = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 0u);
                float jy = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 1u);
                float l1 = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 100u);
                float l2 = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 101u);
                float uu = ((float)x + jx) / (float)w;
                float vv = 1.f - ((float)y + jy) / (float)h;
                V3<float> o, d;
                camera_ray(cam, uu, vv, l1, l2, &o, &d);
                if (cam.dome > 1e-4f && sc.water_material >= 0) {
                    V3<float> pd = o + d * cam.dome;
                    if (pd.y < water_height(sc, pd.x, pd.z)) o = pd - mk(0.f, 1e-3f, 0.f);
                }
                SumHook sum;
                acc = acc + pt_radiance_h<FOG>(sc, ts, o, d, depth, hash3((uint32_t)x, (uint32_t)y, s), &nq, sum);
                rec[(size_t)si * (size_t)w * (size_t)h + (size_t)y * (size_t)w + (size_t)x] =
                    make_float4(sum.L.x, sum.L.y, sum.L.z, 0.f);
            }
            acc = acc * (1.f / (float)spp);
            size_t p = ((size_t)y * (size_t)w + (size_t)x) * 3;
            out.linear[p] = acc.x;
            out.linear[p + 1] = acc.y;
            out.linear[p + 2] = acc.z;
            out.rgb[p] = to_byte(acc.x);
            out.rgb[p + 1] = to_byte(acc.y);
            out.rgb[p + 2] = to_byte(acc.z);
        }
    }
}

hipError_t launch_render_record(const SceneView<float>& sc, const CamView<float>& cam, int w, int h, int spp, int s0,
                                int depth, float4* rec, const LaunchOut& out, hipStream_t st)
{
    int tiles_x = (w + 15) / 16, tiles_y = (h + 15) / 16;
    int ntiles = tiles_x * tiles_y;
    auto kern = sc.fog_kernel ? k_render_record<true> : k_render_record<false>;
    hipLaunchKernelGGL(kern, dim3(ntiles), dim3(256), 0, st, sc, cam, w, h, spp, s0, depth, tiles_x, ntiles, rec, out);
    return hipGetLastError();
}

// Frozen-sampling render: sampling decisions from the material table `f`.
struct FreezeHook : NoHook {
    const DMat<float>* f;
    template <typename R> __device__ const DMat<R>* frozen(int m) const { return &f[m]; }
};

template <bool FOG>
__global__ __launch_bounds__(256) void k_render_frozen(SceneView<float> sc, CamView<float> cam, int w, int h, int spp,
                                                        int s0, int depth, int tiles_x, int ntiles,
                                                        const DMat<float>* frozen, LaunchOut out)
{
    MH_STACK_PRIV(ts)
    int tile = blockIdx.x;
    int nq = 0;
    if (tile < ntiles) {
        int tx = tile % tiles_x, ty = tile / tiles_x;
        int x = tx * 16 + (threadIdx.x & 15);
        int y = ty * 16 + (threadIdx.x >> 4);
        if (x < w && y < h) {
            FreezeHook fh;
            fh.f = frozen;
            V3<float> acc = zero3<float>();
            for (int si = 0; si < spp; ++si) {
                uint32_t s = (uint32_t)(s0 + si);
                float jx = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 0u);
                float jy = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 1u);
                float l1 = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 100u);
                float l2 = rand01<float>((uint32_t)x, (uint32_t)y, s * 2u + 101u);
                float uu = ((float)x + jx) / (float)w;
                float vv = 1.f - ((float)y + jy) / (float)h;
                V3<float> o, d;
                camera_ray(cam, uu, vv, l1, l2, &o, &d);
                if (cam.dome > 1e-4f && sc.water_material >= 0) {
                    V3<float> pd = o + d * cam.dome;
                    if (pd.y < water_height(sc, pd.x, pd.z)) o = pd - mk(0.f, 1e-3f, 0.f);
                }
                acc = acc + pt_radiance_h<FOG>(sc, ts, o, d, depth, hash3((uint32_t)x, (uint32_t)y, s), &nq, fh);
            }
            acc = acc * (1.f / (float)spp);
            size_t p = ((size_t)y * (size_t)w + (size_t)x) * 3;
            out.linear[p] = acc.x;
            out.linear[p + 1] = acc.y;
            out.linear[p + 2] = acc.z;
            out.rgb[p] = to_byte(acc.x);
            out.rgb[p + 1] = to_byte(acc.y);
            out.rgb[p + 2] = to_byte(acc.z);
        }
    }
}

hipError_t launch_render_frozen(const SceneView<float>& sc, const CamView<float>& cam, int w, int h, int spp, int s0,
                                int depth, const void* frozen, const LaunchOut& out, hipStream_t st)
{
    int tiles_x = (w + 15) / 16, tiles_y = (h + 15) / 16;
    int ntiles = tiles_x * tiles_y;
    auto kern = sc.fog_kernel ? k_render_frozen<true> : k_render_frozen<false>;
    hipLaunchKernelGGL(kern, dim3(ntiles), dim3(256), 0, st, sc, cam, w, h, spp, s0, depth, tiles_x, ntiles,
                       (const DMat<float>*)frozen, out);
    return hipGetLastError();
}

hipError_t launch_adjoint(const SceneView<float>& sc, const CamView<float>& cam, int w, int h, int spp, int s0,
                          int depth, const float* dimg, double* grad, const float4* rec, hipStream_t st)
{
    int tiles_x = (w + 15) / 16, tiles_y = (h + 15) / 16;
    int ntiles = tiles_x * tiles_y;
    AdjParams ap{dimg, grad, sc.n_mats, sc.n_mats < MH_ADJ_LDS_MATS ? sc.n_mats : MH_ADJ_LDS_MATS, rec};
    auto kern = rec ? (sc.fog_kernel ? k_adjoint<true, true> : k_adjoint<false, true>)
                    : (sc.fog_kernel ? k_adjoint<true, false> : k_adjoint<false, false>);
    hipLaunchKernelGGL(kern, dim3(ntiles), dim3(256), 0, st, sc, cam, w, h, spp, s0, depth, tiles_x, ntiles, ap);
    return hipGetLastError();
}

} // namespace mh
