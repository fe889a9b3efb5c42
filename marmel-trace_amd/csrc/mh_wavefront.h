// mh_wavefront.h - wavefront (queue-based) path tracer, fp32.
//
// The per-path bounce loop of the reference (src/pathtrace.c:800-1028) is
// split into kernels that each run over a compacted queue:
//
//   gen       start a batch of paths: camera ray, seed, throughput 1
//   extend    nearest hit of every queued path (small register footprint,
//             high occupancy: the traversal is the HBM-bound part)
//   shade     resolve the previous bounce's next-event estimates, add
//             emission / sky, emit this bounce's NEE shadow rays into the
//             shadow queue, sample the BSDF, Russian roulette, and append
//             the survivors to the next extension queue
//   shadow    any-hit test of every queued shadow ray (or the transmissive
//             segment walk), setting one visibility bit per estimate
//   finalize  resolve the last estimates of paths that ended and add the
//             path radiance to its pixel
//
// A batch holds at most one sample per pixel (batch size <= pixel count), and
// batches advance through the samples in increasing order, so each pixel's
// radiance is summed in the reference's sample order without atomics.
// Appends use one atomic per wave: ballot of the predicate, popcount, and the
// lane's rank from mbcnt. Every kernel is a persistent grid that reads its
// queue length from device memory, so the bounce sequence of a batch has
// fixed launch parameters and is captured once into a hipGraph.
//
// Batched engine (MODE WF_ONE or WF_TABLE): the same kernels over a table of jobs (a
// camera view of one scene each, with its own size, spp and depth). Global
// path indices run job by job, sample-major; a batch takes the next `pool`
// of them whatever jobs they belong to, so small jobs fill the machine
// together. Each slot records its job; the kernels look up the job's scene
// (WF_TABLE), or read the one scene every job renders at a launch-uniform
// index (WF_ONE).
// A batch may then hold several samples of one pixel, so pixel sums are
// 64-bit fixed-point atomics (2^-24 units): integer addition is associative,
// and the result does not depend on the order in which paths end.
#pragma once

#include "mh_kernels_impl.h"

namespace mh {

__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

// Block-wide append to Q queues: lane wants want[q] (0/1) entries in queue q;
// returns each wanted entry's position in at[q]. One atomic per queue per
// block and call: a counter that every wave hits serialises in L2 (~7 ns per
// same-address atomic measured). Every thread of the block must call it
// the same number of times (it synchronises); blockDim <= 256.
template <int Q>
__device__ __forceinline__ void block_append(uint32_t* const cnt[Q], const bool want[Q], uint32_t at[Q])
{
    __shared__ uint32_t wn[Q][4];
    __shared__ uint32_t base[Q];
    int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    __syncthreads();  // the previous call's readers are done
    uint64_t m[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        m[q] = __ballot(want[q]);
        if (lane == 0) wn[q][wave] = (uint32_t)__popcll(m[q]);
    }
    __syncthreads();
    int nw = (int)(blockDim.x >> 6);
    if ((int)threadIdx.x < Q) {
        int q = (int)threadIdx.x;
        uint32_t tot = 0;
        for (int w = 0; w < nw; ++w) tot += wn[q][w];
        base[q] = tot ? atomicAdd(cnt[q], tot) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        uint32_t b = base[q];
        for (int w = 0; w < wave; ++w) b += wn[q][w];
        at[q] = b + __builtin_amdgcn_mbcnt_hi((uint32_t)(m[q] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[q], 0u));
    }
}

// block_append<2> for the shade kernel's two queues, plus this thread's
// shadow rays (bit k of `emitted`: light block k, staged in stage[k * bs +
// tid] as direction + tmax, the shared origin in stage[9 * bs + tid])
// appended to the dense shadow queue (counter shc) as records sh0 = origin +
// tmax, sh1 = direction + (slot << 4 | estimate), shq = the scene index, in
// one pass with three barriers (every thread calls it). Within the block's
// range the rays are grouped by light block, so consecutive queue entries
// mostly hold one light's rays from neighbouring paths. S <= 9; the sun is
// block S - 1 and estimate 8.
template <bool TABLE>
__device__ __forceinline__ void shade_append(uint32_t* const cnt[2], const bool want[2], uint32_t at[2],
                                             uint32_t* shc, const WfState& st, const float4* stage,
                                             uint32_t emitted, uint32_t S, uint32_t slot, uint32_t scn)
{
    __shared__ uint32_t wn[2][4];
    __shared__ uint32_t base[2];
    __shared__ uint32_t sn[9][4];
    __shared__ uint32_t off[9][4];
    const int lane = (int)__lane_id(), wave = (int)(threadIdx.x >> 6);
    const int nw = (int)(blockDim.x >> 6);
    __syncthreads();  // the previous call's readers are done
    uint64_t m[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        m[q] = __ballot(want[q]);
        if (lane == 0) wn[q][wave] = (uint32_t)__popcll(m[q]);
    }
    for (uint32_t k = 0; k < S; ++k) {
        uint64_t mk_ = __ballot((emitted >> k) & 1u);
        if (lane == 0) sn[k][wave] = (uint32_t)__popcll(mk_);
    }
    __syncthreads();
    // the queue atomics in wave 0, the shadow atomic in another wave when the
    // block has one, so that their latencies overlap
    const int tsh = blockDim.x > 64 ? 64 : 2;
    if ((int)threadIdx.x < 2) {
        int q = (int)threadIdx.x;
        uint32_t tot = 0;
        for (int w = 0; w < nw; ++w) tot += wn[q][w];
        base[q] = tot ? atomicAdd(cnt[q], tot) : 0u;
    } else if ((int)threadIdx.x == tsh) {
        uint32_t tot = 0;
        for (uint32_t k = 0; k < S; ++k)
            for (int w = 0; w < nw; ++w) tot += sn[k][w];
        uint32_t run = tot ? atomicAdd(shc, tot) : 0u;
        for (uint32_t k = 0; k < S; ++k)
            for (int w = 0; w < nw; ++w) {
                off[k][w] = run;
                run += sn[k][w];
            }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        uint32_t b = base[q];
        for (int w = 0; w < wave; ++w) b += wn[q][w];
        at[q] = b + __builtin_amdgcn_mbcnt_hi((uint32_t)(m[q] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[q], 0u));
    }
    const uint32_t bs = blockDim.x;
    const float4 org = stage[9u * bs + threadIdx.x];
    for (uint32_t k = 0; k < S; ++k) {
        uint64_t mk_ = __ballot((emitted >> k) & 1u);
        if ((emitted >> k) & 1u) {
            const uint32_t p = off[k][wave] + __builtin_amdgcn_mbcnt_hi((uint32_t)(mk_ >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mk_, 0u));
            const float4 w = stage[k * bs + threadIdx.x];
            const uint32_t est = k == S - 1u ? 8u : k;
            st.sh0[p] = make_float4(org.x, org.y, org.z, w.w);
            st.sh1[p] = make_float4(w.x, w.y, w.z, __uint_as_float((slot << 4) | est));
            if (TABLE) st.shq[p] = scn;
        }
    }
}

// Ray-query count of a block: one atomic per block (every thread calls it).
__device__ __forceinline__ void count_rays_block(unsigned long long* ctr, int nq)
{
    __shared__ unsigned long long part[4];
    unsigned long long v = (unsigned long long)nq;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    if (__lane_id() == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        if (t) atomicAdd(ctr, t);
    }
}

// Per-wave sums of per-lane traversal counts into visits[0..n) (COUNT builds).
template <int N> __device__ __forceinline__ void add_visits(unsigned long long* visits, const uint32_t (&c)[N])
{
#pragma unroll
    for (int k = 0; k < N; ++k) {
        unsigned long long v = c[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
        if (__lane_id() == 0 && v) atomicAdd(&visits[k], v);
    }
}


__global__ __launch_bounds__(64) void k_wf_reset(uint32_t* cnt, int which)
{
    // zero the output queue of this bounce plus the shadow / term / slow-list
    // counters
    if (threadIdx.x == 0) {
        cnt[which] = 0;
        cnt[2] = 0;
        cnt[3] = 0;
        cnt[5] = 0;
    }
}

// Job of global path index g: the last job whose first index is <= g.
__device__ __forceinline__ int find_job(const WfMulti& mj, long long g)
{
    int lo = 0, hi = mj.n_jobs - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (mj.jobs[mid].begin <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The scene a path slot renders.
template <int MODE>
__device__ __forceinline__ const SceneView<float>& slot_scene(const SceneView<float>& sc, const WfMulti& mj,
                                                              const WfState& st, uint32_t slot)
{
    return MODE == WF_TABLE ? mj.scenes[mj.jobs[st.jid[slot]].scene]
                            : MODE == WF_ONE ? mj.scenes[mj.one_scene] : sc;
}

// Whether the dynamic-fetch kernels number their waves XCD-major (1):
// workgroup i runs on XCD i mod 8, so the usual numbering (workgroup-major,
// 0) spreads consecutive queue chunks over all eight XCDs; numbered
// XCD-major, the waves of one XCD take consecutive chunks (the rays of
// neighbouring pixels). Measured on the 1080p frame (ms per frame): extend
// 39.3 -> 47.9, shadow 144.8 -> 148.6 with it, so it is off.
#ifndef MH_XCD_ORDER
#define MH_XCD_ORDER 0
#endif

// This wave's index among the grid's waves, wave-uniform: XCD-major when
// MH_XCD_ORDER (a bijection onto [0, waves) for any grid size), else
// workgroup-major.
__device__ __forceinline__ uint32_t grid_wave()
{
    const uint32_t W = blockDim.x >> 6;
    const uint32_t wave = (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (!MH_XCD_ORDER) return blockIdx.x * W + wave;
    const uint32_t G = gridDim.x, q = G >> 3, r = G & 7u;
    const uint32_t x = blockIdx.x & 7u, bx = blockIdx.x >> 3;
    // workgroups on XCDs before x: q each, one more on the first r
    const uint32_t before = x * q + (x < r ? x : r);
    return (before + bx) * W + wave;
}

// The fields a traversal kernel reads from the launch's one scene (WF_SINGLE,
// WF_ONE), held in scalar registers: read once at kernel start, so the walk
// neither re-reads the scene table (which the compiler cannot prove
// unwritten) nor keeps the bases in vector registers.
struct UScene {
    const DNode* nodes;
    const DNode4* nodes4;
    const P32* prims;
    const P32* planes;
    const int* plane_index;
    int n_planes, root_ref, root4, bvh4, has_transmissive;
    V3<float> tlo, thi;
};

__device__ __forceinline__ const void* uniform_ptr(const void* p)
{
    const uint64_t v = (uint64_t)(uintptr_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (const void*)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ UScene uniform_scene(const SceneView<float>& s)
{
    UScene u;
    u.nodes = (const DNode*)uniform_ptr(s.nodes);
    u.nodes4 = (const DNode4*)uniform_ptr(s.nodes4);
    u.prims = (const P32*)uniform_ptr(s.prims);
    u.planes = (const P32*)uniform_ptr(s.planes);
    u.plane_index = (const int*)uniform_ptr(s.plane_index);
    u.n_planes = __builtin_amdgcn_readfirstlane(s.n_planes);
    u.root_ref = __builtin_amdgcn_readfirstlane(s.root_ref);
    u.root4 = __builtin_amdgcn_readfirstlane(s.root4);
    u.bvh4 = __builtin_amdgcn_readfirstlane(s.bvh4);
    u.has_transmissive = __builtin_amdgcn_readfirstlane(s.has_transmissive);
    auto uf = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
    u.tlo = mk(uf(s.tlo.x), uf(s.tlo.y), uf(s.tlo.z));
    u.thi = mk(uf(s.thi.x), uf(s.thi.y), uf(s.thi.z));
    return u;
}

// Whether the segment o + t d (d unit), t in [0, tmax], can meet the box
// [lo, hi] grown by 1e-3 + 1e-5 |coordinate| on every side: an fp32 slab
// test made conservative by the growth and a relative slack on the
// distances. An empty box (lo > hi) is never met, an unbounded one always.
__device__ __forceinline__ bool seg_meets_box(V3<float> o, V3<float> d, float tmax, V3<float> lo, V3<float> hi)
{
    if (!(lo.x <= hi.x && lo.y <= hi.y && lo.z <= hi.z)) return false;
    const float big = 1e30f;
    if (!(lo.x > -big && lo.y > -big && lo.z > -big && hi.x < big && hi.y < big && hi.z < big)) return true;
    float t0 = 0.f, t1 = fminf(tmax, big) * 1.00001f + 2e-3f;
    const float os[3] = {o.x, o.y, o.z}, ds[3] = {d.x, d.y, d.z};
    const float ls[3] = {lo.x, lo.y, lo.z}, hs[3] = {hi.x, hi.y, hi.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float m = 1e-3f + 1e-5f * fmaxf(fabsf(ls[a]), fabsf(hs[a]));
        const float l = ls[a] - m, h = hs[a] + m;
        if (fabsf(ds[a]) < 1e-20f) {
            if (os[a] < l || os[a] > h) return false;
            continue;
        }
        const float inv = 1.f / ds[a];
        const float ta = (l - os[a]) * inv, tb = (h - os[a]) * inv;
        t0 = fmaxf(t0, fminf(ta, tb));
        t1 = fminf(t1, fmaxf(ta, tb));
    }
    return t0 <= t1 * 1.00001f + 1e-4f;
}

// The four children of 4-wide node `ref` of a launch-uniform node array,
// tested from their near and far planes (Slab::nf; near-row offsets nrx,
// nry, nrz from Slab::near_off): hit flags and entry distances, and the
// child refs. Slots past the child count have ref INT_MIN and never hit.
struct Kids4 {
    bool h[4];
    float t[4];
    int r[4];
};
__device__ __forceinline__ Kids4 kids4_nf(const Slab<float>& sl, const DNode4* nodes4, int ref, uint32_t nrx,
                                          uint32_t nry, uint32_t nrz, float tmin, float tmax)
{
    const char* nb = (const char*)nodes4;
    // node arrays stay below 4 GiB: 32-bit byte offsets
    const uint32_t off = (uint32_t)ref << 7;
    const uint32_t vx = off | nrx, vy = off | nry, vz = off | nrz;
    const float4 nx4 = ldg((const float4*)(nb + (size_t)vx));
    const float4 fx4 = ldg((const float4*)(nb + (size_t)(vx ^ 16u)));
    const float4 ny4 = ldg((const float4*)(nb + (size_t)vy + 32));
    const float4 fy4 = ldg((const float4*)(nb + (size_t)(vy ^ 16u) + 32));
    const float4 nz4 = ldg((const float4*)(nb + (size_t)vz + 64));
    const float4 fz4 = ldg((const float4*)(nb + (size_t)(vz ^ 16u) + 64));
    const int4 rf = ldg((const int4*)(nb + (size_t)off + 96));
    Kids4 k;
    k.r[0] = rf.x;
    k.r[1] = rf.y;
    k.r[2] = rf.z;
    k.r[3] = rf.w;
    k.h[0] = sl.nf(nx4.x, fx4.x, ny4.x, fy4.x, nz4.x, fz4.x, tmin, tmax, &k.t[0]);
    k.h[1] = sl.nf(nx4.y, fx4.y, ny4.y, fy4.y, nz4.y, fz4.y, tmin, tmax, &k.t[1]);
    k.h[2] = sl.nf(nx4.z, fx4.z, ny4.z, fy4.z, nz4.z, fz4.z, tmin, tmax, &k.t[2]) && rf.z != INT_MIN;
    k.h[3] = sl.nf(nx4.w, fx4.w, ny4.w, fy4.w, nz4.w, fz4.w, tmin, tmax, &k.t[3]) && rf.w != INT_MIN;
    return k;
}

// The scene of table index `i` (a job's scene) under launch mode MODE.
template <int MODE>
__device__ __forceinline__ const SceneView<float>& mode_scene(const SceneView<float>& sc, const WfMulti& mj, int i)
{
    return MODE == WF_TABLE ? mj.scenes[i] : MODE == WF_ONE ? mj.scenes[mj.one_scene] : sc;
}


// Start global path g (camera ray, seed, unit throughput) in path slot `slot`.
template <int MODE>
__device__ __forceinline__ void start_path(const SceneView<float>& sc0, const CamView<float>& cam0, const WfJob& job,
                                           const WfState& st, const WfMulti& mj, long long g, uint32_t slot)
{
    constexpr bool MULTI = MODE != WF_SINGLE;
    int j = MULTI ? find_job(mj, g) : 0;
    const BJob* bj = MULTI ? &mj.jobs[j] : nullptr;
    const SceneView<float>& sc = mode_scene<MODE>(sc0, mj, MULTI ? bj->scene : 0);
    const CamView<float>& cam = MULTI ? bj->cam : cam0;
    long long l = MULTI ? g - bj->begin : g;
    int npix = MULTI ? bj->npix : job.npix;
    int jw = MULTI ? bj->w : job.w, jh = MULTI ? bj->h : job.h;
    uint32_t pix = tile_pixel((uint32_t)(l % npix), (uint32_t)jw, (uint32_t)jh);
    uint32_t s = (uint32_t)(MULTI ? bj->s0 : job.s0) + (uint32_t)(l / npix);
    uint32_t x = pix % (uint32_t)jw, y = pix / (uint32_t)jw;
    float jx = rand01<float>(x, y, s * 2u + 0u), jy = rand01<float>(x, y, s * 2u + 1u);
    float l1 = rand01<float>(x, y, s * 2u + 100u), l2 = rand01<float>(x, y, s * 2u + 101u);
    float uu = ((float)x + jx) / (float)jw;
    float vv = 1.0f - ((float)y + jy) / (float)jh;
    V3<float> o, d;
    camera_ray(cam, uu, vv, l1, l2, &o, &d);
    if (cam.dome > 1e-4f && sc.water_material >= 0) {
        V3<float> pd = o + d * cam.dome;
        if (pd.y < water_height(sc, pd.x, pd.z)) o = pd - mk(0.f, 1e-3f, 0.f);
    }
    int medium = -1;
    if (sc.water_material >= 0 && sc.water_material < sc.n_mats && o.y < water_height(sc, o.x, o.z))
        medium = sc.water_material;
    d = normalize(d);
    uint32_t seed = hash3(x, y, s);
    uint32_t meta = 0u | (1u << 8) | ((uint32_t)(medium + 1) << 9);
    st.p0[slot] = make_float4(o.x, o.y, o.z, u2f(seed));
    st.p1[slot] = make_float4(d.x, d.y, d.z, u2f(len2(d) > 0.f ? meta : 0xffffffffu));
    st.p2[slot] = make_float4(1.f, 1.f, 1.f, u2f(pix));
    st.p3[slot] = make_float4(0.f, 0.f, 0.f, u2f(s));
    st.npend[slot] = 0;
    if (MULTI) st.jid[slot] = (uint32_t)j;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_wf_gen(SceneView<float> sc0, CamView<float> cam0, WfJob job, WfState st,
                                                WfMulti mj)
{
    constexpr bool MULTI = MODE != WF_SINGLE;
    unsigned long long g0 = st.batch[0];
    long long remain = (MULTI ? mj.total : job.total) - (long long)g0;
    int n = (int)(remain < job.pool ? remain : job.pool);
    if (blockIdx.x == 0 && threadIdx.x == 0) st.cnt[0] = (uint32_t)(n > 0 ? n : 0);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        start_path<MODE>(sc0, cam0, job, st, mj, (long long)g0 + i, (uint32_t)i);
        st.q[0][i] = (uint32_t)i;
    }
}

// ---- path regeneration (batched engine) --------------------------------------
//
// Each bounce step first refills the slots freed by the previous step's ended
// paths with new paths from the global cursor, appending them to the same
// extension queue as the continuing paths: every step's kernels see a full
// pool until the work runs out. cnt[4] counts free slots (listed in freeq).

// Number of new paths the step starts: min(free slots, paths not started).
__device__ __forceinline__ uint32_t refill_count(const WfState& st, const WfMulti& mj)
{
    uint32_t nfree = st.cnt[4];
    long long remain = mj.total - (long long)st.batch[0];
    if (remain <= 0) return 0u;
    return remain < (long long)nfree ? (uint32_t)remain : nfree;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_wf_refill(WfState st, WfMulti mj, int which)
{
    uint32_t nfree = st.cnt[4];
    uint32_t k = refill_count(st, mj);
    unsigned long long g0 = st.batch[0];
    uint32_t n_round = (k + blockDim.x - 1u) / blockDim.x * blockDim.x;  // block-uniform trip count
    SceneView<float> none{};
    CamView<float> cnone{};
    WfJob jnone{};
    uint32_t* const cq[1] = {&st.cnt[which]};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += gridDim.x * blockDim.x) {
        bool active = i < k;
        uint32_t slot = 0;
        if (active) {
            slot = st.freeq[nfree - 1u - i];
            start_path<MODE>(none, cnone, jnone, st, mj, (long long)g0 + i, slot);
        }
        const bool want[1] = {active};
        uint32_t at[1];
        block_append<1>(cq, want, at);
        if (active) st.q[which][at[0]] = slot;
    }
}

// Advance the cursor and pop the used free slots (after k_wf_refill).
__global__ __launch_bounds__(64) void k_wf_refill_commit(WfState st, WfMulti mj)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        uint32_t k = refill_count(st, mj);
        st.batch[0] += (unsigned long long)k;
        st.cnt[4] -= k;
    }
}

// Every slot free, queues empty, cursor at 0.
__global__ __launch_bounds__(256) void k_wf_stream_init(WfState st, int pool)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < pool; i += gridDim.x * blockDim.x)
        st.freeq[i] = (uint32_t)(pool - 1 - i);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st.cnt[0] = st.cnt[1] = st.cnt[2] = st.cnt[3] = 0u;
        st.cnt[4] = (uint32_t)pool;
        st.batch[0] = 0ull;
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void k_wf_extend(SceneView<float> sc0, WfState st, int which, WfMulti mj)
{
#if MH_EXTEND_LDS
    MH_STACK_LDS(ts, MH_EXTEND_LDS, 256)
#else
    MH_STACK_PRIV(ts)
#endif
    uint32_t n = st.cnt[which];
    const uint32_t* q = st.q[which];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t slot = q[i];
        const SceneView<float>& sc = slot_scene<MODE>(sc0, mj, st, slot);
        float4 a = st.p0[slot], b = st.p1[slot];
        float t;
        int prim;
        int rec = closest_hit(sc, ts, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), 1e-4f, 1e30f, &t, &prim);
        st.hit[slot] = make_float2(t, __int_as_float(rec));
    }
}

// Idle lanes of a wave at which the closest-hit extend kernel hands them new
// rays (dynamic fetch); 0 selects k_wf_extend. Measured on the 1080p terrain
// frame (extend ms per frame, binary tree): 0 -> 77.2, 32 -> 61.7,
// 48 -> 59.9, 56 -> 61.0; 4-wide tree, near / far tests, linear stack:
// 16 -> 40.7, 24 -> 39.7, 32 -> 40.0, 48 -> 39.2. The 4K stress scene's
// deeper, less even walks (one 64-spp chunk, s): 16 -> 0.938, 24 -> 0.938,
// 32 -> 0.946, 40 -> 0.964, 48 -> 0.992.
#ifndef MH_EXTEND_REFILL
#define MH_EXTEND_REFILL 24
#endif
// On the 4-wide tree (scenes whose view selects it; the autotuner picks it
// for the 1080p terrain frame): the nearest hit child is visited next from a
// register; the others are stacked 0 = in index order, 1 = nearest last
// (sorted). Extend ms per frame: binary 60.5, 0 -> 59.0, 1 -> 57.1; with
// the near / far child tests: 0 -> 41.2, 1 -> 39.05.
#ifndef MH_EXTEND_W4_SORT
#define MH_EXTEND_W4_SORT 1
#endif
// 1: a leaf's primitives are tested one per step, like nodes (one load per
// lane per step); 0: the whole leaf in one step. Extend ms per frame:
// 0 -> 57.1, 1 -> 55.5.
// 1: the extend kernel's LDS stack is linear (K slots at a 256-lane stride,
// the older half moved to the private overflow when full); 0: a ring.
#ifndef MH_EXTEND_LINEAR
#define MH_EXTEND_LINEAR 1
#endif
// 1: the 4-wide closest-hit walk tests each child from its near and far
// planes (kids4_nf) in every launch mode; 0: the per-scene table mode loads
// whole nodes and tests every plane (children4)
#ifndef MH_EXTEND_NF
#define MH_EXTEND_NF 1
#endif
#ifndef MH_EXTEND_PRIM_STEP
#define MH_EXTEND_PRIM_STEP 1
#endif

// Nearest hit of every queued path with dynamic fetch, on the binary tree:
// the search of closest_hit<float> (planes first, then the tree with the
// nearer hit child held in a register and the farther one stacked with its
// entry distance, strict > culling, lowest prim index on equal t), one node
// or leaf per step per lane. The queue is compacted, so a wave's entries are
// its 64-entry chunks (wave w, w + waves, ...) in order, addressed by a
// wave-uniform position; whenever MH_EXTEND_REFILL lanes have finished and
// entries remain, the idle lanes take the next ones. Same hits as
// k_wf_extend.
template <int MODE, bool COUNT = false>
__global__ __launch_bounds__(256) void k_wf_extend_refill(SceneView<float> sc0, WfState st, int which, WfMulti mj)
{
    constexpr int K = MH_EXTEND_LDS;
    static_assert(K > 0 && (K & (K - 1)) == 0 && K < MH_STACK, "LDS stack depth must be a power of two");
    __shared__ int lds_ref[K * 256];
    __shared__ float lds_t[K * 256];
    int* const myr = lds_ref + threadIdx.x;
    float* const myt = lds_t + threadIdx.x;
    constexpr bool UNI = MODE != WF_TABLE;
#if MH_EXTEND_LINEAR
    // Linear stack of (ref, entry distance): LDS slot k at a 256-lane
    // stride, sp entries in LDS above psp older ones in the private
    // overflow; a full LDS part moves its older half to the overflow.
    constexpr int NO = MH_STACK;
    int ovr[NO];
    float ovt[NO];
    int sp = 0, psp = 0;
    auto spill = [&]() {
        constexpr int H = K / 2;
        for (int j = 0; j < H; ++j) {
            const int o = psp + j < NO ? psp + j : NO - 1;
            ovr[o] = myr[j * 256];
            ovt[o] = myt[j * 256];
        }
        for (int j = H; j < sp; ++j) {
            myr[(j - H) * 256] = myr[j * 256];
            myt[(j - H) * 256] = myt[j * 256];
        }
        psp += H;
        sp -= H;
    };
    auto push = [&](int r, float t) {
        if (sp == K) spill();
        myr[sp * 256] = r;
        myt[sp * 256] = t;
        ++sp;
    };
    auto pop = [&](float* t) -> int {
        // unconditional LDS reads (a slot of this lane either way)
        const int kk = ((sp - 1) & (K - 1)) * 256;
        float tv = lds_read(&myt[kk]);
        int v = lds_read(&myr[kk]);
        if (sp > 0) {
            --sp;
        } else {
            --psp;
            const int j = psp < NO - 1 ? psp : NO - 1;
            tv = ovt[j];
            v = ovr[j];
        }
        *t = tv;
        return v;
    };
#else
    const int bs = (int)blockDim.x;
    int ovr[MH_STACK - K];
    float ovt[MH_STACK - K];
    int sp = 0, base = 0, psp = 0;
    auto push = [&](int r, float t) {
        if (sp == K) {
            int j = psp < MH_STACK - K - 1 ? psp : MH_STACK - K - 1;
            ovr[j] = myr[base * bs];
            ovt[j] = myt[base * bs];
            ++psp;
            base = (base + 1) & (K - 1);
            --sp;
        }
        int k = ((base + sp) & (K - 1)) * bs;
        myr[k] = r;
        myt[k] = t;
        ++sp;
    };
    auto pop = [&](float* t) -> int {
        // unconditional LDS reads (a slot of this lane either way)
        int k = ((base + sp - 1) & (K - 1)) * bs;
        float tv = lds_read(&myt[k]);
        int v = lds_read(&myr[k]);
        if (sp > 0) {
            --sp;
        } else {
            --psp;
            int j = psp < MH_STACK - K - 1 ? psp : MH_STACK - K - 1;
            tv = ovt[j];
            v = ovr[j];
        }
        *t = tv;
        return v;
    };
#endif

    const uint32_t n = st.cnt[which];
    const uint32_t* q = st.q[which];
    const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t wg = grid_wave();
    const int lane = (int)__lane_id();
    // queue index of the wave's p-th entry: chunks wg, wg + nwaves, ...
    auto entry = [&](uint32_t p) -> uint32_t { return (wg + (p >> 6) * nwaves) * 64u + (p & 63u); };
    uint32_t pos = 0;  // wave-uniform
    const float tmin = 1e-4f;
    bool busy = false, held = false;
    Slab<float> sl;
    V3<float> o = zero3<float>(), d = zero3<float>();
    float closest = 0.f;
    int best = INT_MIN, bprim = 0x7fffffff, ref = 0;
    uint32_t slot = 0;
    // the lane's scene arrays (WF_TABLE); UNI: the launch's scene in scalar
    // registers and the lane's near-row offsets (Slab::near_off)
    const DNode* nodes = nullptr;
    const DNode4* nodes4 = nullptr;  // the 4-wide tree when the lane's scene view selects it
    const P32* prims = nullptr;
    const UScene scu = uniform_scene(mode_scene<MODE>(sc0, mj, 0));
    uint32_t nrx = 0, nry = 0, nrz = 0;
    // MH_EXTEND_PRIM_STEP: the current leaf's untested primitives, the next
    // one's index in bits 0..27 (tree primitives number at most 2^27) and
    // how many remain in bits 28..31
    uint32_t leaf = 0;
    uint32_t cnt_q = 0, cnt_n = 0, cnt_t = 0, cnt_o = 0;  // COUNT: queries, nodes, triangles, other prims
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (nidle == 64u && entry(pos) >= n) break;
        if (entry(pos) < n) {
            uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            uint32_t i = entry(pos + rank);
            if (!busy && i < n) {
                slot = q[i];
                const SceneView<float>& sc = slot_scene<MODE>(sc0, mj, st, slot);
                const int n_planes = UNI ? scu.n_planes : sc.n_planes;
                const P32* planes = UNI ? scu.planes : sc.planes;
                const int* plane_index = UNI ? scu.plane_index : sc.plane_index;
                float4 a = st.p0[slot], b = st.p1[slot];
                o = mk(a.x, a.y, a.z);
                d = mk(b.x, b.y, b.z);
                closest = 1e30f;
                best = INT_MIN;
                bprim = 0x7fffffff;
                if (COUNT) {
                    ++cnt_q;
                    cnt_o += (uint32_t)n_planes;
                }
                for (int k = 0; k < n_planes; ++k) {
                    PD<float> p = load_prim(&planes[k]);
                    float t;
                    if (hit_plane<float, false>(p, o, d, tmin, closest, &t, nullptr)) {
                        int pi = plane_index[k];
                        if (best == INT_MIN || t < closest || (t == closest && pi < bprim)) {
                            closest = t;
                            bprim = pi;
                            best = -1 - k;
                        }
                    }
                }
                if ((UNI ? scu.root_ref : sc.root_ref) != INT_MIN) {
                    sl.init(o, d);
#if MH_EXTEND_LINEAR
                    sp = psp = 0;
#else
                    sp = base = psp = 0;
#endif
                    leaf = 0;
                    nrx = Slab<float>::near_off(sl.ix);
                    nry = Slab<float>::near_off(sl.iy);
                    nrz = Slab<float>::near_off(sl.iz);
                    if (UNI) {
                        ref = scu.bvh4 ? scu.root4 : scu.root_ref;
                    } else {
                        nodes = sc.nodes;
                        nodes4 = sc.bvh4 ? sc.nodes4 : nullptr;
                        prims = sc.prims;
                        ref = nodes4 ? sc.root4 : sc.root_ref;
                    }
                    held = true;
                    busy = true;
                } else {
                    st.hit[slot] = make_float2(closest, __int_as_float(best));
                }
            }
            pos += nidle;
        }
        const bool more = entry(pos) < n;  // wave-uniform
        for (;;) {
            if (busy) {
                int pj = -1;  // MH_EXTEND_PRIM_STEP: the primitive this lane tests in this step
                if (MH_EXTEND_PRIM_STEP && (leaf >> 28)) {
                    pj = (int)(leaf & 0x0fffffffu);
                    leaf += 1u - (1u << 28);
                } else {
                bool work = true;
                if (!held) {
                    float te;
                    ref = pop(&te);
                    work = !(best != INT_MIN && te > closest);
                }
                held = false;
                if (work) {
                    if (MH_EXTEND_W4_SORT && MH_EXTEND_NF && ref >= 0 && (UNI ? scu.bvh4 != 0 : nodes4 != nullptr)) {
                        // children from their near and far planes, nearest
                        // first (the sorting network of children4); the
                        // table mode's node base is per lane
                        if (COUNT) ++cnt_n;
                        const Kids4 kd = kids4_nf(sl, UNI ? scu.nodes4 : nodes4, ref, nrx, nry, nrz, tmin, closest);
                        const float inf = __builtin_inff();
                        float kk[4] = {kd.h[0] ? kd.t[0] : inf, kd.h[1] ? kd.t[1] : inf, kd.h[2] ? kd.t[2] : inf,
                                       kd.h[3] ? kd.t[3] : inf};
                        int qq[4] = {kd.r[0], kd.r[1], kd.r[2], kd.r[3]};
#define MH_KSWAP(a_, b_)                                                                                  \
    if (kk[b_] < kk[a_]) {                                                                                \
        float tk_ = kk[a_];                                                                               \
        kk[a_] = kk[b_];                                                                                  \
        kk[b_] = tk_;                                                                                     \
        int tq_ = qq[a_];                                                                                 \
        qq[a_] = qq[b_];                                                                                  \
        qq[b_] = tq_;                                                                                     \
    }
                        MH_KSWAP(0, 1)
                        MH_KSWAP(2, 3)
                        MH_KSWAP(0, 2)
                        MH_KSWAP(1, 3)
                        MH_KSWAP(1, 2)
#undef MH_KSWAP
                        for (int i = 3; i >= 1; --i)
                            if (kk[i] < inf) push(qq[i], kk[i]);
                        ref = qq[0];
                        held = kk[0] < inf;
                    } else if (ref >= 0 && (UNI ? scu.bvh4 != 0 : nodes4 != nullptr)) {
                        DNode4 nd = ldg(&(UNI ? scu.nodes4 : nodes4)[ref]);
                        if (COUNT) ++cnt_n;
                        if (MH_EXTEND_W4_SORT) {
                            int r[4], nh;
                            float t[4];
                            children4(sl, nd, tmin, closest, r, t, &nh);
                            for (int i = 3; i >= 1; --i)
                                if (i < nh) push(r[i], t[i]);
                            ref = r[0];
                            held = nh > 0;
                        } else {
                            int cnt = nd.meta.x;
                            bool h0, h1, h2, h3;
                            float t0, t1, t2, t3;
                            sl.one(nd.lox.x, nd.hix.x, nd.loy.x, nd.hiy.x, nd.loz.x, nd.hiz.x, tmin, closest, &h0, &t0);
                            sl.one(nd.lox.y, nd.hix.y, nd.loy.y, nd.hiy.y, nd.loz.y, nd.hiz.y, tmin, closest, &h1, &t1);
                            sl.one(nd.lox.z, nd.hix.z, nd.loy.z, nd.hiy.z, nd.loz.z, nd.hiz.z, tmin, closest, &h2, &t2);
                            sl.one(nd.lox.w, nd.hix.w, nd.loy.w, nd.hiy.w, nd.loz.w, nd.hiz.w, tmin, closest, &h3, &t3);
                            h2 = h2 && cnt > 2;
                            h3 = h3 && cnt > 3;
                            // the nearest hit child goes on from a register
                            int k = -1;
                            float tk = 0.f;
                            if (h0) { k = 0; tk = t0; }
                            if (h1 && (k < 0 || t1 < tk)) { k = 1; tk = t1; }
                            if (h2 && (k < 0 || t2 < tk)) { k = 2; tk = t2; }
                            if (h3 && (k < 0 || t3 < tk)) { k = 3; tk = t3; }
                            if (h3 && k != 3) push(nd.ref.w, t3);
                            if (h2 && k != 2) push(nd.ref.z, t2);
                            if (h1 && k != 1) push(nd.ref.y, t1);
                            if (h0 && k != 0) push(nd.ref.x, t0);
                            ref = k == 0 ? nd.ref.x : k == 1 ? nd.ref.y : k == 2 ? nd.ref.z : nd.ref.w;
                            held = k >= 0;
                        }
                    } else if (ref >= 0) {
                        DNode nd = ldg(&(UNI ? scu.nodes : nodes)[ref]);
                        if (COUNT) ++cnt_n;
                        bool hl, hr;
                        float tl, tr;
                        sl.test(nd, tmin, closest, &hl, &hr, &tl, &tr);
                        bool lfirst = hl && (!hr || tl <= tr);
                        if (hl && hr) push(lfirst ? nd.c3.y : nd.c3.x, lfirst ? tr : tl);
                        ref = lfirst ? nd.c3.x : nd.c3.y;
                        held = hl || hr;
                    } else {
                        int code = ~ref;
                        int first = code >> 3, cnt = (code & 7) + 1;
                        if (MH_EXTEND_PRIM_STEP) {
                            pj = first;
                            leaf = (uint32_t)(first + 1) | ((uint32_t)(cnt - 1) << 28);
                        } else {
                            for (int j = 0; j < cnt; ++j) {
                                PD<float> p = load_prim(&(UNI ? scu.prims : prims)[first + j]);
                                if (COUNT) {
                                    if (p.kind == K_TRIANGLE) ++cnt_t;
                                    else ++cnt_o;
                                }
                                float t;
                                if (hit_prim<float, false>(p, nullptr, o, d, tmin, closest, &t, nullptr)) {
                                    if (best == INT_MIN || t < closest || (t == closest && p.prim < bprim)) {
                                        closest = t;
                                        bprim = p.prim;
                                        best = first + j;
                                    }
                                }
                            }
                        }
                    }
                }
                }
                if (MH_EXTEND_PRIM_STEP && pj >= 0) {
                    PD<float> p = load_prim(&(UNI ? scu.prims : prims)[pj]);
                    if (COUNT) {
                        if (p.kind == K_TRIANGLE) ++cnt_t;
                        else ++cnt_o;
                    }
                    float t;
                    if (hit_prim<float, false>(p, nullptr, o, d, tmin, closest, &t, nullptr)) {
                        if (best == INT_MIN || t < closest || (t == closest && p.prim < bprim)) {
                            closest = t;
                            bprim = p.prim;
                            best = pj;
                        }
                    }
                }
                if (!held && (leaf >> 28) == 0u && sp == 0 && psp == 0) {
                    st.hit[slot] = make_float2(closest, __int_as_float(best));
                    busy = false;
                }
            }
            const uint64_t b = __ballot(busy);
            if (b == 0ull) break;
            if (more && 64u - (uint32_t)__popcll(b) >= (uint32_t)MH_EXTEND_REFILL) break;
        }
    }
    if (COUNT) {
        const uint32_t c[4] = {cnt_q, cnt_n, cnt_t, cnt_o};
        add_visits<4>(mj.visits, c);
    }
}

// add the visible pending estimates of the previous bounce, in estimate order
__device__ __forceinline__ V3<float> resolve_pending(const WfState& st, uint32_t slot, V3<float> L)
{
    uint32_t np = st.npend[slot];
    if (np == 0) return L;
    uint32_t vis = st.vis[slot];
    int nl = (int)(np & 15u);        // light estimates 0..nl-1
    bool sun = (np >> 4) & 1u;       // sun estimate in slot 8
    V3<float> sum = zero3<float>();
    for (int k = 0; k < nl; ++k)
        if (vis & (1u << k)) {
            float4 c = st.pend[(size_t)slot * 9 + k];
            sum = sum + mk(c.x, c.y, c.z);
        }
    L = L + sum;
    if (sun && (vis & (1u << 8))) {
        float4 c = st.pend[(size_t)slot * 9 + 8];
        L = L + mk(c.x, c.y, c.z);
    }
    return L;
}

template <int MODE, bool FOG>
__global__ __launch_bounds__(256) void k_wf_shade(SceneView<float> sc0, WfState st, WfJob job, int which, int bounce,
                                                   unsigned long long* rays, WfMulti mj)
{
    constexpr bool MULTI = MODE != WF_SINGLE;
    // FOG: fog segments and marched water; the marches trace their sun
    // shadow rays here, with this stack
    MH_STACK_PRIV(ts)
    // the shadow rays a thread casts, staged until their queue positions are
    // known (shade_append): 10 * blockDim.x float4s of dynamic LDS
    extern __shared__ float4 wf_stage[];
    uint32_t n = st.cnt[which];
    const uint32_t* qin = st.q[which];
    uint32_t* qout = st.q[which ^ 1];
    int nq = 0;
    // Shadow rays: S light blocks, the lights (at most 8) and the sun, last;
    // each ray cast is staged in wf_stage[block * blockDim.x + thread] and
    // appended to the dense shadow queue by shade_append.
    const uint32_t S = MULTI ? (uint32_t)mj.sh_stride : (uint32_t)sc0.n_lights + 1u;
    const uint32_t bs = blockDim.x;
    uint32_t* const cq[2] = {&st.cnt[which ^ 1], &st.cnt[3]};
    uint32_t n_round = (n + blockDim.x - 1u) / blockDim.x * blockDim.x;  // block-uniform trip count
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += gridDim.x * blockDim.x) {
        bool active = i < n;
        uint32_t slot = active ? qin[i] : 0u;
        bool alive = false, ended = false;
        uint32_t emitted = 0u;  // shadow rays staged (bit = block)
        uint32_t scn = 0u;      // scene index of the slot (batched engine)
        V3<float> nray_o = zero3<float>(), nray_d = zero3<float>();
        if (active) {
            const BJob* bj = MULTI ? &mj.jobs[st.jid[slot]] : nullptr;
            scn = MULTI ? (uint32_t)bj->scene : 0u;
            const SceneView<float>& sc = mode_scene<MODE>(sc0, mj, (int)scn);
            int depth = MULTI ? bj->depth : job.depth;
            float4 a = st.p0[slot], b = st.p1[slot], c = st.p2[slot], e = st.p3[slot];
            float2 hh = st.hit[slot];
            V3<float> o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
            V3<float> beta = mk(c.x, c.y, c.z), L = mk(e.x, e.y, e.z);
            uint32_t seed = f2u(a.w), meta = f2u(b.w);
            int last_spec = (meta >> 8) & 1u;
            int medium = (int)(meta >> 9) - 1;
            // per-slot bounce under path regeneration (paths of every depth share a step)
            int b_idx = (MULTI && mj.stream) ? (int)(f2u(b.w) & 0xffu) : bounce;
            int rec = __float_as_int(hh.y);
            L = resolve_pending(st, slot, L);
            st.npend[slot] = 0;
            ended = true;
            if (meta == 0xffffffffu) {
                // degenerate primary ray: zero radiance
            } else if (rec == INT_MIN) {
                if (medium >= 0) {
                    if (FOG)
                        pt_water_segment(sc, ts, sc.mats[medium], o, d, 5.f, seed, b_idx, &beta, &L,
                                         (V3<float>*)nullptr, (V3<float>*)nullptr, &nq);
                    else
                        apply_medium(sc.mats[medium], 5.f, &beta, &L);
                }
                V3<float> skyc = sky_sample(d, sc.sky);
                if (FOG && sc.fog.density > 0.f) {
                    float T;
                    V3<float> ins;
                    if (b_idx == 0)
                        pt_fog_segment(sc, ts, o, d, 1e30f, seed, 0, &T, &ins, (V3<float>*)nullptr, &nq);
                    else
                        fog_segment(sc.fog, sc.sky, o, d, 1e30f, &T, &ins, (V3<float>*)nullptr);
                    skyc = skyc * T + ins;
                }
                L = L + vmul(beta, skyc);
            } else {
                DHit<float> h;
                finish_hit(sc, rec, o, d, 1e-4f, hh.x, &h);
                if (medium >= 0) {
                    if (FOG)
                        pt_water_segment(sc, ts, sc.mats[medium], o, d, h.t, seed, b_idx, &beta, &L,
                                         (V3<float>*)nullptr, (V3<float>*)nullptr, &nq);
                    else
                        apply_medium(sc.mats[medium], h.t, &beta, &L);
                } else if (FOG && sc.fog.density > 0.f) {
                    float T;
                    V3<float> ins;
                    if (b_idx == 0)
                        pt_fog_segment(sc, ts, o, d, h.t, seed, 0, &T, &ins, (V3<float>*)nullptr, &nq);
                    else
                        fog_segment(sc.fog, sc.sky, o, d, h.t, &T, &ins, (V3<float>*)nullptr);
                    L = L + vmul(beta, ins);
                    beta = beta * T;
                }
                if (h.mat < 0 || h.mat >= sc.n_mats) {
                    L = L + vmul(beta, sky_sample(d, sc.sky));
                } else {
                    const DMat<float>& m = sc.mats[h.mat];
                    V3<float> P = h.p, N = h.n;
                    if (m.is_water && fabsf(h.n.y) > 0.999f && fabsf(h.n.x) < 1e-4f && fabsf(h.n.z) < 1e-4f)
                        N = water_normal<float>(P.x, P.z);
                    if (dot(N, d) > 0.f) N = -N;
                    N = normalize(N);
                    if (len2(N) > 0.f) {
                        V3<float> V = -d;
                        V3<float> alb = texture_albedo(m, P, h.u, h.v);
                        if ((b_idx == 0 || last_spec) && m.has_emission) L = L + vmul(beta, m.emissive);
                        if (b_idx == 0 && m.has_glow) L = L + vmul(beta, limb_glow(m, N, V, sc.sky.sun_dir));
                        // ---- next-event estimates -> shadow queue
                        uint32_t npend = 0;
                        V3<float> so = P + N * 1e-3f;
                        for (int li = 0; li < sc.n_lights; ++li) {
                            const DLight<float>& lt = sc.lights[li];
                            bool emit = false;
                            V3<float> wi = zero3<float>(), cc = zero3<float>();
                            float tmax = 0.f;
                            if (lt.prim != h.prim) {
                                V3<float> toc = lt.center - P;
                                float dc2 = len2(toc), dc = sqrtf(dc2);
                                if (dc > lt.radius) {
                                    float cmx = sqrtf(1.f - (lt.radius * lt.radius) / dc2);
                                    float omega = 2.f * Consts<float>::pi * (1.f - cmx);
                                    if (omega > 0.f) {
                                        V3<float> w = toc * (1.f / dc);
                                        float e1 = rand01<float>(seed ^ (uint32_t)li, (uint32_t)b_idx, 0xA201u);
                                        float e2 = rand01<float>(seed ^ (uint32_t)li, (uint32_t)b_idx, 0xA202u);
                                        wi = cone_dir(w, cmx, e1, e2);
                                        if (finite3(wi) && dot(N, wi) > 0.f) {
                                            PD<float> lp;
                                            lp.kind = K_SPHERE;
                                            lp.p0 = lt.center;
                                            lp.r0 = lt.radius;
                                            float lt_t;
                                            if (hit_sphere<float, false>(lp, so, wi, 1e-3f, 1e30f, &lt_t, nullptr)) {
                                                emit = true;
                                                tmax = lt_t - 1e-3f;
                                                cc = vmul(beta, shade(m, alb, N, wi, V, lt.emissive) * omega);
                                            }
                                        }
                                    }
                                }
                            }
                            if (emit) {
                                wf_stage[(uint32_t)li * bs + threadIdx.x] = make_float4(wi.x, wi.y, wi.z, tmax);
                                st.pend[(size_t)slot * 9 + li] = make_float4(cc.x, cc.y, cc.z, 0.f);
                                emitted |= 1u << li;
                            }
                        }
                        npend = (uint32_t)sc.n_lights;
                        {
                            bool emit = false;
                            V3<float> Ls = zero3<float>(), cc = zero3<float>();
                            if (b_idx == 0 || (b_idx == 1 && medium >= 0 && last_spec)) {
                                float s1 = rand01<float>(seed, (uint32_t)b_idx, 0xA101u);
                                float s2 = rand01<float>(seed, (uint32_t)b_idx, 0xA102u);
                                Ls = sun_disk_dir(sc.sky.sun_dir, sc.sky.sun_radius, s1, s2);
                                if (finite3(Ls) && dot(N, Ls) > 0.f) {
                                    emit = true;
                                    cc = vmul(beta, shade(m, alb, N, Ls, V, sc.sky.sun_color) * sc.sky.sun_omega);
                                }
                            }
                            if (emit) {
                                wf_stage[(S - 1u) * bs + threadIdx.x] = make_float4(Ls.x, Ls.y, Ls.z, 1e30f);
                                st.pend[(size_t)slot * 9 + 8] = make_float4(cc.x, cc.y, cc.z, 0.f);
                                emitted |= 1u << (S - 1u);
                                npend |= 16u;
                            }
                        }
                        if (emitted) wf_stage[9u * bs + threadIdx.x] = make_float4(so.x, so.y, so.z, 0.f);
                        st.vis[slot] = 0u;
                        st.npend[slot] = npend;
                        // ---- BSDF sample, throughput, roulette
                        uint32_t kb = hash3(seed ^ 0x5054u, (uint32_t)b_idx, 0u);
                        float u1 = rand01<float>(kb, (uint32_t)b_idx, 0xB501u);
                        float u2 = rand01<float>(kb, (uint32_t)b_idx, 0xB502u);
                        float ul = rand01<float>(kb, (uint32_t)b_idx, 0x1001u);
                        float ug = rand01<float>(kb, (uint32_t)b_idx, 0x1002u);
                        Lobe<float> lb;
                        bool ok = m.pbr ? sample_pbr(m, alb, N, d, V, u1, u2, ul, &lb)
                                        : sample_legacy(m, alb, N, d, V, h.front, u1, u2, ul, ug, &lb);
                        if (ok) {
                            last_spec = lb.spec;
                            beta = vmul(beta, lb.w);
                            float mc = fmaxf(fmaxf(beta.x, beta.y), beta.z);
                            if (!(mc > 0.f)) mc = 0.f;
                            if (finite3(beta) && mc > 0.f) {
                                if (lb.refr && (m.is_water || m.beer_lambert)) medium = h.front ? h.mat : -1;
                                bool live = true;
                                if (b_idx >= 2) {
                                    float q = clamp01n(mc);
                                    if (!(q > 0.f) || rand01<float>(kb, (uint32_t)b_idx, 0x5201u) >= q) live = false;
                                    else beta = beta * (1.f / q);
                                }
                                if (live && b_idx < depth) {
                                    V3<float> off = lb.refr ? lb.wi * 1e-4f
                                                            : N * (dot(lb.wi, N) >= 0.f ? 1e-4f : -1e-4f);
                                    nray_o = P + off;
                                    nray_d = lb.wi;
                                    alive = true;
                                    ended = false;
                                }
                            }
                        }
                    }
                }
            }
            uint32_t nmeta = (uint32_t)(b_idx + 1) | ((uint32_t)last_spec << 8) | ((uint32_t)(medium + 1) << 9);
            if (alive) {
                st.p0[slot] = make_float4(nray_o.x, nray_o.y, nray_o.z, a.w);
                st.p1[slot] = make_float4(nray_d.x, nray_d.y, nray_d.z, u2f(nmeta));
            }
            st.p2[slot] = make_float4(beta.x, beta.y, beta.z, c.w);
            st.p3[slot] = make_float4(L.x, L.y, L.z, e.w);
            nq += 1;
        }
        const bool want[2] = {alive, ended};
        uint32_t at[2];
        shade_append<MODE == WF_TABLE>(cq, want, at, &st.cnt[2], st, wf_stage, emitted, S, slot, scn);
        if (alive) qout[at[0]] = slot;
        if (ended) st.term[at[1]] = slot;
    }
    count_rays_block(rays, nq);
}

// TRANS: some scene of the launch has a transmissive material, so a shadow
// ray walks through water / near-clear surfaces (closest-hit steps with
// Beer-Lambert attenuation). Otherwise every first hit is opaque and the
// query is a single any-hit traversal (same answer), with a smaller stack and
// register footprint.
// LIST: only the rays k_wf_shadow_refill<., ., true> listed in st.slowq
// (count in cnt[5]): those whose segment can meet a transmissive surface.
template <int MODE, bool TRANS, bool LIST = false>
__global__ __launch_bounds__(256) void k_wf_shadow(SceneView<float> sc0, WfState st, unsigned long long* rays,
                                                    WfMulti mj)
{
#if MH_SHADOW_LDS
    MH_STACK_LDS(tsw, MH_SHADOW_LDS, 256)
#else
    MH_STACK_PRIV(tsw)
#endif
#if MH_SHADOW_ANY_LDS
    MH_STACK_LDS_REFS(tsa, MH_SHADOW_ANY_LDS, 256)
#else
    MH_STACK_PRIV(tsa)
#endif
    // One lane per ray of the dense shadow queue (LIST: of the slow list).
    const uint32_t n = LIST ? st.cnt[5] : st.cnt[2];
    int nq = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t v = LIST ? st.slowq[i] : i;
        float4 b = st.sh1[v], a = st.sh0[v];
        uint32_t tag = f2u(b.w);
        uint32_t slot = tag >> 4, k = tag & 15u;
        const SceneView<float>& sc = mode_scene<MODE>(sc0, mj, MODE == WF_TABLE ? (int)st.shq[v] : 0);
        V3<float> o = mk(a.x, a.y, a.z), d = mk(b.x, b.y, b.z);
        bool occ;
        if (TRANS) {
            V3<float> att;
            occ = occluded_segment(sc, tsw, o, d, 1e-3f, a.w, &att, &nq);
            if (!occ && k == 8 && att.x < 1e-4f && att.y < 1e-4f && att.z < 1e-4f) occ = true;
            if (!occ && sc.has_transmissive) {
                float4 c = st.pend[(size_t)slot * 9 + k];
                st.pend[(size_t)slot * 9 + k] = make_float4(c.x * att.x, c.y * att.y, c.z * att.z, 0.f);
            }
        } else {
            ++nq;
            occ = any_hit<float, decltype(tsa), 2>(sc, tsa, o, unit_dir(d), 1e-3f, a.w);
        }
        if (!occ) atomicOr(&st.vis[slot], 1u << k);
    }
    count_rays_block(rays, nq);
}

// Idle lanes of a wave at which the any-hit shadow kernel stops traversing to
// hand its idle lanes new rays (dynamic fetch); 0 selects k_wf_shadow's
// one-ray-per-lane loop instead. Swept on the 1080p terrain frame (shadow ms
// per frame), binary tree, tag-scanned queue: 0 -> 278, 8 -> 259, 16 -> 242,
// 24 -> 236, 32 -> 238, 40 -> 240, 48 -> 233, 56 -> 230; 4-wide tree, dense
// queue with prefetched entries (a refill round is about one memory round
// trip): 16 -> 209, 32 -> 194, 48 -> 185.6, 56 -> 186; rays prefetched into
// LDS by direct loads (a round reads LDS only): 16 -> 167.2, 24 -> 160.1,
// 32 -> 158.2, 48 -> 158.3, 56 -> 170.4; with the near / far node test and
// the linear stack: 24 -> 144.2, 32 -> 146.1, 48 -> 149.3.
#ifndef MH_SHADOW_REFILL
#define MH_SHADOW_REFILL 24
#endif
// Tree the any-hit shadow kernel walks: 0 = the binary tree, 1 = the 4-wide
// tree where the scene has one (children in index order), 2 = the 4-wide
// tree, children nearest first. Measured on the 1080p terrain frame (shadow
// ms per frame, node fetches per frame): 0 -> 230, 28.7 G; 1 -> 189, 13.8 G;
// 2 -> 226, 14.2 G (the sorting network costs what the fetches save).
#ifndef MH_SHADOW_W4
#define MH_SHADOW_W4 1
#endif
// 1: a leaf's primitives are tested one per step, like nodes, so every step
// of the wave waits for one load per lane; 0: the whole leaf in one step
// (a load per primitive, one after another). 1080p terrain frame, shadow ms
// per frame: 0 -> 185.6, 1 -> 163.6 (both at 80 VGPRs, 6 waves per SIMD).
// LDS entries per lane of the dynamic-fetch any-hit kernel's stack (the
// linear stack takes any depth from 8).
#ifndef MH_SHADOW_REFILL_LDS
#define MH_SHADOW_REFILL_LDS MH_SHADOW_ANY_LDS
#endif
// 1: the any-hit kernel's LDS stack is linear (K slots at a 256-lane
// stride, the older half moved to the private overflow when full); 0: a
// ring of K slots whose oldest entry moves out when full. Shadow ms per
// 1080p frame: 1 -> 146.06, 0 -> 146.05.
#ifndef MH_SHADOW_LINEAR
#define MH_SHADOW_LINEAR 1
#endif
#ifndef MH_SHADOW_PRIM_STEP
#define MH_SHADOW_PRIM_STEP 1
#endif

// Any-hit shadow queries (no transmissive material) with dynamic fetch. Each
// lane owns one ray at a time and walks the tree one node (or one leaf
// primitive) per step. Each wave reads its own chunks of MH_SHQ_CHUNK = 32
// rays of the dense shadow queue (chunk j at queue position (j * waves +
// wave) * 32) straight into LDS with direct global-to-LDS loads: two chunk
// slots per wave, each [32 origin + tmax records][32 direction + tag
// records], so a chunk is one load instruction. Whenever MH_SHADOW_REFILL of
// its lanes have finished and rays remain, the wave leaves the step loop,
// hands the prefetched rays to its idle lanes (LDS reads: the loads were
// issued at least one round earlier) and refills every chunk slot whose rays
// are all handed out; only the batched engine's scene lookup goes to
// memory. Lanes that finish early therefore take new rays instead of idling
// until the wave's longest traversal ends. Same visibility bits as
// k_wf_shadow<., false> (the query result does not depend on the order).
// TRANS (some scene of the launch has a transmissive material): a ray whose
// segment meets the scene's transmissive box (seg_meets_box) is not walked
// here but listed in st.slowq for k_wf_shadow<., true, true>, whose
// closest-hit steps pass through those surfaces; every other ray meets
// opaque surfaces only, so the any-hit walk gives its answer (and an
// attenuation of 1).
template <int MODE, bool COUNT = false, bool TRANS = false>
__global__ __launch_bounds__(256) void k_wf_shadow_refill(SceneView<float> sc0, WfState st, unsigned long long* rays,
                                                           WfMulti mj)
{
    constexpr bool TABLE = MODE == WF_TABLE;
    // one scene for the whole launch (WF_SINGLE, WF_ONE): its tree and
    // primitive arrays are read at launch-uniform bases, and 4-wide nodes
    // load only each child's near and far planes (Slab::nf)
    constexpr bool UNI = !TABLE;
    // per-lane stack of node refs: the top K in LDS (lane-interleaved, slot
    // k of a lane at my[k * blockDim.x]), older ones in a private overflow
    // array; held in scalars so that restarting it for a new ray is three
    // register writes
    constexpr int K = MH_SHADOW_REFILL_LDS;
    static_assert(K >= 8 && K < MH_STACK && (MH_SHADOW_LINEAR || (K & (K - 1)) == 0),
                  "LDS stack depth: at least 8, a power of two for the ring");
    constexpr uint32_t CH = 32u;  // rays per chunk
    __shared__ int lds_ref[K * 256];
    __shared__ float4 ring[4][2][2 * CH];    // per wave: 2 chunk slots of [origin + tmax][direction + tag]
    __shared__ uint32_t rscn[4][2][CH];      // WF_TABLE: the rays' scene indices
    int* const my = lds_ref + threadIdx.x;
#if MH_SHADOW_LINEAR
    // Linear stack: LDS slot k of this lane at my[k * 256] (a constant
    // stride: one address op per push or pop), sp entries in LDS above
    // psp older ones in the private overflow array; when the LDS part is
    // full its older half moves to the overflow (rare: the walk rarely
    // holds more than a dozen entries).
    constexpr int NO = MH_STACK;
    int ovf[NO];
    int sp = 0, psp = 0;
    auto spill = [&]() {
        constexpr int H = K / 2;
        for (int j = 0; j < H; ++j) ovf[psp + j < NO ? psp + j : NO - 1] = my[j * 256];
        for (int j = H; j < sp; ++j) my[(j - H) * 256] = my[j * 256];
        psp += H;
        sp -= H;
    };
    auto push = [&](int r) {
        if (sp == K) spill();
        my[sp * 256] = r;
        ++sp;
    };
    // room for the four pushes of a 4-wide node
    auto room4 = [&]() {
        if (sp > K - 4) spill();
    };
    // each child's ref goes to the next free slot, which it keeps only if hit
    auto put = [&](bool h, int r) {
        my[sp * 256] = r;
        sp += (int)h;
    };
    auto pop = [&]() -> int {
        // the LDS read is unconditional (a slot of this lane either way)
        int v = lds_read(&my[(sp > 0 ? sp - 1 : 0) * 256]);
        if (sp > 0) {
            --sp;
        } else {
            --psp;
            v = ovf[psp < NO - 1 ? psp : NO - 1];
        }
        return v;
    };
#else
    const int bs = (int)blockDim.x;
    // capacity MH_STACK: K - 1 ring slots (push4 keeps one free) + the overflow
    constexpr int NO = MH_STACK - K + 1;
    int ovf[NO];
    int sp = 0, base = 0, psp = 0;
    auto spill = [&]() {
        ovf[psp < NO - 1 ? psp : NO - 1] = my[base * bs];
        ++psp;
        base = (base + 1) & (K - 1);
        --sp;
    };
    auto push = [&](int r) {
        if (sp == K) spill();
        my[((base + sp) & (K - 1)) * bs] = r;
        ++sp;
    };
    // Pushes of a 4-wide node's hit children without a branch per child:
    // room4() makes the ring hold four more entries plus one free slot
    // (oldest entries to the overflow, rare); then put() writes each child's
    // ref to the next free slot and advances past it only if the child was
    // hit, so a missed child's write lands in a free slot.
    auto room4 = [&]() {
        while (sp > K - 5) spill();
    };
    auto put = [&](bool h, int r) {
        my[((base + sp) & (K - 1)) * bs] = r;
        sp += (int)h;
    };
    auto pop = [&]() -> int {
        // the LDS read is unconditional (a slot of this lane either way)
        int v = lds_read(&my[((base + sp - 1) & (K - 1)) * bs]);
        if (sp > 0) {
            --sp;
        } else {
            --psp;
            v = ovf[psp < NO - 1 ? psp : NO - 1];
        }
        return v;
    };
#endif

    const uint32_t n = st.cnt[2];
    // the wave index through readfirstlane, so that everything derived from
    // it (chunk cursors, LDS slot addresses) is known to be wave-uniform
    const int lane = (int)__lane_id(), wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // queue positions in 32 bits: n <= 9 * pool <= 9 * 2^26, and a chunk
    // start is computed only while the previous one was below n, so every
    // start stays below n + 32 * waves < 2^32
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);             // waves in the grid
    const uint32_t gw = grid_wave();
    uint32_t jn = 0;          // next chunk of the wave's sequence to load (wave-uniform)
    uint32_t rd = 0, wr = 0;  // rays of the sequence handed out / loaded so far (wave-uniform)
    bool seq_more = gw * CH < n;  // chunks of the sequence remain unloaded (wave-uniform)
    // Load chunk jn into slot jn & 1: lanes 0..31 the origin records, lanes
    // 32..63 the direction records, each lane's 16 bytes at slot + 16 * lane
    // (positions past the queue end read its last ray and are never handed out).
    auto load_chunk = [&]() {
        const uint32_t start = (jn * nw + gw) * CH;
        const uint32_t valid = n - start < CH ? n - start : CH;
        uint32_t p = start + ((uint32_t)lane & (CH - 1u));
        if (p > n - 1u) p = n - 1u;
        auto* dst = (__attribute__((address_space(3))) void*)&ring[wave][jn & 1u][0];
        if (lane < (int)CH) {
            __builtin_amdgcn_global_load_lds((const void*)(st.sh0 + p), dst, 16, 0, 0);
            if (TABLE)
                __builtin_amdgcn_global_load_lds((const void*)(st.shq + p),
                                                 (__attribute__((address_space(3))) void*)&rscn[wave][jn & 1u][0], 4,
                                                 0, 0);
        } else {
            __builtin_amdgcn_global_load_lds((const void*)(st.sh1 + p), dst, 16, 0, 0);
        }
        wr += valid;
        ++jn;
        seq_more = valid == CH && (jn * nw + gw) * CH < n;
    };
    int nq = 0;
    bool busy = false;
    Slab<float> sl;
    V3<float> o = zero3<float>(), d = zero3<float>();
    float tmax = 0.f;
    uint32_t tag = 0;  // slot << 4 | estimate of the lane's ray
    // the lane's tree (WF_TABLE): the 4-wide one when MH_SHADOW_W4 and the
    // scene has one (flagged in bit 0: node arrays are 256-byte aligned),
    // else the binary one
    uintptr_t tree = 0;
    const P32* prims = nullptr;
    // UNI: the launch's scene, and the lane's near-row offsets (Slab::near_off)
    const UScene scu = uniform_scene(mode_scene<MODE>(sc0, mj, 0));
    uint32_t nrx = 0, nry = 0, nrz = 0;
    // MH_SHADOW_PRIM_STEP: the current leaf's untested primitives, the next
    // one's index in bits 0..27 (tree primitives number at most 2^27) and
    // how many remain in bits 28..31
    uint32_t leaf = 0;
    uint32_t cnt_q = 0, cnt_n = 0, cnt_t = 0, cnt_o = 0;  // COUNT: queries, nodes, triangles, other prims
    for (;;) {
        const uint64_t idle = __ballot(!busy);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        const uint32_t avail = wr - rd;
        if (nidle == 64u && avail == 0u && !seq_more) break;
        // the loaded rays to the idle lanes, in lane order
        const uint32_t take = avail < nidle ? avail : nidle;
        bool got = false, slow = false;
        uint32_t scn = 0u, qpos = 0u;
        if (take > 0u) {
            // the chunk loads of earlier rounds have landed in LDS: s_waitcnt
            // vmcnt(0) (the traversal steps leave no other load in flight)
            __builtin_amdgcn_s_waitcnt(0x0f70);
            uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                      __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
            if (!busy && rank < take) {
                const uint32_t s = rd + rank;
                const float4 a = ring[wave][(s / CH) & 1u][s & (CH - 1u)];
                const float4 b = ring[wave][(s / CH) & 1u][CH + (s & (CH - 1u))];
                if (TABLE) scn = rscn[wave][(s / CH) & 1u][s & (CH - 1u)];
                o = mk(a.x, a.y, a.z);
                tmax = a.w;
                d = mk(b.x, b.y, b.z);
                tag = f2u(b.w);
                got = true;
                // the ray's queue position (its chunk's start + its place)
                qpos = ((s / CH) * nw + gw) * CH + (s & (CH - 1u));
            }
            rd += take;
        }
        // every LDS read of this round is done before a chunk load may
        // overwrite its slot: s_waitcnt lgkmcnt(0)
        __builtin_amdgcn_s_waitcnt(0xc07f);
        // a slot is free once every ray of the chunk loaded two before into it
        // is handed out
        if (seq_more && (jn < 2u || rd >= (jn - 1u) * CH)) load_chunk();
        if (seq_more && (jn < 2u || rd >= (jn - 1u) * CH)) load_chunk();
        if (got) {
            const SceneView<float>& sc = mode_scene<MODE>(sc0, mj, (int)scn);
            d = unit_dir(d);
            if (TRANS)
                slow = (UNI ? scu.has_transmissive : sc.has_transmissive) &&
                       seg_meets_box(o, d, tmax, UNI ? scu.tlo : sc.tlo, UNI ? scu.thi : sc.thi);
        }
        if (TRANS) {
            // the slow rays to the list: one atomic per wave
            const uint64_t sm = __ballot(slow);
            if (sm) {
                const int first = __builtin_ctzll(sm);
                uint32_t at = 0u;
                if (lane == first) at = atomicAdd(&st.cnt[5], (uint32_t)__popcll(sm));
                at = (uint32_t)__shfl((int)at, first);
                if (slow)
                    st.slowq[at + __builtin_amdgcn_mbcnt_hi((uint32_t)(sm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)sm, 0u))] = qpos;
            }
        }
        if (got && !slow) {
            const SceneView<float>& sc = mode_scene<MODE>(sc0, mj, (int)scn);
            const int n_planes = UNI ? scu.n_planes : sc.n_planes;
            const P32* planes = UNI ? scu.planes : sc.planes;
            const int root_ref = UNI ? scu.root_ref : sc.root_ref;
            ++nq;
            // planes first (never in the tree), as any_hit does
            bool occ = false;
            for (int pl = 0; pl < n_planes && !occ; ++pl) {
                occ = occ_prim(load_prim(&planes[pl]), o, d, 1e-3f, tmax);
                if (COUNT) ++cnt_o;
            }
            if (COUNT) ++cnt_q;
            if (!occ && root_ref != INT_MIN) {
                sl.init(o, d);
#if MH_SHADOW_LINEAR
                sp = psp = 0;
#else
                sp = base = psp = 0;
#endif
                leaf = 0;
                if (UNI) {
                    nrx = Slab<float>::near_off(sl.ix);
                    nry = Slab<float>::near_off(sl.iy);
                    nrz = Slab<float>::near_off(sl.iz);
                } else {
                    prims = sc.prims;
                }
                if (UNI) {
                    push(MH_SHADOW_W4 && scu.nodes4 ? scu.root4 : scu.root_ref);
                } else if (MH_SHADOW_W4 && sc.nodes4) {
                    tree = (uintptr_t)sc.nodes4 | 1u;
                    push(sc.root4);
                } else {
                    tree = (uintptr_t)sc.nodes;
                    push(sc.root_ref);
                }
                busy = true;
            } else if (!occ) {
                atomicOr(&st.vis[tag >> 4], 1u << (tag & 15u));
            }
        }
        const bool more = wr > rd || seq_more;  // wave-uniform
        // traversal steps until the wave is done, or enough lanes are idle
        // while work remains
        for (;;) {
            if (busy) {
                bool occ = false;
                int pj = -1;  // MH_SHADOW_PRIM_STEP: the primitive this lane tests in this step
                if (MH_SHADOW_PRIM_STEP && (leaf >> 28)) {
                    pj = (int)(leaf & 0x0fffffffu);
                    leaf += 1u - (1u << 28);
                } else {
                int ref = pop();
                const bool wide = UNI ? (MH_SHADOW_W4 && scu.nodes4 != nullptr) : (tree & 1u) != 0u;
                if (MH_SHADOW_W4 && ref >= 0 && wide) {
                    if (COUNT) ++cnt_n;
                    if (UNI && MH_SHADOW_W4 == 1) {
                        // each child's near and far planes: rows picked by
                        // address (node arrays stay below 4 GiB: 32-bit offsets)
                        const char* nb = (const char*)scu.nodes4;
                        const uint32_t off = (uint32_t)ref << 7;
                        const uint32_t vx = off | nrx, vy = off | nry, vz = off | nrz;
                        const float4 nx4 = ldg((const float4*)(nb + (size_t)vx));
                        const float4 fx4 = ldg((const float4*)(nb + (size_t)(vx ^ 16u)));
                        const float4 ny4 = ldg((const float4*)(nb + (size_t)vy + 32));
                        const float4 fy4 = ldg((const float4*)(nb + (size_t)(vy ^ 16u) + 32));
                        const float4 nz4 = ldg((const float4*)(nb + (size_t)vz + 64));
                        const float4 fz4 = ldg((const float4*)(nb + (size_t)(vz ^ 16u) + 64));
                        const int4 rf = ldg((const int4*)(nb + (size_t)off + 96));
                        room4();
                        put(sl.nf(nx4.w, fx4.w, ny4.w, fy4.w, nz4.w, fz4.w, 1e-3f, tmax) && rf.w != INT_MIN, rf.w);
                        put(sl.nf(nx4.z, fx4.z, ny4.z, fy4.z, nz4.z, fz4.z, 1e-3f, tmax) && rf.z != INT_MIN, rf.z);
                        put(sl.nf(nx4.y, fx4.y, ny4.y, fy4.y, nz4.y, fz4.y, 1e-3f, tmax), rf.y);
                        put(sl.nf(nx4.x, fx4.x, ny4.x, fy4.x, nz4.x, fz4.x, 1e-3f, tmax), rf.x);
                    } else if (MH_SHADOW_W4 == 2) {
                        DNode4 nd = ldg(&(UNI ? scu.nodes4 : (const DNode4*)(tree & ~(uintptr_t)1))[ref]);
                        int r[4], nh;
                        float t[4];
                        children4(sl, nd, 1e-3f, tmax, r, t, &nh);
                        for (int i = 3; i >= 0; --i)
                            if (i < nh) push(r[i]);
                    } else {
                        // boxes and refs only: a slot past the child count has ref INT_MIN
                        DNode4Box nd = ldg(
                            (const DNode4Box*)&(UNI ? scu.nodes4 : (const DNode4*)(tree & ~(uintptr_t)1))[ref]);
                        bool h;
                        float t;
                        room4();
                        sl.one(nd.lox.w, nd.hix.w, nd.loy.w, nd.hiy.w, nd.loz.w, nd.hiz.w, 1e-3f, tmax, &h, &t);
                        put(h && nd.ref.w != INT_MIN, nd.ref.w);
                        sl.one(nd.lox.z, nd.hix.z, nd.loy.z, nd.hiy.z, nd.loz.z, nd.hiz.z, 1e-3f, tmax, &h, &t);
                        put(h && nd.ref.z != INT_MIN, nd.ref.z);
                        sl.one(nd.lox.y, nd.hix.y, nd.loy.y, nd.hiy.y, nd.loz.y, nd.hiz.y, 1e-3f, tmax, &h, &t);
                        put(h, nd.ref.y);
                        sl.one(nd.lox.x, nd.hix.x, nd.loy.x, nd.hiy.x, nd.loz.x, nd.hiz.x, 1e-3f, tmax, &h, &t);
                        put(h, nd.ref.x);
                    }
                } else if (ref >= 0) {
                    DNode nd = ldg(&(UNI ? scu.nodes : (const DNode*)tree)[ref]);
                    if (COUNT) ++cnt_n;
                    bool hl, hr;
                    float tl, tr;
                    sl.test(nd, 1e-3f, tmax, &hl, &hr, &tl, &tr);
                    if (hl && hr) {
                        bool lfirst = tl <= tr;
                        push(lfirst ? nd.c3.y : nd.c3.x);
                        push(lfirst ? nd.c3.x : nd.c3.y);
                    } else if (hl) {
                        push(nd.c3.x);
                    } else if (hr) {
                        push(nd.c3.y);
                    }
                } else {
                    int code = ~ref;
                    int first = code >> 3, cnt = (code & 7) + 1;
                    if (MH_SHADOW_PRIM_STEP) {
                        pj = first;
                        leaf = (uint32_t)(first + 1) | ((uint32_t)(cnt - 1) << 28);
                    } else {
                        for (int j = 0; j < cnt && !occ; ++j) {
                            PD<float> pr = load_prim(&(UNI ? scu.prims : prims)[first + j]);
                            occ = occ_prim(pr, o, d, 1e-3f, tmax);
                            if (COUNT) {
                                if (pr.kind == K_TRIANGLE) ++cnt_t;
                                else ++cnt_o;
                            }
                        }
                    }
                }
                }
                if (MH_SHADOW_PRIM_STEP && pj >= 0) {
                    PD<float> pr = load_prim(&(UNI ? scu.prims : prims)[pj]);
                    occ = occ_prim(pr, o, d, 1e-3f, tmax);
                    if (COUNT) {
                        if (pr.kind == K_TRIANGLE) ++cnt_t;
                        else ++cnt_o;
                    }
                }
                if (occ || ((leaf >> 28) == 0u && sp == 0 && psp == 0)) {
                    if (!occ) atomicOr(&st.vis[tag >> 4], 1u << (tag & 15u));
                    busy = false;
                }
            }
            const uint64_t b = __ballot(busy);
            if (b == 0ull) break;
            if (more && 64u - (uint32_t)__popcll(b) >= (uint32_t)MH_SHADOW_REFILL) break;
        }
    }
    if (COUNT) {
        const uint32_t c[4] = {cnt_q, cnt_n, cnt_t, cnt_o};
        add_visits<4>(mj.visits + 4, c);
    }
    count_rays_block(rays, nq);
}

// Radiance in 2^-24 fixed point. A sample is capped at 2^26 (no physical
// estimate here comes near it), i.e. 2^50 units, so a pixel sum of up to
// MH_BATCH_MAX_SPP = 2^14 - 1 samples stays below 2^64 (mh_render_batch
// rejects jobs of more samples).
__device__ __forceinline__ unsigned long long to_fixed(float v)
{
    return __float2ull_rn(fminf(v, 67108864.0f) * 16777216.0f);
}

// defer (single-scene engine, batches of whole samples): each path's
// radiance stays in its slot for k_wf_single_sum.
template <int MODE>
__global__ __launch_bounds__(256) void k_wf_finalize(WfState st, WfMulti mj, int defer)
{
    constexpr bool MULTI = MODE != WF_SINGLE;
    uint32_t n = st.cnt[3];
    // path regeneration appends the freed slots: block-uniform trip count
    uint32_t n_round = (MULTI && mj.stream) ? (n + blockDim.x - 1u) / blockDim.x * blockDim.x : n;
    uint32_t* const cq[1] = {&st.cnt[4]};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_round; i += gridDim.x * blockDim.x) {
        bool active = i < n;
        uint32_t slot = 0;
        if (active) {
            slot = st.term[i];
            float4 c = st.p2[slot], e = st.p3[slot];
            V3<float> L = resolve_pending(st, slot, mk(e.x, e.y, e.z));
            st.npend[slot] = 0;
            if (!finite3(L)) L = zero3<float>();
            L.x = L.x < 0.f ? 0.f : L.x;
            L.y = L.y < 0.f ? 0.f : L.y;
            L.z = L.z < 0.f ? 0.f : L.z;
            uint32_t pix = f2u(c.w);
            if ((MULTI && !mj.stream) || (!MULTI && defer)) {
                // fixed batches: the slot keeps the path's radiance until
                // k_wf_batch_sum / k_wf_single_sum adds the batch's samples
                // of each pixel
                st.p3[slot] = make_float4(L.x, L.y, L.z, e.w);
            } else if (MULTI) {
                size_t p = ((size_t)mj.jobs[st.jid[slot]].acc_off + pix) * 3;
                atomicAdd(&mj.acc64[p], to_fixed(L.x));
                atomicAdd(&mj.acc64[p + 1], to_fixed(L.y));
                atomicAdd(&mj.acc64[p + 2], to_fixed(L.z));
            } else {
                // one sample per pixel per batch: no other thread touches this pixel
                st.acc[(size_t)pix * 3] += L.x;
                st.acc[(size_t)pix * 3 + 1] += L.y;
                st.acc[(size_t)pix * 3 + 2] += L.z;
            }
        }
        if (MULTI && mj.stream) {
            // the slot is free for the next step's new paths
            const bool want[1] = {active};
            uint32_t at[1];
            block_append<1>(cq, want, at);
            if (active) st.freeq[at[0]] = slot;
        }
    }
}

// Fixed batches of the batched engine: slot i holds global path g0 + i for
// the whole batch, and k_wf_finalize leaves each path's final radiance in its
// slot (every path has ended after the last bounce). One thread per pixel of
// each job the batch covers adds the pixel's samples of the batch, each in
// 2^-24 fixed point as k_wf_finalize's atomics would, to its sums: the same
// integers, with no atomics (a pixel's samples in one batch go to one thread).
__global__ __launch_bounds__(256) void k_wf_batch_sum(WfState st, WfMulti mj, int pool)
{
    const long long g0 = (long long)st.batch[0];
    const long long g1 = g0 + pool < mj.total ? g0 + pool : mj.total;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (int j = 0; j < mj.n_jobs; ++j) {
        const BJob& bj = mj.jobs[j];
        const long long npix = bj.npix;
        const long long jb = bj.begin, je = bj.begin + npix * bj.spp;
        // the job's paths in this batch: job-local indices [lo, hi)
        const long long lo = (g0 > jb ? g0 : jb) - jb, hi = (g1 < je ? g1 : je) - jb;
        if (lo >= hi) continue;
        for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < npix; t += stride) {
            // samples s of tile-order pixel t with lo <= s * npix + t < hi
            long long l = lo > t ? t + (lo - t + npix - 1) / npix * npix : t;
            if (l >= hi) continue;
            unsigned long long f0 = 0ull, f1 = 0ull, f2 = 0ull;
            for (; l < hi; l += npix) {
                const float4 e = st.p3[(size_t)(jb + l - g0)];
                f0 += to_fixed(e.x);
                f1 += to_fixed(e.y);
                f2 += to_fixed(e.z);
            }
            const size_t p = ((size_t)bj.acc_off + tile_pixel((uint32_t)t, (uint32_t)bj.w, (uint32_t)bj.h)) * 3;
            mj.acc64[p] += f0;
            mj.acc64[p + 1] += f1;
            mj.acc64[p + 2] += f2;
        }
    }
}

// Single-scene engine, a batch of k whole samples (pool = k * npix): each
// pixel adds its k samples' radiance to its fp32 sum in sample order (slot
// j * npix + t holds sample j of tile-order pixel t), the same additions in
// the same order as k batches of one sample each.
__global__ __launch_bounds__(256) void k_wf_single_sum(WfState st, WfJob job)
{
    const long long g0 = (long long)st.batch[0];
    const long long remain = job.total - g0;
    const long long n = remain < (long long)job.pool ? remain : (long long)job.pool;
    const int k = (int)(n / job.npix);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < job.npix; t += gridDim.x * blockDim.x) {
        if (k <= 0) break;
        const uint32_t pix = f2u(st.p2[t].w);
        float sx = st.acc[(size_t)pix * 3], sy = st.acc[(size_t)pix * 3 + 1], sz = st.acc[(size_t)pix * 3 + 2];
        for (int j = 0; j < k; ++j) {
            const float4 e = st.p3[(size_t)j * (size_t)job.npix + (size_t)t];
            sx += e.x;
            sy += e.y;
            sz += e.z;
        }
        st.acc[(size_t)pix * 3] = sx;
        st.acc[(size_t)pix * 3 + 1] = sy;
        st.acc[(size_t)pix * 3 + 2] = sz;
    }
}

__global__ __launch_bounds__(256) void k_wf_next_batch(WfState st, int pool)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) st.batch[0] += (unsigned long long)pool;
}

__global__ __launch_bounds__(256) void k_wf_clear(WfState st, int npix)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npix * 3; i += gridDim.x * blockDim.x) st.acc[i] = 0.f;
    if (blockIdx.x == 0 && threadIdx.x == 0) st.batch[0] = 0;
}

__global__ __launch_bounds__(256) void k_wf_resolve(WfState st, int npix, int spp, LaunchOut out)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        float k = 1.0f / (float)spp;
        float r = st.acc[(size_t)i * 3] * k, g = st.acc[(size_t)i * 3 + 1] * k, b = st.acc[(size_t)i * 3 + 2] * k;
        out.linear[(size_t)i * 3] = r;
        out.linear[(size_t)i * 3 + 1] = g;
        out.linear[(size_t)i * 3 + 2] = b;
        out.rgb[(size_t)i * 3] = to_byte(r);
        out.rgb[(size_t)i * 3 + 1] = to_byte(g);
        out.rgb[(size_t)i * 3 + 2] = to_byte(b);
    }
}

// Batched engine: zero the fixed-point sums and the batch cursor.
__global__ __launch_bounds__(256) void k_wf_multi_clear(WfState st, WfMulti mj)
{
    size_t n = (size_t)mj.total_pix * 3;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        mj.acc64[i] = 0ull;
    // the batched megakernel runs without the wavefront state
    if (st.batch && blockIdx.x == 0 && threadIdx.x == 0) st.batch[0] = 0;
}

// Batched engine: per-pixel means and tone map of every job, written to the
// concatenated output (job order, each job top-down RGB).
// ---- batched megakernel ------------------------------------------------------
//
// The same jobs as the batched wavefront engine, as work units of one
// 16x16-pixel tile and a range of samples of one job (BUnit), one
// workgroup per unit in a single launch: every lane runs whole paths
// (pt_radiance, the megakernel's estimator) and the units of all jobs fill
// the machine together. Each sample's radiance goes to 2^-24 fixed point
// before it is summed, so the result does not depend on how samples are cut
// into units or in which order units run.
template <bool FOG>
__global__ __launch_bounds__(256) void k_batch_mega(WfMulti mj, const BUnit* units, int n_units,
                                                    unsigned long long* rays)
{
    MH_STACK_PRIV(ts)
    int nq = 0;
    if ((int)blockIdx.x < n_units) {
        BUnit u = units[blockIdx.x];
        const BJob& job = mj.jobs[u.job];
        const SceneView<float>& sc = mj.scenes[job.scene];
        const CamView<float>& cam = job.cam;
        int w = job.w, h = job.h;
        int tiles_x = (w + 15) / 16;
        int x = (u.tile % tiles_x) * 16 + (int)(threadIdx.x & 15);
        int y = (u.tile / tiles_x) * 16 + (int)(threadIdx.x >> 4);
        if (x < w && y < h) {
            unsigned long long f[3] = {0ull, 0ull, 0ull};
            for (int si = u.s_begin; si < u.s_end; ++si) {
                uint32_t s = (uint32_t)(job.s0 + si);
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
                V3<float> L = pt_radiance<FOG>(sc, ts, o, d, job.depth, hash3((uint32_t)x, (uint32_t)y, s), &nq);
                f[0] += to_fixed(L.x);
                f[1] += to_fixed(L.y);
                f[2] += to_fixed(L.z);
            }
            size_t p = ((size_t)job.acc_off + (size_t)y * (size_t)w + (size_t)x) * 3;
            atomicAdd(&mj.acc64[p], f[0]);
            atomicAdd(&mj.acc64[p + 1], f[1]);
            atomicAdd(&mj.acc64[p + 2], f[2]);
        }
    }
    count_rays(rays, nq);
}

__global__ __launch_bounds__(256) void k_wf_multi_resolve(WfMulti mj, LaunchOut out)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < mj.total_pix;
         i += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = mj.n_jobs - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (mj.jobs[mid].acc_off <= i) lo = mid;
            else hi = mid - 1;
        }
        double k = 1.0 / (16777216.0 * (double)mj.jobs[lo].spp);
        for (int c = 0; c < 3; ++c) {
            float v = (float)((double)mj.acc64[i * 3 + c] * k);
            out.linear[i * 3 + c] = v;
            out.rgb[i * 3 + c] = to_byte(v);
        }
    }
}

} // namespace mh
