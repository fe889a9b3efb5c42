// mh_build.hip - BVH build on the device and the device scene layout.
//
// Binned SAH, top-down, with the split decisions of the host builder
// (mt_scene_build_bvh, host/mt_scene.c): node bounds and centroid bounds,
// `bins` bins on every axis of nonzero centroid extent, the sweep in the
// host's order with its strict comparisons, a leaf when count <= max_leaf and
// count <= 1 + cost / area. Each double operation is the host's (this file is
// compiled without FMA contraction), each sum is exact (counts) or an
// order-independent min / max, so every node splits its primitive SET as the
// host's does. What may differ: the order of prims inside a leaf (stable
// partitions here, a two-ended swap there) and, where all centroids of a node
// of more than 8 prims coincide, which half the object-median split takes.
//
// Large nodes (> SMALL prims) are split level by level: a node's slots are cut
// into chunks of at most CH (work items); one block reduces a chunk in LDS and
// merges it into the node's accumulators with 64-bit atomics (doubles mapped
// to order-preserving integer keys); one thread per node decides; a stable
// partition through per-chunk prefix sums moves the slots. Subtrees of at
// most SMALL prims are built by one thread each. Node numbers, task lists and
// output positions all come from prefix sums, so the result is the same on
// every run.
#include "mh_build.h"
#include "mh_pack.h"

#include <cfloat>
#include <climits>
#include <cmath>

namespace mh {
namespace {

constexpr int CH = 2048;   // slots per work item
constexpr int SMALL = 64;  // subtrees of at most this many prims: one thread each
constexpr int MAXB = 64;   // bin count limit (the host's)
constexpr int BT = 256;    // block size of the per-chunk kernels

struct Task {
    int node, begin, end, depth;
};
struct STask {
    int node, begin, end, depth, rbase, pad0, pad1, pad2;
};
struct Work {
    int task, begin, end, pad;
};
// kind 0: leaf, 1: SAH split (axis, bin), 2: object-median split at mid
struct Dec {
    int kind, axis, bin, mid;
};
struct BinMap {
    double lo[3], scale[3];
    int valid[3], pad;
};
enum { C_NT = 0, C_NODES, C_NSMALL, C_RESERVE, C_DEPTH, C_LEAVES, C_BAD, C_N };

// ---- doubles as order-preserving 64-bit keys (exact min / max by atomics)

__device__ __forceinline__ unsigned long long dkey(double x)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dval(unsigned long long k)
{
    unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}
constexpr unsigned long long KMIN0 = ~0ull;  // identity of min
constexpr unsigned long long KMAX0 = 0ull;   // identity of max

// ---- the host builder's box arithmetic

__device__ void prim_box(const mh_prim& p, double* mn, double* mx)
{
    const mh_vec3 &a = p.p0, &c = p.p1, &e = p.p2;
    switch (p.kind) {
    case MH_PRIM_SPHERE:
        mn[0] = a.x - p.r0; mn[1] = a.y - p.r0; mn[2] = a.z - p.r0;
        mx[0] = a.x + p.r0; mx[1] = a.y + p.r0; mx[2] = a.z + p.r0;
        break;
    case MH_PRIM_BOX:
        mn[0] = a.x - fabs(c.x); mn[1] = a.y - fabs(c.y); mn[2] = a.z - fabs(c.z);
        mx[0] = a.x + fabs(c.x); mx[1] = a.y + fabs(c.y); mx[2] = a.z + fabs(c.z);
        break;
    case MH_PRIM_TRIANGLE:
        mn[0] = fmin(a.x, fmin(c.x, e.x)); mx[0] = fmax(a.x, fmax(c.x, e.x));
        mn[1] = fmin(a.y, fmin(c.y, e.y)); mx[1] = fmax(a.y, fmax(c.y, e.y));
        mn[2] = fmin(a.z, fmin(c.z, e.z)); mx[2] = fmax(a.z, fmax(c.z, e.z));
        break;
    case MH_PRIM_CYLINDER: {
        double r = p.r0 > p.r1 ? p.r0 : p.r1;
        if (r < 0.0) r = 0.0;
        mn[0] = fmin(a.x, c.x) - r; mn[1] = fmin(a.y, c.y) - r; mn[2] = fmin(a.z, c.z) - r;
        mx[0] = fmax(a.x, c.x) + r; mx[1] = fmax(a.y, c.y) + r; mx[2] = fmax(a.z, c.z) + r;
        break;
    }
    default:
        for (int k = 0; k < 3; ++k) {
            mn[k] = DBL_MAX;
            mx[k] = -DBL_MAX;
        }
    }
}

__device__ __forceinline__ double box_area(const double* mn, const double* mx)
{
    double dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
    if (!(dx > 0.0)) dx = 0.0;
    if (!(dy > 0.0)) dy = 0.0;
    if (!(dz > 0.0)) dz = 0.0;
    return 2.0 * (dx * dy + dy * dz + dz * dx);
}

__device__ __forceinline__ void box_empty(double* mn, double* mx)
{
    for (int k = 0; k < 3; ++k) {
        mn[k] = DBL_MAX;
        mx[k] = -DBL_MAX;
    }
}

__device__ __forceinline__ void box_grow(double* mn, double* mx, const double* omn, const double* omx)
{
    for (int k = 0; k < 3; ++k) {
        if (omn[k] < mn[k]) mn[k] = omn[k];
        if (omx[k] > mx[k]) mx[k] = omx[k];
    }
}

__device__ __forceinline__ float round_down(double x)
{
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
}
__device__ __forceinline__ float round_up(double x)
{
    float f = (float)x;
    if ((double)f < x) f = nextafterf(f, INFINITY);
    return f;
}

__device__ __forceinline__ int bin_of(double c, double lo, double scale, int nbins)
{
    int b = (int)((c - lo) * scale);
    if (b >= nbins) b = nbins - 1;
    if (b < 0) b = 0;
    return b;
}

__device__ void write_bounds(mh_bvh_node* nd, const double* mn, const double* mx)
{
    for (int k = 0; k < 3; ++k) {
        nd->bmin[k] = round_down(mn[k] - 1e-7);
        nd->bmax[k] = round_up(mx[k] + 1e-7);
    }
}

// The host's sweep of one axis: `bc` / `bmn` / `bmx` are the bins. Updates
// the best split (strictly smaller cost wins, axes and bins in order).
__device__ void sweep_axis(int ax, int nbins, const int* bc, const double (*bmn)[3], const double (*bmx)[3],
                           double* best_cost, int* best_axis, int* best_bin)
{
    double ra[MAXB];
    int rc[MAXB];
    double amn[3], amx[3];
    box_empty(amn, amx);
    int n = 0;
    for (int b = nbins - 1; b > 0; --b) {
        if (bc[b]) {
            box_grow(amn, amx, bmn[b], bmx[b]);
            n += bc[b];
        }
        ra[b] = n ? box_area(amn, amx) : 0.0;
        rc[b] = n;
    }
    box_empty(amn, amx);
    n = 0;
    for (int b = 0; b < nbins - 1; ++b) {
        if (bc[b]) {
            box_grow(amn, amx, bmn[b], bmx[b]);
            n += bc[b];
        }
        if (n == 0 || rc[b + 1] == 0) continue;
        double c = box_area(amn, amx) * (double)n + ra[b + 1] * (double)rc[b + 1];
        if (c < *best_cost) {
            *best_cost = c;
            *best_axis = ax;
            *best_bin = b;
        }
    }
}

// ---- prefix sums ------------------------------------------------------------

__global__ __launch_bounds__(BT) void k_scan_block(const int* in, int* out, int n, int* bsum, int* total)
{
    __shared__ int sh[BT];
    int base = blockIdx.x * (BT * 4) + threadIdx.x * 4;
    int v[4], s = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0;
        s += v[k];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < BT; off <<= 1) {
        int t = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += t;
        __syncthreads();
    }
    int run = sh[threadIdx.x] - s;
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == BT - 1) {
        bsum[blockIdx.x] = sh[BT - 1];
        if (total) *total = sh[BT - 1];
    }
}

__global__ void k_scan_add(int* out, int n, const int* boff)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += boff[i / (BT * 4)];
}

__global__ void k_set_int(int* p, int v) { *p = v; }

// Exclusive prefix sum of in[0, n) into out (in == out allowed); the sum of
// all into *total (device). scratch: >= 2 * ceil(n / 1024) + 64 ints.
hipError_t scan_excl(const int* in, int* out, int n, int* total, int* scratch, hipStream_t st)
{
    if (n <= 0) {
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, total, 0);
        return hipGetLastError();
    }
    int nb = (n + BT * 4 - 1) / (BT * 4);
    int* bsum = scratch;
    int* boff = scratch + nb;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(BT), 0, st, in, out, n, bsum, nb == 1 ? total : nullptr);
    if (nb > 1) {
        hipError_t e = scan_excl(bsum, boff, nb, total, scratch + 2 * nb, st);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_add, dim3((n + 255) / 256), dim3(256), 0, st, out, n, boff);
    }
    return hipGetLastError();
}

// ---- primitives -------------------------------------------------------------

__global__ void k_not_plane(const mh_prim* prims, int n, int* flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = prims[i].kind != MH_PRIM_PLANE ? 1 : 0;
}

// non-plane prims in file order -> idx, planes in file order -> planes
__global__ void k_split_lists(const mh_prim* prims, int n, const int* pos, int* idx, int* planes)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (prims[i].kind != MH_PRIM_PLANE) idx[pos[i]] = i;
    else planes[i - pos[i]] = i;
}

__global__ void k_boxes(const mh_prim* prims, const int* idx, int nt, double* pb, double* cen, int* slot)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    double mn[3], mx[3];
    prim_box(prims[idx[j]], mn, mx);
    for (int k = 0; k < 3; ++k) {
        pb[6 * (size_t)j + k] = mn[k];
        pb[6 * (size_t)j + 3 + k] = mx[k];
        cen[3 * (size_t)j + k] = 0.5 * (mn[k] + mx[k]);
    }
    slot[j] = j;
}

// ---- large nodes, level by level -----------------------------------------------

__global__ void k_chunks(const Task* tasks, int nT, int* nch)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nT) nch[t] = (tasks[t].end - tasks[t].begin + CH - 1) / CH;
}

__global__ void k_emit_work(const Task* tasks, int nT, const int* wbase, const int* nch, Work* work)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    Task tk = tasks[t];
    for (int k = 0; k < nch[t]; ++k) {
        int b = tk.begin + k * CH;
        int e = b + CH < tk.end ? b + CH : tk.end;
        work[wbase[t] + k] = Work{t, b, e, 0};
    }
}

// accumulators per task: 12 bound keys (node min, node max, centroid min,
// centroid max), then per axis and bin a count and 6 box keys
__global__ void k_clear_acc(unsigned long long* kb, unsigned* bc, unsigned long long* bk, int nT, int nbins)
{
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t nk = (size_t)nT * 12, nc = (size_t)nT * 3 * nbins, nbk = nc * 6;
    if (i < nk) kb[i] = (i % 12) % 6 < 3 ? KMIN0 : KMAX0;
    if (i < nc) bc[i] = 0;
    if (i < nbk) bk[i] = (i % 6) < 3 ? KMIN0 : KMAX0;
}

__global__ __launch_bounds__(BT) void k_bounds(const Work* work, const int* slot, const double* pb,
                                               const double* cen, unsigned long long* kb)
{
    __shared__ unsigned long long sk[12];
    if (threadIdx.x < 12) sk[threadIdx.x] = (threadIdx.x % 6) < 3 ? KMIN0 : KMAX0;
    __syncthreads();
    Work w = work[blockIdx.x];
    double mn[6], mx[6];
    for (int k = 0; k < 6; ++k) {
        mn[k] = DBL_MAX;
        mx[k] = -DBL_MAX;
    }
    bool any = false;
    for (int j = w.begin + threadIdx.x; j < w.end; j += BT) {
        int q = slot[j];
        any = true;
        for (int k = 0; k < 3; ++k) {
            double a = pb[6 * (size_t)q + k], b = pb[6 * (size_t)q + 3 + k], c = cen[3 * (size_t)q + k];
            if (a < mn[k]) mn[k] = a;
            if (b > mx[k]) mx[k] = b;
            if (c < mn[3 + k]) mn[3 + k] = c;
            if (c > mx[3 + k]) mx[3 + k] = c;
        }
    }
    if (any) {
        for (int k = 0; k < 3; ++k) {
            atomicMin(&sk[k], dkey(mn[k]));
            atomicMax(&sk[3 + k], dkey(mx[k]));
            atomicMin(&sk[6 + k], dkey(mn[3 + k]));
            atomicMax(&sk[9 + k], dkey(mx[3 + k]));
        }
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        unsigned long long* dst = &kb[(size_t)w.task * 12 + threadIdx.x];
        if ((threadIdx.x % 6) < 3) atomicMin(dst, sk[threadIdx.x]);
        else atomicMax(dst, sk[threadIdx.x]);
    }
}

__device__ void load_bounds(const unsigned long long* kb, int t, double* nmn, double* nmx, double* cmn, double* cmx)
{
    const unsigned long long* k = kb + (size_t)t * 12;
    for (int a = 0; a < 3; ++a) {
        nmn[a] = dval(k[a]);
        nmx[a] = dval(k[3 + a]);
        cmn[a] = dval(k[6 + a]);
        cmx[a] = dval(k[9 + a]);
    }
}

// The host's bin mapping per axis: lo = centroid min, scale = bins / extent,
// axes of zero extent skipped.
__global__ void k_binmap(const unsigned long long* kb, int nT, int nbins, BinMap* bm)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    double nmn[3], nmx[3], cmn[3], cmx[3];
    load_bounds(kb, t, nmn, nmx, cmn, cmx);
    BinMap m;
    for (int ax = 0; ax < 3; ++ax) {
        double lo = cmn[ax], ext = cmx[ax] - cmn[ax];
        m.valid[ax] = ext > 0.0 ? 1 : 0;
        m.lo[ax] = lo;
        m.scale[ax] = m.valid[ax] ? (double)nbins / ext : 0.0;
    }
    m.pad = 0;
    bm[t] = m;
}

__global__ __launch_bounds__(BT) void k_bins(const Work* work, const int* slot, const double* pb, const double* cen,
                                             const BinMap* bm, int nbins, unsigned* bc, unsigned long long* bk)
{
    extern __shared__ unsigned long long lds[];
    unsigned long long* lk = lds;                       // 3 * nbins * 6 keys
    unsigned* lc = (unsigned*)(lds + 3 * nbins * 6);    // 3 * nbins counts
    int nb3 = 3 * nbins;
    for (int i = threadIdx.x; i < nb3 * 6; i += BT) lk[i] = (i % 6) < 3 ? KMIN0 : KMAX0;
    for (int i = threadIdx.x; i < nb3; i += BT) lc[i] = 0u;
    __syncthreads();
    Work w = work[blockIdx.x];
    BinMap m = bm[w.task];
    for (int j = w.begin + threadIdx.x; j < w.end; j += BT) {
        int q = slot[j];
        double b6[6];
        for (int k = 0; k < 6; ++k) b6[k] = pb[6 * (size_t)q + k];
        for (int ax = 0; ax < 3; ++ax) {
            if (!m.valid[ax]) continue;
            int b = bin_of(cen[3 * (size_t)q + ax], m.lo[ax], m.scale[ax], nbins);
            int e = ax * nbins + b;
            atomicAdd(&lc[e], 1u);
            for (int k = 0; k < 3; ++k) {
                atomicMin(&lk[e * 6 + k], dkey(b6[k]));
                atomicMax(&lk[e * 6 + 3 + k], dkey(b6[3 + k]));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb3; i += BT) {
        if (!lc[i]) continue;
        size_t g = (size_t)w.task * nb3 + i;
        atomicAdd(&bc[g], lc[i]);
        for (int k = 0; k < 3; ++k) {
            atomicMin(&bk[g * 6 + k], lk[i * 6 + k]);
            atomicMax(&bk[g * 6 + 3 + k], lk[i * 6 + 3 + k]);
        }
    }
}

__global__ void k_decide(const Task* tasks, int nT, const unsigned long long* kb, const unsigned* bc,
                         const unsigned long long* bk, const BinMap* bm, int nbins, int max_leaf, mh_bvh_node* nodes,
                         Dec* dec, int* ctr)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    Task tk = tasks[t];
    int count = tk.end - tk.begin;
    double nmn[3], nmx[3], cmn[3], cmx[3];
    load_bounds(kb, t, nmn, nmx, cmn, cmx);
    mh_bvh_node* nd = &nodes[tk.node];
    write_bounds(nd, nmn, nmx);
    double area = box_area(nmn, nmx);
    atomicMax(&ctr[C_DEPTH], tk.depth);
    int best_axis = -1, best_bin = -1;
    double best_cost = DBL_MAX;
    if (count > 1) {
        for (int ax = 0; ax < 3; ++ax) {
            if (!bm[t].valid[ax]) continue;
            int cnt[MAXB];
            double bmn[MAXB][3], bmx[MAXB][3];
            size_t g = ((size_t)t * 3 + ax) * nbins;
            for (int b = 0; b < nbins; ++b) {
                cnt[b] = (int)bc[g + b];
                for (int k = 0; k < 3; ++k) {
                    bmn[b][k] = dval(bk[(g + b) * 6 + k]);
                    bmx[b][k] = dval(bk[(g + b) * 6 + 3 + k]);
                }
            }
            sweep_axis(ax, nbins, cnt, bmn, bmx, &best_cost, &best_axis, &best_bin);
        }
    }
    double split_cost = best_axis >= 0 ? 1.0 + best_cost / (area > 0.0 ? area : 1.0) : DBL_MAX;
    bool leaf = count <= 1 || (count <= max_leaf && (double)count <= split_cost);
    Dec d{0, best_axis, best_bin, -1};
    if (!leaf) {
        if (best_axis >= 0) d.kind = 1;
        else if (count <= 8) leaf = true;
        else {
            d.kind = 2;
            d.mid = tk.begin + count / 2;
        }
    }
    if (leaf) {
        d.kind = 0;
        nd->left = nd->right = -1;
        nd->first = tk.begin;
        nd->count = count;
        atomicAdd(&ctr[C_LEAVES], 1);
    }
    dec[t] = d;
}

__global__ __launch_bounds__(BT) void k_count_left(const Work* work, const Dec* dec, const BinMap* bm,
                                                   const int* slot, const double* cen, int nbins, int* cl)
{
    __shared__ int sh[BT];
    Work w = work[blockIdx.x];
    Dec d = dec[w.task];
    int n = 0;
    if (d.kind == 1) {
        double lo = bm[w.task].lo[d.axis], scale = bm[w.task].scale[d.axis];
        for (int j = w.begin + threadIdx.x; j < w.end; j += BT)
            n += bin_of(cen[3 * (size_t)slot[j] + d.axis], lo, scale, nbins) <= d.bin ? 1 : 0;
    }
    sh[threadIdx.x] = n;
    __syncthreads();
    for (int off = BT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < (unsigned)off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) cl[blockIdx.x] = sh[0];
}

__global__ void k_chunk_offsets(const Task* tasks, int nT, const int* wbase, const int* nch, const int* cl,
                                int* lofs, Dec* dec)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT || dec[t].kind != 1) return;
    int run = 0;
    for (int k = 0; k < nch[t]; ++k) {
        lofs[wbase[t] + k] = run;
        run += cl[wbase[t] + k];
    }
    dec[t].mid = tasks[t].begin + run;
}

// Stable partition of one chunk into tmp: the left elements of the node in
// slot order from node.begin, the right ones from mid.
__global__ __launch_bounds__(BT) void k_scatter(const Work* work, const Task* tasks, const Dec* dec, const BinMap* bm,
                                                const int* slot, const double* cen, int nbins, const int* lofs,
                                                int* tmp)
{
    constexpr int E = CH / BT;
    __shared__ int sh[BT];
    Work w = work[blockIdx.x];
    Dec d = dec[w.task];
    if (d.kind != 1) return;
    Task tk = tasks[w.task];
    double lo = bm[w.task].lo[d.axis], scale = bm[w.task].scale[d.axis];
    int j0 = w.begin + threadIdx.x * E;
    int q[E];
    bool left[E];
    int nl = 0;
    for (int k = 0; k < E; ++k) {
        int j = j0 + k;
        left[k] = false;
        q[k] = 0;
        if (j < w.end) {
            q[k] = slot[j];
            left[k] = bin_of(cen[3 * (size_t)q[k] + d.axis], lo, scale, nbins) <= d.bin;
            nl += left[k] ? 1 : 0;
        }
    }
    sh[threadIdx.x] = nl;
    __syncthreads();
    for (int off = 1; off < BT; off <<= 1) {
        int v = threadIdx.x >= (unsigned)off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    int lr = sh[threadIdx.x] - nl;                     // left elements before this thread's
    int lbase = tk.begin + lofs[blockIdx.x];            // this chunk's first left position
    int rbase = d.mid + (w.begin - tk.begin - lofs[blockIdx.x]);
    for (int k = 0; k < E; ++k) {
        int j = j0 + k;
        if (j >= w.end) break;
        int o = j - w.begin;                            // offset in the chunk
        if (left[k]) tmp[lbase + lr++] = q[k];
        else tmp[rbase + (o - lr)] = q[k];
    }
}

__global__ void k_copy_back(const Work* work, const Dec* dec, const int* tmp, int* slot)
{
    Work w = work[blockIdx.x];
    if (dec[w.task].kind != 1) return;
    for (int j = w.begin + threadIdx.x; j < w.end; j += blockDim.x) slot[j] = tmp[j];
}

// One block: children of every split node get consecutive numbers (left,
// right) in task order; children larger than SMALL become the next level's
// tasks, the others small tasks with a reserve of 2 count - 1 node slots.
__global__ __launch_bounds__(1024) void k_next_level(const Task* tasks, int nT, const Dec* dec, mh_bvh_node* nodes,
                                                     Task* next, STask* small, int* ctr)
{
    __shared__ int sh[4][1024];
    __shared__ int carry[4];
    int node0 = ctr[C_NODES], small0 = ctr[C_NSMALL], res0 = ctr[C_RESERVE];
    if (threadIdx.x < 4) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int c0 = 0; c0 < nT; c0 += 1024) {
        int t = c0 + threadIdx.x;
        Task tk{0, 0, 0, 0};
        Dec d{0, -1, -1, -1};
        if (t < nT) {
            tk = tasks[t];
            d = dec[t];
        }
        bool split = t < nT && d.kind != 0;
        int cL = split ? d.mid - tk.begin : 0, cR = split ? tk.end - d.mid : 0;
        bool bigL = split && cL > SMALL, bigR = split && cR > SMALL;
        bool smL = split && !bigL, smR = split && !bigR;
        int v[4] = {split ? 1 : 0, (bigL ? 1 : 0) + (bigR ? 1 : 0), (smL ? 1 : 0) + (smR ? 1 : 0),
                    (smL ? 2 * cL - 1 : 0) + (smR ? 2 * cR - 1 : 0)};
        for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = v[k];
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {
            int u[4];
            for (int k = 0; k < 4; ++k) u[k] = threadIdx.x >= (unsigned)off ? sh[k][threadIdx.x - off] : 0;
            __syncthreads();
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += u[k];
            __syncthreads();
        }
        int ex[4];
        for (int k = 0; k < 4; ++k) ex[k] = carry[k] + sh[k][threadIdx.x] - v[k];
        if (split && (cL <= 0 || cR <= 0)) atomicAdd(&ctr[C_BAD], 1);  // never: both sides hold prims
        if (split && cL > 0 && cR > 0) {
            int l = node0 + 2 * ex[0], r = l + 1;
            mh_bvh_node* nd = &nodes[tk.node];
            nd->left = l;
            nd->right = r;
            nd->first = 0;
            nd->count = 0;
            int li = ex[1], si = small0 + ex[2], rb = res0 + ex[3];
            if (bigL) next[li++] = Task{l, tk.begin, d.mid, tk.depth + 1};
            else {
                small[si++] = STask{l, tk.begin, d.mid, tk.depth + 1, rb, 0, 0, 0};
                rb += 2 * cL - 1;
            }
            if (bigR) next[li++] = Task{r, d.mid, tk.end, tk.depth + 1};
            else small[si++] = STask{r, d.mid, tk.end, tk.depth + 1, rb, 0, 0, 0};
        }
        __syncthreads();
        if (threadIdx.x == 1023)
            for (int k = 0; k < 4; ++k) carry[k] += sh[k][1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ctr[C_NT] = carry[1];
        ctr[C_NODES] = node0 + 2 * carry[0];
        ctr[C_NSMALL] = small0 + carry[2];
        ctr[C_RESERVE] = res0 + carry[3];
    }
}

// ---- small subtrees: one thread each, the host loop --------------------------

template <int NB>
__global__ __launch_bounds__(64) void k_small(const STask* small, int n_small, int* slot, const double* pb,
                                              const double* cen, mh_bvh_node* reserve, int nbins, int max_leaf,
                                              int* used, int* ctr)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_small) return;
    STask t = small[i];
    mh_bvh_node* R = reserve + t.rbase;   // local node j at R[j]; 0 = the subtree root
    struct LT {
        int node, begin, end, depth;
    };
    LT stk[SMALL + 2];
    int sp = 0, nn = 1, maxd = 0, leaves = 0;
    stk[sp++] = LT{0, t.begin, t.end, t.depth};
    while (sp > 0) {
        LT k = stk[--sp];
        int count = k.end - k.begin;
        double nmn[3], nmx[3], cmn[3], cmx[3];
        box_empty(nmn, nmx);
        box_empty(cmn, cmx);
        for (int j = k.begin; j < k.end; ++j) {
            int q = slot[j];
            box_grow(nmn, nmx, &pb[6 * (size_t)q], &pb[6 * (size_t)q + 3]);
            box_grow(cmn, cmx, &cen[3 * (size_t)q], &cen[3 * (size_t)q]);
        }
        mh_bvh_node* nd = &R[k.node];
        write_bounds(nd, nmn, nmx);
        double area = box_area(nmn, nmx);
        if (k.depth > maxd) maxd = k.depth;
        int best_axis = -1, best_bin = -1;
        double best_cost = DBL_MAX;
        if (count > 1) {
            for (int ax = 0; ax < 3; ++ax) {
                double lo = cmn[ax], ext = cmx[ax] - cmn[ax];
                if (!(ext > 0.0)) continue;
                double scale = (double)nbins / ext;
                int cnt[NB];
                double bmn[NB][3], bmx[NB][3];
                for (int b = 0; b < nbins; ++b) {
                    box_empty(bmn[b], bmx[b]);
                    cnt[b] = 0;
                }
                for (int j = k.begin; j < k.end; ++j) {
                    int q = slot[j];
                    int b = bin_of(cen[3 * (size_t)q + ax], lo, scale, nbins);
                    cnt[b]++;
                    box_grow(bmn[b], bmx[b], &pb[6 * (size_t)q], &pb[6 * (size_t)q + 3]);
                }
                sweep_axis(ax, nbins, cnt, bmn, bmx, &best_cost, &best_axis, &best_bin);
            }
        }
        double split_cost = best_axis >= 0 ? 1.0 + best_cost / (area > 0.0 ? area : 1.0) : DBL_MAX;
        bool leaf = count <= 1 || (count <= max_leaf && (double)count <= split_cost);
        int mid = -1;
        if (!leaf) {
            if (best_axis >= 0) {
                double lo = cmn[best_axis], scale = (double)nbins / (cmx[best_axis] - cmn[best_axis]);
                int a = k.begin, b = k.end - 1;
                while (a <= b) {
                    int q = slot[a];
                    if (bin_of(cen[3 * (size_t)q + best_axis], lo, scale, nbins) <= best_bin) {
                        ++a;
                    } else {
                        slot[a] = slot[b];
                        slot[b] = q;
                        --b;
                    }
                }
                mid = a;
            }
            if (mid <= k.begin || mid >= k.end) {
                if (count <= 8 && best_axis < 0) leaf = true;
                else mid = k.begin + count / 2;
            }
        }
        if (leaf) {
            nd->left = nd->right = -1;
            nd->first = k.begin;
            nd->count = count;
            ++leaves;
            continue;
        }
        int l = nn++, r = nn++;
        nd->left = l;
        nd->right = r;
        nd->first = 0;
        nd->count = 0;
        stk[sp++] = LT{r, mid, k.end, k.depth + 1};
        stk[sp++] = LT{l, k.begin, mid, k.depth + 1};
    }
    used[i] = nn;
    atomicMax(&ctr[C_DEPTH], maxd);
    atomicAdd(&ctr[C_LEAVES], leaves);
}

__global__ void k_used_extra(const int* used, int n, int* extra)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) extra[i] = used[i] - 1;
}

// Move each subtree from its reserve into the final array: local 0 -> the
// subtree root's number, local j -> node0 + base + j - 1.
__global__ void k_small_place(const STask* small, int n_small, const int* used, const int* base, int node0,
                              const mh_bvh_node* reserve, mh_bvh_node* nodes)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_small) return;
    STask t = small[i];
    int b = node0 + base[i];
    for (int j = 0; j < used[i]; ++j) {
        mh_bvh_node nd = reserve[t.rbase + j];
        if (nd.count == 0) {
            nd.left = b + nd.left - 1;
            nd.right = b + nd.right - 1;
        }
        nodes[j == 0 ? t.node : b + j - 1] = nd;
    }
}

// ---- device layout ------------------------------------------------------------

__global__ void k_inner_flag(const mh_bvh_node* nodes, int n, int* flag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = nodes[i].count == 0 ? 1 : 0;
}

__device__ __forceinline__ int ref_of(const mh_bvh_node* nodes, const int* dev_index, int c)
{
    const mh_bvh_node& nd = nodes[c];
    if (nd.count > 0) return ~((nd.first << 3) | (nd.count - 1));
    return dev_index[c];
}

__global__ void k_dnodes(const mh_bvh_node* nodes, int n, const int* dev_index, DNode* out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || nodes[i].count != 0) return;
    const mh_bvh_node& L = nodes[nodes[i].left];
    const mh_bvh_node& Rn = nodes[nodes[i].right];
    DNode d;
    d.c0 = make_float4(L.bmin[0], L.bmax[0], L.bmin[1], L.bmax[1]);
    d.c1 = make_float4(L.bmin[2], L.bmax[2], Rn.bmin[0], Rn.bmax[0]);
    d.c2 = make_float4(Rn.bmin[1], Rn.bmax[1], Rn.bmin[2], Rn.bmax[2]);
    d.c3 = make_int4(ref_of(nodes, dev_index, nodes[i].left), ref_of(nodes, dev_index, nodes[i].right), 0, 0);
    out[dev_index[i]] = d;
}

__device__ __forceinline__ double area4(const mh_bvh_node& b)
{
    double dx = (double)b.bmax[0] - b.bmin[0], dy = (double)b.bmax[1] - b.bmin[1], dz = (double)b.bmax[2] - b.bmin[2];
    return dx * dy + dy * dz + dz * dx;
}

// 4-wide collapse of one binary node (the host's greedy rule: replace the
// largest-area interior child by its two children, in place, until four).
__device__ int collapse4(const mh_bvh_node* nodes, int n2, int* c)
{
    int m = 2;
    c[0] = nodes[n2].left;
    c[1] = nodes[n2].right;
    while (m < 4) {
        int best = -1;
        double ba = -1.0;
        for (int i = 0; i < m; ++i)
            if (nodes[c[i]].count == 0 && area4(nodes[c[i]]) > ba) {
                ba = area4(nodes[c[i]]);
                best = i;
            }
        if (best < 0) break;
        const mh_bvh_node& e = nodes[c[best]];
        for (int i = m; i > best + 1; --i) c[i] = c[i - 1];
        c[best] = e.left;
        c[best + 1] = e.right;
        ++m;
    }
    return m;
}

__global__ void k4_count(const mh_bvh_node* nodes, const int* frontier, int nF, int* ninner)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    int c[4];
    int m = collapse4(nodes, frontier[f], c);
    int k = 0;
    for (int i = 0; i < m; ++i) k += nodes[c[i]].count == 0 ? 1 : 0;
    ninner[f] = k;
}

__global__ void k4_write(const mh_bvh_node* nodes, const int* frontier, int nF, const int* off, int base, DNode4* out,
                         int* next)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nF) return;
    int c[4];
    int m = collapse4(nodes, frontier[f], c);
    float lo[3][4], hi[3][4];
    int ref[4];
    int k = 0;
    for (int i = 0; i < 4; ++i) {
        for (int a = 0; a < 3; ++a) {
            lo[a][i] = 0.f;
            hi[a][i] = 0.f;
        }
        ref[i] = INT_MIN;
        if (i >= m) continue;
        const mh_bvh_node& ch = nodes[c[i]];
        for (int a = 0; a < 3; ++a) {
            lo[a][i] = ch.bmin[a];
            hi[a][i] = ch.bmax[a];
        }
        if (ch.count > 0) {
            ref[i] = ~((ch.first << 3) | (ch.count - 1));
        } else {
            ref[i] = base + nF + off[f] + k;
            next[off[f] + k] = c[i];
            ++k;
        }
    }
    DNode4 o;
    o.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    o.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    o.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    o.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    o.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    o.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    o.ref = make_int4(ref[0], ref[1], ref[2], ref[3]);
    o.meta = make_int4(m, 0, 0, 0);
    out[base + f] = o;
}

__global__ void k_gather(const mh_prim* prims, const mh_prim_attr* attrs, bool any_attr, const int* idx,
                         const int* slot, int nt, int32_t* leaf_prims, P32* p32, P64* p64, DAttr<float>* a32,
                         DAttr<double>* a64)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nt) return;
    int pi = idx[slot[i]];
    leaf_prims[i] = pi;
    mh_prim q = prims[pi];
    if (!attrs) q.flags = 0;
    P32 r = pack32(q);
    set_prim_index32(r, pi);
    p32[i] = r;
    p64[i] = pack64(q, pi);
    if (any_attr) {
        a32[i] = conv_attr<float>(attrs[pi]);
        a64[i] = conv_attr<double>(attrs[pi]);
    }
}

__global__ void k_gather_planes(const mh_prim* prims, const int32_t* planes, int np, P32* pl32, P64* pl64)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    int pi = planes[i];
    P32 r = pack32(prims[pi]);
    set_prim_index32(r, pi);
    pl32[i] = r;
    pl64[i] = pack64(prims[pi], pi);
}

inline dim3 grid_of(long long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }

} // namespace

#define BTRY(expr)                                                                                    \
    do {                                                                                              \
        hipError_t e_ = (expr);                                                                       \
        if (e_ != hipSuccess) return (int)e_;                                                         \
    } while (0)

int gpu_build_scene(const mh_prim* d_prims, const mh_prim_attr* d_attrs, int n, bool any_attr, int max_leaf,
                    int bins, hipStream_t st, GpuBvh* out, int* err_depth)
{
    *err_depth = 0;
    max_leaf = max_leaf > 0 ? (max_leaf > 8 ? 8 : max_leaf) : 4;
    int nbins = bins > 1 ? (bins > MAXB ? MAXB : bins) : 16;
    std::vector<void*> temp;
    auto dalloc = [&](void** p, size_t bytes, bool keep) -> hipError_t {
        *p = nullptr;
        if (bytes == 0) return hipSuccess;
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) return e;
        if (keep) {
            out->owned.push_back(*p);
            out->bytes += bytes;
        } else {
            temp.push_back(*p);
        }
        return hipSuccess;
    };
    struct Freer {
        std::vector<void*>* v;
        ~Freer()
        {
            for (void* p : *v) (void)hipFree(p);
        }
    } freer{&temp};
    auto tmp_int = [&](int** p, size_t count) { return dalloc((void**)p, count * sizeof(int), false); };

    int *ctr, *scan_tmp;
    BTRY(tmp_int(&ctr, C_N + 4));
    // scans run over at most 2 n entries (tree nodes)
    BTRY(tmp_int(&scan_tmp, 4 * ((size_t)n / 1024 + 2) + 256));
    std::vector<int> hc(C_N + 4, 0);
    auto read_ctr = [&]() -> hipError_t {
        hipError_t e = hipMemcpyAsync(hc.data(), ctr, sizeof(int) * (C_N + 4), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return e;
    };

    // prims: non-plane list (file order) and planes
    int *flag, *pos;
    BTRY(tmp_int(&flag, (size_t)n + 1));
    BTRY(tmp_int(&pos, (size_t)n + 1));
    if (n > 0) hipLaunchKernelGGL(k_not_plane, grid_of(n), dim3(256), 0, st, d_prims, n, flag);
    BTRY(scan_excl(flag, pos, n, ctr + C_N, scan_tmp, st));
    BTRY(read_ctr());
    int nt = hc[C_N], np = n - nt;
    int *idx;
    BTRY(tmp_int(&idx, (size_t)nt + 1));
    BTRY(dalloc((void**)&out->planes, sizeof(int32_t) * (size_t)np, true));
    if (n > 0) hipLaunchKernelGGL(k_split_lists, grid_of(n), dim3(256), 0, st, d_prims, n, pos, idx, out->planes);
    out->n_planes = np;

    // per-prim boxes / centroids (the host's arithmetic) and the slot order
    double *pb, *cen;
    int *slot, *tmp;
    BTRY(dalloc((void**)&pb, sizeof(double) * 6 * (size_t)nt, false));
    BTRY(dalloc((void**)&cen, sizeof(double) * 3 * (size_t)nt, false));
    BTRY(tmp_int(&slot, (size_t)nt + 1));
    BTRY(tmp_int(&tmp, (size_t)nt + 1));
    if (nt > 0) hipLaunchKernelGGL(k_boxes, grid_of(nt), dim3(256), 0, st, d_prims, idx, nt, pb, cen, slot);

    mh_bvh_node* nodes = nullptr;
    BTRY(dalloc((void**)&nodes, sizeof(mh_bvh_node) * (size_t)(nt > 0 ? 2 * nt : 1), true));
    out->nodes = nodes;
    int n_nodes = 0, n_small = 0;
    STask* small = nullptr;
    mh_bvh_node* reserve = nullptr;
    if (nt > 0) {
        BTRY(dalloc((void**)&small, sizeof(STask) * (size_t)nt, false));
        BTRY(dalloc((void**)&reserve, sizeof(mh_bvh_node) * 2 * (size_t)nt, false));
        int max_tasks = nt / (SMALL + 1) + 2;
        Task *ta, *tb;
        BTRY(dalloc((void**)&ta, sizeof(Task) * (size_t)max_tasks, false));
        BTRY(dalloc((void**)&tb, sizeof(Task) * (size_t)max_tasks, false));
        std::vector<int> init(C_N + 4, 0);
        int nT = 0;
        Task root{0, 0, nt, 1};
        STask sroot{0, 0, nt, 1, 0, 0, 0, 0};
        if (nt > SMALL) {
            BTRY(hipMemcpyAsync(ta, &root, sizeof root, hipMemcpyHostToDevice, st));
            nT = 1;
        } else {
            BTRY(hipMemcpyAsync(small, &sroot, sizeof sroot, hipMemcpyHostToDevice, st));
            init[C_NSMALL] = 1;
            init[C_RESERVE] = 2 * nt - 1;
        }
        init[C_NODES] = 1;
        BTRY(hipMemcpyAsync(ctr, init.data(), sizeof(int) * (C_N + 4), hipMemcpyHostToDevice, st));
        BTRY(hipStreamSynchronize(st));

        // level buffers, sized for the largest level
        int max_work = nt / CH + max_tasks + 1;
        int *nch, *wbase, *cl, *lofs;
        Work* work;
        Dec* dec;
        BinMap* bm;
        unsigned long long *kb, *bk;
        unsigned* bc;
        BTRY(tmp_int(&nch, (size_t)max_tasks));
        BTRY(tmp_int(&wbase, (size_t)max_tasks));
        BTRY(tmp_int(&cl, (size_t)max_work));
        BTRY(tmp_int(&lofs, (size_t)max_work));
        BTRY(dalloc((void**)&work, sizeof(Work) * (size_t)max_work, false));
        BTRY(dalloc((void**)&dec, sizeof(Dec) * (size_t)max_tasks, false));
        BTRY(dalloc((void**)&bm, sizeof(BinMap) * (size_t)max_tasks, false));
        BTRY(dalloc((void**)&kb, sizeof(unsigned long long) * 12 * (size_t)max_tasks, false));
        BTRY(dalloc((void**)&bc, sizeof(unsigned) * 3 * nbins * (size_t)max_tasks, false));
        BTRY(dalloc((void**)&bk, sizeof(unsigned long long) * 18 * nbins * (size_t)max_tasks, false));
        size_t bins_lds = sizeof(unsigned long long) * 18 * nbins + sizeof(unsigned) * 3 * nbins;
        int levels = 0;

        while (nT > 0) {
            hipLaunchKernelGGL(k_chunks, grid_of(nT), dim3(256), 0, st, ta, nT, nch);
            BTRY(scan_excl(nch, wbase, nT, ctr + C_N, scan_tmp, st));
            BTRY(read_ctr());
            int nW = hc[C_N];
            if (nW > max_work) return (int)hipErrorUnknown;
            hipLaunchKernelGGL(k_emit_work, grid_of(nT), dim3(256), 0, st, ta, nT, wbase, nch, work);
            size_t nclr = (size_t)nT * 18 * nbins;
            hipLaunchKernelGGL(k_clear_acc, grid_of((long long)nclr), dim3(256), 0, st, kb, bc, bk, nT, nbins);
            hipLaunchKernelGGL(k_bounds, dim3(nW), dim3(BT), 0, st, work, slot, pb, cen, kb);
            hipLaunchKernelGGL(k_binmap, grid_of(nT), dim3(256), 0, st, kb, nT, nbins, bm);
            hipLaunchKernelGGL(k_bins, dim3(nW), dim3(BT), bins_lds, st, work, slot, pb, cen, bm, nbins, bc, bk);
            hipLaunchKernelGGL(k_decide, grid_of(nT, 64), dim3(64), 0, st, ta, nT, kb, bc, bk, bm, nbins, max_leaf,
                               nodes, dec, ctr);
            hipLaunchKernelGGL(k_count_left, dim3(nW), dim3(BT), 0, st, work, dec, bm, slot, cen, nbins, cl);
            hipLaunchKernelGGL(k_chunk_offsets, grid_of(nT), dim3(256), 0, st, ta, nT, wbase, nch, cl, lofs, dec);
            hipLaunchKernelGGL(k_scatter, dim3(nW), dim3(BT), 0, st, work, ta, dec, bm, slot, cen, nbins, lofs, tmp);
            hipLaunchKernelGGL(k_copy_back, dim3(nW), dim3(BT), 0, st, work, dec, tmp, slot);
            hipLaunchKernelGGL(k_next_level, dim3(1), dim3(1024), 0, st, ta, nT, dec, nodes, tb, small, ctr);
            BTRY(hipGetLastError());
            BTRY(read_ctr());
            nT = hc[C_NT];
            // children are strictly smaller than their parent, so the levels
            // end; a count out of bounds means a broken invariant
            if (nT > max_tasks || hc[C_BAD] != 0 || ++levels > 200) return (int)hipErrorUnknown;
            std::swap(ta, tb);
        }
        BTRY(read_ctr());
        n_small = hc[C_NSMALL];
        int node0 = hc[C_NODES];
        if (hc[C_RESERVE] > 2 * nt || n_small > nt) return (int)hipErrorUnknown;
        int *used, *extra;
        BTRY(tmp_int(&used, (size_t)n_small + 1));
        BTRY(tmp_int(&extra, (size_t)n_small + 1));
        if (nbins <= 16)
            hipLaunchKernelGGL(k_small<16>, grid_of(n_small, 64), dim3(64), 0, st, small, n_small, slot, pb, cen,
                               reserve, nbins, max_leaf, used, ctr);
        else
            hipLaunchKernelGGL(k_small<MAXB>, grid_of(n_small, 64), dim3(64), 0, st, small, n_small, slot, pb, cen,
                               reserve, nbins, max_leaf, used, ctr);
        hipLaunchKernelGGL(k_used_extra, grid_of(n_small), dim3(256), 0, st, used, n_small, extra);
        BTRY(scan_excl(extra, extra, n_small, ctr + C_N, scan_tmp, st));
        hipLaunchKernelGGL(k_small_place, grid_of(n_small, 64), dim3(64), 0, st, small, n_small, used, extra, node0,
                           reserve, nodes);
        BTRY(hipGetLastError());
        BTRY(read_ctr());
        n_nodes = node0 + hc[C_N];
        out->max_depth = hc[C_DEPTH];
        out->n_leaves = hc[C_LEAVES];
        if (out->max_depth > 60) {
            *err_depth = out->max_depth;
            return 0;
        }
    }
    out->n_nodes = n_nodes;
    out->n_leaf = nt;

    // device layout: interior nodes numbered by a prefix sum over the tree's
    // array, refs as the kernels read them
    if (n_nodes > 0) {
        int *dev_index, *iflag;
        BTRY(tmp_int(&dev_index, (size_t)n_nodes + 1));
        BTRY(tmp_int(&iflag, (size_t)n_nodes + 1));
        hipLaunchKernelGGL(k_inner_flag, grid_of(n_nodes), dim3(256), 0, st, nodes, n_nodes, iflag);
        BTRY(scan_excl(iflag, dev_index, n_nodes, ctr + C_N, scan_tmp, st));
        BTRY(read_ctr());
        out->n_inner = hc[C_N];
        BTRY(dalloc((void**)&out->dnodes, sizeof(DNode) * (size_t)out->n_inner, true));
        hipLaunchKernelGGL(k_dnodes, grid_of(n_nodes), dim3(256), 0, st, nodes, n_nodes, dev_index, out->dnodes);
        mh_bvh_node root;
        BTRY(hipMemcpyAsync(&root, nodes, sizeof root, hipMemcpyDeviceToHost, st));
        BTRY(hipStreamSynchronize(st));
        out->root_ref = root.count > 0 ? ~((root.first << 3) | (root.count - 1)) : 0;
        out->root4 = out->root_ref;
        // 4-wide tree, breadth first
        if (root.count == 0) {
            BTRY(dalloc((void**)&out->dnodes4, sizeof(DNode4) * (size_t)out->n_inner, true));
            int *fa, *fb, *ninner;
            BTRY(tmp_int(&fa, (size_t)out->n_inner + 1));
            BTRY(tmp_int(&fb, (size_t)out->n_inner + 1));
            BTRY(tmp_int(&ninner, (size_t)out->n_inner + 1));
            int zero = 0;
            BTRY(hipMemcpyAsync(fa, &zero, sizeof zero, hipMemcpyHostToDevice, st));
            int nF = 1, base = 0, levels = 0;
            while (nF > 0) {
                hipLaunchKernelGGL(k4_count, grid_of(nF), dim3(256), 0, st, nodes, fa, nF, ninner);
                BTRY(scan_excl(ninner, ninner, nF, ctr + C_N, scan_tmp, st));
                hipLaunchKernelGGL(k4_write, grid_of(nF), dim3(256), 0, st, nodes, fa, nF, ninner, base, out->dnodes4,
                                   fb);
                BTRY(hipGetLastError());
                BTRY(read_ctr());
                base += nF;
                nF = hc[C_N];
                ++levels;
                if (base + nF > out->n_inner) return (int)hipErrorUnknown;
                std::swap(fa, fb);
            }
            out->n_nodes4 = base;
            out->depth4 = levels;
            out->root4 = 0;
        }
    }

    // leaf-ordered primitive records, planes
    BTRY(dalloc((void**)&out->leaf_prims, sizeof(int32_t) * (size_t)nt, true));
    BTRY(dalloc((void**)&out->p32, sizeof(P32) * (size_t)nt, true));
    BTRY(dalloc((void**)&out->p64, sizeof(P64) * (size_t)nt, true));
    if (any_attr) {
        BTRY(dalloc((void**)&out->a32, sizeof(DAttr<float>) * (size_t)nt, true));
        BTRY(dalloc((void**)&out->a64, sizeof(DAttr<double>) * (size_t)nt, true));
    }
    if (nt > 0)
        hipLaunchKernelGGL(k_gather, grid_of(nt), dim3(256), 0, st, d_prims, d_attrs, any_attr, idx, slot, nt,
                           out->leaf_prims, out->p32, out->p64, out->a32, out->a64);
    BTRY(dalloc((void**)&out->pl32, sizeof(P32) * (size_t)np, true));
    BTRY(dalloc((void**)&out->pl64, sizeof(P64) * (size_t)np, true));
    if (np > 0)
        hipLaunchKernelGGL(k_gather_planes, grid_of(np), dim3(256), 0, st, d_prims, out->planes, np, out->pl32,
                           out->pl64);
    BTRY(hipGetLastError());
    BTRY(hipStreamSynchronize(st));
    return 0;
}

} // namespace mh
