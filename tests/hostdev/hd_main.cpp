// hd_main.cpp - the library's render code compiled for the CPU (test
// infrastructure only): reads a flattened scene and a render request written
// by tests/hostdev/hostdev.py, builds the device layout with the library's
// own mh_layout.h, renders every pixel with the megakernels' per-pixel
// functions (pt_pixel / whitted_pixel, traversal on the private stack), and
// writes the linear means, the 8-bit image and the ray count. Built plain,
// under AddressSanitizer + UndefinedBehaviorSanitizer, and with every local
// pre-filled by pattern or zero (an image that differs between the two
// reads an uninitialised value) by tests/hostdev/Makefile.
//
//   hd_render <request file> <output file>
#include "mh_layout.h"
#include "mh_render.h"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

int mh::layout_fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vfprintf(stderr, fmt, ap);
    va_end(ap);
    fputc('\n', stderr);
    return code;
}

using namespace mh;

namespace {

struct Request {
    int32_t n_prims, has_attrs, n_materials, n_nodes, n_leaf, n_planes;
    int32_t w, h, spp, depth, mode, precision, s0, threads;
    double water_level;
    int32_t water_material, width;  // width 2: the binary tree, else the 4-wide one when built
    mh_sky sky;
    mh_fog fog;
    mh_ocean ocean;
    mh_camera cam;
};

template <typename T> bool read_n(FILE* f, std::vector<T>* v, size_t n)
{
    v->resize(n);
    return n == 0 || fread(v->data(), sizeof(T), n, f) == n;
}

template <typename R>
void render(const mh_scene_soa& soa, const Request& q, std::vector<float>* lin, std::vector<unsigned char>* rgb,
            long long* rays)
{
    HostLayout hl;
    if (build_host_layout(&soa, &hl)) exit(3);
    Layout lay;
    lay.nodes = hl.nodes.data();
    lay.nodes4 = hl.nodes4.data();
    lay.n_nodes = (int)hl.nodes.size();
    lay.root_ref = hl.root_ref;
    lay.root4 = hl.root4;
    lay.bvh4 = hl.bvh4;
    lay.p32 = hl.p32.data();
    lay.p64 = hl.p64.data();
    lay.pl32 = hl.pl32.data();
    lay.pl64 = hl.pl64.data();
    lay.plidx = soa.planes;
    lay.a32 = hl.a32.empty() ? nullptr : hl.a32.data();
    lay.a64 = hl.a64.empty() ? nullptr : hl.a64.data();
    lay.m32 = hl.m32.data();
    lay.m64 = hl.m64.data();
    lay.n_planes = soa.n_planes;
    DLight<float> l32[8];
    DLight<double> l64[8];
    int nl = collect_lights(light_spheres(&soa), soa.materials, soa.n_materials, l32, l64);
    SceneView<R> v{};
    int has_water = any_water(soa.materials, soa.n_materials);
    int tr = any_transmissive(soa.materials, soa.n_materials);
    float tlo[3], thi[3];
    const std::vector<double> mb = material_boxes(&soa);
    transmissive_box(mb, soa.materials, soa.n_materials, tlo, thi);
    const int split = trans_split_pays(mb, tlo, thi);
    if constexpr (sizeof(R) == 4) fill_view(v, lay, &soa, has_water, l32, nl, tr, tlo, thi, split);
    else fill_view(v, lay, &soa, has_water, l64, nl, tr, tlo, thi, split);
    if (q.width == 2) v.bvh4 = 0;
    CamView<R> cam = conv_cam<R>(q.cam);
    const int w = q.w, h = q.h;
    lin->assign((size_t)w * h * 3, 0.f);
    rgb->assign((size_t)w * h * 3, 0);
    std::atomic<int> next{0};
    std::atomic<long long> nrays{0};
    const int cap = 17 * (q.depth + 2) + 4;
    int g = (int)(sqrt((double)q.spp) + 0.5);
    bool strat = g > 0 && g * g == q.spp;
    auto work = [&]() {
        PrivStack ts;
        std::vector<WItem<R>> stk((size_t)cap);
        int nq = 0, ovf = 0;
        for (int y; (y = next.fetch_add(1)) < h;) {
            for (int x = 0; x < w; ++x) {
                V3<R> acc;
                if (q.mode == MH_MODE_WHITTED)
                    acc = whitted_pixel(v, ts, cam, x, y, w, h, q.spp, q.s0, q.depth, g, strat, stk.data(), cap, &ovf,
                                        &nq);
                else if (v.fog_kernel)
                    acc = pt_pixel<true>(v, ts, cam, x, y, w, h, q.spp, q.s0, q.depth, &nq);
                else
                    acc = pt_pixel<false>(v, ts, cam, x, y, w, h, q.spp, q.s0, q.depth, &nq);
                size_t p = ((size_t)y * w + x) * 3;
                (*lin)[p] = (float)acc.x;
                (*lin)[p + 1] = (float)acc.y;
                (*lin)[p + 2] = (float)acc.z;
                (*rgb)[p] = to_byte(acc.x);
                (*rgb)[p + 1] = to_byte(acc.y);
                (*rgb)[p + 2] = to_byte(acc.z);
            }
        }
        if (ovf) {
            fprintf(stderr, "Whitted work stack overflow\n");
            exit(6);
        }
        nrays += nq;
    };
    int nt = q.threads > 0 ? q.threads : 1;
    std::vector<std::thread> pool;
    for (int i = 0; i < nt; ++i) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    *rays = nrays;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        fprintf(stderr, "usage: %s <request> <output>\n", argv[0]);
        return 1;
    }
    FILE* f = fopen(argv[1], "rb");
    if (!f) return 2;
    Request q;
    std::vector<mh_prim> prims;
    std::vector<mh_prim_attr> attrs;
    std::vector<mh_material> mats;
    std::vector<mh_bvh_node> nodes;
    std::vector<int32_t> leaf, planes;
    bool ok = fread(&q, sizeof q, 1, f) == 1 && read_n(f, &prims, q.n_prims) &&
              read_n(f, &attrs, q.has_attrs ? q.n_prims : 0) && read_n(f, &mats, q.n_materials) &&
              read_n(f, &nodes, q.n_nodes) && read_n(f, &leaf, q.n_leaf) && read_n(f, &planes, q.n_planes);
    fclose(f);
    if (!ok) return 2;
    mh_scene_soa soa{};
    soa.abi_version = MH_ABI_VERSION;
    soa.n_prims = q.n_prims;
    soa.prims = prims.data();
    soa.attrs = q.has_attrs ? attrs.data() : nullptr;
    soa.n_materials = q.n_materials;
    soa.materials = mats.data();
    soa.sky = q.sky;
    soa.water_level = q.water_level;
    soa.water_material = q.water_material;
    soa.n_nodes = q.n_nodes;
    soa.nodes = nodes.data();
    soa.n_leaf_prims = q.n_leaf;
    soa.leaf_prims = leaf.data();
    soa.n_planes = q.n_planes;
    soa.planes = planes.data();
    soa.fog = q.fog;
    soa.ocean = q.ocean;
    std::vector<float> lin;
    std::vector<unsigned char> rgb;
    long long rays = 0;
    if (q.precision == MH_PREC_FP64) render<double>(soa, q, &lin, &rgb, &rays);
    else render<float>(soa, q, &lin, &rgb, &rays);
    FILE* o = fopen(argv[2], "wb");
    if (!o) return 2;
    fwrite(lin.data(), sizeof(float), lin.size(), o);
    fwrite(rgb.data(), 1, rgb.size(), o);
    fwrite(&rays, sizeof rays, 1, o);
    fclose(o);
    return 0;
}
