"""The device BVH build (mh_scene_create_gpu_bvh).

  * the device tree IS the host builder's tree (mt_scene_build_bvh, same
    max_leaf / bins): walked from both roots in lock step, every node has
    the same fp32 bounds, the same shape and every leaf the same set of
    prims, on scenes of spheres + plane, terrain, displaced icospheres, the
    forest-and-pond workload, with and without the small-subtree path,
  * renders and queries through the device-built tree equal those through
    the host tree bit for bit (fp64 and fp32, both node widths),
  * degenerate inputs: planes only, one prim, coincident centroids (where
    the object-median split may take a different half: hits still equal),
  * the 2.1 M-triangle terrain of the benchmark: same tree, build time.
"""
import ctypes as C
import time

import numpy as np
import pytest

import scenes
from marmel_amd import GpuScene, HostScene, abi, native

pytestmark = pytest.mark.gpu

FP32, FP64 = abi.PREC_FP32, abi.PREC_FP64


def host_tree(hs):
    s = hs.soa()
    if s.n_nodes == 0:
        return np.zeros(0, abi.NODE_DTYPE), np.zeros(0, np.int32)
    nodes = np.ctypeslib.as_array(C.cast(s.nodes, C.POINTER(C.c_uint8)),
                                  (s.n_nodes * abi.NODE_DTYPE.itemsize,)).view(abi.NODE_DTYPE).copy()
    leaf = np.ctypeslib.as_array(C.cast(s.leaf_prims, C.POINTER(C.c_int32)), (s.n_leaf_prims,)).copy() \
        if s.n_leaf_prims else np.zeros(0, np.int32)
    return nodes, leaf


def leaf_sets(nodes, leaf, idx):
    """(k, 8) sorted prim ids of leaves idx, padded with -1."""
    first, count = nodes["first"][idx], nodes["count"][idx]
    off = np.arange(8)
    pos = first[:, None] + off[None, :]
    ok = off[None, :] < count[:, None]
    out = np.where(ok, leaf[np.where(ok, pos, 0)], np.iinfo(np.int32).max)
    return np.sort(out, axis=1)


def assert_same_tree(hn, hl, dn, dl):
    """Lock-step walk of both trees from node 0; returns (nodes, max depth)."""
    assert len(hn) == len(dn), (len(hn), len(dn))
    if len(hn) == 0:
        return 0, 0
    pairs = np.zeros((1, 2), np.int64)
    seen, depth = 0, 0
    while len(pairs):
        depth += 1
        a, b = hn[pairs[:, 0]], dn[pairs[:, 1]]
        seen += len(pairs)
        assert np.array_equal(a["bmin"], b["bmin"]) and np.array_equal(a["bmax"], b["bmax"]), "bounds differ"
        assert np.array_equal(a["count"], b["count"]), "shape differs"
        lf = a["count"] > 0
        if lf.any():
            assert np.array_equal(leaf_sets(hn, hl, pairs[lf, 0]), leaf_sets(dn, dl, pairs[lf, 1])), "leaf sets"
        inner = ~lf
        pairs = np.concatenate([np.stack([a["left"][inner], b["left"][inner]], 1),
                                np.stack([a["right"][inner], b["right"][inner]], 1)])
    assert seen == len(hn)
    return seen, depth


CASES = {
    "spheres": (native.GEN_SPHERES, dict(), 1),
    "terrain": (native.GEN_TERRAIN, dict(grid=96, n_spheres=40), 3),
    "icosphere": (native.GEN_ICOSPHERE, dict(subdiv=5, n_spheres=120), 7),
    "materials": (native.GEN_MATERIALS, dict(), 4),
    "forest": (native.GEN_FOREST, dict(grid=4, subdiv=3), 5),
}


def build_both(kind, kw, seed, **opts):
    hs = HostScene()
    hs.generate(kind, seed=seed, **kw)
    st = hs.build_bvh(**opts)
    gs_h = GpuScene(hs.soa())
    gs_d = GpuScene(hs.soa(bvh=False), gpu_bvh=True, max_leaf=opts.get("max_leaf", 0), bins=opts.get("bins", 0))
    return hs, st, gs_h, gs_d


@pytest.mark.parametrize("name", list(CASES))
def test_device_tree_is_the_host_tree(name):
    kind, kw, seed = CASES[name]
    print("case", name, flush=True)
    hs, st, gs_h, gs_d = build_both(kind, kw, seed)
    try:
        info = gs_d.bvh_info()
        dn, dl = gs_d.bvh_export()
        hn, hl = host_tree(hs)
        seen, depth = assert_same_tree(hn, hl, dn, dl)
        print(name, "nodes", seen, "depth", depth, "host depth", st.max_depth, "device", info.max_depth,
              "build %.2f ms" % info.build_ms, "host %.1f ms" % (st.build_seconds * 1e3))
        assert info.device_built == 1 and info.n_nodes == len(hn) and info.max_depth == st.max_depth
        assert info.width4 == gs_h.bvh_info().width4
        # renders and queries through either tree are identical
        w, h = 64, 48
        cam = hs.camera(w / h)
        for prec in (FP64, FP32):
            for bw in (2, 4):
                a, la = gs_h.render(cam, w, h, 2, 6, precision=prec, bvh_width=bw)
                b, lb = gs_d.render(cam, w, h, 2, 6, precision=prec, bvh_width=bw)
                assert np.array_equal(a, b) and np.array_equal(la, lb), (prec, bw)
        a, _ = gs_h.render(cam, w, h, 2, 4, mode=abi.MODE_WHITTED, precision=FP64)
        b, _ = gs_d.render(cam, w, h, 2, 4, mode=abi.MODE_WHITTED, precision=FP64)
        assert np.array_equal(a, b)
        import checkers as ck
        rays = ck.seeded_rays(3000, 5, [-12, 0.05, -12], [12, 6, 12])
        for prec in (FP64, FP32):
            ha, hb = gs_h.trace_closest(rays, precision=prec), gs_d.trace_closest(rays, precision=prec)
            assert np.array_equal(ha["prim_index"], hb["prim_index"]) and np.array_equal(ha["t"], hb["t"])
            assert np.array_equal(gs_h.trace_any(rays, 1e-3, 1e30, prec), gs_d.trace_any(rays, 1e-3, 1e30, prec))
    finally:
        gs_h.close()
        gs_d.close()


@pytest.mark.parametrize("opts", [dict(max_leaf=1, bins=8), dict(max_leaf=8, bins=32), dict(max_leaf=2, bins=64)])
def test_device_tree_follows_build_options(opts):
    kind, kw, seed = CASES["icosphere"]
    print("options", opts, flush=True)
    hs, st, gs_h, gs_d = build_both(kind, kw, seed, **opts)
    try:
        hn, hl = host_tree(hs)
        dn, dl = gs_d.bvh_export()
        assert_same_tree(hn, hl, dn, dl)
    finally:
        gs_h.close()
        gs_d.close()


def test_device_build_degenerate_inputs():
    m = [scenes.material(albedo=(0.5, 0.5, 0.5))]
    sk = scenes.sky(horizon=(0.6, 0.7, 0.9), zenith=(0.2, 0.3, 0.8), sun_color=(1, 1, 1), sun_dir=(0.3, 0.8, -0.5))
    plane = scenes.prim(abi.PRIM_PLANE, 0, p0=(0, 0, 0), p1=(0, 1, 0))
    cases = {
        "planes only": [plane, scenes.prim(abi.PRIM_PLANE, 0, p0=(0, 5, 0), p1=(0, -1, 0))],
        "one sphere": [plane, scenes.prim(abi.PRIM_SPHERE, 0, p0=(0, 1, 0), r0=1.0)],
        # 40 concentric spheres: every centroid coincides (object-median splits)
        "concentric": [plane] + [scenes.prim(abi.PRIM_SPHERE, 0, p0=(0, 2, 0), r0=0.2 + 0.05 * i) for i in range(40)],
        "70 spheres": [plane] + [scenes.prim(abi.PRIM_SPHERE, 0, p0=(0.5 * (i % 10) - 2.5, 0.6, 0.5 * (i // 10) - 2),
                                             r0=0.2) for i in range(70)],
    }
    for name, prims in cases.items():
        print("degenerate case", name, flush=True)
        hs = scenes.build(prims, m, sk)
        gh = GpuScene(hs.soa())
        gd = GpuScene(hs.soa(bvh=False), gpu_bvh=True)
        try:
            cam = hs.camera(4 / 3, eye=[0, 3, 9], at=[0, 1, 0], up=[0, 1, 0], vfov=40)
            a, la = gh.render(cam, 48, 36, 2, 4, precision=FP64)
            b, lb = gd.render(cam, 48, 36, 2, 4, precision=FP64)
            assert np.array_equal(la, lb), name
            if name != "concentric":
                hn, hl = host_tree(hs)
                dn, dl = gd.bvh_export()
                assert_same_tree(hn, hl, dn, dl)
        finally:
            gh.close()
            gd.close()


def test_device_build_benchmark_terrain():
    """The 1920x1080 benchmark scene's 2,093,058 triangles + 200 spheres."""
    hs = HostScene()
    hs.generate(native.GEN_TERRAIN, seed=2, grid=1024)
    t0 = time.time()
    st = hs.build_bvh()
    t_host = time.time() - t0
    gd = GpuScene(hs.soa(bvh=False), gpu_bvh=True)
    gd.close()  # first build warms the runtime (module load, allocator)
    t0 = time.time()
    gd = GpuScene(hs.soa(bvh=False), gpu_bvh=True)
    t_create = time.time() - t0
    try:
        info = gd.bvh_info()
        print("terrain 2.09M triangles: host build %.2f s, device build %.1f ms (create incl. upload %.1f ms), "
              "nodes %d depth %d" % (t_host, info.build_ms, t_create * 1e3, info.n_nodes, info.max_depth))
        hn, hl = host_tree(hs)
        dn, dl = gd.bvh_export()
        assert_same_tree(hn, hl, dn, dl)
        assert info.max_depth == st.max_depth
    finally:
        gd.close()
