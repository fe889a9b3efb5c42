"""CPU tests of the product host side and the C ABI boundary (no GPU calls):
the libraries load and export every entry point the headers declare, the host
BVH is structurally sound, the generators are deterministic, the camera and
the BMP writer reproduce the reference's bytes."""
import ctypes as C
import hashlib
import os
import re

import numpy as np
import pytest

import checkers as ck
from marmel_amd import HostScene, abi, encode_bmp, native

INCLUDE = os.path.join(ck.REPO, "include")


def declared(header, prefix):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,prefix,lib", [("marmel_hip.h", "mh", "libmarmel_hip.so"),
                                               ("marmel_host.h", "mt", "libmarmel_host.so")])
def test_library_exports_every_declared_symbol(header, prefix, lib):
    names = declared(header, prefix)
    assert len(names) > 5
    L = C.CDLL(os.path.join(native.LIB_DIR, lib))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_gpu_entry_points_reject_null_arguments():
    """Reference codes for bad arguments are returned before any device call."""
    L = native.hip()
    assert L.mh_scene_create(None, None) == 1
    assert L.mh_render(None, None, None, None, None) == 1
    assert L.mh_render_noalloc(None, None, None, None, None) == 1
    assert L.mh_pathtrace_render(None, None, 10, 10, 1, 1, None) == 1
    assert L.mh_render_image(None, None, 10, 10, 1, 1, None) == 1
    assert L.mh_batch_create(None) == 1
    assert L.mh_render_batch(None, None, 0, None, 0, None, None, None) == 1
    assert L.mh_batch_last_tune(None, None) == 1
    assert L.mh_last_kernel_times(None) == 1
    assert L.mh_scene_set_materials(None, None, 0) == 1
    assert L.mh_scene_set_sky(None, None) == 1
    assert L.mh_scene_set_fog(None, None) == 1
    assert L.mh_scene_create_gpu_bvh(None, 0, 0, None) == 1
    assert L.mh_scene_bvh_info(None, None) == 1
    assert L.mh_scene_bvh_export(None, None, None) == 1
    p = abi.RenderExParams(samples_per_pixel=4, max_depth=5, adaptive=1)
    assert L.mh_render_image_ex(None, None, 10, 10, C.byref(p), None) == 1
    assert L.mh_render_image_ex(None, None, 10, 10, None, None) == 1
    assert L.mh_render_adaptive(None, None, 10, 10, C.byref(p), abi.PREC_FP64, None, None, None) == 1
    f = abi.flare_default()
    assert L.mh_flare_apply(None, 10, 10, C.byref(f)) == 1
    assert L.mh_scene_flare(None, C.byref(f)) == 1


def test_flare_entry_validates_and_skips_disabled_on_the_host():
    """The reference's argument checks (src/flare.c:62-68) come first; a
    disabled flare or intensity <= 0 changes nothing and needs no device."""
    L = native.hip()
    img = np.full((4, 6, 3), 200, np.uint8)
    f = abi.flare_default()
    assert L.mh_flare_apply(abi.ptr(img), 0, 4, C.byref(f)) == 1
    assert L.mh_flare_apply(abi.ptr(img), 6, -1, C.byref(f)) == 1
    assert L.mh_flare_apply(abi.ptr(img), 6, 4, None) == 1
    for q in (abi.flare_default(enabled=0), abi.FlareParams(enabled=1, intensity=0.0),
              abi.FlareParams(enabled=1, intensity=-2.0)):
        assert L.mh_flare_apply(abi.ptr(img), 6, 4, C.byref(q)) == 0
        assert (img == 200).all()


def test_sample_stats_need_no_device():
    tot, mx = C.c_ulonglong(7), C.c_int(7)
    assert native.hip().mh_last_sample_stats(C.byref(tot), C.byref(mx)) == 0
    assert native.hip().mh_last_sample_stats(None, None) == 0


def bvh_arrays(hs):
    s = hs.soa()
    nodes = np.ctypeslib.as_array(C.cast(s.nodes, C.POINTER(C.c_uint8)),
                                  (s.n_nodes * abi.NODE_DTYPE.itemsize,)).view(abi.NODE_DTYPE)
    leaf = np.ctypeslib.as_array(C.cast(s.leaf_prims, C.POINTER(C.c_int32)), (s.n_leaf_prims,))
    planes = np.ctypeslib.as_array(C.cast(s.planes, C.POINTER(C.c_int32)), (s.n_planes,)) \
        if s.n_planes else np.zeros(0, np.int32)
    return s, nodes.copy(), leaf.copy(), planes.copy()


def vec(p, f):
    return np.array(p[f].tolist(), np.float64)


def prim_bounds(p):
    k = p["kind"]
    if k == abi.PRIM_SPHERE:
        c, r = vec(p, "p0"), p["r0"]
        return c - r, c + r
    if k == abi.PRIM_TRIANGLE:
        v = np.array([vec(p, "p0"), vec(p, "p1"), vec(p, "p2")])
        return v.min(0), v.max(0)
    if k == abi.PRIM_BOX:
        c, hh = vec(p, "p0"), np.abs(vec(p, "p1"))
        return c - hh, c + hh
    if k == abi.PRIM_CYLINDER:
        a, b = vec(p, "p0"), vec(p, "p1")
        r = max(p["r0"], p["r1"], 0.0)
        return np.minimum(a, b) - r, np.maximum(a, b) + r
    raise AssertionError(k)


@pytest.mark.parametrize("kind,kw", [(native.GEN_SPHERES, {}), (native.GEN_TERRAIN, dict(grid=64)),
                                     (native.GEN_ICOSPHERE, dict(subdiv=3))])
def test_host_bvh_is_sound(kind, kw):
    hs = HostScene()
    hs.generate(kind, seed=2, **kw)
    st = hs.build_bvh()
    prims, _, _, _, _, _ = hs.arrays()
    s, nodes, leaf, planes = bvh_arrays(hs)
    is_plane = prims["kind"] == abi.PRIM_PLANE
    assert sorted(planes.tolist()) == np.nonzero(is_plane)[0].tolist()
    assert sorted(leaf.tolist()) == np.nonzero(~is_plane)[0].tolist()
    covered = np.zeros(len(leaf), int)
    stack = [(0, 1)]
    depth = 0
    while stack:
        n, dpt = stack.pop()
        depth = max(depth, dpt)
        nd = nodes[n]
        lo, hi = np.array(nd["bmin"], np.float64), np.array(nd["bmax"], np.float64)
        if nd["count"] > 0:
            assert 1 <= nd["count"] <= 8
            for i in range(nd["first"], nd["first"] + nd["count"]):
                covered[i] += 1
                plo, phi = prim_bounds(prims[leaf[i]])
                assert (lo <= plo).all() and (phi <= hi).all()
        else:
            for c in (nd["left"], nd["right"]):
                clo, chi = np.array(nodes[c]["bmin"], np.float64), np.array(nodes[c]["bmax"], np.float64)
                assert (lo <= clo).all() and (chi <= hi).all()
                stack.append((c, dpt + 1))
    assert (covered == 1).all()
    assert depth == st.max_depth <= 60


def test_generators_are_deterministic_and_shaped():
    a, b = HostScene(), HostScene()
    ia = a.generate(native.GEN_TERRAIN, seed=4, grid=33)
    b.generate(native.GEN_TERRAIN, seed=4, grid=33)
    a.build_bvh()
    b.build_bvh()
    pa, _, ma, _, _, _ = a.arrays()
    pb, _, mb, _, _, _ = b.arrays()
    assert pa.tobytes() == pb.tobytes() and ma.tobytes() == mb.tobytes()
    assert ia.n_triangles == 2 * 32 * 32 and ia.n_spheres == 200 and ia.n_emissive == 8
    em = [i for i in range(len(pa)) if pa[i]["kind"] == abi.PRIM_SPHERE
          and ma[pa[i]["material"]]["pbr"] and any(ma[pa[i]["material"]]["emissive"].tolist())]
    assert em == list(range(8))  # the emitters lead the file order
    tri = pa[pa["kind"] == abi.PRIM_TRIANGLE]
    for f in ("p0", "p1", "p2"):
        v = np.array(tri[f].tolist())
        assert np.array_equal(v.astype(np.float32).astype(np.float64), v)  # fp32-exact vertices
    c = HostScene()
    c.generate(native.GEN_TERRAIN, seed=5, grid=33)
    c.build_bvh()
    pc, _, _, _, _, _ = c.arrays()
    assert pc.tobytes() != pa.tobytes()


def test_camera_matches_reference_camera_create(golden):
    g = golden["default_scene"]
    for w, h in ((320, 180), (160, 90)):
        hs = HostScene()
        cam = hs.camera(w / h, eye=[-18, 6, 22], at=[0, 5, -20], up=[0, 1, 0], vfov=40.0)
        assert bytes(cam) == g[f"cam_{w}x{h}"].tobytes()


def test_bmp_writer_reproduces_reference_hash(golden):
    rgb = golden["default_scene"]["whitted_320x180_s4_d4"]
    assert hashlib.sha256(encode_bmp(rgb)).hexdigest() == \
        "f9f4551a12c4c73f85c7954ce7e00adebd768798c2471def03df0e97c818ffcb"


def test_ppm_writer_layout(tmp_path):
    from marmel_amd.scene import native as nat
    rgb = np.arange(2 * 3 * 3, dtype=np.uint8).reshape(2, 3, 3)
    p = str(tmp_path / "x.ppm")
    assert nat.host().mt_write_ppm(p.encode(), abi.ptr(np.ascontiguousarray(rgb)), 3, 2) == 0
    assert open(p, "rb").read() == b"P6\n3 2\n255\n" + rgb.tobytes()


def test_tree_deeper_than_the_traversal_stack_is_refused():
    """Every traversal stack holds 64 entries and the 4-wide tree is used only
    when 3 * its depth + 1 fits, so mh_scene_create refuses a binary tree of
    more than 60 levels (code 5) before any device call; a tree of exactly 60
    levels passes the check (it fails later only for want of a device)."""
    L = native.hip()
    for levels, refused in ((61, True), (60, False)):
        n = levels  # a chain: interior node k has leaf k and interior k + 1
        prims = np.zeros(n, abi.PRIM_DTYPE)
        for i in range(n):
            prims[i]["kind"] = abi.PRIM_SPHERE
            prims[i]["p0"] = (float(i), 0.0, 0.0)
            prims[i]["r0"] = 0.4
        mats = np.zeros(1, abi.MATERIAL_DTYPE)
        mats[0]["albedo"] = (0.5, 0.5, 0.5)
        nodes = np.zeros(2 * n - 1, abi.NODE_DTYPE)
        for k in range(n - 1):
            nodes[k]["left"], nodes[k]["right"] = n - 1 + k, k + 1
            nodes[n - 1 + k]["first"], nodes[n - 1 + k]["count"] = k, 1
        nodes[n - 2]["right"] = 2 * n - 2   # the deepest interior node ends in two leaves
        nodes[2 * n - 2]["first"], nodes[2 * n - 2]["count"] = n - 1, 1
        for nd in nodes:
            nd["bmin"] = (-1.0, -1.0, -1.0)
            nd["bmax"] = (float(n), 1.0, 1.0)
        leaf = np.arange(n, dtype=np.int32)
        s = abi.SceneSoa(abi_version=abi.ABI_VERSION, n_prims=n, prims=abi.ptr(prims), attrs=None,
                         n_materials=1, materials=abi.ptr(mats), sky=abi.Sky(), water_level=0.0,
                         water_material=-1, n_nodes=len(nodes), nodes=abi.ptr(nodes), n_leaf_prims=n,
                         leaf_prims=abi.ptr(leaf), n_planes=0, planes=None, fog=abi.fog_default(),
                         ocean=abi.Ocean())
        h = C.c_void_p()
        rc = L.mh_scene_create(C.byref(s), C.byref(h))
        msg = L.mh_last_error().decode()
        if refused:
            assert rc == 5 and "deeper than 60" in msg, (rc, msg)
        else:
            assert rc != 5, msg
            if rc == 0:
                L.mh_scene_destroy(h)
