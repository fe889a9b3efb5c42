"""Geometry ingest (mt_scene_add_obj, mt_scene_add_ocean) against the
reference's loaders, and the ocean's water-height surface in the oracle.

  * OBJ: the triangles (vertices, normals, uvs, flags, material) equal the
    reference loader's (src/obj.c) bit for bit on seeded OBJ text mixing
    face syntaxes, polygons, negative indices, comments, CRLF and number
    formats, under the default placement and under rotation + scale +
    translation, auto-centre + auto-scale and flat shading; on 1 and 8
    threads alike; pinned by golden hashes where the reference is absent,
  * ocean: the Gerstner surface, skirts and floor equal the reference's
    (src/scene.c scene_grow_ocean, read from a reference scene text with an
    ocean block), and so do the water level, water material and ocean
    parameters the scene records,
  * the oracle's water height under an ocean (which decides whether a camera
    ray starts under water) equals the reference's: images and radiance of
    rays starting below and above the waves are identical.
"""
import ctypes as C
import hashlib

import numpy as np
import pytest

import checkers as ck
import ingest_data
import scenes
from marmel_amd import HostScene, abi

XFORMS = {
    "default": dict(),
    "rotated": dict(position=[1.5, -0.25, 3.0], scale=[1.2, 0.8, 1.5], rotate_deg=[20.0, -35.0, 70.0]),
    "auto": dict(auto_center=1, auto_scale=4.0, rotate_deg=[0.0, 90.0, 0.0]),
    "flat": dict(smooth=0, scale=[2.0, 2.0, 2.0]),
}


def ours(text, xf, threads=0, material=3):
    hs = HostScene()
    hs.add_materials(np.array([scenes.material(albedo=(0.5, 0.5, 0.5))] * 4, dtype=abi.MATERIAL_DTYPE))
    n = hs.add_obj(text, material, HostScene.obj_xform(**xf), threads=threads)
    s = hs.soa(bvh=False)
    prims = np.ctypeslib.as_array(C.cast(s.prims, C.POINTER(C.c_uint8)),
                                  (s.n_prims * abi.PRIM_DTYPE.itemsize,)).view(abi.PRIM_DTYPE).copy()
    attrs = np.ctypeslib.as_array(C.cast(s.attrs, C.POINTER(C.c_uint8)),
                                  (s.n_prims * abi.ATTR_DTYPE.itemsize,)).view(abi.ATTR_DTYPE).copy() \
        if s.attrs else np.zeros(s.n_prims, abi.ATTR_DTYPE)
    return n, prims, attrs


def reference_obj(text, xf, material=3, cap=1 << 20):
    x = HostScene.obj_xform(**xf)
    prims = np.zeros(cap, abi.PRIM_DTYPE)
    attrs = np.zeros(cap, abi.ATTR_DTYPE)
    n = ck.ref().mtref_obj_prims(text, len(text), material, ck.d3(x.position), ck.d3(x.scale),
                                 ck.d3(x.rotate_deg), x.smooth, x.auto_center, x.auto_scale,
                                 abi.ptr(prims), abi.ptr(attrs), cap)
    return n, prims[:n], attrs[:n]


def digest(prims, attrs):
    return hashlib.sha256(prims.tobytes() + attrs.tobytes()).hexdigest()


def test_obj_ingest_matches_golden(golden):
    g = golden["ingest"]
    text = ingest_data.torus_obj()
    assert hashlib.sha256(text).hexdigest() == g["torus_text_sha"].tobytes().decode()
    for name, xf in XFORMS.items():
        n, p, a = ours(text, xf)
        assert n == int(g[f"{name}_count"]), name
        assert digest(p, a) == g[f"{name}_sha"].tobytes().decode(), name


@pytest.mark.parametrize("name", list(XFORMS))
def test_obj_ingest_is_thread_count_independent(name):
    text = ingest_data.torus_obj(nu=30, nv=20, seed=11)
    n1, p1, a1 = ours(text, XFORMS[name], threads=1)
    n8, p8, a8 = ours(text, XFORMS[name], threads=8)
    assert n1 == n8 and np.array_equal(p1, p8) and np.array_equal(a1, a8)


@pytest.mark.skipif(not ck.have_ref(), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("name", list(XFORMS))
def test_obj_ingest_equals_live_reference(name):
    for seed, crlf in ((7, 7), (12, 0)):
        text = ingest_data.torus_obj(nu=36, nv=22, seed=seed, crlf_every=crlf)
        n, p, a = ours(text, XFORMS[name], threads=4)
        nr, pr, ar = reference_obj(text, XFORMS[name])
        assert n == nr > 0
        assert np.array_equal(p, pr), name
        assert np.array_equal(a, ar), name
        # every face syntax occurs: smooth+uv, smooth only, uv only, neither
        fl = p["flags"]
        if XFORMS[name].get("smooth", 1):
            assert {0, abi.PRIM_SMOOTH, abi.PRIM_UV, abi.PRIM_SMOOTH | abi.PRIM_UV} <= set(fl.tolist())


OCEAN_TEXT = b"""
material sea {
    type = water
    albedo = 0.02 0.06 0.08
    transparency = 1.0
    ior = 1.333
    is_water = 1
}
ocean {
    material = sea
    center = 1.5 -0.2 2.0
    size = 30.0 24.0
    resolution = 37 29
    amplitude = 0.3
    wavelength = 5.5
    direction = 0.8 0.6
    steepness = 0.7
    chop = 0.1
    chop_wavelength = 1.7
    depth = 3.0
    seed = 8844
}
"""

OCEAN_DESC = dict(center=[1.5, -0.2, 2.0], size=[30.0, 24.0], res_x=37, res_z=29, amplitude=0.3, wavelength=5.5,
                  direction=[0.8, 0.0, 0.6], steepness=0.7, chop=0.1, chop_wavelength=1.7, depth=3.0, seed=8844)


def ocean_scene(mat_index):
    hs = HostScene()
    mats = [scenes.material(albedo=(0.5, 0.5, 0.5))] * mat_index
    mats.append(scenes.material(albedo=(0.02, 0.06, 0.08), transparency=1.0, ior=1.333, is_water=1))
    hs.add_materials(np.array(mats, dtype=abi.MATERIAL_DTYPE))
    n = hs.add_ocean(HostScene.ocean_desc(**OCEAN_DESC), mat_index)
    return hs, n


@pytest.mark.skipif(not ck.have_ref(), reason="oracle/_ref not built (reference absent)")
def test_ocean_equals_live_reference():
    r = ck.RefScene(text=OCEAN_TEXT)
    rc, pr, ar, mats, sky, wl, wm, fog = r.export()
    ro = abi.Ocean()
    assert ck.ref().mtref_ocean(r.h, C.byref(ro)) == 1
    hs, n = ocean_scene(wm)
    prims, attrs, _, _, wl2, wm2 = hs.arrays()
    assert n == len(pr) == 2 * 36 * 28 + 4 * 36 + 4 * 28 + 2
    assert np.array_equal(prims, pr)
    assert np.array_equal(attrs, ar)
    assert (wl2, wm2) == (wl, wm)
    assert bytes(hs.ocean()) == bytes(ro)


def test_ocean_matches_golden(golden):
    g = golden["ingest"]
    hs, n = ocean_scene(int(g["ocean_material"]))
    prims, attrs, _, _, wl, wm = hs.arrays()
    assert digest(prims, attrs) == g["ocean_sha"].tobytes().decode()
    assert bytes(hs.ocean()) == g["ocean_params"].tobytes()


def ocean_world():
    """Ocean + a few spheres above and below the waves, sun and sky."""
    hs, _ = ocean_scene(1)
    hs.add_prims(np.array([scenes.prim(abi.PRIM_SPHERE, 0, p0=(x, y, z), r0=0.8)
                           for x, y, z in ((0, 1.0, 0), (3, -1.5, 4), (-4, 0.3, 6))], dtype=abi.PRIM_DTYPE))
    hs.set_sky(scenes.sky(horizon=(0.7, 0.8, 0.9), zenith=(0.2, 0.35, 0.8), sun_color=(3, 2.8, 2.5),
                          sun_dir=(0.3, 0.8, -0.5), glow=0.3))
    hs.build_bvh()
    return hs


def test_ocean_water_height_oracle_equals_reference():
    """Rays from just above and just below the moving surface: the oracle
    decides 'starts in water' as the reference does (live when built, else
    against the closed form of the surface)."""
    hs = ocean_world()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm, ocean=hs.ocean())
    o = ck.OracleScene(hold.soa)
    rng = np.random.default_rng(5)
    x, z = rng.uniform(-10, 10, 400), rng.uniform(-8, 10, 400)
    oc = hs.ocean()
    # the water surface, restated in numpy
    d = np.array([oc.direction.x, oc.direction.z]) / np.hypot(oc.direction.x, oc.direction.z)
    wlen, amp = oc.wavelength, oc.amplitude
    lam = [wlen, wlen * 0.52, oc.chop_wavelength, oc.chop_wavelength * 0.40]
    A = [amp, amp * 0.38, oc.chop, oc.chop * 0.32]
    ang = [0.0, 0.628, -0.488, 1.134]
    phi = [0.0, 1.7 + (oc.seed % 100) * 0.03, 3.1 + ((oc.seed * 7) % 2**32 % 100) * 0.03,
           0.8 + ((oc.seed * 13) % 2**32 % 100) * 0.03]
    y = np.full_like(x, oc.center.y)
    for h in range(4):
        dx = np.cos(ang[h]) * d[0] - np.sin(ang[h]) * d[1]
        dz = np.sin(ang[h]) * d[0] + np.cos(ang[h]) * d[1]
        y += A[h] * np.cos(2 * np.pi / lam[h] * (dx * x + dz * z) + phi[h])
    rays = np.zeros(800, abi.RAY_DTYPE)
    rays["ox"], rays["oz"] = np.concatenate([x, x]), np.concatenate([z, z])
    rays["oy"] = np.concatenate([y - 0.02, y + 0.02])
    dd = rng.normal(size=(800, 3))
    dd[:, 1] = np.abs(dd[:, 1])   # looking up: under water the ray first meets the surface
    dd /= np.linalg.norm(dd, axis=1, keepdims=True)
    rays["dx"], rays["dy"], rays["dz"] = dd[:, 0], dd[:, 1], dd[:, 2]
    rad = o.radiance(rays, 6, 900)
    # without the ocean surface the starting medium is decided by the flat level
    flat = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa).radiance(rays, 6, 900)
    changed = np.any(rad != flat, axis=1)
    assert changed.any()
    if ck.have_ref():
        r = ck.RefScene(soa=hold.soa)
        assert np.array_equal(rad, r.radiance(rays, 6, 900))
        cam = hs.camera(4 / 3, eye=[0, 0.4, -9], at=[0, 0, 4], up=[0, 1, 0], vfov=50)
        for mode in (0, 1):
            a, _ = r.render(cam, 64, 48, 2, 6, mode)
            b, _, _, _ = o.render(cam, 64, 48, 2, 6, mode=mode)
            assert np.array_equal(a, b)
