"""Generate the golden fixtures under tests/golden/ from the REFERENCE itself.

Runs in the build container only, where oracle/_ref/libmtref.so was compiled
from the reference sources (oracle/Makefile). Writes data only: the default
scene exported in the flattened ABI layout, reference-rendered images, and
reference input -> output vectors for the per-function checks. Nothing of the
reference's source text is copied.

    python tests/golden/make_golden.py
"""
import ctypes as C
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import checkers as ck  # noqa: E402
from marmel_amd import abi, native, HostScene, encode_bmp  # noqa: E402

REF_SCENE = "/root/reference/scenes/default.scene"


def sbytes(struct):
    return np.frombuffer(bytes(struct), np.uint8).copy()


def default_scene():
    text = open(REF_SCENE, "rb").read()
    r = ck.RefScene(text=text)
    rc, prims, attrs, mats, sky, wl, wm = r.export()
    assert rc == 0
    out = {}
    for (w, h, spp, depth) in [(320, 180, 4, 4), (160, 90, 16, 6)]:
        cam = r.camera(w / h)
        out[f"cam_{w}x{h}"] = sbytes(cam)
        for mode, tag in [(1, "whitted"), (0, "pt")]:
            rgb, _ = r.render(cam, w, h, spp, depth, mode, threads=8)
            key = f"{tag}_{w}x{h}_s{spp}_d{depth}"
            out[key] = rgb
            out[key + "_sha_bmp"] = np.frombuffer(
                hashlib.sha256(encode_bmp(rgb)).hexdigest().encode(), np.uint8).copy()
    np.savez_compressed(os.path.join(HERE, "default_scene.npz"), prims=prims, attrs=attrs,
                        mats=mats, sky=sbytes(sky), water_level=np.float64(wl),
                        water_material=np.int32(wm), **out)
    return r, prims, attrs, mats, sky, wl, wm


def host_scene_images():
    """Reference renders of the host-generated benchmark scenes (small sizes)."""
    out = {}
    cases = [("spheres", native.GEN_SPHERES, {}, (160, 120, 4, 6)),
             ("terrain", native.GEN_TERRAIN, dict(grid=48), (120, 68, 2, 8)),
             ("icosphere", native.GEN_ICOSPHERE, dict(subdiv=2), (96, 72, 2, 6))]
    for name, kind, kw, (w, h, spp, depth) in cases:
        hs = HostScene()
        info = hs.generate(kind, seed=1, **kw)
        hs.build_bvh()
        prims, attrs, mats, sky, wl, wm = hs.arrays()
        hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm)
        r = ck.RefScene(soa=hold.soa)
        cam = hs.camera(w / h)
        out[f"{name}_cam"] = sbytes(cam)
        out[f"{name}_prims_sha"] = np.frombuffer(
            hashlib.sha256(prims.tobytes() + mats.tobytes()).hexdigest().encode(), np.uint8).copy()
        for mode, tag in [(0, "pt"), (1, "whitted")]:
            rgb, _ = r.render(cam, w, h, spp, depth, mode, threads=8)
            out[f"{name}_{tag}_{w}x{h}_s{spp}_d{depth}"] = rgb
        rays = ck.seeded_rays(2000, 11, [-12, 0.05, -12], [12, 6, 12])
        hits = r.intersect(rays)
        out[f"{name}_rays_sha"] = np.frombuffer(
            hashlib.sha256(rays.tobytes()).hexdigest().encode(), np.uint8).copy()
        out[f"{name}_hits_prim"] = hits["prim_index"]
        out[f"{name}_hits_t"] = hits["t"]
    np.savez_compressed(os.path.join(HERE, "host_scenes.npz"), **out)


FOG_SCENE = (native.GEN_FOREST, dict(grid=3, subdiv=2), 5)
FOG_IMG = (64, 48, 2, 6)


def fog_scenes():
    """Reference radiance (fp64, seed 500 + i for ray i, depth 8) and images
    of the forest-and-pond workload under each fog of scenes.FOG_SETS, and
    3D gradient noise / fBm vectors."""
    import scenes
    kind, kw, seed = FOG_SCENE
    hs = HostScene()
    hs.generate(kind, seed=seed, **kw)
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    out = {"prims_sha": np.frombuffer(hashlib.sha256(prims.tobytes() + mats.tobytes()).hexdigest()
                                      .encode(), np.uint8).copy()}
    rays = scenes.fog_rays()
    out["rays"] = rays
    w, h, spp, depth = FOG_IMG
    cam = hs.camera(w / h)
    out["cam"] = sbytes(cam)
    for name, f in scenes.FOG_SETS.items():
        hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm, fog=f)
        r = ck.RefScene(soa=hold.soa)
        out[f"{name}_fog"] = sbytes(f)
        out[f"{name}_rad"] = r.radiance(rays, 8, 500)
        for mode, tag in [(0, "pt"), (1, "whitted")]:
            rgb, _ = r.render(cam, w, h, spp, depth, mode, threads=8)
            out[f"{name}_{tag}"] = rgb
    rng = np.random.default_rng(31)
    xyz = np.ascontiguousarray(np.concatenate([rng.uniform(-40, 40, (400, 3)),
                                               np.round(rng.uniform(-5, 5, (100, 3)))]))
    per, fbm = np.zeros(len(xyz)), np.zeros(len(xyz))
    ck.ref().mtref_noise3(abi.ptr(xyz), len(xyz), 1337, abi.ptr(per), abi.ptr(fbm))
    out["noise_xyz"], out["perlin3"], out["fbm3"] = xyz, per, fbm
    np.savez_compressed(os.path.join(HERE, "fog.npz"), **out)


def ingest():
    """Reference OBJ triangles of tests/ingest_data.torus_obj() under each
    placement of test_ingest.XFORMS (hashes), and the reference ocean of
    test_ingest.OCEAN_TEXT (hash, material, ocean parameters)."""
    import ingest_data
    import test_ingest as ti
    text = ingest_data.torus_obj()
    out = {"torus_text_sha": np.frombuffer(hashlib.sha256(text).hexdigest().encode(), np.uint8).copy()}
    for name, xf in ti.XFORMS.items():
        n, p, a = ti.reference_obj(text, xf)
        out[f"{name}_count"] = np.array(n)
        out[f"{name}_sha"] = np.frombuffer(ti.digest(p, a).encode(), np.uint8).copy()
    r = ck.RefScene(text=ti.OCEAN_TEXT)
    rc, pr, ar, mats, sky, wl, wm, fog = r.export()
    ro = abi.Ocean()
    ck.ref().mtref_ocean(r.h, C.byref(ro))
    out["ocean_material"] = np.array(wm)
    out["ocean_sha"] = np.frombuffer(ti.digest(pr, ar).encode(), np.uint8).copy()
    out["ocean_params"] = sbytes(ro)
    np.savez_compressed(os.path.join(HERE, "ingest.npz"), **out)


def adaptive():
    """Reference render_image_ex (adaptive Whitted) of the default scene at
    96x54, depth 5, under test_oracle.ADAPTIVE_CASES: images, total and
    maximum per-pixel sample counts."""
    import test_oracle as to
    text = open(REF_SCENE, "rb").read()
    r = ck.RefScene(text=text)
    cam = r.camera(96 / 54)
    out = {"cam": sbytes(cam)}
    for i, (n0, nmax, tau) in enumerate(to.ADAPTIVE_CASES):
        rgb, tot, mx = r.render_ex(cam, 96, 54, n0, 5, 1, nmax, tau)
        out[f"rgb_{i}"], out[f"total_{i}"], out[f"max_{i}"] = rgb, np.array(tot), np.array(mx)
    np.savez_compressed(os.path.join(HERE, "adaptive.npz"), **out)


def flare():
    """Reference flare_apply on tests/flare_cases.py's images (stored with
    the outputs) and on the default scene's 320x180 Whitted image, under
    each parameter set."""
    import flare_cases as fc
    out = {}
    imgs = {k: f() for k, f in fc.IMAGES.items()}
    d = np.load(os.path.join(HERE, "default_scene.npz"))
    imgs["whitted_320x180"] = d["whitted_320x180_s4_d4"]
    for name, img in imgs.items():
        if name != "whitted_320x180":
            out[f"in_{name}"] = img
        for pn, p in fc.PARAMS.items():
            if name == "whitted_320x180" and pn not in ("default", "strong", "thresh_low"):
                continue
            o = np.ascontiguousarray(img.copy())
            rc = ck.ref().mtref_flare(abi.ptr(o), img.shape[1], img.shape[0], C.byref(p))
            assert rc == 0
            out[f"out_{name}_{pn}"] = o
    np.savez_compressed(os.path.join(HERE, "flare.npz"), **out)


LENS_IMG = {"pt": (96, 72, 8, 5), "whitted": (96, 72, 4, 4)}

# Cameras of the per-function lens vectors: (aperture, focus distance,
# squeeze). Pinhole and negative aperture (the exact pinhole ray), the round
# lens, squeezes above and below 1, squeeze <= 0 and within 1e-12 of 1 (both
# the round lens), focus distance 0 (the unit fallback).
LENS_VEC_CAMS = [(0.0, 8.4, 1.0), (-3.0, 8.4, 1.0), (0.15, 8.4, 1.0), (0.25, 6.0, 1.8),
                 (0.2, 10.0, 0.55), (0.3, 5.0, -1.0), (0.3, 5.0, 1.0 + 1e-13), (0.1, 0.0, 1.3)]
# Texture materials of the per-function texture vectors: (kind, scale).
TEX_VEC_MATS = [(0, 1.0), (1, 1.0), (1, 0.5), (1, 0.0), (1, -2.0), (1, float("nan")), (1, 2.3),
                (7, 1.0), (7, 4.0), (7, 0.0), (7, -1.0), (7, 0.37)]


def lens_inputs():
    """(u, v, r1, r2) quadruples and texture points / uvs of the lens and
    texture vectors (seeded; edges r1 = 0, r1 = 1 and cell boundaries)."""
    rng = np.random.default_rng(20261204)
    uvr = rng.random((400, 4))
    uvr[:20, 2] = 0.0
    uvr[20:40, 2] = 1.0
    uvr[40:60, 3] = 0.0
    p = rng.uniform(-20, 20, (600, 3))
    p[:100] = np.round(p[:100] * 2) / 2          # on cell faces of scales 1 and 0.5
    p[100:150] *= 5e4                             # large coordinates
    uv = rng.uniform(-3, 3, (600, 2))
    uv[:100] = np.round(uv[:100] * 4) / 4
    return np.ascontiguousarray(uvr), np.ascontiguousarray(p), np.ascontiguousarray(uv)


def lens_texture():
    """Reference camera_ray_dof and texture_albedo_uv vectors, and PT / Whitted
    images of tests/scenes.lens_texture_scene under each thin-lens camera of
    scenes.LENS_CAMS."""
    import scenes
    L = ck.ref()
    out = {}
    uvr, p, uv = lens_inputs()
    from marmel_amd import make_camera
    for i, (ap, fd, sq) in enumerate(LENS_VEC_CAMS):
        cam = make_camera(scenes.LENS_EYE, scenes.LENS_AT, [0, 1, 0], scenes.LENS_VFOV, 4 / 3)
        cam.aperture, cam.focus_distance, cam.anamorphic_squeeze = ap, fd, sq
        o = np.zeros((len(uvr), 6))
        L.mtref_camera_ray_dof(C.byref(cam), abi.ptr(uvr), len(uvr), abi.ptr(o))
        out[f"cam{i}"] = sbytes(cam)
        out[f"cam{i}_rays"] = o
    out["uvr"], out["tex_p"], out["tex_uv"] = uvr, p, uv
    for i, (kind, scale) in enumerate(TEX_VEC_MATS):
        m = abi.Material()
        m.albedo = abi.vec3((0.5, 1.0, 0.25))
        m.texture_kind, m.texture_scale = kind, scale
        m.texture_color_a = abi.vec3((0.9, 0.1, 0.3))
        m.texture_color_b = abi.vec3((0.2, 0.7, 0.8))
        o = np.zeros((len(p), 3))
        L.mtref_texture_albedo(C.byref(m), abi.ptr(p), abi.ptr(uv), len(p), abi.ptr(o))
        out[f"tex{i}_mat"] = sbytes(m)
        out[f"tex{i}_out"] = o
    hs = scenes.lens_texture_scene()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    out["prims_sha"] = np.frombuffer(hashlib.sha256(prims.tobytes() + attrs.tobytes() + mats.tobytes())
                                     .hexdigest().encode(), np.uint8).copy()
    r = ck.RefScene(soa=ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa)
    for name in scenes.LENS_CAMS:
        for tag, mode in (("pt", 0), ("whitted", 1)):
            w, h, spp, depth = LENS_IMG[tag]
            cam = scenes.lens_camera(hs, name, w / h)
            out[f"{name}_cam"] = sbytes(cam)
            rgb, _ = r.render(cam, w, h, spp, depth, mode, threads=8)
            out[f"{name}_{tag}"] = rgb
    np.savez_compressed(os.path.join(HERE, "lens_texture.npz"), **out)


def vectors(r, sky):
    """Reference outputs of the per-function routines on seeded inputs."""
    L = ck.ref()
    rng = np.random.default_rng(20261203)
    out = {}
    # sky: directions over the upper and lower hemisphere incl. near the sun
    d = rng.normal(size=(3000, 3))
    sun = np.array([sky.sun_dir.x, sky.sun_dir.y, sky.sun_dir.z])
    d[:300] = sun + 0.02 * rng.normal(size=(300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = np.ascontiguousarray(d)
    s = np.zeros_like(d)
    L.mtref_sky_sample(r.h, abi.ptr(d), len(d), abi.ptr(s))
    out["sky_dirs"], out["sky_out"] = d, s
    xz = np.ascontiguousarray(rng.uniform(-60, 60, size=(1000, 2)))
    wn = np.zeros((1000, 3))
    L.mtref_water_normal(abi.ptr(xz), 1000, abi.ptr(wn))
    out["water_xz"], out["water_n"] = xz, wn
    xy = np.ascontiguousarray(rng.uniform(-300, 300, size=(2000, 2)))
    per = np.zeros(2000)
    fbm = np.zeros(2000)
    L.mtref_noise(abi.ptr(xy), 2000, 5, 1337, abi.ptr(per), abi.ptr(fbm))
    out["noise_xy"], out["perlin2"], out["fbm2_o5"] = xy, per, fbm
    u = np.ascontiguousarray(rng.random(size=(1000, 2)))
    out["u2"] = u
    for name, fn, args in [("sun_disk", L.mtref_sun_disk_dir, (ck.d3(sun), 0.5)),
                           ("light_cone", L.mtref_light_cone_dir, (ck.d3([0.3, 0.8, -0.52]), 0.93)),
                           ("ggx_half", L.mtref_ggx_half, (ck.d3([0.0, 0.6, 0.8]), 0.09))]:
        o = np.zeros((1000, 3))
        fn(args[0], args[1], abi.ptr(u), 1000, abi.ptr(o))
        out[name] = o
    o = np.zeros((1000, 3))
    L.mtref_cosine_dir(ck.d3([0.2, 0.9, -0.38]), abi.ptr(u), 1000, abi.ptr(o))
    out["cosine"] = o
    ci = rng.uniform(-0.1, 1.1, size=500)
    eta = rng.uniform(0.5, 1.6, size=500)
    out["fresnel_in"] = np.stack([ci, eta], 1)
    out["fresnel_out"] = np.array([L.mtref_fresnel_dielectric(a, b) for a, b in zip(ci, eta)])
    ps_in = rng.uniform(-0.1, 1.1, size=(500, 3))
    out["pspec_in"] = ps_in
    out["pspec_out"] = np.array([L.mtref_specular_probability(*x) for x in ps_in])
    tb = rng.uniform(-0.2, 1.2, size=2000)
    out["tobyte_in"] = tb
    out["tobyte_out"] = np.array([L.mtref_to_byte(x) for x in tb], np.uint8)
    np.savez_compressed(os.path.join(HERE, "vectors.npz"), **out)


def main():
    if not ck.have_ref():
        raise SystemExit("oracle/_ref/libmtref.so missing: make -C oracle ref")
    if len(sys.argv) > 1:       # regenerate the named fixtures only
        for name in sys.argv[1:]:
            globals()[name]()
        return
    r, prims, attrs, mats, sky, wl, wm = default_scene()
    vectors(r, sky)
    host_scene_images()
    fog_scenes()
    ingest()
    lens_texture()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
