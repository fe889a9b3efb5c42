"""GPU parity tests: libmarmel_hip through its C ABI against the oracle and the
reference-generated golden fixtures.

Bars:
  * nearest-hit prim ids equal to the CPU BVH except where hit distances tie
    within 1e-5 relative (fp64: also equal t),
  * fixed-spp Whitted (--no-pathtrace --no-adaptive) within 1/255 per 8-bit
    channel on >= 99.9 % of pixels,
  * path-traced images: per-pixel means consistent with a converged FP64 CPU
    render within 3 sigma of the Monte Carlo error, relative MSE falling as
    1/spp; closed forms (furnace, point sun, sphere light) hold.
"""
import numpy as np
import pytest

import checkers as ck
import scenes
from marmel_amd import GpuScene, HostScene, abi, device_info, native

pytestmark = pytest.mark.gpu

FP32, FP64 = abi.PREC_FP32, abi.PREC_FP64


def cam_of(arr):
    return abi.Camera.from_buffer_copy(arr.tobytes())


def host_from_arrays(prims, attrs, mats, sky, wl, wm):
    hs = HostScene()
    hs.add_materials(mats)
    hs.add_prims(prims, attrs)
    hs.set_sky(sky)
    hs.set_water(wl, wm)
    hs.build_bvh()
    return hs


@pytest.fixture(scope="module")
def default_scene(golden):
    g = golden["default_scene"]
    sky = abi.Sky.from_buffer_copy(g["sky"].tobytes())
    hs = host_from_arrays(g["prims"], g["attrs"], g["mats"], sky, float(g["water_level"]),
                          int(g["water_material"]))
    hold = ck.SoaHolder(g["prims"], g["attrs"], g["mats"], sky, float(g["water_level"]),
                        int(g["water_material"]))
    gs = GpuScene(hs.soa())
    yield g, hs, gs, ck.OracleScene(hold.soa)
    gs.close()


HOST = {"spheres": (native.GEN_SPHERES, {}), "terrain": (native.GEN_TERRAIN, dict(grid=48)),
        "icosphere": (native.GEN_ICOSPHERE, dict(subdiv=2))}


def host_scene(name, seed=1):
    kind, kw = HOST[name]
    hs = HostScene()
    hs.generate(kind, seed=seed, **kw)
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    return hs, GpuScene(hs.soa()), ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa)


def within(a, b, tol=1):
    return (np.abs(a.astype(int) - b.astype(int)) <= tol).all(axis=-1).mean()


def test_device_is_gfx950():
    name, cus, mem = device_info()
    print(name, cus, mem)
    assert "gfx950" in name


# ---- nearest hit / any hit ------------------------------------------------------

def tie_mask(t_a, t_b):
    return np.abs(t_a - t_b) <= 1e-5 * np.maximum(np.abs(t_b), 1e-30)


@pytest.mark.parametrize("name", list(HOST))
def test_closest_hit_ids_match_reference(golden, name):
    g = golden["host_scenes"]
    hs, gs, o = host_scene(name)
    rays = ck.seeded_rays(2000, 11, [-12, 0.05, -12], [12, 6, 12])
    ref_p, ref_t = g[f"{name}_hits_prim"], g[f"{name}_hits_t"]
    h64 = gs.trace_closest(rays, precision=FP64)
    same = h64["prim_index"] == ref_p
    assert (same | tie_mask(h64["t"], ref_t)).all()
    hit = same & (ref_p >= 0)
    assert np.allclose(h64["t"][hit], ref_t[hit], rtol=1e-12, atol=0)
    h32 = gs.trace_closest(rays, precision=FP32)
    ok32 = (h32["prim_index"] == ref_p) | tie_mask(h32["t"], ref_t)
    print(name, "fp32 unexplained mismatches", int((~ok32).sum()), "of", len(rays))
    assert ok32.mean() >= 0.999
    gs.close()


def test_closest_hit_default_scene_camera_rays(default_scene):
    g, hs, gs, o = default_scene
    cam = cam_of(g["cam_320x180"])
    rays = ck.camera_rays(cam, 160, 90, seed=3)
    ho = o.intersect(rays)
    for prec, bar in ((FP64, 1.0), (FP32, 0.999)):
        hg = gs.trace_closest(rays, precision=prec)
        ok = (hg["prim_index"] == ho["prim_index"]) | tie_mask(hg["t"], ho["t"])
        print("prec", prec, "unexplained", int((~ok).sum()), "of", len(rays))
        assert ok.mean() >= bar


@pytest.mark.parametrize("name", list(HOST))
def test_any_hit_matches_oracle(name):
    hs, gs, o = host_scene(name)
    rays = ck.seeded_rays(4000, 21, [-12, 0.05, -12], [12, 6, 12])
    occ = o.occluded(rays, 1e-3, 1e30)
    assert np.array_equal(gs.trace_any(rays, 1e-3, 1e30, FP64), occ)
    assert (gs.trace_any(rays, 1e-3, 1e30, FP32) == occ).mean() >= 0.999
    # bounded segments: shadow rays stopping short of their target
    seg = o.occluded(rays, 1e-3, 5.0)
    assert np.array_equal(gs.trace_any(rays, 1e-3, 5.0, FP64), seg)
    gs.close()


# ---- images -----------------------------------------------------------------------

@pytest.mark.parametrize("prec", [FP64, FP32])
def test_whitted_default_scene_within_1_255(default_scene, prec):
    g, hs, gs, o = default_scene
    for (w, h, spp, depth) in ((320, 180, 4, 4), (160, 90, 16, 6)):
        rgb, _ = gs.render(cam_of(g[f"cam_{w}x{h}"]), w, h, spp, depth, mode=abi.MODE_WHITTED,
                           precision=prec)
        frac = within(rgb, g[f"whitted_{w}x{h}_s{spp}_d{depth}"])
        print("whitted", w, h, "prec", prec, "within 1/255:", frac)
        # the drop-in Whitted entry runs fp64 (the 99.9 % bar); the fp32 option
        # is held to 99.5 % (rounding flips a few near-threshold pixels)
        assert frac >= (0.999 if prec == FP64 else 0.995)


def test_whitted_drop_in_entry(default_scene):
    g, hs, gs, o = default_scene
    rc, rgb = gs.render_image(cam_of(g["cam_320x180"]), 320, 180, 4, 4)
    assert rc == 0
    assert within(rgb, g["whitted_320x180_s4_d4"]) >= 0.999


def test_whitted_work_stack_overflow_is_reported(default_scene, monkeypatch):
    """Code 6 (the Whitted work stack overflowed) comes back from the drop-in
    entry when the stack is too small for the scene's reflected and refracted
    rays: MARMEL_WHITTED_STACK_CAP lowers the capacity, and every push checks
    it before writing. With the capacity restored the same call renders."""
    g, hs, gs, o = default_scene
    cam = cam_of(g["cam_320x180"])
    monkeypatch.setenv("MARMEL_WHITTED_STACK_CAP", "1")
    rc, _ = gs.render_image(cam, 320, 180, 4, 4)
    assert rc == 6
    monkeypatch.delenv("MARMEL_WHITTED_STACK_CAP")
    rc, rgb = gs.render_image(cam, 320, 180, 4, 4)
    assert rc == 0
    assert within(rgb, g["whitted_320x180_s4_d4"]) >= 0.999


def test_pathtrace_fp64_matches_reference_image(default_scene):
    g, hs, gs, o = default_scene
    for (w, h, spp, depth) in ((320, 180, 4, 4), (160, 90, 16, 6)):
        rgb, _ = gs.render(cam_of(g[f"cam_{w}x{h}"]), w, h, spp, depth, precision=FP64)
        frac = within(rgb, g[f"pt_{w}x{h}_s{spp}_d{depth}"])
        print("pt fp64", w, h, "within 1/255:", frac)
        assert frac >= 0.999


@pytest.mark.parametrize("name", list(HOST))
def test_host_scene_images(golden, name):
    g = golden["host_scenes"]
    hs, gs, o = host_scene(name)
    size = {"spheres": (160, 120, 4, 6), "terrain": (120, 68, 2, 8), "icosphere": (96, 72, 2, 6)}[name]
    w, h, spp, depth = size
    cam = cam_of(g[f"{name}_cam"])
    for tag, mode in (("pt", abi.MODE_PATHTRACE), ("whitted", abi.MODE_WHITTED)):
        ref = g[f"{name}_{tag}_{w}x{h}_s{spp}_d{depth}"]
        rgb64, _ = gs.render(cam, w, h, spp, depth, mode=mode, precision=FP64)
        f64 = within(rgb64, ref)
        rgb32, _ = gs.render(cam, w, h, spp, depth, mode=mode, precision=FP32)
        f32 = within(rgb32, ref)
        print(name, tag, "fp64", f64, "fp32", f32)
        assert f64 >= 0.999
        if mode == abi.MODE_WHITTED:
            assert f32 >= 0.995
    gs.close()


def test_pathtrace_fp32_statistically_consistent():
    """GPU fp32 per-pixel means vs a converged FP64 oracle render: z-scores of
    the per-pixel difference follow the Monte Carlo error."""
    hs, gs, o = host_scene("spheres")
    w, h = 48, 36
    cam = hs.camera(w / h)
    _, ref_mean, ref_var, _ = o.render(cam, w, h, 1024, 6, want_var=True, sample_offset=5000)
    n = 64
    _, lin = gs.render(cam, w, h, n, 6, precision=FP32)
    ck.assert_mc_consistent(lin, n, ref_mean, ref_var, 1024, "fp32 vs converged")
    gs.close()


def test_relmse_falls_as_one_over_spp():
    hs, gs, o = host_scene("spheres")
    w, h = 48, 36
    cam = hs.camera(w / h)
    _, ref_mean, _, _ = o.render(cam, w, h, 2048, 6, sample_offset=100000)
    errs = []
    for spp in (4, 16, 64):
        _, lin = gs.render(cam, w, h, spp, 6, precision=FP32)
        errs.append(np.mean((lin - ref_mean) ** 2 / (ref_mean ** 2 + 1e-2)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print("relMSE", errs, "ratios", r1, r2)
    assert 2.5 < r1 < 6.5 and 2.5 < r2 < 6.5
    gs.close()


# ---- closed forms on the GPU -------------------------------------------------------

def test_gpu_white_furnace():
    hs = scenes.white_furnace()
    gs = GpuScene(hs.soa())
    cam = hs.camera(1.0, eye=[0, 0, 6], at=[0, 0, 0], up=[0, 1, 0], vfov=20.0)
    for prec in (FP64, FP32):
        _, lin = gs.render(cam, 32, 32, 8, 8, precision=prec)
        tol = 0 if prec == FP64 else 2e-6
        assert np.abs(lin - 1.0).max() <= tol
    gs.close()


def test_gpu_point_sun_and_sphere_light():
    hs, s = scenes.sun_plane()
    gs = GpuScene(hs.soa())
    cam = hs.camera(1.0, eye=[0.3, 2.0, 0.2], at=[0.3, 0.0, 0.25], up=[0, 0, 1], vfov=10.0)
    expect = 0.6 * np.array([1.0, 0.9, 0.8]) * s[1]
    for prec in (FP64, FP32):
        _, lin = gs.render(cam, 16, 16, 2, 0, precision=prec)
        assert np.allclose(lin, expect, rtol=1e-5)
    gs.close()
    h, R, le, alb = 4.0, 0.5, 10.0, 0.5
    hs = scenes.lamp_over_plane(h, R, le, alb)
    gs = GpuScene(hs.soa())
    # a tiny patch around the foot point, many samples per pixel
    cam = hs.camera(1.0, eye=[0.0, 1.0, 0.001], at=[0.0, 0.0, 0.0], up=[0, 0, 1], vfov=0.01)
    _, lin = gs.render(cam, 8, 8, 4096, 0, precision=FP32)
    expect = alb * le * np.pi * R * R / (h * h)
    assert abs(lin[..., 0].mean() / expect - 1) < 0.01
    gs.close()


def test_drop_in_error_codes(default_scene):
    g, hs, gs, o = default_scene
    cam = cam_of(g["cam_320x180"])
    assert gs.pathtrace_render(cam, 0, 10, 1, 1)[0] == 2
    assert gs.pathtrace_render(cam, 10, -1, 1, 1)[0] == 2
    assert gs.pathtrace_render(cam, 10, 10, 0, 1)[0] == 3
    assert gs.pathtrace_render(cam, 10, 10, 1, -1)[0] == 3
    assert gs.render_image(cam, 10, 10, 0, 1)[0] == 3
    rc, rgb = gs.pathtrace_render(cam, 32, 18, 2, 3)
    assert rc == 0 and rgb.any()


def test_progressive_chunks_average_to_full_render():
    hs, gs, o = host_scene("spheres")
    cam = hs.camera(4 / 3)
    _, full = gs.render(cam, 40, 30, 8, 6, precision=FP64)
    _, a = gs.render(cam, 40, 30, 4, 6, precision=FP64, sample_offset=0)
    _, b = gs.render(cam, 40, 30, 4, 6, precision=FP64, sample_offset=4)
    assert np.allclose(full, 0.5 * (a + b), rtol=1e-5, atol=1e-6)
    gs.close()


# ---- wavefront engine ----------------------------------------------------------------

WF = abi.ENGINE_WAVEFRONT


@pytest.mark.parametrize("name", list(HOST))
def test_wavefront_matches_megakernel(name):
    """Same estimator, same random streams: the wavefront image equals the
    megakernel image up to fp32 rounding, and is deterministic."""
    hs, gs, o = host_scene(name)
    w, h = 96, 64
    cam = hs.camera(w / h)
    rgb_m, lin_m = gs.render(cam, w, h, 8, 8, precision=FP32)
    rgb_w, lin_w = gs.render(cam, w, h, 8, 8, precision=FP32, engine=WF, pool=1000, grid_blocks=64)
    rgb_w2, lin_w2 = gs.render(cam, w, h, 8, 8, precision=FP32, engine=WF, pool=1000,
                               grid_blocks=64, use_graph=True)
    assert np.array_equal(lin_w, lin_w2)  # graph replay == direct launches, bit for bit
    frac = within(rgb_w, rgb_m)
    print(name, "wavefront vs megakernel within 1/255:", frac,
          "mean |d lin|", np.abs(lin_w - lin_m).mean())
    assert frac >= 0.995
    assert np.abs(lin_w - lin_m).mean() < 2e-3
    gs.close()


def test_wavefront_sums_do_not_depend_on_the_batch_size():
    """The single-scene wavefront engine sums each pixel's samples in fp32 in
    sample order whether a batch holds part of one sample of every pixel,
    exactly one, or several whole samples (their radiance added in sample
    order at the batch end): the same bits for every pool, with and without
    graph replay, also when the last batch holds fewer samples."""
    hs, gs, o = host_scene("terrain")
    w, h, spp = 96, 64, 11
    cam = hs.camera(w / h)
    npix = w * h
    ref = None
    for pool, graph in ((1000, False), (npix, False), (2 * npix, False), (4 * npix, True), (5 * npix + 7, False),
                        (0, True)):
        _, lin = gs.render(cam, w, h, spp, 8, precision=FP32, engine=WF, pool=pool, grid_blocks=64,
                           use_graph=graph)
        if ref is None:
            ref = lin
        assert np.array_equal(lin, ref), (pool, graph)
    gs.close()


def test_wavefront_matches_oracle_statistics():
    hs, gs, o = host_scene("terrain")
    w, h = 48, 36
    cam = hs.camera(w / h)
    _, ref_mean, ref_var, _ = o.render(cam, w, h, 512, 8, want_var=True, sample_offset=7000)
    _, lin = gs.render(cam, w, h, 64, 8, engine=WF)
    # heavy-tailed NEE from small bright emitters: a few more single pixels
    # beyond 3 sigma than a normal law gives; 4x4 blocks and the image sum are
    # held to the normal law
    ck.assert_mc_consistent(lin, 64, ref_mean, ref_var, 512, "wavefront terrain")
    gs.close()


# ---- BVH node width --------------------------------------------------------------------

@pytest.mark.parametrize("name", list(HOST))
def test_bvh_width_gives_identical_images(name):
    """The 4-wide tree is a regrouping of the binary one; with strict culling
    and the lowest-index tie rule the nearest hit does not depend on the
    visit order, so every image is bit-identical across widths."""
    hs, gs, o = host_scene(name)
    w, h = 80, 60
    cam = hs.camera(w / h)
    for kw in (dict(precision=FP64), dict(precision=FP32), dict(precision=FP32, engine=WF, pool=2000)):
        r2, l2 = gs.render(cam, w, h, 4, 6, bvh_width=2, **kw)
        r4, l4 = gs.render(cam, w, h, 4, 6, bvh_width=4, **kw)
        assert np.array_equal(l2, l4), kw
        assert np.array_equal(r2, r4), kw
    rgb, _ = gs.render(cam, w, h, 2, 3, mode=abi.MODE_WHITTED, precision=FP64, bvh_width=2)
    rgb4, _ = gs.render(cam, w, h, 2, 3, mode=abi.MODE_WHITTED, precision=FP64, bvh_width=4)
    assert np.array_equal(rgb, rgb4)
    gs.close()


def test_megakernels_are_deterministic():
    """Run to run, every engine and precision gives the same bits (a read of
    stale or uninitialised memory would show here)."""
    hs, gs, o = host_scene("terrain")
    w, h = 120, 68
    cam = hs.camera(w / h)
    for kw in (dict(precision=FP64), dict(precision=FP32), dict(precision=FP32, engine=WF)):
        a = gs.render(cam, w, h, 2, 8, **kw)[1]
        b = gs.render(cam, w, h, 2, 8, **kw)[1]
        c = gs.render(cam, w, h, 2, 8, **kw)[1]
        assert np.array_equal(a, b) and np.array_equal(a, c), kw
    gs.close()


def test_fp32_matches_fp64_far_from_the_camera():
    """fp32 and fp64 on the same sample streams, on the 1080p terrain's layout
    (80-unit half extent, camera ~100 units out) at 240 x 135: per band of
    image rows the fp32 mean stays within 0.1 % of the fp64 mean. A hit point
    formed as o + t d in fp32 lies ~1e-4 off the surface there, the size of
    the next ray's offset, and bounce rays re-hit their own surface: that
    made the far terrain 1.6 % too dark (also with the sun and emitters off)."""
    hs = HostScene()
    hs.generate(native.GEN_TERRAIN, seed=2, grid=256)
    hs.build_bvh()
    gs = GpuScene(hs.soa())
    cam = hs.camera(1920 / 1080)
    w, h = 240, 135
    _, l64 = gs.render(cam, w, h, 32, 8, precision=FP64)
    _, l32 = gs.render(cam, w, h, 32, 8, precision=FP32)
    for y0 in range(0, h, 15):
        a, b = l32[y0:y0 + 15].astype(np.float64).mean(), l64[y0:y0 + 15].mean()
        print("rows", y0, "fp64 mean", round(b, 5), "fp32 relative difference", (a - b) / b)
        assert abs(a - b) <= 1e-3 * b
    gs.close()
