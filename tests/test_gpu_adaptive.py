"""Adaptive Whitted on the GPU (mh_render_image_ex / mh_render_adaptive)
against the reference's render_image_ex (golden fixture) and the oracle.

  * default scene, four (n0, cap, tau) settings: fp64 images within 1/255 of
    the reference's on >= 99.9 % of pixels, per-pixel sample counts equal to
    the oracle's on >= 99.5 % (a count differs only where a last-bit
    difference of a Whitted sample moves a pixel across tau), totals within
    0.5 %; the drop-in entry point equals mh_render_adaptive's fp64 bytes,
  * what the definition fixes: cap = n0 is the fixed render (bit for bit);
    adaptive off is mh_render_image; every count lies in [n0, cap]; the
    statistics equal the count image's sum and maximum; a tau that no pixel
    exceeds stops every pixel at n0,
  * generated scenes (spheres, forest with transmissive water) and a fog
    scene: the same bars against the oracle,
  * fp32: within 1/255 on >= 99 % (its refine decisions are made on fp32
    samples),
  * the reference's error codes,
  * a 1920x1080 render at n0 = 4, cap 16 (timing printed).
"""
import time

import numpy as np
import pytest

import checkers as ck
import scenes
from marmel_amd import GpuScene, HostScene, abi, native

pytestmark = pytest.mark.gpu

FP32, FP64 = abi.PREC_FP32, abi.PREC_FP64
CASES = [(4, 0, 0.0), (4, 16, 0.05), (3, 10, 0.01), (1, 5, 0.03)]


def cam_of(arr):
    return abi.Camera.from_buffer_copy(arr.tobytes())


def within(a, b, tol=1):
    return (np.abs(a.astype(int) - b.astype(int)) <= tol).all(axis=-1).mean()


def host_from_golden(g):
    sky = abi.Sky.from_buffer_copy(g["sky"].tobytes())
    hs = HostScene()
    hs.add_materials(g["mats"])
    hs.add_prims(g["prims"], g["attrs"])
    hs.set_sky(sky)
    hs.set_water(float(g["water_level"]), int(g["water_material"]))
    hs.build_bvh()
    hold = ck.SoaHolder(g["prims"], g["attrs"], g["mats"], sky, float(g["water_level"]),
                        int(g["water_material"]))
    return hs, hold


@pytest.fixture(scope="module")
def default_scene(golden):
    hs, hold = host_from_golden(golden["default_scene"])
    gs = GpuScene(hs.soa())
    yield golden["adaptive"], gs, ck.OracleScene(hold.soa)
    gs.close()


def check_case(gs, o, cam, w, h, n0, depth, nmax, tau, ref=None, prec=FP64):
    rgb, lin, ns, tot, mx = gs.render_adaptive(cam, w, h, n0, depth, nmax, tau, precision=prec)
    orgb, omean, ons, otot, omx = o.render_adaptive(cam, w, h, n0, depth, nmax, tau)
    cap = nmax if nmax > 0 else 4 * n0
    cap = max(cap, n0)
    f_img = within(rgb, ref if ref is not None else orgb)
    f_ns = (ns == ons).mean()
    print("adaptive", (w, h, n0, nmax, tau), "prec", prec, "image within 1/255:", f_img, "counts equal:", f_ns,
          "total", tot, "oracle", otot, "max", mx)
    assert ns.min() >= n0 and ns.max() <= cap
    assert tot == int(ns.sum()) and mx == int(ns.max())
    return f_img, f_ns, tot, otot


@pytest.mark.parametrize("i", range(len(CASES)))
def test_default_scene_matches_reference(default_scene, i):
    a, gs, o = default_scene
    n0, nmax, tau = CASES[i]
    cam = cam_of(a["cam"])
    f_img, f_ns, tot, otot = check_case(gs, o, cam, 96, 54, n0, 5, nmax, tau, ref=a[f"rgb_{i}"])
    assert f_img >= 0.999
    assert f_ns >= 0.995
    assert abs(tot - int(a[f"total_{i}"])) <= 0.005 * int(a[f"total_{i}"])
    # the drop-in entry point: fp64, the same bytes
    p = abi.RenderExParams(samples_per_pixel=n0, max_depth=5, adaptive=1, adaptive_max_spp=nmax, adaptive_tau=tau)
    rc, rgb = gs.render_image_ex(cam, 96, 54, p)
    assert rc == 0
    rgb2, _, _, _, _ = gs.render_adaptive(cam, 96, 54, n0, 5, nmax, tau)
    assert np.array_equal(rgb, rgb2)


def test_definition_closed_forms(default_scene):
    a, gs, o = default_scene
    cam = cam_of(a["cam"])
    w, h = 96, 54
    fixed, flin = gs.render(cam, w, h, 4, 5, mode=abi.MODE_WHITTED, precision=FP64)
    rgb, lin, ns, tot, mx = gs.render_adaptive(cam, w, h, 4, 5, 4, 0.02)
    assert np.array_equal(rgb, fixed) and np.array_equal(lin, flin)
    assert (ns == 4).all() and tot == w * h * 4 and mx == 4
    # adaptive off: the fixed render, and the fixed render's statistics
    p = abi.RenderExParams(samples_per_pixel=4, max_depth=5, adaptive=0)
    rc, rgb0 = gs.render_image_ex(cam, w, h, p)
    assert rc == 0 and np.array_equal(rgb0, fixed)
    assert GpuScene.last_sample_stats() == (w * h * 4, 4)
    # a tolerance no pixel exceeds: every pixel stops after the first pass
    rgb, lin, ns, tot, mx = gs.render_adaptive(cam, w, h, 4, 5, 64, 1e9)
    assert (ns == 4).all() and np.array_equal(rgb, fixed)
    # a tiny tolerance: many pixels go to the cap, none beyond it
    rgb, lin, ns, tot, mx = gs.render_adaptive(cam, w, h, 2, 5, 7, 1e-9)
    assert mx == 7 and ns.min() >= 2 and (ns == 7).mean() > 0.3


@pytest.mark.parametrize("kind,kw,seed", [(native.GEN_SPHERES, {}, 7), (native.GEN_FOREST, dict(grid=3, subdiv=2), 5)])
def test_generated_scenes_match_oracle(kind, kw, seed):
    hs = HostScene()
    hs.generate(kind, seed=seed, **kw)
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    o = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa)
    gs = GpuScene(hs.soa())
    try:
        cam = hs.camera(4 / 3)
        for n0, nmax, tau in CASES[:3]:
            f_img, f_ns, tot, otot = check_case(gs, o, cam, 80, 60, n0, 5, nmax, tau)
            assert f_img >= 0.999 and f_ns >= 0.995
            assert abs(tot - otot) <= 0.005 * otot
        f_img, f_ns, tot, otot = check_case(gs, o, cam, 80, 60, 4, 5, 16, 0.05, prec=FP32)
        assert f_img >= 0.99
    finally:
        gs.close()


def test_fog_scene_matches_oracle():
    fog = scenes.FOG_SETS["ashenvale"]
    hs = HostScene()
    hs.generate(native.GEN_SPHERES, seed=3)
    hs.set_fog(fog)
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    o = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm, fog=fog).soa)
    gs = GpuScene(hs.soa())
    try:
        f_img, f_ns, tot, otot = check_case(gs, o, hs.camera(4 / 3), 80, 60, 4, 5, 16, 0.02)
        assert f_img >= 0.999 and f_ns >= 0.995
    finally:
        gs.close()


def test_error_codes(default_scene):
    a, gs, o = default_scene
    cam = cam_of(a["cam"])
    p = abi.RenderExParams(samples_per_pixel=4, max_depth=5, adaptive=1)
    assert gs.render_image_ex(cam, 32, 18, None)[0] == 1
    assert gs.render_image_ex(cam, 0, 18, p)[0] == 2
    assert gs.render_image_ex(cam, 32, -1, p)[0] == 2
    for bad in (dict(samples_per_pixel=0), dict(max_depth=-1)):
        q = abi.RenderExParams(samples_per_pixel=4, max_depth=5, adaptive=1)
        for k, v in bad.items():
            setattr(q, k, v)
        assert gs.render_image_ex(cam, 32, 18, q)[0] == 3
    rc, rgb = gs.render_image_ex(cam, 32, 18, p)
    assert rc == 0 and rgb.any()


def test_full_hd_timing(default_scene):
    a, gs, o = default_scene
    cam = cam_of(a["cam"])
    gs.render_adaptive(cam, 1920, 1080, 4, 5, 16, 0.02)
    t0 = time.time()
    rgb, lin, ns, tot, mx = gs.render_adaptive(cam, 1920, 1080, 4, 5, 16, 0.02)
    dt = time.time() - t0
    t = GpuScene.last_timing()
    _, flin = gs.render(cam, 1920, 1080, 4, 5, mode=abi.MODE_WHITTED, precision=FP64)
    tf = GpuScene.last_timing()
    print("1920x1080 adaptive n0=4 cap=16: %.1f ms kernels (%.1f ms call), %d samples (%.2f per pixel), "
          "%.1f Msamples/s; fixed 4 spp %.1f ms" % (t.kernel_ms, dt * 1e3, tot, tot / (1920 * 1080),
                                                  tot / t.kernel_ms * 1e-3, tf.kernel_ms))
    assert mx <= 16 and ns.min() == 4 and np.isfinite(lin).all()
