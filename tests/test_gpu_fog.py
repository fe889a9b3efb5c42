"""GPU fog: the forest-and-pond workload under the fog blocks of the
reference's own scene files (tests/scenes.py FOG_SETS: closed-form segments,
shadow-marched camera segments with 3D fBm turbulence, marched underwater
segments, the step clamps).

  * fp64 path tracer and Whitted images equal the reference's images
    (tests/golden/fog.npz) within 1/255 on >= 99.9 % of pixels, and the
    fp64 per-pixel means equal the oracle's to fp32 storage precision,
  * fp32 megakernel, wavefront and batched engines are statistically
    consistent with a converged FP64 oracle render (3 sigma),
  * fog set through the host scene (mt_scene_set_fog) and on the device
    (mh_scene_set_fog) render the same; density 0 renders the fogless image,
  * the fp32 adjoint under fog equals the oracle's FP64 adjoint (which is
    pinned by finite differences under fog, tests/test_adjoint_oracle.py).
"""
import os

import numpy as np
import pytest

import checkers as ck
import scenes
from marmel_amd import Batch, GpuScene, HostScene, abi, native

pytestmark = pytest.mark.gpu

FP32, FP64 = abi.PREC_FP32, abi.PREC_FP64


def within(a, b, tol=1):
    return (np.abs(a.astype(int) - b.astype(int)) <= tol).all(axis=-1).mean()


@pytest.fixture(scope="module")
def world(golden):
    g = golden["fog"]
    hs = HostScene()
    hs.generate(native.GEN_FOREST, seed=5, grid=3, subdiv=2)
    hs.build_bvh()
    arrs = hs.arrays()
    gs = GpuScene(hs.soa())
    yield g, hs, gs, arrs
    gs.close()


def dump(name, a):
    """Keep an image for offline inspection when MARMEL_TEST_DUMP names a folder."""
    d = os.environ.get("MARMEL_TEST_DUMP")
    if d:
        os.makedirs(d, exist_ok=True)
        np.save(os.path.join(d, name + ".npy"), a)


# Bars of the marched fogs, calibrated with the oracle in place of the GPU
# (64 vs 1024 samples): few march steps through turbulence correlate
# neighbouring pixels (4x4 block z^2 1.65-1.68 with the oracle against
# itself), 128 steps give heavier single-pixel tails (frac3 0.969-0.973).
FOG_BARS = {"steps3": dict(bfrac3=0.95, bz2=2.0), "steps200": dict(frac3=0.955, bfrac3=0.96, bz2=1.5)}


def oracle_for(arrs, fog):
    return ck.OracleScene(ck.SoaHolder(*arrs, fog=fog).soa)


@pytest.mark.parametrize("name", list(scenes.FOG_SETS))
def test_fog_fp64_matches_reference_images(world, name):
    g, hs, gs, arrs = world
    fog = scenes.FOG_SETS[name]
    gs.set_fog(fog)
    cam = abi.Camera.from_buffer_copy(g["cam"].tobytes())
    fr = {}
    for tag, mode in (("pt", abi.MODE_PATHTRACE), ("whitted", abi.MODE_WHITTED)):
        rgb, lin = gs.render(cam, 64, 48, 2, 6, mode=mode, precision=FP64)
        dump(f"fog_{name}_{tag}_rgb", rgb)
        dump(f"fog_{name}_{tag}_lin", lin)
        ref = g[f"{name}_{tag}"]
        fr[tag] = within(rgb, ref)
        dd = np.abs(rgb.astype(int) - ref.astype(int)).max(axis=-1)
        print(name, tag, "fp64 within 1/255 of the reference:", fr[tag], "pixels off", int((dd > 1).sum()),
              "at", [tuple(int(v) for v in p) for p in np.argwhere(dd > 1)][:8], "max", int(dd.max()))
    # linear means and images against the oracle at a larger size
    o = oracle_for(arrs, fog)
    w, h = 120, 90
    cam2 = hs.camera(w / h)
    big = {}
    for tag, mode in (("pt", 0), ("whitted", 1)):
        ref, mean, _, _ = o.render(cam2, w, h, 2, 8 if mode == 0 else 6, mode=mode)
        rgb, lin = gs.render(cam2, w, h, 2, 8 if mode == 0 else 6, mode=mode, precision=FP64)
        dump(f"fog_{name}_{tag}_big_rgb", rgb)
        big[tag] = within(rgb, ref)
        close = np.isclose(lin, mean, rtol=2e-6, atol=1e-7).all(axis=-1).mean()
        print(name, tag, w, "x", h, "fp64 within 1/255 of the oracle:", big[tag],
              "linear means equal to fp32 precision:", close)
        if mode == 0:
            assert close >= 0.999
    assert fr["pt"] >= 0.999 and big["pt"] >= 0.999
    assert fr["whitted"] >= 0.999 and big["whitted"] >= 0.999


@pytest.mark.parametrize("name", ["ashenvale", "reef", "nordic"])
def test_fog_fp32_engines_statistically_consistent(world, name):
    g, hs, gs, arrs = world
    fog = scenes.FOG_SETS[name]
    gs.set_fog(fog)
    o = oracle_for(arrs, fog)
    w, h = 40, 30
    cam = hs.camera(w / h)
    _, ref_mean, ref_var, _ = o.render(cam, w, h, 1024, 8, want_var=True, sample_offset=7000)
    n = 64
    outs = {
        "megakernel": gs.render(cam, w, h, n, 8, precision=FP32)[1],
        "wavefront": gs.render(cam, w, h, n, 8, precision=FP32, engine=abi.ENGINE_WAVEFRONT)[1],
    }
    b = Batch()
    try:
        outs["batched"] = b.render([gs], [dict(scene=0, cam=cam, width=w, height=h, spp=n, depth=8)])[0][1]
    finally:
        b.close()
    for eng, lin in outs.items():
        lin = np.asarray(lin, np.float64).reshape(h, w, 3)
        assert np.isfinite(lin).all()
        ck.assert_mc_consistent(lin, n, ref_mean, ref_var, 1024, f"{name} {eng}", **FOG_BARS.get(name, {}))


def test_fog_host_and_device_setters_agree(world):
    g, hs, gs, arrs = world
    fog = scenes.FOG_SETS["ashenvale"]
    hs2 = HostScene()
    hs2.generate(native.GEN_FOREST, seed=5, grid=3, subdiv=2)
    hs2.set_fog(fog)
    hs2.build_bvh()
    assert bytes(hs2.soa().fog) == bytes(fog)
    gs2 = GpuScene(hs2.soa())
    try:
        cam = hs.camera(4 / 3)
        gs.set_fog(fog)
        a, la = gs.render(cam, 48, 36, 2, 6, precision=FP64)
        b, lb = gs2.render(cam, 48, 36, 2, 6, precision=FP64)
        assert np.array_equal(a, b) and np.array_equal(la, lb)
        # density 0 is no fog, whatever the other fields hold (steps 0: no
        # marched water either)
        off = scenes.fog(0.0, (0.9, 0.1, 0.1), inscatter=2.0, steps=0)
        gs.set_fog(off)
        c, lc = gs.render(cam, 48, 36, 2, 6, precision=FP64)
        gs.set_fog(abi.fog_default())
        d, ld = gs.render(cam, 48, 36, 2, 6, precision=FP64)
        assert np.array_equal(lc, ld)
        assert not np.array_equal(la, ld)
    finally:
        gs2.close()


def test_fog_rejects_bad_parameters(world):
    g, hs, gs, arrs = world
    bad = scenes.fog(-0.1, (0.5, 0.5, 0.5))
    with pytest.raises(Exception):
        gs.set_fog(bad)
    nan = scenes.fog(float("nan"), (0.5, 0.5, 0.5))
    with pytest.raises(Exception):
        gs.set_fog(nan)


def test_fog_adjoint_matches_fp64_oracle_adjoint():
    fog = scenes.fog(0.05, (0.3, 0.35, 0.4), 1.0, 0.1, 0.8, 0.6, 0.2, 0.3, 8)
    hs = HostScene()
    hs.generate(native.GEN_MATERIALS, seed=4)
    hs.set_fog(fog)
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    gs = GpuScene(hs.soa())
    try:
        o = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm, fog=fog).soa)
        w, h, spp, depth = 64, 48, 4, 4
        cam = hs.camera(w / h)
        rng = np.random.default_rng(11)
        target = rng.uniform(0.0, 0.5, (h, w, 3))
        _, lin = gs.render(cam, w, h, spp, depth)
        dimg = 2.0 * (lin.astype(np.float64) - target) / lin.size
        gm, gsun = gs.adjoint(cam, w, h, spp, depth, dimg.astype(np.float32))
        om, osun = o.adjoint(cam, w, h, spp, depth, dimg, len(mats))
        g = np.concatenate([np.asarray(gm, np.float64).reshape(-1), np.asarray(gsun, np.float64)])
        r = np.concatenate([np.asarray(om, np.float64).reshape(-1), np.asarray(osun, np.float64)])
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        rs = np.linalg.norm(np.asarray(gsun) - osun) / np.linalg.norm(osun)
        print("fog adjoint: relative L2", rel, "sun part", rs)
        assert rel < 0.02
        assert rs < 0.02
    finally:
        gs.close()
