"""The flare post-pass on the GPU (mh_flare_apply, mh_scene_flare) against
the reference's flare_apply:

  * every golden case (seeded images, the default scene's Whitted image,
    the defaults and every clamp): bytes equal the reference's,
  * rows too wide for LDS (the global-scratch variant) and 1080p rows equal
    the oracle (itself equal to the reference, tests/test_flare.py),
  * on the device framebuffer after a render: the same bytes as the host
    entry point on the downloaded image,
  * the reference's error codes and no-op parameters.
"""
import time

import numpy as np
import pytest

import checkers as ck
import flare_cases as fc
from marmel_amd import GpuScene, HostScene, abi, flare_apply, native
from test_flare import oracle_flare

pytestmark = pytest.mark.gpu


def test_golden_cases_bit_exact(golden):
    g = golden["flare"]
    n = 0
    for key in g.files:
        if not key.startswith("out_"):
            continue
        rest = key[4:]
        pn = next(p for p in fc.PARAMS if rest.endswith("_" + p))
        name = rest[: -len(pn) - 1]
        img = golden["default_scene"]["whitted_320x180_s4_d4"] if name == "whitted_320x180" else g[f"in_{name}"]
        out = np.ascontiguousarray(img.copy())
        assert flare_apply(out, fc.PARAMS[pn]) == 0
        same = np.array_equal(out, g[key])
        if not same:
            d = np.abs(out.astype(int) - g[key].astype(int))
            print(key, "bytes off", int((d > 0).sum()), "max", int(d.max()))
        assert same, key
        n += 1
    print("flare golden cases equal:", n)


@pytest.mark.parametrize("w,h", [(6000, 24), (1920, 1080), (2730, 8), (2731, 8)])
def test_wide_and_full_hd_rows_equal_oracle(w, h):
    img = fc.speckle(w, h, w + h)
    for pn in ("default", "strong"):
        p = fc.PARAMS[pn]
        a = np.ascontiguousarray(img.copy())
        assert flare_apply(a, p) == 0
        rc, b = oracle_flare(img, p)
        assert np.array_equal(a, b), (w, h, pn)
    a = np.ascontiguousarray(img.copy())
    t0 = time.time()
    flare_apply(a, fc.PARAMS["default"])
    print("flare %dx%d: %.2f ms per call (host buffer in and out)" % (w, h, (time.time() - t0) * 1e3))


def test_device_framebuffer_after_render():
    hs = HostScene()
    hs.generate(native.GEN_SPHERES, seed=7)
    hs.build_bvh()
    gs = GpuScene(hs.soa())
    try:
        cam = hs.camera(16 / 9)
        for mode in (abi.MODE_WHITTED, abi.MODE_PATHTRACE):
            rgb, _ = gs.render(cam, 320, 180, 4, 5, mode=mode, precision=abi.PREC_FP64)
            p = fc.params(threshold=0.6, intensity=1.5)
            gs.flare(p)
            dev, _ = gs.download(320, 180)
            rc, ref = oracle_flare(rgb, p)
            host = np.ascontiguousarray(rgb.copy())
            assert flare_apply(host, p) == 0
            print("mode", mode, "pixels changed by the flare:", (dev != rgb).any(-1).mean())
            assert np.array_equal(dev, ref) and np.array_equal(host, ref)
        gs.render_adaptive(cam, 320, 180, 2, 5, 8, 0.02)
        gs.flare(fc.PARAMS["default"])
    finally:
        gs.close()


def test_error_codes_and_noops():
    p = fc.params()
    img = np.full((4, 8, 3), 100, np.uint8)
    assert flare_apply(None, p) == 1
    assert flare_apply(img, None) == 1
    L = native.hip()
    assert L.mh_flare_apply(abi.ptr(img), 0, 4, None) == 1
    for name, q in fc.NOOP.items():
        x = fc.speckle(30, 5, 1)
        y = x.copy()
        assert flare_apply(x, q) == 0 and np.array_equal(x, y), name
