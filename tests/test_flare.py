"""The flare oracle (oracle/or_flare.c, the CPU restatement of the
reference's flare_apply) pinned before the GPU post-pass is checked with it:

  * bytes equal the reference's golden outputs (tests/golden/flare.npz) on
    seeded images and the default scene's Whitted image, under the defaults
    and every clamp (threshold, streak length, tint, radii of narrow rows),
  * bytes equal the live reference (when oracle/_ref was built here) on
    images the fixtures do not hold,
  * the reference's own expectations (tests/test_anamorphic.c:131-172),
  * closed form: a lone highlight in a black row spreads as the two-scale
    symmetric exponential sum_s 0.045 w_s e alpha_s^|k| (numpy, double),
  * disabled / zero intensity leave the image as it is.
"""
import ctypes as C

import numpy as np
import pytest

import checkers as ck
import flare_cases as fc
from marmel_amd import abi


def oracle_flare(img, p):
    o = np.ascontiguousarray(img.copy())
    rc = ck.oracle().or_flare_apply(abi.ptr(o), o.shape[1], o.shape[0], C.byref(p) if p is not None else None)
    return rc, o


def test_oracle_matches_reference_golden(golden):
    g = golden["flare"]
    n = 0
    for key in g.files:
        if not key.startswith("out_"):
            continue
        rest = key[4:]
        pn = next(p for p in fc.PARAMS if rest.endswith("_" + p))
        name = rest[: -len(pn) - 1]
        img = golden["default_scene"]["whitted_320x180_s4_d4"] if name == "whitted_320x180" else g[f"in_{name}"]
        rc, out = oracle_flare(img, fc.PARAMS[pn])
        assert rc == 0
        assert np.array_equal(out, g[key]), key
        n += 1
    assert n >= 35


def test_fixture_inputs_are_the_seeded_images(golden):
    g = golden["flare"]
    for name, f in fc.IMAGES.items():
        assert np.array_equal(f(), g[f"in_{name}"]), name


@pytest.mark.skipif(not ck.have_ref(), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_oracle_equals_live_reference(seed):
    rng = np.random.default_rng(seed)
    w, h = int(rng.integers(1, 400)), int(rng.integers(1, 40))
    img = fc.speckle(w, h, seed)
    p = fc.params(intensity=float(rng.uniform(0.1, 5)), threshold=float(rng.uniform(-0.2, 1.2)),
                  streak=float(rng.uniform(0.0, 1.5)), tint=tuple(rng.uniform(0, 1, 3)))
    rc, a = oracle_flare(img, p)
    b = np.ascontiguousarray(img.copy())
    assert ck.ref().mtref_flare(abi.ptr(b), w, h, C.byref(p)) == rc == 0
    assert np.array_equal(a, b)
    changed = (a != img).any(-1).mean()
    print("seed", seed, (w, h), "pixels changed", changed)


def test_reference_expectations():
    """tests/test_anamorphic.c:131-172, restated on the oracle."""
    p = fc.params()
    dummy = np.zeros((1, 1, 3), np.uint8)
    L = ck.oracle()
    assert L.or_flare_apply(None, 100, 100, C.byref(p)) != 0
    assert L.or_flare_apply(abi.ptr(dummy), 0, 100, C.byref(p)) != 0
    assert L.or_flare_apply(abi.ptr(dummy), 100, -5, C.byref(p)) != 0
    assert L.or_flare_apply(abi.ptr(dummy), 100, 100, None) != 0
    grey = np.full((16, 64, 3), 100, np.uint8)
    rc, out = oracle_flare(grey, p)
    assert rc == 0 and (out == 100).all()
    rc, out = oracle_flare(fc.reference_pattern(), p)
    left, right = int(out[8, 25, 2]), int(out[8, 39, 2])
    assert left > 100 and right > 100 and abs(left - right) <= 1
    assert (out[:8] == 100).all() and (out[9:] == 100).all()


@pytest.mark.parametrize("w,x0,streak", [(200, 77, 0.25), (64, 10, 0.9), (9, 4, 0.01)])
def test_lone_highlight_closed_form(w, x0, streak):
    img = np.zeros((3, w, 3), np.uint8)
    img[1, x0] = [255, 240, 200]
    p = fc.params(intensity=1.7, threshold=0.6, streak=streak, tint=(0.3, 0.6, 0.9))
    rc, out = oracle_flare(img, p)
    c = img[1, x0] / 255.0
    lum = 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    m = 0.5 * (lum + c.max())
    e = (m - 0.6) / 0.4 * 1.7
    src = (0.25 * c + 0.75 * np.array([0.3, 0.6, 0.9])) * e
    r1, r2 = max(streak * w * 0.12, 2.0), max(streak * w, 6.0)
    k = np.abs(np.arange(w) - x0)
    resp = 0.45 * 0.045 * np.exp(-1 / r1) ** k + 0.55 * 0.045 * np.exp(-1 / r2) ** k
    expect = np.clip(np.floor(img[1].astype(float) + 255 * resp[:, None] * src[None, :] + 0.5), 0, 255)
    assert np.abs(out[1].astype(int) - expect.astype(int)).max() <= 1
    assert (out[0] == 0).all() and (out[2] == 0).all()


def test_noop_parameters_leave_the_image():
    img = fc.speckle(50, 10, 3)
    for name, p in fc.NOOP.items():
        rc, out = oracle_flare(img, p)
        assert rc == 0 and np.array_equal(out, img), name
