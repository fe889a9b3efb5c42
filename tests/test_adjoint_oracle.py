"""Pins the FP64 path-replay adjoint of the oracle (CPU, no GPU).

The adjoint differentiates the path tracer's estimator with every sampling
decision held fixed. The oracle's frozen-sampling mode renders exactly that
function of the material / sun parameters (decisions taken from frozen
materials, BSDF values from the current ones), so central finite
differences of the frozen render's loss must equal the adjoint path for path,
to FD accuracy, for every kind of parameter. A dropped term, a wrong sign, a
transposed channel or a wrong lobe derivative fails this check.
"""
import numpy as np
import pytest

import checkers as ck
import scenes
from marmel_amd import HostScene, abi, native

FIELDS = [("albedo", "x"), ("albedo", "y"), ("albedo", "z"), ("roughness", None), ("metallic", None),
          ("emissive", "x"), ("emissive", "y"), ("emissive", "z")]


def set_param(mats, m, k, v):
    f, c = FIELDS[k]
    if c is None:
        mats[m][f] = v
    else:
        mats[m][f][c] = v


def get_param(mats, m, k):
    f, c = FIELDS[k]
    return float(mats[m][f] if c is None else mats[m][f][c])


def make_scene(kind=native.GEN_MATERIALS, kw=None, seed=4, fog=None):
    hs = HostScene()
    hs.generate(kind, seed=seed, **(kw or {}))
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm, fog=fog)
    return hs, hold, mats


@pytest.fixture(scope="module")
def scene():
    return make_scene()


# The differentiable workload; the same under marched fog (shadow-marched
# camera segments, closed-form later segments: in-scatter linear in the sun
# colour); the forest-and-pond workload under fog whose steps also march the
# underwater segments.
FD_CASES = {
    "materials": (dict(), 30),
    "materials_fog": (dict(fog=scenes.fog(0.05, (0.3, 0.35, 0.4), 1.0, 0.1, 0.8, 0.6, 0.2, 0.3, 8)), 30),
    "forest_pond_fog": (dict(kind=native.GEN_FOREST, kw=dict(grid=3, subdiv=2), seed=5,
                             fog=scenes.fog(0.03, (0.5, 0.6, 0.7), 0.5, 0.2, 0.7, 0.7, 0.0, 0.0, 12)), 10),
}


def loss_and_dimg(mean, target):
    d = mean - target
    return float((d * d).mean()), 2.0 * d / d.size


@pytest.mark.parametrize("case", list(FD_CASES))
def test_adjoint_equals_frozen_finite_differences(case):
    kw, min_checked = FD_CASES[case]
    hs, hold, mats0 = make_scene(**kw)
    o = ck.OracleScene(hold.soa)
    w, h, spp, depth = 28, 20, 2, 4
    cam = hs.camera(w / h)
    rng = np.random.default_rng(7)
    target = rng.uniform(0.0, 0.6, size=(h, w, 3))
    _, mean, _, _ = o.render(cam, w, h, spp, depth)
    loss0, dimg = loss_and_dimg(mean, target)
    nm = len(mats0)
    gm, gs = o.adjoint(cam, w, h, spp, depth, dimg, nm)
    assert np.isfinite(gm).all() and np.isfinite(gs).all()

    # materials seen by many paths: the ground, a light, and the biggest-
    # gradient spheres; every parameter kind of each
    score = np.abs(gm).sum(axis=1)
    picks = [0, 1] + [int(i) for i in np.argsort(-score)[:4] if i not in (0, 1)]
    o.set_freeze(mats0)
    eps = 1e-5
    checked = 0
    try:
        for m in picks:
            for k in range(8):
                v0 = get_param(mats0, m, k)
                if k in (3, 4) and v0 in (0.0, 1.0):
                    continue  # on the clamp's kink: no two-sided derivative
                # Emission can take a path whose radiance is exactly 0 below
                # zero, where the final clamp kinks: the adjoint gives the
                # derivative from above, so emission is checked by forward
                # differences (O(eps) truncation), the rest centrally.
                steps = (1e-6, 0.0) if k >= 5 else (eps, -eps)
                vals = []
                for st in steps:
                    mats = mats0.copy()
                    set_param(mats, m, k, v0 + st)
                    o.set_materials(mats)
                    _, mp, _, _ = o.render(cam, w, h, spp, depth)
                    vals.append(loss_and_dimg(mp, target)[0])
                fd = (vals[0] - vals[1]) / (steps[0] - steps[1])
                g = gm[m, k]
                tol = 1e-6 * np.abs(gm).max() + 1e-4 * abs(g)
                assert abs(g - fd) <= tol, (m, k, g, fd)
                checked += int(abs(g) > 1e-9)
        o.set_materials(mats0)
        sky0 = hold.sky
        for c, comp in enumerate("xyz"):
            vals = []
            for sgn in (1, -1):
                sky = abi.Sky.from_buffer_copy(bytes(sky0))
                setattr(sky.sun_color, comp, getattr(sky0.sun_color, comp) + sgn * eps)
                o.set_sky(sky)
                _, mp, _, _ = o.render(cam, w, h, spp, depth)
                vals.append(loss_and_dimg(mp, target)[0])
            fd = (vals[0] - vals[1]) / (2 * eps)
            assert abs(gs[c] - fd) <= 1e-6 * np.abs(gs).max() + 1e-4 * abs(gs[c]), (c, gs[c], fd)
        o.set_sky(sky0)
    finally:
        o.set_freeze(None)
        o.set_materials(mats0)
    assert checked >= min_checked  # most picked parameters carry real gradient


def test_frozen_render_at_the_frozen_point_is_the_plain_render(scene):
    """Freezing at the current parameters changes nothing."""
    hs, hold, mats0 = scene
    o = ck.OracleScene(hold.soa)
    cam = hs.camera(1.3)
    rgb0, m0, _, _ = o.render(cam, 26, 20, 2, 4)
    o.set_freeze(mats0)
    try:
        rgb1, m1, _, _ = o.render(cam, 26, 20, 2, 4)
    finally:
        o.set_freeze(None)
    assert np.array_equal(rgb0, rgb1) and np.array_equal(m0, m1)


def test_adjoint_is_linear_in_the_image_gradient(scene):
    hs, hold, mats0 = scene
    o = ck.OracleScene(hold.soa)
    w, h = 16, 12
    cam = hs.camera(w / h)
    rng = np.random.default_rng(3)
    a, b = rng.normal(size=(h, w, 3)), rng.normal(size=(h, w, 3))
    ga, sa = o.adjoint(cam, w, h, 2, 4, a, len(mats0))
    gb, sb = o.adjoint(cam, w, h, 2, 4, b, len(mats0))
    gc, sc = o.adjoint(cam, w, h, 2, 4, 2.0 * a - b, len(mats0))
    assert np.allclose(gc, 2.0 * ga - gb, rtol=1e-9, atol=1e-12)
    assert np.allclose(sc, 2.0 * sa - sb, rtol=1e-9, atol=1e-12)
