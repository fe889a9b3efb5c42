"""Edge cases on every renderer, against the oracle (which equals the
reference, tests/test_oracle.py):

  * a scene with no primitives (sky with sun, clouds and glow only), a scene
    of planes only (no tree), a one-sphere scene (the tree is one leaf),
  * images of 1x1, 1x7, 7x1 and 13x5 pixels, spp 1 and depth 0,
  * engines: fp64 path tracer and Whitted (the 1/255 bar, here exact),
    adaptive Whitted (bytes and sample counts), fp32 megakernel and
    wavefront (close to the fp64 image), the batched engine with all three
    scenes in one launch sequence, and trees built on the device for the
    empty and one-leaf scenes.
"""
import numpy as np
import pytest

import checkers as ck
import scenes
from marmel_amd import Batch, GpuScene, HostScene, abi

pytestmark = pytest.mark.gpu

FP32, FP64 = abi.PREC_FP32, abi.PREC_FP64
SIZES = [(1, 1), (1, 7), (7, 1), (13, 5)]


def sky():
    return scenes.sky(horizon=(0.7, 0.8, 0.9), zenith=(0.2, 0.35, 0.8), sun_color=(3, 2.8, 2.5),
                      sun_dir=(0.3, 0.8, -0.5), glow=0.4)


def make(kind):
    mats = [scenes.material(albedo=(0.6, 0.5, 0.4)), scenes.material(albedo=(0.3, 0.6, 0.3), reflectivity=0.4)]
    prims = []
    if kind == "planes":
        prims = [scenes.prim(abi.PRIM_PLANE, 0, p0=(0, -1, 0), p1=(0, 1, 0)),
                 scenes.prim(abi.PRIM_PLANE, 1, p0=(0, 0, 6), p1=(0, 0, -1))]
    elif kind == "sphere":
        prims = [scenes.prim(abi.PRIM_SPHERE, 1, p0=(0, 0, 0), r0=1.0)]
    hs = HostScene()
    hs.add_materials(np.array(mats, dtype=abi.MATERIAL_DTYPE))
    if prims:
        hs.add_prims(np.array(prims, dtype=abi.PRIM_DTYPE))
    hs.set_sky(sky())
    hs.build_bvh()
    return hs


KINDS = ["empty", "planes", "sphere"]


def oracle_of(hs):
    prims, attrs, mats, sk, wl, wm = hs.arrays()
    return ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sk, wl, wm).soa)


def camera(hs, w, h):
    return hs.camera(w / h, eye=[0, 0.5, -5], at=[0, 0, 0], up=[0, 1, 0], vfov=50)


@pytest.mark.parametrize("kind", KINDS)
def test_fp64_renderers_equal_oracle(kind):
    hs = make(kind)
    o = oracle_of(hs)
    gs = GpuScene(hs.soa())
    try:
        for (w, h) in SIZES:
            cam = camera(hs, w, h)
            for spp, depth in ((1, 0), (3, 4)):
                for mode in (0, 1):
                    ref, _, _, _ = o.render(cam, w, h, spp, depth, mode=mode)
                    rgb, _ = gs.render(cam, w, h, spp, depth, mode=mode, precision=FP64)
                    assert np.array_equal(rgb, ref), (kind, w, h, spp, depth, mode)
                a, _, ns, tot, mx = gs.render_adaptive(cam, w, h, spp, depth, 4 * spp, 0.02)
                b, _, ons, otot, omx = o.render_adaptive(cam, w, h, spp, depth, 4 * spp, 0.02)
                assert np.array_equal(a, b) and np.array_equal(ns, ons) and (tot, mx) == (otot, omx)
    finally:
        gs.close()


@pytest.mark.parametrize("kind", KINDS)
def test_fp32_engines_close_to_oracle(kind):
    hs = make(kind)
    o = oracle_of(hs)
    gs = GpuScene(hs.soa())
    try:
        for (w, h) in SIZES:
            cam = camera(hs, w, h)
            _, m, _, _ = o.render(cam, w, h, 4, 4)
            for eng in (abi.ENGINE_MEGAKERNEL, abi.ENGINE_WAVEFRONT):
                _, lin = gs.render(cam, w, h, 4, 4, precision=FP32, engine=eng)
                err = np.abs(lin - m) / (np.abs(m) + 1e-3)
                print(kind, (w, h), "engine", eng, "max rel. diff", err.max())
                # the same samples in fp32 (measured: max 4.5e-6 relative)
                assert err.max() < 1e-3
    finally:
        gs.close()


def test_batch_with_edge_scenes():
    hss = [make(k) for k in KINDS]
    gss = [GpuScene(hs.soa()) for hs in hss]
    b = Batch()
    try:
        jobs = []
        for i, (hs, (w, h)) in enumerate(zip(hss, [(1, 1), (13, 5), (7, 1)])):
            jobs.append(dict(scene=i, cam=camera(hs, w, h), width=w, height=h, spp=3, depth=4))
        res = b.render(gss, jobs)
        for j, (rgb, lin) in zip(jobs, res):
            _, lin1 = gss[j["scene"]].render(j["cam"], j["width"], j["height"], 3, 4, precision=FP32,
                                               engine=abi.ENGINE_WAVEFRONT)
            assert np.abs(lin.reshape(lin1.shape) - lin1).max() < 1e-3
    finally:
        b.close()
        for g in gss:
            g.close()


@pytest.mark.parametrize("kind", ["empty", "sphere"])
def test_device_tree_of_tiny_scenes(kind):
    hs = make(kind)
    gd = GpuScene(hs.soa(bvh=False), gpu_bvh=True)
    gh = GpuScene(hs.soa())
    try:
        cam = camera(hs, 13, 5)
        a, la = gd.render(cam, 13, 5, 2, 4, precision=FP64)
        b, lb = gh.render(cam, 13, 5, 2, 4, precision=FP64)
        assert np.array_equal(la, lb)
        print(kind, "device tree", gd.bvh_info().n_nodes, "nodes")
    finally:
        gd.close()
        gh.close()
