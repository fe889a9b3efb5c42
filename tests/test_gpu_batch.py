"""GPU tests of the batched engine (mh_render_batch): many jobs over several
scenes in one persistent wavefront launch sequence.

  * every job's image is the same estimator as a single render of that job
    (same sample streams): equal to the megakernel image up to fp32 rounding,
  * the fixed-point pixel sums make the result independent of the launch
    settings (pool, grid, block, graph, autotune): bit for bit,
  * a job's per-pixel means are consistent with a converged FP64 oracle
    render (3 sigma), and bad jobs return the reference's error codes.
"""
import numpy as np
import pytest

import checkers as ck
from marmel_amd import Batch, GpuScene, HostScene, MarmelError, abi, native

pytestmark = pytest.mark.gpu

HOST = {"spheres": (native.GEN_SPHERES, {}), "terrain": (native.GEN_TERRAIN, dict(grid=48)),
        "icosphere": (native.GEN_ICOSPHERE, dict(subdiv=2))}


@pytest.fixture(scope="module")
def world():
    out = {}
    for name, (kind, kw) in HOST.items():
        hs = HostScene()
        hs.generate(kind, seed=1, **kw)
        hs.build_bvh()
        out[name] = (hs, GpuScene(hs.soa()))
    yield out
    for hs, gs in out.values():
        gs.close()


def jobs_of(world, names):
    """Uneven jobs: sizes, spp, depths and sample offsets differ (pools far
    below and far above the path count; regeneration and fixed batches)."""
    spec = [("spheres", 64, 48, 8, 6, 0), ("terrain", 80, 45, 4, 8, 0), ("icosphere", 40, 40, 16, 5, 0),
            ("spheres", 33, 17, 12, 3, 100), ("terrain", 24, 56, 2, 10, 7)]
    jobs = []
    for sc, w, h, spp, depth, off in spec:
        hs, _ = world[sc]
        jobs.append(dict(scene=names.index(sc), cam=hs.camera(w / h), width=w, height=h, spp=spp,
                         depth=depth, sample_offset=off))
    return jobs


NAMES = ["spheres", "terrain", "icosphere"]


def within(a, b, tol=1):
    return (np.abs(a.astype(int) - b.astype(int)) <= tol).all(axis=-1).mean()


def test_batch_jobs_equal_single_renders(world):
    b = Batch()
    scenes = [world[n][1] for n in NAMES]
    jobs = jobs_of(world, NAMES)
    res = b.render(scenes, jobs, pool=4096, grid_blocks=64)
    for j, (rgb, lin) in zip(jobs, res):
        gs = scenes[j["scene"]]
        rgb_m, lin_m = gs.render(j["cam"], j["width"], j["height"], j["spp"], j["depth"],
                                 sample_offset=j["sample_offset"])
        frac = within(rgb, rgb_m)
        d = np.abs(lin - lin_m).mean()
        print(NAMES[j["scene"]], j["width"], j["height"], "within 1/255:", frac, "mean |d|", d)
        assert frac >= 0.995
        assert d < 2e-3
    b.close()


def test_batched_megakernel_equals_single_megakernel_renders(world):
    """The batched megakernel (one launch of tile x 4-sample units over all
    jobs) runs the megakernel's estimator: each job equals its single
    megakernel render up to fp32 rounding (and the 2^-24 fixed-point sums);
    deterministic; the
    autotuner's engine choice keeps the image of that engine."""
    b = Batch()
    scenes = [world[n][1] for n in NAMES]
    jobs = jobs_of(world, NAMES)
    res = b.render(scenes, jobs, engine=abi.BATCH_MEGAKERNEL)
    res2 = b.render(scenes, jobs, engine=abi.BATCH_MEGAKERNEL)
    for j, (rgb, lin), (rgb2, lin2) in zip(jobs, res, res2):
        assert np.array_equal(lin, lin2)
        gs = scenes[j["scene"]]
        rgb_m, lin_m = gs.render(j["cam"], j["width"], j["height"], j["spp"], j["depth"],
                                 sample_offset=j["sample_offset"])
        d = np.abs(lin - lin_m)
        print(NAMES[j["scene"]], j["width"], j["height"], "batched megakernel max |d|", d.max(),
              "within 1/255", within(rgb, rgb_m))
        # same code in another kernel: last-bit differences of contracted
        # arithmetic can flip an occasional path decision
        assert d.mean() < 1e-4
        assert within(rgb, rgb_m) >= 0.995
    res3 = b.render(scenes, jobs, autotune=True)
    t = b.last_tune()
    print("autotuned engine", t["engine"])
    ref = res if t["engine"] == abi.BATCH_MEGAKERNEL else b.render(scenes, jobs, pool=t["pool"],
                                                                   fixed_batches=bool(t["fixed_batches"]))
    for (r3, l3), (r, l) in zip(res3, ref):
        assert np.array_equal(l3, l)
    b.close()


def test_batch_is_independent_of_launch_settings(world):
    b = Batch()
    scenes = [world[n][1] for n in NAMES]
    jobs = jobs_of(world, NAMES)
    ref = b.render(scenes, jobs, pool=1 << 21, grid_blocks=2048, block=256, use_graph=True)
    for kw in (dict(pool=1000, grid_blocks=64, block=128, use_graph=False),
               dict(pool=777, grid_blocks=300, block=64, use_graph=True),
               dict(pool=50000, grid_blocks=512, block=256, use_graph=True, kernel_clock=True),
               dict(pool=4096, grid_blocks=64, fixed_batches=True),
               dict(pool=3000, grid_blocks=128, fixed_batches=True, use_graph=False),
               dict(pool=1 << 21, bvh_width=2),
               dict(pool=999, grid_blocks=96, bvh_width=4, use_graph=False)):
        res = b.render(scenes, jobs, **kw)
        for (r0, l0), (r1, l1) in zip(ref, res):
            assert np.array_equal(l0, l1), kw
            assert np.array_equal(r0, r1), kw
    # graph replay again after other settings: same bits
    res = b.render(scenes, jobs, pool=1 << 21, grid_blocks=2048, block=256, use_graph=True)
    for (r0, l0), (r1, l1) in zip(ref, res):
        assert np.array_equal(l0, l1)
    b.close()


def test_batch_autotune_keeps_the_image(world):
    b = Batch()
    scenes = [world[n][1] for n in NAMES]
    jobs = jobs_of(world, NAMES)
    ref = b.render(scenes, jobs)
    res = b.render(scenes, jobs, autotune=True)
    t = b.last_tune()
    print("autotuned", t)
    assert t["pool"] in tuple(1 << k for k in range(19, 27))
    assert t["fixed_batches"] in (0, 1)
    assert t["grid_blocks"] in (1024, 2048, 4096, 8192, 16384) and t["block"] in (128, 256)
    assert t["bvh_width"] in (2, 4)
    for (r0, l0), (r1, l1) in zip(ref, res):
        assert np.array_equal(l0, l1)
    res2 = b.render(scenes, jobs, autotune=True)  # cached choice
    assert b.last_tune() == t
    for (r0, l0), (r1, l1) in zip(ref, res2):
        assert np.array_equal(l0, l1)
    b.close()


def test_batch_job_statistics_vs_oracle(world):
    hs, gs = world["terrain"]
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    o = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa)
    w, h = 48, 36
    cam = hs.camera(w / h)
    _, ref_mean, ref_var, _ = o.render(cam, w, h, 512, 8, want_var=True, sample_offset=9000)
    hs2, gs2 = world["spheres"]
    b = Batch()
    jobs = [dict(scene=1, cam=hs2.camera(1.0), width=16, height=16, spp=4, depth=4),
            dict(scene=0, cam=cam, width=w, height=h, spp=64, depth=8)]
    (_, _), (_, lin) = b.render([gs, gs2], jobs)
    ck.assert_mc_consistent(lin, 64, ref_mean, ref_var, 512, "batched terrain")
    b.close()


@pytest.mark.parametrize("patch,code", [(dict(scene=5), 5), (dict(width=0), 2), (dict(height=-3), 2),
                                        (dict(spp=0), 3), (dict(depth=-1), 3), (dict(depth=300), 4),
                                        (dict(spp=16384), 4),
                                        (dict(sample_offset=-1), 3)])
def test_batch_error_codes(world, patch, code):
    hs, gs = world["spheres"]
    job = dict(scene=0, cam=hs.camera(1.0), width=8, height=8, spp=1, depth=2)
    job.update(patch)
    b = Batch()
    with pytest.raises(MarmelError) as e:
        b.render([gs], [job])
    assert e.value.code == code
    b.close()


def test_autotune_stops_at_the_largest_pool_that_fits(world):
    """With most of the device's memory taken, a pool whose path state does
    not fit (2^26 paths are tried: the job has 66 M) ends the pool search,
    allocation failure is not an error, and the image is the same bits as
    a render on the chosen engine without the autotuner."""
    import torch
    hs, gs = world["spheres"]
    jobs = [dict(scene=0, cam=hs.camera(16 / 9), width=1920, height=1080, spp=32, depth=2)]
    b = Batch()
    free, _ = torch.cuda.mem_get_info()
    keep = 24 << 30  # room for pools up to 2^25 paths (~20 GB), not 2^26 (~40 GB)
    blocker = torch.empty(max(free - keep, 0), dtype=torch.uint8, device="cuda") if free > keep else None
    try:
        res = b.render([gs], jobs, autotune=True)
        t = b.last_tune()
    finally:
        del blocker
        torch.cuda.empty_cache()
    print("free before the blocker %.1f GiB, autotuned %s" % (free / 2**30, t))
    assert t["pool"] <= 1 << 25
    ref = b.render([gs], jobs, engine=t["engine"])
    assert np.array_equal(ref[0][1], res[0][1])
    b.close()


def _water_plane_scene():
    import scenes
    hs = HostScene()
    hs.generate(native.GEN_SPHERES, seed=3)
    wm = hs.add_materials(np.array([scenes.material(albedo=(0.02, 0.06, 0.08), transparency=1.0, ior=1.333,
                                                    is_water=1)], dtype=abi.MATERIAL_DTYPE))
    hs.add_prims(np.array([scenes.prim(abi.PRIM_PLANE, wm, p0=(0, 0.3, 0), p1=(0, 1, 0))], dtype=abi.PRIM_DTYPE))
    hs.build_bvh()
    return hs, hs.camera(1.0)


def test_transmissive_split_keeps_the_image(monkeypatch):
    """Shadow queries of scenes with water / near-clear materials: a ray
    whose segment misses the box around those primitives takes the any-hit
    walk, the others the closest-hit steps through them (src/pathtrace.c:
    504-532). The image with the queue split (MARMEL_SHADOW_SPLIT=1) equals
    the one with every ray through the closest-hit steps (=0), bit for bit:
    on the forest with its water box, on the ocean world (water triangles)
    and with a water plane (no finite box: every ray takes the steps); and on
    the spheres scene after mh_scene_set_materials turns it from all opaque
    to near-clear spheres over an opaque ground (the box must follow the
    materials: a stale empty box would take every sphere for an opaque
    occluder)."""
    from test_ingest import ocean_world
    hf = HostScene()
    hf.generate(native.GEN_FOREST, seed=5, grid=6, subdiv=3)
    hf.build_bvh()
    ho = ocean_world()
    ho.build_bvh()
    hw, cw = _water_plane_scene()
    cases = [("forest", hf, hf.camera(16 / 9), 320, 180),
             ("ocean", ho, ho.camera(1.0, eye=[0.0, 0.25, -9.0], at=[0.0, -0.2, 4.0], up=[0, 1, 0], vfov=55), 160, 160),
             ("water plane", hw, cw, 128, 128)]
    b = Batch()
    for name, hs, cam, w, h in cases:
        gs = GpuScene(hs.soa())
        job = [dict(scene=0, cam=cam, width=w, height=h, spp=16, depth=8)]
        for fixed in (True, False):
            monkeypatch.setenv("MARMEL_SHADOW_SPLIT", "1")
            (_, split), = b.render([gs], job, pool=1 << 16, fixed_batches=fixed)
            monkeypatch.setenv("MARMEL_SHADOW_SPLIT", "0")
            (_, steps), = b.render([gs], job, pool=1 << 16, fixed_batches=fixed)
            print(name, "fixed batches" if fixed else "path regeneration", "image mean", split.mean())
            assert split.mean() > 0
            assert np.array_equal(split, steps), (name, fixed)
        gs.close()
    # materials changed after upload
    hs = HostScene()
    hs.generate(native.GEN_SPHERES, seed=3)
    hs.build_bvh()
    prims, _, mats, _, _, _ = hs.arrays()
    ground = int(prims["material"][prims["kind"] == abi.PRIM_PLANE][0])
    opaque = mats.copy()
    opaque["transparency"] = 0.0
    opaque["is_water"] = 0
    clear = opaque.copy()
    clear["transparency"] = 0.9
    clear["transparency"][ground] = 0.0
    gs = GpuScene(hs.soa())
    job = [dict(scene=0, cam=hs.camera(1.0), width=128, height=128, spp=16, depth=6)]
    means = []
    for m in (opaque, clear):
        gs.set_materials(m)
        monkeypatch.setenv("MARMEL_SHADOW_SPLIT", "1")
        (_, split), = b.render([gs], job, pool=1 << 16, fixed_batches=True)
        monkeypatch.setenv("MARMEL_SHADOW_SPLIT", "0")
        (_, steps), = b.render([gs], job, pool=1 << 16, fixed_batches=True)
        assert np.array_equal(split, steps)
        means.append(float(split.mean()))
    print("spheres: opaque image mean", means[0], "clear spheres", means[1])
    assert abs(means[1] - means[0]) > 0.01 * means[0]
    gs.close()
    b.close()
