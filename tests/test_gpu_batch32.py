"""GPU test of the 32-scene batched workload at its full size: 32 seeded
scenes of 1,280 .. 409,600 triangles (displaced icospheres, stratified over
1k .. 500k) with 20 .. 300 spheres, 256^2 .. 1024^2 pixels, 4 .. 256 spp,
depth 4 .. 10 (703 Msamples), rendered in ONE mh_render_batch call with the
trees built on the device.

  * the batch is independent of how paths are scheduled: path regeneration
    and fixed batches give the same bits, and every job equals the same job
    rendered alone through the batched engine, bit for bit (fixed-point pixel
    sums of the same per-sample estimates),
  * every job equals its plain megakernel render up to fp32 rounding
    (>= 99.5 % of pixels within 1/255),
  * three jobs, cut down to 48 x 48 pixels, against converged FP64 oracle
    renders (Monte Carlo consistency of pixels, 4x4 blocks, image sums).

The reference renders one image per call (src/pathtrace.c:1256-1363); the
batch has no reference counterpart beyond that per-image estimator.
"""
import time

import numpy as np
import pytest

import checkers as ck
from marmel_amd import Batch, GpuScene
from marmel_amd.workloads import batch32_scenes, batch32_spec

pytestmark = pytest.mark.gpu


def within(a, b, tol=1):
    return (np.abs(a.astype(int) - b.astype(int)) <= tol).all(axis=-1).mean()


@pytest.fixture(scope="module")
def batch32():
    spec = batch32_spec()
    hss = batch32_scenes(spec, host_bvh=False)
    gss = [GpuScene(hs.soa(bvh=False), gpu_bvh=True) for hs in hss]
    jobs = [dict(scene=i, cam=hs.camera(1.0), width=r, height=r, spp=spp, depth=dp)
            for i, (hs, (_, _, _, _, r, spp, dp)) in enumerate(zip(hss, spec))]
    tris = [hs.info.n_triangles for hs in hss]
    print("batch32: triangles", min(tris), "..", max(tris), "total", sum(tris),
          "Msamples", sum(j["width"] * j["height"] * j["spp"] for j in jobs) / 1e6, flush=True)
    yield spec, hss, gss, jobs
    for g in gss:
        g.close()


def test_batch32_full_size_matches_single_renders(batch32):
    spec, hss, gss, jobs = batch32
    assert min(hs.info.n_triangles for hs in hss) >= 1000
    assert max(hs.info.n_triangles for hs in hss) >= 400000
    b = Batch()
    try:
        t0 = time.perf_counter()
        res = b.render(gss, jobs)
        print("batch32 one call (path regeneration):", round(time.perf_counter() - t0, 2), "s", flush=True)
        fixed = b.render(gss, jobs, fixed_batches=True, pool=1 << 22)
        for (r0, l0), (r1, l1) in zip(res, fixed):
            assert np.array_equal(l0, l1)
            assert np.array_equal(r0, r1)
        worst = 1.0
        for k, (j, (rgb, lin)) in enumerate(zip(jobs, res)):
            (rgb1, lin1), = b.render([gss[j["scene"]]], [dict(j, scene=0)])
            assert np.array_equal(lin, lin1), k
            rgb_m, lin_m = gss[j["scene"]].render(j["cam"], j["width"], j["height"], j["spp"], j["depth"])
            frac = within(rgb, rgb_m)
            worst = min(worst, frac)
            assert frac >= 0.995, (k, frac)
            assert np.isfinite(lin).all()
        print("batch32: worst job within 1/255 of its megakernel render:", worst, flush=True)
    finally:
        b.close()


@pytest.mark.parametrize("which", ["smallest", "median", "largest"])
def test_batch32_jobs_vs_converged_oracle(batch32, which):
    spec, hss, gss, jobs = batch32
    order = np.argsort([h.info.n_triangles for h in hss], kind="stable")
    k = int({"smallest": order[0], "median": order[len(order) // 2], "largest": order[-1]}[which])
    hs = hss[k]
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    o = ck.OracleScene(ck.SoaHolder(prims, attrs, mats, sky, wl, wm).soa)
    j = dict(jobs[k], width=48, height=48, spp=64)
    _, ref_mean, ref_var, _ = o.render(j["cam"], 48, 48, 512, j["depth"], want_var=True, sample_offset=20000)
    b = Batch()
    try:
        (rgb, lin), = b.render([gss[k]], [dict(j, scene=0)])
    finally:
        b.close()
    ck.assert_mc_consistent(lin, 64, ref_mean, ref_var, 512,
                            f"batch32 job {k} ({hs.info.n_triangles} triangles, depth {j['depth']})")
