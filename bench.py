#!/usr/bin/env python3
"""Benchmark of the GPU path tracer on the 1080p terrain workload: batched
engine vs the plain GPU kernel at equal RMSE.

Workload (one GPU): fBm heightfield terrain on a 1024 x 1024 vertex grid
(2,093,058 fp32 triangles) + 200 seeded spheres (8 emissive, radiance
U[2,20]), GGX PBR materials, 1920 x 1080, 64 spp, depth 8, NEE to the sun and
the 8 emitters, Russian roulette. One step = one full frame (scene already
resident in HBM), by default rendered as one job of the batched engine
(autotuned once, cached); the plain megakernel renders the same frame and
sample streams for the ratio.

Prints ONE JSON line. Also reports:
  roofline      for the dominant kernel, k_wf_shadow (the any-hit shadow
                queries): algorithmic bytes per launch over its average launch
                duration, both from this run. The duration comes from HIP
                event pairs the library records around every launch on its
                stream during one clocked frame (kernel_clock). The
                bytes are the shadow queries' traversal counts from the CPU
                oracle (run with shadow segments as one any-hit query, which
                is what the GPU does and gives the identical image) on a
                downscaled render of the same scene, priced per query as
                36 B (32 B ray in + 4 B visibility) + 64 B per interior node
                + 48 B per triangle + 32 B per sphere tested, scaled to the
                frame's samples and divided by the launches.
                peak = MI355X HBM3E 8 TB/s. traffic = measured HBM bytes per
                launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on a
                copy kernel) when known, else null.
  cpu_baseline  the oracle (FP64 C restatement of the reference renderer,
                byte-identical to it) on the host cores, on a downscaled
                (1/3 per axis) render of the same scene at the same spp/depth.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "marmel-trace_amd"))

HBM_PEAK_GBS = 8000.0
# Measured HBM bytes per k_wf_shadow_refill launch of the batched engine
# (the default main engine) on this workload: rocprofv3 --pmc FETCH_SIZE
# (6,086,904.7 KiB, x 2.0) + WRITE_SIZE (425,176.0 KiB, x 1.0) per dispatch,
# averaged over the clocked frame's 18 dispatches (pool 2^26 paths, 16384
# workgroups), the factors being what the same counters read for a 1 GiB
# copy (bench.py --hbm-calib 1024). Re-measure when the shadow kernel or its
# launch changes.
SHADOW_TRAFFIC_PER_LAUNCH = 12901.4e6
# What the same kernel's counters show limits it (8.01 ms per launch): VALU
# instructions x 4 cycles fill 63 % of the SIMDs' issue cycles at 2.35 GHz
# (SQ_INSTS_VALU 3.03e9, SQ_BUSY_CYCLES) with 56 % of the lanes active
# (VALUUtilization); 23 of 32 wave slots per CU occupied (SQ_WAVE_CYCLES,
# 78 VGPRs, 24.5 KiB LDS: 6 waves per SIMD); L2 hit 69 %; HBM 1.61 TB/s.
SHADOW_LIMITER = ("dependent node fetches and VALU issue (63 % of SIMD cycles, 56 % of lanes active: "
                  "divergent traversal steps) at 6 of 8 waves per SIMD; L2 hit 69 %; HBM at 1.6 TB/s, 20 % of peak")
W, H, SPP, DEPTH = 1920, 1080, 64, 8
GRID = 1024
STRESS_MAX_LEAF, STRESS_BINS = 8, 32  # device tree build options of the stress scene
# the autotuner's choice for the terrain frame, which the stress leg reuses
STRESS_TUNE = dict(pool=1 << 26, grid_blocks=16384, block=256, bvh_width=4, fixed_batches=1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_scene(grid):
    from marmel_amd import HostScene, native
    hs = HostScene()
    info = hs.generate(native.GEN_TERRAIN, seed=2, grid=grid)
    st = hs.build_bvh()
    return hs, info, st


# Per-unit algorithmic bytes (fixed fp32 record sizes of the device layout).
B_NODE, B_TRI, B_SPH, B_PLANE = 64, 48, 32, 16
B_NODE4 = 112                       # 4-wide node as the shadow kernel reads it (SoA child boxes + refs)
B_CLOSEST_IO, B_ANY_IO = 48, 36     # ray in (32 B) + hit (16 B) / visibility (4 B)
B_BOUNCE, B_SAMPLE = 192, 12        # path state + material per bounce; accumulator per sample


def shadow_bytes(st):
    """Bytes of the any-hit shadow queries the stats cover (k_wf_shadow's work)."""
    return (st["any"] * B_ANY_IO + st["a_inner"] * B_NODE + st["a_tris"] * B_TRI
            + (st["a_prims"] - st["a_tris"]) * B_SPH)


def frame_bytes(st, samples):
    """Bytes of all kernels of the frame."""
    return (shadow_bytes(st) + st["closest"] * B_CLOSEST_IO + st["c_inner"] * B_NODE
            + st["c_tris"] * B_TRI + (st["c_prims"] - st["c_tris"]) * B_SPH
            + st["planes"] * B_PLANE + st["ext"] * B_BOUNCE + samples * B_SAMPLE)


def oracle_scene(hs):
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import checkers as ck
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm)
    return hold, ck.OracleScene(hold.soa)


def traversal_stats(o, cam, scale, threads):
    """GPU-algorithm traversal counts per sample on a downscaled frame."""
    w, h = W // scale, H // scale
    o.set_anyhit_shadow(True)
    try:
        o.render(cam, w, h, SPP, DEPTH, mode=0, threads=threads)
    finally:
        o.set_anyhit_shadow(False)
    return o.stats(), w * h * SPP, f"{w}x{h}"


def cpu_baseline(o, cam, scale, threads):
    """Oracle render (reference algorithm) of a downscaled frame."""
    w, h = W // scale, H // scale
    _, _, _, sec = o.render(cam, w, h, SPP, DEPTH, mode=0, threads=threads)
    samples = w * h * SPP
    return samples / sec / 1e6, f"{w}x{h} downscaled frame of the same scene/camera, {SPP} spp, depth {DEPTH}", sec


def reference_baseline(hold, cam, scale, threads):
    """The reference renderer itself (its pathtrace_render compiled from its
    own sources into oracle/_ref by oracle/Makefile, pthread tile pool) on the
    same downscaled frame; None when that library was not built. Its render
    clock (CLOCK_MONOTONIC around the tile pool, reference
    src/pathtrace.c:1270-1340) excludes the scene and BVH build."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import checkers as ck
    if not ck.have_ref():
        return None
    t0 = time.time()
    r = ck.RefScene(soa=hold.soa)
    log(f"bench: reference scene + BVH built in {time.time() - t0:.1f}s")
    w, h = W // scale, H // scale
    rgb, sec = r.render(cam, w, h, SPP, DEPTH, 0, threads=threads)
    del r
    return w * h * SPP / sec / 1e6, sec


def host_cpus():
    """nproc (what the reference sizes its pool by), the CPUs this process may
    run on, and the cgroup CPU quota (cores) when one is set."""
    nproc = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = nproc
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    return nproc, aff, quota


def adjoint_leg(reps):
    """Differentiable workload: 512 x 512, 32 spp, depth 4, 64 materials, L2
    loss vs the image of hidden seeded parameters; ms per gradient."""
    import numpy as np
    from marmel_amd import GpuScene
    from marmel_amd.workloads import hidden_params, materials_scene
    hs = materials_scene()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    g = GpuScene(hs.soa())
    w = h = 512
    spp, depth = 32, 4
    cam = hs.camera(1.0)
    g.set_materials(hidden_params(mats))
    _, target = g.render(cam, w, h, spp, depth, sample_offset=1 << 20)
    g.set_materials(mats)
    _, lin = g.render(cam, w, h, spp, depth)
    fwd_ms = g.last_timing().kernel_ms
    d = (2.0 * (lin.astype(np.float64) - target) / lin.size).astype(np.float32)
    g.adjoint(cam, w, h, spp, depth, d)  # warm-up
    ks, t0 = [], time.perf_counter()
    for _ in range(reps):
        g.adjoint(cam, w, h, spp, depth, d)
        ks.append(g.last_timing().kernel_ms)
    wall2 = (time.perf_counter() - t0) / reps * 1e3
    # a gradient step as the torch wrapper runs it: a recording forward
    # render (the loss needs the image anyway), then the adjoint, which
    # replays each path once
    rk, kr, t0 = [], [], time.perf_counter()
    for _ in range(reps):
        g.render_record(cam, w, h, spp, depth)
        rk.append(g.last_timing().kernel_ms)
        g.adjoint(cam, w, h, spp, depth, d)
        kr.append(g.last_timing().kernel_ms)
    step = (time.perf_counter() - t0) / reps * 1e3
    g.close()
    return {"workload": "512x512, 32 spp, depth 4, 64 PBR materials (513 parameters), L2 loss",
            "ms_per_grad": round(sum(kr) / len(kr), 2),
            "ms_per_step_forward_and_grad": round(step, 2),
            "forward_record_kernel_ms": round(sum(rk) / len(rk), 2),
            "ms_per_grad_without_record": round(wall2, 2),
            "kernel_ms_per_grad_without_record": round(sum(ks) / len(ks), 2),
            "forward_kernel_ms": round(fwd_ms, 2)}


def timed_reps(fn, reps):
    """Wall seconds of `reps` calls of fn after one untimed call."""
    fn()
    out = []
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        out.append(time.perf_counter() - t)
    return out


def batch_leg(reps):
    """Batched workload: 32 uneven scenes (1,280-409,600 triangles) in one
    persistent wavefront launch sequence (autotuned), against rendering each
    job alone with the plain GPU kernel and with the single-scene wavefront
    engine; every rate is the mean over `reps` timed repetitions (min / max
    reported)."""
    from marmel_amd import Batch, GpuScene, abi
    from marmel_amd.workloads import batch32_scenes, batch32_spec
    t0 = time.time()
    spec = batch32_spec()
    hss = batch32_scenes(spec, host_bvh=False)
    t_gen = time.time() - t0
    t1 = time.time()
    gss = [GpuScene(hs.soa(bvh=False), gpu_bvh=True) for hs in hss]  # trees built on the device
    t_up = time.time() - t1
    dev_build_ms = sum(g.bvh_info().build_ms for g in gss)
    jobs = [dict(scene=i, cam=hs.camera(1.0), width=r, height=r, spp=spp, depth=dp)
            for i, (hs, (_, _, _, _, r, spp, dp)) in enumerate(zip(hss, spec))]
    tris = [hs.info.n_triangles for hs in hss]
    samples = sum(j["width"] * j["height"] * j["spp"] for j in jobs)
    log(f"bench: batch32 scenes ({min(tris)}-{max(tris)} triangles, {sum(tris)} in all) generated in "
        f"{t_gen:.1f}s, uploaded with device-built trees in {t_up:.2f}s (device builds {dev_build_ms:.0f} ms), "
        f"{samples / 1e6:.1f} Msamples")
    b = Batch()
    b.render(gss, jobs, autotune=True, fetch=False)  # tunes once, cached on the handle
    tune = b.last_tune()

    def rate(secs):
        r = [samples / s / 1e6 for s in secs]
        return round(sum(r) / len(r), 1), round(min(r), 1), round(max(r), 1)

    batched = rate(timed_reps(lambda: b.render(gss, jobs, autotune=True, fetch=False), reps))
    log(f"bench: batch32 batched {batched} Msamples/s (mean, min, max of {reps}), tune {tune}")
    # each batched engine on its own (the tuner keeps the faster)
    eng = {}
    for name, e in (("wavefront", abi.BATCH_WAVEFRONT), ("megakernel", abi.BATCH_MEGAKERNEL)):
        kw = dict(pool=tune["pool"], grid_blocks=tune["grid_blocks"], block=tune["block"],
                  bvh_width=tune["bvh_width"] if e == abi.BATCH_WAVEFRONT else 0,
                  fixed_batches=bool(tune["fixed_batches"]), engine=e)
        eng[name] = rate(timed_reps(lambda: b.render(gss, jobs, fetch=False, **kw), reps))[0]
    log(f"bench: batch32 batched engines {eng} Msamples/s")

    def each(**kw):
        for j in jobs:
            gss[j["scene"]].render(j["cam"], j["width"], j["height"], j["spp"], j["depth"], fetch=False, **kw)

    plain = rate(timed_reps(each, reps))
    single_wf = rate(timed_reps(lambda: each(engine=abi.ENGINE_WAVEFRONT, use_graph=True), reps))
    log(f"bench: batch32 plain megakernel {plain}, single-scene wavefront {single_wf} Msamples/s")
    b.close()
    for gsc in gss:
        gsc.close()
    return {"workload": f"32 seeded scenes: displaced icospheres {min(tris):,}-{max(tris):,} triangles "
                        "(stratified over 1k-500k) + 20-300 spheres, 256^2-1024^2, 4-256 spp, depth 4-10",
            "msamples": round(samples / 1e6, 1), "triangles_total": sum(tris), "reps": reps,
            "batched_msamples_per_s": batched[0], "batched_min_max": batched[1:],
            "plain_msamples_per_s": plain[0], "plain_min_max": plain[1:],
            "single_scene_wavefront_msamples_per_s": single_wf[0],
            "batched_vs_plain": round(batched[0] / plain[0], 3),
            "batched_vs_plain_worst_case": round(batched[1] / plain[2], 3), "tune": tune,
            "batched_engines_msamples_per_s": eng,
            "device_bvh_build_ms": round(dev_build_ms, 1)}


def rmse_leg(g, cam, o, batch, tune, ref_spp, rows, oracle_spp, threads):
    """Equal RMSE on the 1080p frame: a converged reference rendered on the
    GPU (batched engine, ref_spp samples on streams disjoint from the frames'),
    cross-checked against the FP64 oracle on `rows` full-width rows of the
    frame (Monte Carlo consistency), then the RMSE of the batched engine's and
    the megakernel's 64-spp frames (the same sample streams 0..63) against it.
    Both engines evaluate the same estimator, so their RMSEs must agree to
    within the fp32 rounding of individual paths."""
    import numpy as np
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import checkers as ck
    from marmel_amd import abi
    kw = {}
    if tune:
        kw = dict(pool=tune["pool"], grid_blocks=tune["grid_blocks"], block=tune["block"],
                  bvh_width=tune["bvh_width"], fixed_batches=bool(tune["fixed_batches"]))
    t0 = time.perf_counter()
    (_, ref), = batch.render([g], [dict(scene=0, cam=cam, width=W, height=H, spp=ref_spp, depth=DEPTH,
                                        sample_offset=1 << 20)], **kw)
    t_ref = time.perf_counter() - t0
    (_, lin_b), = batch.render([g], [dict(scene=0, cam=cam, width=W, height=H, spp=SPP, depth=DEPTH)], **kw)
    _, lin_m = g.render(cam, W, H, SPP, DEPTH, engine=abi.ENGINE_MEGAKERNEL)
    ref = ref.astype(np.float64)
    rmse_b = float(np.sqrt(np.mean((lin_b - ref) ** 2)))
    rmse_m = float(np.sqrt(np.mean((lin_m - ref) ** 2)))
    # the same paths in both frames: their difference is rounding, far below the noise
    engines_rms = float(np.sqrt(np.mean((lin_b.astype(np.float64) - lin_m) ** 2)))
    y0 = H // 2 - rows // 2
    t1 = time.perf_counter()
    _, om, ov, _ = o.render(cam, W, H, oracle_spp, DEPTH, sample_offset=1 << 24, y0=y0, y1=y0 + rows,
                            threads=threads, want_var=True)
    t_or = time.perf_counter() - t1
    s = ck.mc_consistency(ref[y0:y0 + rows], ref_spp, om[y0:y0 + rows], ov[y0:y0 + rows], oracle_spp)
    # 256 oracle samples of heavy-tailed NEE terms: single pixels are held to
    # 95 %, 4x4 blocks and the row sums to the normal law
    ok = s["frac3"] >= 0.95 and s["bfrac3"] >= 0.97 and s["bz2"] < 1.35 and bool((np.abs(s["Z"]) < 4).all())
    log(f"bench: rmse reference {ref_spp} spp in {t_ref:.1f}s; oracle rows {y0}..{y0 + rows} at {oracle_spp} spp "
        f"in {t_or:.1f}s: {s}; rmse batched {rmse_b:.5f} megakernel {rmse_m:.5f} (engines differ by "
        f"{engines_rms:.2e} rms)")
    return {"reference": f"GPU batched engine, {ref_spp} spp on disjoint sample streams",
            "reference_vs_oracle_rows": {"rows": [y0, y0 + rows], "oracle_spp": oracle_spp,
                                         "frac3": round(s["frac3"], 4), "block_z2": round(s["bz2"], 3),
                                         "image_z": [round(float(z), 2) for z in s["Z"]], "consistent": ok},
            "batched_rmse": round(rmse_b, 6), "megakernel_rmse": round(rmse_m, 6),
            "rmse_ratio": round(rmse_b / rmse_m, 5), "engines_rms_difference": engines_rms,
            "equal_within_noise": bool(abs(rmse_b / rmse_m - 1.0) < 0.01)}


def stress_leg(tune, seconds, chunk=64, target=1024):
    """Stress workload (3840 x 2160, ~9.9 M triangles of seeded trees and a
    refractive pond with Beer-Lambert water, sun disk 0.5 deg, depth 12),
    progressive: chunks of `chunk` spp with increasing sample offsets through
    the batched engine, accumulated into the running mean, until `target` spp
    or `seconds` of rendering; the time for `target` spp is extrapolated
    linearly from the chunks' mean when the budget ends first."""
    import numpy as np
    from marmel_amd import Batch, GpuScene, HostScene, native
    t0 = time.time()
    hs = HostScene()
    info = hs.generate(native.GEN_FOREST, seed=5, grid=200, subdiv=6)
    # the deep, mixed tree of this scene walks faster with leaves of up to 8
    # primitives and 32 SAH bins (one 64-spp chunk: 1.210 s vs 1.243 s with
    # the defaults, the same image); the terrain frame and the 32-scene batch
    # gain nothing from them and keep the defaults
    gs = GpuScene(hs.soa(bvh=False), gpu_bvh=True, max_leaf=STRESS_MAX_LEAF, bins=STRESS_BINS)
    bi = gs.bvh_info()
    w, h, depth = 3840, 2160, 12
    cam = hs.camera(w / h)
    log(f"bench: stress scene {info.n_triangles} triangles, generated and uploaded in {time.time() - t0:.1f}s "
        f"(device BVH {bi.build_ms:.0f} ms)")
    kw = {}
    if tune:
        kw = dict(pool=tune["pool"], grid_blocks=tune["grid_blocks"], block=tune["block"],
                  bvh_width=tune["bvh_width"], fixed_batches=bool(tune["fixed_batches"]))
    b = Batch()
    acc = np.zeros((h, w, 3), np.float64)
    secs = []
    done = 0
    t_start = time.perf_counter()
    while done < target and (not secs or time.perf_counter() - t_start < seconds):
        t1 = time.perf_counter()
        (_, lin), = b.render([gs], [dict(scene=0, cam=cam, width=w, height=h, spp=chunk, depth=depth,
                                         sample_offset=done)], **kw)
        secs.append(time.perf_counter() - t1)
        acc += lin
        done += chunk
        log(f"bench: stress chunk {done // chunk}: {done} spp, {secs[-1]:.2f} s")
    b.close()
    gs.close()
    mean = acc / (done // chunk)
    per_chunk = sum(secs[1:]) / max(len(secs) - 1, 1) if len(secs) > 1 else secs[0]
    msps = w * h * chunk / per_chunk / 1e6
    return {"workload": f"3840x2160, {info.n_triangles:,} triangles of seeded trees + refractive pond "
                        f"(Beer-Lambert), sun radius 0.5 deg, depth {depth}, progressive {chunk}-spp chunks",
            "tree": {"max_leaf": STRESS_MAX_LEAF, "bins": STRESS_BINS, "nodes": bi.n_nodes, "depth": bi.max_depth},
            "spp_rendered": done, "chunks": len(secs), "msamples_per_s": round(msps, 1),
            "seconds_per_chunk": round(per_chunk, 3),
            "seconds_for_%d_spp" % target: round(per_chunk * target / chunk, 1),
            "extrapolated": done < target, "image_mean": round(float(mean.mean()), 5),
            "device_bvh_build_ms": round(bi.build_ms, 1)}


def post_leg(g, cam, reps):
    """Steps after the framebuffer on the same scene and camera: the adaptive
    Whitted sampler (fp64, n0 = 4, cap 16, tau 0.02, depth 5; the drop-in's
    precision) against the fixed 4-spp Whitted render, then the streak flare
    (defaults, enabled) on the device framebuffer of that frame."""
    from marmel_amd import GpuScene, abi
    from marmel_amd.abi import flare_default
    g.render_adaptive(cam, W, H, 4, 5, 16, 0.02)  # warm-up (allocates the state)
    ks, tots = [], []
    for _ in range(reps):
        _, _, _, tot, mx = g.render_adaptive(cam, W, H, 4, 5, 16, 0.02)
        ks.append(GpuScene.last_timing().kernel_ms)
        tots.append(tot)
    ad_ms = sum(ks) / len(ks)
    g.render(cam, W, H, 4, 5, mode=abi.MODE_WHITTED, precision=abi.PREC_FP64, fetch=False)
    fixed_ms = GpuScene.last_timing().kernel_ms
    fl = flare_default()
    g.flare(fl)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.flare(fl)
    flare_ms = (time.perf_counter() - t0) / reps * 1e3
    return {"workload": f"{W}x{H} Whitted fp64 depth 5, adaptive n0=4 cap=16 tau=0.02; flare defaults",
            "adaptive_kernel_ms": round(ad_ms, 2), "samples": tots[-1],
            "samples_per_pixel": round(tots[-1] / (W * H), 3),
            "adaptive_msamples_per_s": round(tots[-1] / ad_ms * 1e-3, 1),
            "fixed_4spp_kernel_ms": round(fixed_ms, 2), "flare_ms": round(flare_ms, 3)}


def hbm_calibration(mib):
    """Two plain kernels of known traffic for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings: a device copy (torch copy_) and an elementwise
    multiply (y = 2x), each reading and writing exactly `mib` MiB."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    y.copy_(x)
    torch.cuda.synchronize()
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    log(f"bench: calibration copy and multiply, each read {n * 4} bytes and wrote {n * 4} bytes")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-scale", type=int, default=3)
    ap.add_argument("--stats-scale", type=int, default=8,
                    help="downscale of the oracle render that counts traversal work")
    ap.add_argument("--no-graph", action="store_true",
                    help="time clocked direct launches instead of hipGraph replays (per-kernel times from "
                         "the timed frames themselves)")
    ap.add_argument("--engine", choices=["batched", "wavefront", "megakernel"], default="batched",
                    help="batched: the frame as one job of the batched engine (autotuned, cached); "
                         "wavefront: the single-scene wavefront render; megakernel: the plain kernel")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="frames of the other engine timed for the batched-vs-plain ratio")
    ap.add_argument("--pool", type=int, default=0, help="wavefront paths per batch (0 = default)")
    ap.add_argument("--grid-blocks", type=int, default=0)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint ms/grad leg")
    ap.add_argument("--adjoint-reps", type=int, default=3)
    ap.add_argument("--no-batch32", action="store_true", help="skip the 32-scene batched leg")
    ap.add_argument("--batch-reps", type=int, default=3, help="timed repetitions of each 32-scene rate")
    ap.add_argument("--no-rmse", action="store_true", help="skip the equal-RMSE leg (converged reference)")
    ap.add_argument("--no-stress", action="store_true", help="skip the 4K / 9.9M-triangle progressive leg")
    ap.add_argument("--stress-seconds", type=float, default=40.0,
                    help="render budget of the stress leg (1024 spp are extrapolated past it)")
    ap.add_argument("--rmse-ref-spp", type=int, default=2048, help="samples of the converged reference")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the adaptive Whitted + flare leg (steps after the framebuffer)")
    ap.add_argument("--only-post", action="store_true",
                    help="run only the adaptive Whitted + flare leg (for profiling its kernels)")
    ap.add_argument("--only-batch32", action="store_true",
                    help="run only the 32-scene batched leg (for profiling its kernels)")
    ap.add_argument("--only-stress", action="store_true",
                    help="run only the stress leg, with the launch settings the terrain frame is tuned to "
                         "(for profiling its kernels; --stress-seconds bounds it)")
    ap.add_argument("--hbm-calib", type=int, default=0, metavar="MIB",
                    help="only run the copy / multiply kernels of MIB MiB that calibrate the HBM counters")
    args = ap.parse_args()
    if args.hbm_calib > 0:
        hbm_calibration(args.hbm_calib)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path would shard by image tiles or sample ranges with one reduction", file=sys.stderr)
        sys.exit(2)

    from marmel_amd import Batch, GpuScene, abi, device_info

    if args.only_stress:
        log(f"bench: stress {stress_leg(STRESS_TUNE, args.stress_seconds)}")
        return
    if args.only_batch32:
        log(f"bench: batch32 {batch_leg(args.batch_reps)}")
        return
    t0 = time.time()
    hs, info, st = build_scene(args.grid)
    log(f"bench: scene {info.n_prims} prims ({info.n_triangles} triangles, {info.n_spheres} spheres), "
        f"BVH {st.n_nodes} nodes depth {st.max_depth} built in {st.build_seconds:.2f}s "
        f"(setup {time.time() - t0:.1f}s)")
    name, cus, mem = device_info()
    g = GpuScene(hs.soa())
    log(f"bench: device {name}, scene resident: {g.device_bytes / 2**20:.1f} MiB")
    cam = hs.camera(W / H)

    if args.only_post:
        log(f"bench: post {post_leg(g, cam, 3)}")
        return

    # Timed frames replay each batch's bounce sequence as a hipGraph (the
    # production path); per-kernel times come from one more frame launched
    # directly with a HIP-event pair around every kernel on the render stream.
    # The batched engine renders the frame as one job (autotuned once, the
    # choice cached on the handle; paths of several samples of a pixel in
    # flight, fixed-point pixel sums).
    wf = dict(engine=abi.ENGINE_WAVEFRONT, pool=args.pool, grid_blocks=args.grid_blocks)
    batch = Batch()
    job = dict(scene=0, cam=cam, width=W, height=H, spp=SPP, depth=DEPTH)

    def frame(engine, off, clocked=False):
        if engine == "batched":
            if clocked:
                t = batch.last_tune()
                batch.render([g], [dict(job, sample_offset=off)], fetch=False, pool=t["pool"],
                             grid_blocks=t["grid_blocks"], block=t["block"], bvh_width=t["bvh_width"],
                             fixed_batches=bool(t["fixed_batches"]), engine=t["engine"], use_graph=False,
                             kernel_clock=True)
            else:
                batch.render([g], [dict(job, sample_offset=off)], fetch=False, autotune=True,
                             use_graph=not args.no_graph, kernel_clock=args.no_graph)
        elif engine == "wavefront":
            kw = dict(wf, use_graph=False, kernel_clock=True) if clocked or args.no_graph else dict(wf, use_graph=True)
            g.render(cam, W, H, SPP, DEPTH, fetch=False, sample_offset=off, **kw)
        else:
            g.render(cam, W, H, SPP, DEPTH, fetch=False, sample_offset=off, engine=abi.ENGINE_MEGAKERNEL)
        return g.last_timing()

    for i in range(args.warmup):
        frame(args.engine, 0)
        log(f"bench: warmup {i + 1}/{args.warmup} done")
    kernel_ms, rays = [], 0
    ktimes = {}
    t_start = time.perf_counter()
    for i in range(args.steps):
        t = frame(args.engine, i * SPP)
        kernel_ms.append(t.kernel_ms)
        rays += t.rays
        if args.no_graph and args.engine != "megakernel":
            for k, (ms, n) in g.last_kernel_times().items():
                a = ktimes.setdefault(k, [0.0, 0])
                a[0] += ms
                a[1] += n
        log(f"bench: step {i + 1}/{args.steps} {args.engine} {t.kernel_ms:.1f} ms")
    wall = time.perf_counter() - t_start
    ms_step = wall / args.steps * 1e3
    samples = W * H * SPP
    value = samples * args.steps / wall / 1e6
    mean_kernel_ms = sum(kernel_ms) / len(kernel_ms)
    k_frames = args.steps
    tune = batch.last_tune() if args.engine == "batched" else None
    if args.engine != "megakernel" and not args.no_graph:
        # the same kernels launched directly, clocked (not part of value)
        tc = frame(args.engine, args.steps * SPP, clocked=True)
        for k, (ms, n) in g.last_kernel_times().items():
            ktimes[k] = [ms, n]
        k_frames = 1
        log(f"bench: clocked frame {tc.kernel_ms:.1f} ms (per-kernel times)")
    gpu_counts = None
    if args.engine == "batched":
        # the same frame once more with the counting kernels: the traversal
        # work on the tree the GPU actually walks (not timed)
        t = batch.last_tune()
        batch.set_counting(True)
        batch.render([g], [dict(job, sample_offset=args.steps * SPP)], fetch=False, pool=t["pool"],
                     grid_blocks=t["grid_blocks"], block=t["block"], bvh_width=t["bvh_width"],
                     fixed_batches=bool(t["fixed_batches"]), engine=t["engine"], use_graph=False)
        gpu_counts = batch.last_counts()
        batch.set_counting(False)
        log(f"bench: GPU traversal counts of one frame: {gpu_counts}")
    single_wf = None
    if args.engine == "batched" and args.plain_steps > 0:
        frame("wavefront", 0)
        t0 = time.perf_counter()
        for i in range(args.plain_steps):
            frame("wavefront", i * SPP)
        single_wf = samples * args.plain_steps / (time.perf_counter() - t0) / 1e6
        log(f"bench: single-scene wavefront {single_wf:.1f} Msamples/s")

    # the plain GPU kernel on the same frame and sample streams (equal RMSE:
    # both engines evaluate the identical estimator sample for sample)
    plain = None
    other = "megakernel" if args.engine != "megakernel" else "batched"
    if args.plain_steps > 0:
        frame(other, 0)
        t0 = time.perf_counter()
        for i in range(args.plain_steps):
            tt = frame(other, i * SPP)
            log(f"bench: {other} step {i + 1}/{args.plain_steps} {tt.kernel_ms:.1f} ms")
        plain = samples * args.plain_steps / (time.perf_counter() - t0) / 1e6

    adj = None if args.no_adjoint else adjoint_leg(args.adjoint_reps)
    if adj:
        log(f"bench: adjoint {adj}")
    b32 = None if args.no_batch32 else batch_leg(args.batch_reps)
    c5 = None if args.no_stress else stress_leg(tune, args.stress_seconds)
    if c5:
        log(f"bench: stress {c5}")
    eq = None
    if not args.no_rmse:
        hold_r, o_r = oracle_scene(hs)
        eq = rmse_leg(g, cam, o_r, batch, tune, args.rmse_ref_spp, 8, 256, min(os.cpu_count() or 1, 64))
    post = None if args.no_post else post_leg(g, cam, 3)
    if post:
        log(f"bench: post {post}")

    # host threads as the reference sizes its pool: min(nproc, 64)
    # (reference src/pathtrace.c:1210-1234)
    nproc, affinity, quota = host_cpus()
    threads = min(nproc, 64)
    cpu = None
    roof = None
    frame_gbs = None
    hold, o = oracle_scene(hs)
    st, st_samples, st_size = traversal_stats(o, cam, args.stats_scale, threads)
    log(f"bench: traversal counts from a {st_size} oracle render: {st}")
    if not args.no_cpu:
        port_msps, sample, sec = cpu_baseline(o, cam, args.cpu_scale, threads)
        log(f"bench: oracle (port) {port_msps:.3f} Msamples/s on {threads} threads ({sec:.1f}s); "
            f"nproc {nproc}, affinity {affinity}, cgroup quota {quota}")
        ref = reference_baseline(hold, cam, args.cpu_scale, threads)
        cpu = {"value": round(port_msps, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
               "sample": sample, "nproc": nproc, "affinity_cpus": affinity, "cgroup_cpu_quota": quota,
               "port_msamples_per_s": round(port_msps, 4)}
        if ref is not None:
            ref_msps, rsec = ref
            log(f"bench: reference renderer {ref_msps:.3f} Msamples/s on {threads} threads ({rsec:.1f}s)")
            cpu.update(value=round(ref_msps, 4), kind="reference",
                       reference_msamples_per_s=round(ref_msps, 4),
                       port_over_reference=round(port_msps / ref_msps, 3))
    scale = samples / st_samples
    frame_gbs = frame_bytes(st, st_samples) * scale / (mean_kernel_ms * 1e-3) / 1e9
    sh_ms, sh_n = ktimes.get("shadow", (0.0, 0))
    if sh_n > 0:
        per_launch = shadow_bytes(st) * scale * k_frames / sh_n
        avg_s = sh_ms / sh_n * 1e-3
        achieved = per_launch / avg_s / 1e9
        batched = args.engine == "batched"
        roof = {"bound": "hbm", "kernel": "k_wf_shadow_refill<true> (batched engine)" if batched
                else "k_wf_shadow_refill<false>",
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": SHADOW_TRAFFIC_PER_LAUNCH if batched else None,
                "bytes_per_launch": round(per_launch),
                "avg_launch_us": round(avg_s * 1e6, 1), "launches": sh_n,
                "kernel_share": round(sh_ms / sum(v[0] for v in ktimes.values()), 3),
                "achieved_basis": "oracle traversal counts on the reference-shaped tree (12-bin SAH, widest axis)"}
        if batched:
            roof["traffic_frac"] = round(SHADOW_TRAFFIC_PER_LAUNCH / avg_s / 1e9 / HBM_PEAK_GBS, 4)
            # what actually limits the kernel: dependent node fetches and lane
            # divergence (rocprofv3 counters of this kernel, see profiles/)
            roof["limiter"] = SHADOW_LIMITER
        if gpu_counts:
            c = gpu_counts
            gb = (c["shadow_queries"] * B_ANY_IO + c["shadow_nodes"] * B_NODE4 + c["shadow_triangles"] * B_TRI
                  + c["shadow_other"] * B_SPH) / sh_n
            roof["bytes_per_launch_gpu_tree"] = round(gb)
            roof["achieved_gpu_tree"] = round(gb / avg_s / 1e9, 1)
            roof["frac_gpu_tree"] = round(gb / avg_s / 1e9 / HBM_PEAK_GBS, 4)
            roof["gpu_nodes_per_query"] = round(c["shadow_nodes"] / max(c["shadow_queries"], 1), 2)
            roof["oracle_nodes_per_query"] = round(st["a_inner"] / max(st["any"], 1), 2)
    line = {
        "metric": "Msamples/s at 1080p, batched vs plain GPU at equal RMSE; adjoint ms/grad",
        "value": round(value, 2), "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "GPU path tracer at scale: 1024x1024-grid fBm terrain (2,093,058 tris) "
                               "+ 200 spheres (8 emissive), 1920x1080, 64 spp, depth 8, NEE + RR",
                   "engine": args.engine, "width": W, "height": H, "spp": SPP, "depth": DEPTH,
                   "grid": args.grid},
        "mrays_per_s": round(rays / (sum(kernel_ms) * 1e-3) / 1e6, 1),
        "frame_algorithmic_gbs": None if frame_gbs is None else round(frame_gbs, 1),
        "kernel_ms_per_frame": {k: round(v[0] / k_frames, 2) for k, v in ktimes.items() if v[1]},
        f"{other}_msamples_per_s": None if plain is None else round(plain, 2),
        "batched_vs_plain": None if plain is None else round(
            (value / plain) if args.engine != "megakernel" else (plain / value), 3),
        "single_scene_wavefront_msamples_per_s": None if single_wf is None else round(single_wf, 2),
        "tune": tune,
        "adjoint": adj,
        "batch32": b32,
        "equal_rmse": eq,
        "stress": c5,
        "after_framebuffer": post,
        "roofline": roof,
        "cpu_baseline": cpu,
        "gpu_traversal_counts_per_frame": gpu_counts,
        "device": name,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
