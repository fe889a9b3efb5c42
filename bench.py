#!/usr/bin/env python3
"""Benchmark of the plain GPU path tracer on the 1080p terrain workload.

Workload (one GPU): fBm heightfield terrain on a 1024 x 1024 vertex grid
(2,093,058 fp32 triangles) + 200 seeded spheres (8 emissive, radiance
U[2,20]), GGX PBR materials, 1920 x 1080, 64 spp, depth 8, NEE to the sun and
the 8 emitters, Russian roulette. One step = one full frame rendered into the
scene's device framebuffer (scene already resident in HBM).

Prints ONE JSON line. Also reports:
  roofline      algorithmic bytes per frame (traversal counts of the CPU oracle
                on a downscaled render of the same scene, priced with the fixed
                record sizes: 64 B node, 48 B triangle, 32 B sphere, 16 B plane
                test, 48 B ray+hit per closest query, 36 B per any-hit query,
                192 B path state + material per bounce, 12 B per sample) over
                the mean kernel time measured with HIP events on the render
                stream; peak = MI355X HBM3E 8 TB/s.
  cpu_baseline  the oracle (FP64 C restatement of the reference renderer,
                byte-identical to it) on the host cores, on a downscaled
                (1/8 per axis) render of the same scene at the same spp/depth.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "marmel-trace_amd"))

HBM_PEAK_GBS = 8000.0
W, H, SPP, DEPTH = 1920, 1080, 64, 8
GRID = 1024


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_scene(grid):
    from marmel_amd import HostScene, native
    hs = HostScene()
    info = hs.generate(native.GEN_TERRAIN, seed=2, grid=grid)
    st = hs.build_bvh()
    return hs, info, st


def algorithmic_bytes(stats, samples):
    closest, anyq, nodes, prims, planes, tris, ext = stats
    sph = prims - tris
    return (closest * 48 + anyq * 36 + nodes * 64 + tris * 48 + sph * 32 + planes * 16
            + ext * 192 + samples * 12)


def cpu_baseline(hs, cam_fn, scale, threads):
    """Oracle render of a downscaled frame; returns (Msamples/s, bytes/sample, sample, cores)."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import checkers as ck
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm)
    o = ck.OracleScene(hold.soa)
    w, h = W // scale, H // scale
    cam = cam_fn(W / H)
    _, _, _, sec = o.render(cam, w, h, SPP, DEPTH, mode=0, threads=threads)
    st = o.stats()
    samples = w * h * SPP
    return samples / sec / 1e6, algorithmic_bytes(st, samples) / samples, \
        f"{w}x{h} downscaled frame of the same scene/camera, {SPP} spp, depth {DEPTH}", sec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--grid", type=int, default=GRID)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-scale", type=int, default=3)
    ap.add_argument("--engine", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--plain-steps", type=int, default=2,
                    help="frames of the other engine timed for the batched-vs-plain ratio")
    ap.add_argument("--pool", type=int, default=0, help="wavefront paths per batch (0 = default)")
    ap.add_argument("--grid-blocks", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by image tiles / samples, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from marmel_amd import GpuScene, abi, device_info

    t0 = time.time()
    hs, info, st = build_scene(args.grid)
    log(f"bench: scene {info.n_prims} prims ({info.n_triangles} triangles, {info.n_spheres} spheres), "
        f"BVH {st.n_nodes} nodes depth {st.max_depth} built in {st.build_seconds:.2f}s "
        f"(setup {time.time() - t0:.1f}s)")
    name, cus, mem = device_info()
    g = GpuScene(hs.soa())
    log(f"bench: device {name}, scene resident: {g.device_bytes / 2**20:.1f} MiB")
    cam = hs.camera(W / H)

    engines = {"wavefront": dict(engine=abi.ENGINE_WAVEFRONT, pool=args.pool, grid_blocks=args.grid_blocks,
                                 use_graph=True),
               "megakernel": dict(engine=abi.ENGINE_MEGAKERNEL)}
    main_kw = engines[args.engine]
    for i in range(args.warmup):
        g.render(cam, W, H, SPP, DEPTH, fetch=False, **main_kw)
        log(f"bench: warmup {i + 1}/{args.warmup} done")
    kernel_ms, rays = [], 0
    t_start = time.perf_counter()
    for i in range(args.steps):
        g.render(cam, W, H, SPP, DEPTH, fetch=False, sample_offset=i * SPP, **main_kw)
        t = g.last_timing()
        kernel_ms.append(t.kernel_ms)
        rays += t.rays
        log(f"bench: step {i + 1}/{args.steps} {args.engine} {t.kernel_ms:.1f} ms")
    wall = time.perf_counter() - t_start
    ms_step = wall / args.steps * 1e3
    samples = W * H * SPP
    value = samples * args.steps / wall / 1e6
    mean_kernel_ms = sum(kernel_ms) / len(kernel_ms)

    # the plain GPU kernel on the same frame and sample streams (equal RMSE:
    # both engines evaluate the identical estimator sample for sample)
    plain = None
    other = "megakernel" if args.engine == "wavefront" else "wavefront"
    if args.plain_steps > 0:
        g.render(cam, W, H, SPP, DEPTH, fetch=False, **engines[other])
        t0 = time.perf_counter()
        for i in range(args.plain_steps):
            g.render(cam, W, H, SPP, DEPTH, fetch=False, sample_offset=i * SPP, **engines[other])
            log(f"bench: {other} step {i + 1}/{args.plain_steps} {g.last_timing().kernel_ms:.1f} ms")
        plain = samples * args.plain_steps / (time.perf_counter() - t0) / 1e6

    cpu = None
    bytes_per_sample = None
    if not args.no_cpu:
        threads = min(os.cpu_count() or 1, 64)
        cpu_msps, bytes_per_sample, sample, sec = cpu_baseline(hs, hs.camera, args.cpu_scale, threads)
        log(f"bench: cpu baseline {cpu_msps:.3f} Msamples/s on {threads} threads ({sec:.1f}s)")
        cpu = {"value": round(cpu_msps, 4), "unit": "Msamples/s", "cores": threads, "kind": "port",
               "sample": sample}
    roof = None
    if bytes_per_sample is not None:
        achieved = bytes_per_sample * samples / (mean_kernel_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_sample": round(bytes_per_sample, 1), "kernel_ms": round(mean_kernel_ms, 2)}
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
        f"{other}_msamples_per_s": None if plain is None else round(plain, 2),
        "batched_vs_plain": None if plain is None else round(
            (value / plain) if args.engine == "wavefront" else (plain / value), 3),
        "roofline": roof,
        "cpu_baseline": cpu,
        "device": name,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
