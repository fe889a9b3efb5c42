"""Seeded workloads of the benchmark and the tests (all synthetic).

terrain    1024 x 1024-vertex fBm terrain (2,093,058 triangles) + 200 spheres
           (8 emissive, radiance U[2,20]), GGX PBR; 1920 x 1080, 64 spp, depth 8.
batch32    32 scenes of uneven size: a triangle budget stratified
           log-uniform in [1,000, 500,000] met by 1..15 randomly displaced icospheres of one
           subdivision level (1,280 .. 491,520 triangles per scene), 20..300
           spheres, square resolutions 256..1024, spp log-uniform in
           [4, 256], depth 4..10, all drawn from seed 3.
materials  the differentiable workload: plane + 63 spheres, 64 GGX PBR
           materials; 512 x 512, 32 spp, depth 4; the target image is rendered
           from hidden parameters drawn from seed 4.
"""
import numpy as np

from . import native
from .diff import mats_to_vector, vector_to_mats
from .scene import HostScene


def terrain(grid=1024, seed=2):
    hs = HostScene()
    info = hs.generate(native.GEN_TERRAIN, seed=seed, grid=grid)
    st = hs.build_bvh()
    return hs, info, st


def batch32_spec(seed=3, n=32):
    """[(scene seed, subdiv, icospheres, n_spheres, res, spp, depth)] of the
    batched workload. The triangle budgets T are stratified log-uniform over
    [1e3, 5e5] (one in each of n equal log-strata, in shuffled order, so the
    batch spans the whole range): icospheres of 20 * 4^subdiv triangles with
    subdiv = max(3, floor(log4(T / 20)) - 1), as many as fit in T (at least
    one)."""
    rng = np.random.default_rng(seed)
    strata = rng.permutation(n)
    out = []
    for i in range(n):
        u = (strata[i] + rng.uniform()) / n
        budget = float(np.exp(np.log(1e3) + u * (np.log(5e5) - np.log(1e3))))
        sub = max(3, int(np.floor(np.log(budget / 20.0) / np.log(4.0))) - 1)
        count = max(1, int(budget // (20 * 4 ** sub)))
        ns = int(rng.integers(20, 301))
        res = int(rng.integers(256, 1025))
        spp = int(np.round(np.exp(rng.uniform(np.log(4), np.log(256)))))
        depth = int(rng.integers(4, 11))
        out.append((1000 + i, sub, count, ns, res, spp, depth))
    return out


def batch32_scenes(spec, host_bvh=True):
    """The 32 scenes; host_bvh=False leaves the trees to the device build
    (GpuScene(hs.soa(bvh=False), gpu_bvh=True))."""
    scenes = []
    for sseed, sub, count, ns, res, spp, depth in spec:
        hs = HostScene()
        hs.generate(native.GEN_ICOSPHERE, seed=sseed, subdiv=sub, n_spheres=ns, count=count)
        if host_bvh:
            hs.build_bvh()
        scenes.append(hs)
    return scenes


def materials_scene(seed=4):
    hs = HostScene()
    hs.generate(native.GEN_MATERIALS, seed=seed)
    hs.build_bvh()
    return hs


def hidden_params(mats, seed=4):
    """The hidden material parameters the differentiable workload's target
    image is rendered from: the scene's, perturbed by seeded noise."""
    rng = np.random.default_rng(seed)
    v = mats_to_vector(mats).reshape(len(mats), 8)
    v[:, 0:3] = np.clip(v[:, 0:3] + rng.uniform(-0.15, 0.15, (len(mats), 3)), 0.05, 0.95)
    v[:, 3] = np.clip(v[:, 3] + rng.uniform(-0.2, 0.2, len(mats)), 0.05, 0.95)
    v[:, 4] = np.clip(v[:, 4] + rng.uniform(-0.2, 0.2, len(mats)), 0.02, 0.98)
    em = v[:, 5:8] > 0
    v[:, 5:8] = np.where(em, v[:, 5:8] * rng.uniform(0.7, 1.3, (len(mats), 3)), 0.0)
    return vector_to_mats(v.reshape(-1), mats)
