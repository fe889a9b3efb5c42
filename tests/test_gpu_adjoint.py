"""GPU tests of the path-replay adjoint (mh_adjoint) and its torch wrapper.

  * the GPU adjoint (fp32) equals the oracle's FP64 adjoint on the same
    sample streams (the oracle's adjoint is pinned exactly by finite
    differences of its frozen-sampling render, tests/test_adjoint_oracle.py),
  * on the differentiable workload (512 x 512, 32 spp, depth 4, 64
    materials, L2 loss against an image of hidden seeded parameters) the
    adjoint matches central finite differences with common random numbers
    (eps 1e-3) within 2 % relative over the parameters that carry gradient,
  * the torch.autograd.Function returns the adjoint (with the sun-intensity
    chain rule), and a few gradient steps lower the loss.
"""
import numpy as np
import pytest
import torch

import checkers as ck
from marmel_amd import GpuScene, HostScene, abi, native
from marmel_amd.diff import DiffScene, mats_to_vector, vector_to_mats
from marmel_amd.workloads import hidden_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mscene():
    hs = HostScene()
    hs.generate(native.GEN_MATERIALS, seed=4)
    hs.build_bvh()
    prims, attrs, mats, sky, wl, wm = hs.arrays()
    gs = GpuScene(hs.soa())
    hold = ck.SoaHolder(prims, attrs, mats, sky, wl, wm)
    yield hs, gs, ck.OracleScene(hold.soa), mats, sky, hold
    gs.close()


def flat(gm, gs):
    return np.concatenate([np.asarray(gm, np.float64).reshape(-1), np.asarray(gs, np.float64)])


def test_gpu_adjoint_matches_fp64_oracle_adjoint(mscene):
    hs, gs, o, mats, sky, hold = mscene
    gs.set_materials(mats)
    gs.set_sky(sky)
    w, h, spp, depth = 64, 48, 4, 4
    cam = hs.camera(w / h)
    rng = np.random.default_rng(11)
    target = rng.uniform(0.0, 0.5, (h, w, 3))
    _, lin = gs.render(cam, w, h, spp, depth)
    dimg = 2.0 * (lin.astype(np.float64) - target) / lin.size
    gm, gsun = gs.adjoint(cam, w, h, spp, depth, dimg.astype(np.float32))
    om, osun = o.adjoint(cam, w, h, spp, depth, dimg, len(mats))
    g, r = flat(gm, gsun), flat(om, osun)
    rel = np.linalg.norm(g - r) / np.linalg.norm(r)
    big = np.abs(r) > 0.05 * np.abs(r).max()
    worst = np.max(np.abs(g - r)[big] / np.abs(r)[big])
    print("gpu vs fp64 adjoint: relative L2", rel, "worst large entry", worst, "n large", big.sum())
    assert rel < 0.02
    assert worst < 0.05


def test_gpu_adjoint_is_linear_and_deterministic(mscene):
    hs, gs, o, mats, sky, hold = mscene
    gs.set_materials(mats)
    gs.set_sky(sky)
    w, h = 32, 24
    cam = hs.camera(w / h)
    rng = np.random.default_rng(5)
    a = rng.normal(size=(h, w, 3)).astype(np.float32)
    g1 = flat(*gs.adjoint(cam, w, h, 4, 4, a))
    g2 = flat(*gs.adjoint(cam, w, h, 4, 4, a))
    g3 = flat(*gs.adjoint(cam, w, h, 4, 4, (2.0 * a).astype(np.float32)))
    assert np.allclose(g1, g2, rtol=1e-5, atol=1e-9)
    assert np.allclose(g3, 2.0 * g1, rtol=1e-4, atol=1e-8)


def test_recorded_forward_gives_the_same_adjoint(mscene):
    """mh_render_record stores each sample's sum; the adjoint that follows
    replays each path once and returns the two-pass adjoint (up to the order
    of float atomic sums), also under fog; an update drops the record (the
    next adjoint traces twice again, same result), and a record of another
    view is not used."""
    hs, gs, o, mats, sky, hold = mscene
    gs.set_materials(mats)
    gs.set_sky(sky)
    w, h, spp, depth = 96, 64, 4, 4
    cam = hs.camera(w / h)
    rng = np.random.default_rng(8)
    a = rng.normal(size=(h, w, 3)).astype(np.float32)
    ref = flat(*gs.adjoint(cam, w, h, spp, depth, a))
    t2 = GpuScene.last_timing().kernel_ms
    rgb, lin = gs.render_record(cam, w, h, spp, depth)
    # the adjoint's code: the frozen-sampling render at the current point,
    # bit for bit; the megakernel (built with FMA contraction) agrees up to
    # fp32 rounding flipping an occasional path decision
    assert np.array_equal(lin, gs.render_frozen(cam, w, h, spp, depth, mats))
    _, lin_m = gs.render(cam, w, h, spp, depth)
    print("record vs megakernel: mean |d|", np.abs(lin - lin_m).mean())
    assert np.abs(lin - lin_m).mean() < 1e-3
    g1 = flat(*gs.adjoint(cam, w, h, spp, depth, a))
    t1 = GpuScene.last_timing().kernel_ms
    print("adjoint with record %.2f ms, without %.2f ms" % (t1, t2))

    def same(g, r):
        # float atomic sums in a different order: equal to float precision
        rel = np.linalg.norm(g - r) / np.linalg.norm(r)
        big = np.abs(r) > 1e-3 * np.abs(r).max()
        worst = np.max(np.abs(g - r)[big] / np.abs(r)[big])
        print("  relative L2", rel, "worst entry above 1e-3 of max", worst)
        return rel < 1e-5 and worst < 1e-4
    assert same(g1, ref)
    assert t1 < t2
    gs.set_sky(sky)  # any update drops the record
    g2 = flat(*gs.adjoint(cam, w, h, spp, depth, a))
    assert same(g2, ref)
    gs.render_record(cam, w, h, spp, depth, sample_offset=4)  # another sample range
    g3 = flat(*gs.adjoint(cam, w, h, spp, depth, a))
    assert same(g3, ref)
    import scenes
    gs.set_fog(scenes.FOG_SETS["ashenvale"])
    ref_f = flat(*gs.adjoint(cam, w, h, spp, depth, a))
    gs.render_record(cam, w, h, spp, depth)
    g4 = flat(*gs.adjoint(cam, w, h, spp, depth, a))
    assert same(g4, ref_f)
    assert not same(ref_f, ref)
    gs.set_fog(abi.fog_default())


def c4_setup(mscene):
    hs, gs, o, mats, sky, hold = mscene
    w, h, spp, depth = 512, 512, 32, 4
    cam = hs.camera(1.0)
    gs.set_sky(sky)
    gs.set_materials(hidden_params(mats))
    _, target = gs.render(cam, w, h, spp, depth, sample_offset=1 << 20)  # independent samples
    gs.set_materials(mats)
    return hs, gs, mats, sky, cam, w, h, spp, depth, target.astype(np.float64)


def loss_of(gs, cam, w, h, spp, depth, target):
    _, lin = gs.render(cam, w, h, spp, depth)
    d = lin.astype(np.float64) - target
    return float((d * d).mean()), d


def frozen_loss(gs, cam, w, h, spp, depth, target, frozen):
    lin = gs.render_frozen(cam, w, h, spp, depth, frozen)
    d = lin.astype(np.float64) - target
    return float((d * d).mean())


def test_frozen_render_at_the_frozen_point_is_the_plain_render(mscene):
    hs, gs, o, mats, sky, hold = mscene
    gs.set_materials(mats)
    gs.set_sky(sky)
    cam = hs.camera(1.0)
    rgb, lin = gs.render(cam, 96, 96, 8, 4)
    lf = gs.render_frozen(cam, 96, 96, 8, 4, mats)
    close = (np.abs(lin - lf) <= 1e-3 * (1.0 + np.abs(lin))).all(axis=-1).mean()
    print("frozen vs plain render, pixels within 1e-3 relative:", close)
    assert close >= 0.995


def test_adjoint_matches_central_differences_c4(mscene):
    """C4: 512x512, 32 spp, depth 4, 64 materials, L2 loss vs hidden params.

    Central differences (eps 1e-3) with common random numbers of the render
    with sampling frozen at the current parameters -- the function the
    adjoint differentiates -- over the 16 parameters with the largest
    gradient: within 2 % relative. Plain-render central differences are a
    noisy estimate wherever a parameter moves sampling probabilities (lobe
    switches); for the sampling-independent parameters (emission, sun
    intensity) they must agree within 2 % too. A parameter on the boundary
    of its domain (emission 0: lowering it would make light negative) is
    differenced one-sidedly, from inside."""
    hs, gs, mats, sky, cam, w, h, spp, depth, target = c4_setup(mscene)
    L0, d = loss_of(gs, cam, w, h, spp, depth, target)
    gm, gsun = gs.adjoint(cam, w, h, spp, depth, (2.0 * d / d.size).astype(np.float32))
    g = gm.astype(np.float64).reshape(-1)
    v0 = mats_to_vector(mats)
    picks = list(np.argsort(-np.abs(g))[:16])
    eps = 1e-3
    fd, ad = [], []
    def steps(idx):
        return (eps, 0.0) if idx % 8 >= 5 and v0[idx] == 0.0 else (eps, -eps)

    try:
        for idx in picks:
            vals = []
            for st in steps(idx):
                v = v0.copy()
                v[idx] += st
                gs.set_materials(vector_to_mats(v, mats))
                vals.append(frozen_loss(gs, cam, w, h, spp, depth, target, mats))
            fd.append((vals[0] - vals[1]) / (steps(idx)[0] - steps(idx)[1]))
            ad.append(g[idx])
            print("param", idx // 8, "slot", idx % 8, "adjoint %.6e frozen-fd %.6e" % (ad[-1], fd[-1]))
        fd, ad = np.array(fd), np.array(ad)
        rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        print("C4 adjoint vs frozen central FD (eps 1e-3, CRN), 16 largest: relative L2", rel)
        assert rel < 0.02

        # plain-render central differences: emission of the lights and the sun
        em = [i for i in np.argsort(-np.abs(g)) if i % 8 >= 5][:6]
        fd2, ad2 = [], []
        for idx in em:
            vals = []
            for st in steps(idx):
                v = v0.copy()
                v[idx] += st
                gs.set_materials(vector_to_mats(v, mats))
                vals.append(loss_of(gs, cam, w, h, spp, depth, target)[0])
            fd2.append((vals[0] - vals[1]) / (steps(idx)[0] - steps(idx)[1]))
            ad2.append(g[idx])
        gs.set_materials(mats)
        base = np.array([sky.sun_color.x, sky.sun_color.y, sky.sun_color.z])
        vals = []
        for sgn in (1, -1):
            sk = abi.Sky.from_buffer_copy(bytes(sky))
            c = (1.0 + sgn * eps) * base
            sk.sun_color.x, sk.sun_color.y, sk.sun_color.z = float(c[0]), float(c[1]), float(c[2])
            gs.set_sky(sk)
            vals.append(loss_of(gs, cam, w, h, spp, depth, target)[0])
        gs.set_sky(sky)
        fd2.append((vals[0] - vals[1]) / (2 * eps))
        ad2.append(float(np.dot(base, gsun)))
        fd2, ad2 = np.array(fd2), np.array(ad2)
        rel2 = np.linalg.norm(ad2 - fd2) / np.linalg.norm(fd2)
        print("C4 emission + sun intensity vs plain central FD:", list(zip(ad2, fd2)), "relative L2", rel2)
        assert rel2 < 0.02
    finally:
        gs.set_materials(mats)
        gs.set_sky(sky)


def test_torch_function_gradient_and_descent(mscene):
    hs, gs, o, mats, sky, hold = mscene
    w, h, spp, depth = 96, 96, 8, 4
    cam = hs.camera(1.0)
    gs.set_sky(sky)
    gs.set_materials(hidden_params(mats))
    _, target = gs.render(cam, w, h, spp, depth, sample_offset=1 << 20)
    target = torch.from_numpy(target.astype(np.float32))
    ds = DiffScene(gs, mats, sky, cam, w, h, spp, depth)
    p = ds.initial_params().requires_grad_(True)
    img = ds.render(p)
    loss = ((img - target) ** 2).mean()
    loss.backward()
    # the wrapper's gradient is the adjoint's, sun colour mapped to intensity
    ds.apply(p.detach())
    d = (2.0 * (img.detach() - target) / img.numel()).numpy()
    gm, gsun = gs.adjoint(cam, w, h, spp, depth, d)
    ref = np.concatenate([gm.reshape(-1), [np.dot(ds.base_sun, gsun)]])
    assert np.allclose(p.grad.numpy(), ref, rtol=1e-4, atol=1e-10)
    # gradient descent on albedos / emission only (bounded, well conditioned)
    mask = torch.zeros_like(p)
    mask.view(-1)[:-1].view(len(mats), 8)[:, [0, 1, 2, 5, 6, 7]] = 1.0
    losses = [float(loss.detach())]
    q = p.detach().clone()
    for it in range(6):
        q.requires_grad_(True)
        l2 = ((ds.render(q) - target) ** 2).mean()
        l2.backward()
        with torch.no_grad():
            q = q - 2.0 * mask * q.grad / (q.grad.abs().max() + 1e-12) * 0.02
            q[:-1].view(len(mats), 8)[:, 0:3].clamp_(0.01, 0.99)
        losses.append(float(l2.detach()))
    print("descent losses", losses)
    assert losses[-1] < losses[0]
    gs.set_materials(mats)
