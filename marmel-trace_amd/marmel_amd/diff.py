"""Differentiable rendering: the GPU path tracer as a torch.autograd.Function.

Parameters (one flat float64 CPU tensor): for each material
[albedo r, g, b, roughness, metallic, emissive r, g, b], then the sun
intensity s (the sky's sun colour is s times a fixed base colour). 64
materials give 64 * 8 + 1 = 513 parameters.

forward  writes the parameters into the uploaded scene (mh_scene_set_materials,
         mh_scene_set_sky) and renders the fp32 path tracer; returns the
         linear image (h, w, 3) as a float32 CPU tensor.
backward runs the path-replay adjoint kernel (mh_adjoint) on the same sample
         streams with the incoming image gradient and maps the sun-colour
         gradient to the intensity by the chain rule (d/ds = base . d/dcolour).

torch here is plumbing for the optimiser; the GPU work is the library's.
"""
import numpy as np
import torch

from . import abi

FIELDS = (("albedo", "x"), ("albedo", "y"), ("albedo", "z"), ("roughness", None), ("metallic", None),
          ("emissive", "x"), ("emissive", "y"), ("emissive", "z"))


def mats_to_vector(mats):
    out = np.zeros((len(mats), 8))
    for k, (f, c) in enumerate(FIELDS):
        out[:, k] = mats[f] if c is None else mats[f][c]
    return out.reshape(-1)


def vector_to_mats(v, mats_template):
    mats = mats_template.copy()
    v = np.asarray(v, np.float64).reshape(len(mats), 8)
    for k, (f, c) in enumerate(FIELDS):
        if c is None:
            mats[f] = v[:, k]
        else:
            mats[f][c] = v[:, k]
    return mats


class DiffScene:
    """A GPU scene whose materials and sun intensity are optimisable."""

    def __init__(self, gpu_scene, mats, sky, cam, width, height, spp, depth, sample_offset=0):
        self.gs = gpu_scene
        self.mats0 = np.array(mats, copy=True)
        self.sky0 = abi.Sky.from_buffer_copy(bytes(sky))
        self.base_sun = np.array([sky.sun_color.x, sky.sun_color.y, sky.sun_color.z])
        self.cam, self.w, self.h, self.spp, self.depth = cam, width, height, spp, depth
        self.sample_offset = sample_offset
        self.n = len(mats)

    def initial_params(self):
        return torch.tensor(np.concatenate([mats_to_vector(self.mats0), [1.0]]), dtype=torch.float64)

    def apply(self, params):
        """Write a parameter vector into the GPU scene (skipped when it is
        already there: an update would drop the forward's recorded samples)."""
        v = params.detach().cpu().numpy().astype(np.float64)
        last = getattr(self, "_applied", None)
        if last is not None and last[1] == getattr(self.gs, "updates", 0) and np.array_equal(last[0], v):
            return
        self.gs.set_materials(vector_to_mats(v[:-1], self.mats0))
        sky = abi.Sky.from_buffer_copy(bytes(self.sky0))
        sc = v[-1] * self.base_sun
        sky.sun_color.x, sky.sun_color.y, sky.sun_color.z = float(sc[0]), float(sc[1]), float(sc[2])
        self.gs.set_sky(sky)
        self._applied = (v.copy(), getattr(self.gs, "updates", 0))

    def render(self, params):
        return RenderFunction.apply(params, self)


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, ds):
        ds.apply(params)
        # recording render: the backward pass then replays each path once
        _, lin = ds.gs.render_record(ds.cam, ds.w, ds.h, ds.spp, ds.depth, sample_offset=ds.sample_offset)
        ctx.ds = ds
        ctx.save_for_backward(params.detach().clone())
        return torch.from_numpy(np.ascontiguousarray(lin, dtype=np.float32))

    @staticmethod
    def backward(ctx, grad_out):
        (params,) = ctx.saved_tensors
        ds = ctx.ds
        ds.apply(params)  # the adjoint differentiates at the forward's parameters
        d = grad_out.detach().cpu().numpy().astype(np.float32)
        gm, gsun = ds.gs.adjoint(ds.cam, ds.w, ds.h, ds.spp, ds.depth, d, sample_offset=ds.sample_offset)
        g = np.concatenate([gm.reshape(-1).astype(np.float64), [float(np.dot(ds.base_sun, gsun))]])
        return torch.from_numpy(g).to(params.dtype), None
