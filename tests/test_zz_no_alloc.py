"""GPU test of mh_render_noalloc, the entry point that allocates no device
memory: through the C ABI, into device buffers the caller owns, its images
equal mh_render's (the same megakernel) bit for bit, on two calls in a row
with the second image smaller, with and without a linear output buffer
(without one the library's fixed scratch holds the linear framebuffer), in
fp32 and fp64. Sizes beyond the scratch and the other renderers are refused
with the library's codes before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from marmel_amd import GpuScene, HostScene, abi, native

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scene():
    hs = HostScene()
    hs.generate(native.GEN_SPHERES, seed=1)
    hs.build_bvh()
    gs = GpuScene(hs.soa())
    yield hs, gs
    gs.close()


def noalloc(gs, cam, w, h, spp, depth, prec, with_linear):
    rgb = torch.full((h, w, 3), 7, dtype=torch.uint8, device="cuda")
    lin = torch.full((h, w, 3), -1.0, dtype=torch.float32, device="cuda") if with_linear else None
    torch.cuda.synchronize()
    rc = gs.render_noalloc(cam, w, h, spp, depth, rgb.data_ptr(), lin.data_ptr() if with_linear else None,
                           precision=prec)
    assert rc == 0, native.hip().mh_last_error().decode()
    return rgb.cpu().numpy(), (lin.cpu().numpy() if with_linear else None)


@pytest.mark.parametrize("prec", [abi.PREC_FP32, abi.PREC_FP64])
def test_noalloc_equals_render_twice_smaller_second(scene, prec):
    hs, gs = scene
    print("noalloc precision", prec, flush=True)
    for (w, h, spp, depth), with_linear in (((160, 120, 4, 6), False), ((61, 37, 3, 5), True),
                                            ((160, 120, 4, 6), True), ((45, 23, 2, 4), False)):
        cam = hs.camera(w / h)
        rgb_ref, lin_ref = gs.render(cam, w, h, spp, depth, precision=prec)
        rgb, lin = noalloc(gs, cam, w, h, spp, depth, prec, with_linear)
        print(" ", w, h, spp, depth, "linear buffer" if with_linear else "scratch linear", flush=True)
        assert np.array_equal(rgb, rgb_ref)
        if with_linear:
            assert np.array_equal(lin, lin_ref)


def test_noalloc_refuses_what_it_cannot_do(scene):
    hs, gs = scene
    cam = hs.camera(1.0)
    rgb = torch.zeros(16, dtype=torch.uint8, device="cuda")
    assert gs.render_noalloc(cam, 1449, 1449, 1, 1, rgb.data_ptr()) == 2   # 1449^2 > 2^21 pixels
    assert gs.render_noalloc(cam, 0, 4, 1, 1, rgb.data_ptr()) == 2
    assert gs.render_noalloc(cam, 4, 4, 0, 1, rgb.data_ptr()) == 3
    L = native.hip()
    for mode, engine in ((abi.MODE_WHITTED, abi.ENGINE_MEGAKERNEL), (abi.MODE_PATHTRACE, abi.ENGINE_WAVEFRONT)):
        p = abi.RenderParams(width=4, height=4, spp=1, max_depth=1, mode=mode, engine=engine)
        assert L.mh_render_noalloc(gs._h, C.byref(cam), C.byref(p), C.c_void_p(rgb.data_ptr()), None) == 4
    assert torch.count_nonzero(rgb).item() == 0
