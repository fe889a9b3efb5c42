"""Python side of the CPU build of the library's render code (test
infrastructure only): builds tests/hostdev/Makefile's variants and renders a
flattened scene through one of them (hd_main.cpp), returning the linear
means, the 8-bit image and the ray count."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

from marmel_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = ("hd_render", "hd_render_asan", "hd_render_pattern", "hd_render_zero",
            "hd_render_anysort0", "hd_render_heldfp64", "hd_render_nobundle")


class Request(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("n_prims", "has_attrs", "n_materials", "n_nodes", "n_leaf",
                                         "n_planes", "w", "h", "spp", "depth", "mode", "precision", "s0",
                                         "threads")] + \
               [("water_level", C.c_double), ("water_material", C.c_int32), ("width", C.c_int32),
                ("sky", abi.Sky), ("fog", abi.Fog), ("ocean", abi.Ocean), ("cam", abi.Camera)]


def build(jobs=6):
    """Builds every variant (make, under a file lock so that parallel test
    workers do not write the same binaries); raises with the compiler output
    on failure."""
    import fcntl
    os.makedirs(os.path.join(HERE, "out"), exist_ok=True)
    with open(os.path.join(HERE, "out", ".lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        r = subprocess.run(["make", "-s", "-C", HERE, f"-j{jobs}"], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("host build of the render code failed:\n" + r.stdout + r.stderr)


def _bytes(ptr, n, ctype):
    return C.string_at(ptr, n * C.sizeof(ctype)) if n else b""


def render(variant, soa, cam, w, h, spp, depth, mode=abi.MODE_PATHTRACE, precision=abi.PREC_FP64, s0=0,
           threads=8, width=4, env=None):
    """Renders through tests/hostdev/out/<variant>; soa must carry a BVH. width 2 walks the
    binary tree, 4 the 4-wide one (when the scene's tree is shallow enough to build it)."""
    assert soa.n_nodes > 0, "the host build traverses the flattened scene's own BVH"
    q = Request(n_prims=soa.n_prims, has_attrs=1 if soa.attrs else 0, n_materials=soa.n_materials,
                n_nodes=soa.n_nodes, n_leaf=soa.n_leaf_prims, n_planes=soa.n_planes, w=w, h=h, spp=spp,
                depth=depth, mode=mode, precision=precision, s0=s0, threads=threads, width=width,
                water_level=soa.water_level, water_material=soa.water_material, sky=soa.sky, fog=soa.fog,
                ocean=soa.ocean, cam=cam)
    blob = (bytes(q) + _bytes(soa.prims, soa.n_prims, abi.Prim) +
            (_bytes(soa.attrs, soa.n_prims, abi.PrimAttr) if soa.attrs else b"") +
            _bytes(soa.materials, soa.n_materials, abi.Material) +
            _bytes(soa.nodes, soa.n_nodes, abi.BvhNode) + _bytes(soa.leaf_prims, soa.n_leaf_prims, C.c_int32) +
            _bytes(soa.planes, soa.n_planes, C.c_int32))
    with tempfile.TemporaryDirectory() as d:
        req, out = os.path.join(d, "req.bin"), os.path.join(d, "out.bin")
        with open(req, "wb") as f:
            f.write(blob)
        e = dict(os.environ)
        e.update(env or {})
        r = subprocess.run([os.path.join(HERE, "out", variant), req, out], capture_output=True, text=True, env=e)
        if r.returncode:
            raise RuntimeError(f"{variant} exited {r.returncode}:\n{r.stderr[-4000:]}")
        raw = open(out, "rb").read()
    n = w * h * 3
    lin = np.frombuffer(raw, np.float32, n).reshape(h, w, 3)
    rgb = np.frombuffer(raw, np.uint8, n, offset=4 * n).reshape(h, w, 3)
    rays = int(np.frombuffer(raw, np.int64, 1, offset=5 * n)[0])
    return rgb, lin, rays
