"""Host scene (libmarmel_host) and device scene (libmarmel_hip) wrappers.

HostScene  -- the C11 host scene: materials, primitives in file order, sky,
              BVH build, flattened mh_scene_soa view, look-at camera.
GpuScene   -- a scene uploaded to the GPU; render / trace entry points.

This mirrors the reference's render interface (reference src/pathtrace.h:78-80
pathtrace_render, src/render.h:29-31 render_image, src/scene.h:162-169
scene_intersect / scene_occluded) with the same return codes.
"""
import ctypes as C

import numpy as np

from . import abi
from . import native


class MarmelError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"code {code}: {msg}")
        self.code = code


class HostScene:
    def __init__(self):
        self._L = native.host()
        self._h = C.c_void_p(self._L.mt_scene_new())
        if not self._h:
            raise MemoryError("mt_scene_new failed")
        self.info = None
        self.bvh_stats = None
        self._keep = []

    def close(self):
        if self._h:
            self._L.mt_scene_free(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _err(self, rc):
        raise MarmelError(rc, self._L.mt_last_error().decode())

    # ---- construction -------------------------------------------------
    def generate(self, kind, seed=1, grid=0, n_spheres=0, n_emissive=0, subdiv=0, count=0):
        p = native.GenParams(kind=kind, seed=seed, grid=grid, n_spheres=n_spheres,
                             n_emissive=n_emissive, subdiv=subdiv, count=count)
        info = native.GenInfo()
        rc = self._L.mt_scene_generate(self._h, C.byref(p), C.byref(info))
        if rc:
            raise MarmelError(rc, "scene generation failed")
        self.info = info
        return info

    def add_material(self, mat):
        idx = self._L.mt_scene_add_material(self._h, C.byref(mat))
        if idx < 0:
            raise MemoryError("add_material failed")
        return idx

    def add_materials(self, mats):
        """mats: numpy array of MATERIAL_DTYPE."""
        mats = np.ascontiguousarray(mats, dtype=abi.MATERIAL_DTYPE)
        first = None
        for i in range(len(mats)):
            m = abi.Material.from_buffer_copy(mats[i:i + 1].tobytes())
            idx = self.add_material(m)
            first = idx if first is None else first
        return first

    def add_prims(self, prims, attrs=None):
        prims = np.ascontiguousarray(prims, dtype=abi.PRIM_DTYPE)
        if attrs is not None:
            attrs = np.ascontiguousarray(attrs, dtype=abi.ATTR_DTYPE)
        first = self._L.mt_scene_add_prims(self._h, abi.ptr(prims), abi.ptr(attrs), len(prims))
        if first < 0:
            self._err(-1)
        return first

    def set_sky(self, sky):
        self._L.mt_scene_set_sky(self._h, C.byref(sky))

    def set_water(self, level, material):
        self._L.mt_scene_set_water(self._h, level, material)

    def set_fog(self, fog):
        """Scene fog (abi.Fog; density 0 disables it)."""
        self._L.mt_scene_set_fog(self._h, C.byref(fog))

    @staticmethod
    def obj_xform(**kw):
        """mt_obj_xform with the defaults, fields overridden by keyword."""
        x = native.ObjXform()
        native.host().mt_obj_xform_default(C.byref(x))
        for k, v in kw.items():
            if isinstance(v, (list, tuple)):
                getattr(x, k)[:] = [float(t) for t in v]
            else:
                setattr(x, k, v)
        return x

    def add_obj(self, text, material, xform=None, threads=0):
        """Add the triangles of Wavefront OBJ text (bytes or str); returns the count."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        n = self._L.mt_scene_add_obj(self._h, data, len(data), material,
                                     C.byref(xform) if xform is not None else None, threads)
        if n < 0:
            raise MarmelError(int(n), self._L.mt_last_error().decode())
        return n

    def add_obj_file(self, path, material, xform=None, threads=0):
        n = self._L.mt_scene_add_obj_file(self._h, str(path).encode(), material,
                                          C.byref(xform) if xform is not None else None, threads)
        if n < 0:
            raise MarmelError(int(n), self._L.mt_last_error().decode())
        return n

    @staticmethod
    def ocean_desc(**kw):
        """mt_ocean_desc with the reference block's defaults, overridden by keyword."""
        o = native.OceanDesc()
        native.host().mt_ocean_default(C.byref(o))
        for k, v in kw.items():
            if isinstance(v, (list, tuple)):
                getattr(o, k)[:] = [float(t) for t in v]
            else:
                setattr(o, k, v)
        return o

    def add_ocean(self, desc, material, threads=0):
        n = self._L.mt_scene_add_ocean(self._h, C.byref(desc), material, threads)
        if n < 0:
            raise MarmelError(int(n), self._L.mt_last_error().decode())
        return n

    def counts(self):
        n, m = C.c_int(), C.c_int()
        self._L.mt_scene_counts(self._h, C.byref(n), C.byref(m))
        return n.value, m.value

    def build_bvh(self, max_leaf=4, bins=16):
        opts = native.BvhOptions(max_leaf=max_leaf, bins=bins, threads=0)
        st = native.BvhStats()
        rc = self._L.mt_scene_build_bvh(self._h, C.byref(opts), C.byref(st))
        if rc:
            self._err(rc)
        self.bvh_stats = st
        return st

    def soa(self, bvh=True):
        """Flattened scene; bvh=False leaves the tree out (for a device build)."""
        s = abi.SceneSoa()
        rc = (self._L.mt_scene_soa if bvh else self._L.mt_scene_soa_no_bvh)(self._h, C.byref(s))
        if rc:
            self._err(rc)
        return s

    def camera(self, aspect, eye=None, at=None, up=None, vfov=None):
        info = self.info
        eye = eye if eye is not None else list(info.eye)
        at = at if at is not None else list(info.at)
        up = up if up is not None else list(info.up)
        vfov = vfov if vfov is not None else info.vfov
        return make_camera(eye, at, up, vfov, aspect)

    # ---- numpy views of the flattened scene (for checkers) ----------------
    def arrays(self):
        s = self.soa(bvh=False)
        n, m = s.n_prims, s.n_materials

        def view(ptr, count, dtype):
            if count == 0 or not ptr:
                return np.zeros(0, dtype)
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)),
                                         (count * dtype.itemsize,)).view(dtype).copy()
        prims = view(s.prims, n, abi.PRIM_DTYPE)
        mats = view(s.materials, m, abi.MATERIAL_DTYPE)
        attrs = None
        if s.attrs and n:
            attrs = np.ctypeslib.as_array(C.cast(s.attrs, C.POINTER(C.c_uint8)),
                                          (n * abi.ATTR_DTYPE.itemsize,)).view(abi.ATTR_DTYPE).copy()
        return prims, attrs, mats, s.sky, s.water_level, s.water_material

    def ocean(self):
        """The scene's mh_ocean (enabled once an ocean was added)."""
        return abi.Ocean.from_buffer_copy(bytes(self.soa(bvh=False).ocean))


def make_camera(eye, at, up, vfov, aspect):
    L = native.host()
    cam = abi.Camera()
    e = (C.c_double * 3)(*eye)
    a = (C.c_double * 3)(*at)
    u = (C.c_double * 3)(*up)
    L.mt_camera_create(e, a, u, vfov, aspect, C.byref(cam))
    return cam


def write_bmp(path, rgb):
    h, w, _ = rgb.shape
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    rc = native.host().mt_write_bmp(path.encode(), abi.ptr(rgb), w, h)
    if rc:
        raise MarmelError(rc, native.host().mt_last_error().decode())


def encode_bmp(rgb):
    h, w, _ = rgb.shape
    rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
    L = native.host()
    n = L.mt_encode_bmp(abi.ptr(rgb), w, h, None, 0)
    buf = np.zeros(n, np.uint8)
    L.mt_encode_bmp(abi.ptr(rgb), w, h, abi.ptr(buf), n)
    return buf.tobytes()


class GpuScene:
    """A scene resident in GPU memory (mh_scene_create), or, with
    gpu_bvh=True, uploaded without a tree and its BVH built on the device
    (mh_scene_create_gpu_bvh with max_leaf / bins)."""

    def __init__(self, soa, gpu_bvh=False, max_leaf=0, bins=0):
        self._L = native.hip()
        h = C.c_void_p()
        if gpu_bvh:
            rc = self._L.mh_scene_create_gpu_bvh(C.byref(soa), max_leaf, bins, C.byref(h))
        else:
            rc = self._L.mh_scene_create(C.byref(soa), C.byref(h))
        if rc:
            raise MarmelError(rc, self._L.mh_last_error().decode())
        self._h = h
        self.n_materials = soa.n_materials

    def bvh_info(self):
        info = abi.BvhInfo()
        self._check(self._L.mh_scene_bvh_info(self._h, C.byref(info)))
        return info

    def bvh_export(self):
        """(nodes, leaf_prims) of a device-built tree (numpy)."""
        info = self.bvh_info()
        nodes = np.zeros(info.n_nodes, abi.NODE_DTYPE)
        leaf = np.zeros(info.n_leaf, np.int32)
        self._check(self._L.mh_scene_bvh_export(self._h, abi.ptr(nodes), abi.ptr(leaf)))
        return nodes, leaf

    def close(self):
        if self._h:
            self._L.mh_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise MarmelError(rc, self._L.mh_last_error().decode())

    @property
    def device_bytes(self):
        return self._L.mh_scene_device_bytes(self._h)

    def render(self, cam, width, height, spp, depth, mode=abi.MODE_PATHTRACE,
               precision=abi.PREC_FP32, sample_offset=0, fetch=True,
               engine=abi.ENGINE_MEGAKERNEL, pool=0, grid_blocks=0, use_graph=False,
               kernel_clock=False, bvh_width=0):
        p = abi.RenderParams(width=width, height=height, spp=spp, max_depth=depth, mode=mode,
                             precision=precision, sample_offset=sample_offset, engine=engine,
                             pool=pool, grid_blocks=grid_blocks, use_graph=int(use_graph),
                             kernel_clock=int(kernel_clock), bvh_width=bvh_width)
        rgb = lin = None
        if fetch:
            rgb = np.zeros((height, width, 3), np.uint8)
            lin = np.zeros((height, width, 3), np.float32)
        rc = self._L.mh_render(self._h, C.byref(cam), C.byref(p), abi.ptr(rgb), abi.ptr(lin))
        self._check(rc)
        return rgb, lin

    def render_noalloc(self, cam, width, height, spp, depth, d_rgb, d_linear=None,
                       precision=abi.PREC_FP32, sample_offset=0, bvh_width=0):
        """mh_render_noalloc: the path-tracing megakernel into the caller's
        device buffers (integer addresses; d_linear may be None). Returns the
        code; the library allocates no device memory for the call."""
        p = abi.RenderParams(width=width, height=height, spp=spp, max_depth=depth,
                             mode=abi.MODE_PATHTRACE, precision=precision,
                             sample_offset=sample_offset, engine=abi.ENGINE_MEGAKERNEL,
                             bvh_width=bvh_width)
        return self._L.mh_render_noalloc(self._h, C.byref(cam), C.byref(p), C.c_void_p(d_rgb),
                                         C.c_void_p(d_linear) if d_linear else None)

    def pathtrace_render(self, cam, width, height, spp, depth):
        """Drop-in form of reference pathtrace_render: returns (code, rgb)."""
        rgb = np.zeros((max(height, 1), max(width, 1), 3), np.uint8)
        rc = self._L.mh_pathtrace_render(self._h, C.byref(cam), width, height, spp, depth,
                                         abi.ptr(rgb))
        return rc, rgb

    def render_image(self, cam, width, height, spp, depth):
        """Drop-in form of reference render_image (fixed-spp Whitted)."""
        rgb = np.zeros((max(height, 1), max(width, 1), 3), np.uint8)
        rc = self._L.mh_render_image(self._h, C.byref(cam), width, height, spp, depth,
                                     abi.ptr(rgb))
        return rc, rgb

    def render_image_ex(self, cam, width, height, params):
        """Drop-in form of reference render_image_ex: returns (code, rgb)."""
        rgb = np.zeros((max(height, 1), max(width, 1), 3), np.uint8)
        rc = self._L.mh_render_image_ex(self._h, C.byref(cam), width, height,
                                        C.byref(params) if params is not None else None, abi.ptr(rgb))
        return rc, rgb

    def render_adaptive(self, cam, width, height, spp, depth, max_spp=0, tau=0.0,
                        precision=abi.PREC_FP64, adaptive=True):
        """Adaptive Whitted: (rgb, linear means, samples per pixel, total, max)."""
        p = abi.RenderExParams(samples_per_pixel=spp, max_depth=depth, adaptive=int(adaptive),
                               adaptive_max_spp=max_spp, adaptive_tau=tau)
        rgb = np.zeros((height, width, 3), np.uint8)
        lin = np.zeros((height, width, 3), np.float32)
        ns = np.zeros((height, width), np.int32)
        self._check(self._L.mh_render_adaptive(self._h, C.byref(cam), width, height, C.byref(p), precision,
                                               abi.ptr(rgb), abi.ptr(lin), abi.ptr(ns)))
        tot, mx = last_sample_stats()
        return rgb, lin, ns, tot, mx

    def flare(self, params):
        """Streak flare on the device framebuffer of the last render (then
        download() returns the composited bytes)."""
        self._check(self._L.mh_scene_flare(self._h, C.byref(params)))

    def download(self, width, height):
        """(rgb, linear) of the device framebuffer of the last render (its size)."""
        rgb = np.zeros((height, width, 3), np.uint8)
        lin = np.zeros((height, width, 3), np.float32)
        self._check(self._L.mh_framebuffer_download(self._h, abi.ptr(rgb), abi.ptr(lin)))
        return rgb, lin

    def trace_closest(self, rays, tmin=1e-4, tmax=1e30, precision=abi.PREC_FP32):
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        out = np.zeros(len(rays), abi.HIT_DTYPE)
        self._check(self._L.mh_trace_closest(self._h, abi.ptr(rays), len(rays), tmin, tmax,
                                             precision, abi.ptr(out)))
        return out

    def trace_any(self, rays, tmin=1e-3, tmax=1e30, precision=abi.PREC_FP32):
        rays = np.ascontiguousarray(rays, dtype=abi.RAY_DTYPE)
        out = np.zeros(len(rays), np.uint8)
        self._check(self._L.mh_trace_any(self._h, abi.ptr(rays), len(rays), tmin, tmax,
                                         precision, abi.ptr(out)))
        return out

    def set_materials(self, mats):
        """Replace every material (numpy MATERIAL_DTYPE array, same count)."""
        mats = np.ascontiguousarray(mats, dtype=abi.MATERIAL_DTYPE)
        self.updates = getattr(self, "updates", 0) + 1
        self._check(self._L.mh_scene_set_materials(self._h, abi.ptr(mats), len(mats)))

    def set_sky(self, sky):
        self.updates = getattr(self, "updates", 0) + 1
        self._check(self._L.mh_scene_set_sky(self._h, C.byref(sky)))

    def set_fog(self, fog):
        """Replace the uploaded scene's fog (abi.Fog)."""
        self.updates = getattr(self, "updates", 0) + 1
        self._check(self._L.mh_scene_set_fog(self._h, C.byref(fog)))

    def render_frozen(self, cam, width, height, spp, depth, frozen_mats, sample_offset=0):
        """fp32 path tracer with sampling decisions taken from frozen_mats; returns linear (h, w, 3)."""
        f = np.ascontiguousarray(frozen_mats, dtype=abi.MATERIAL_DTYPE)
        p = abi.RenderParams(width=width, height=height, spp=spp, max_depth=depth,
                             mode=abi.MODE_PATHTRACE, precision=abi.PREC_FP32,
                             sample_offset=sample_offset)
        lin = np.zeros((height, width, 3), np.float32)
        self._check(self._L.mh_render_frozen(self._h, C.byref(cam), C.byref(p), abi.ptr(f), len(f), None,
                                             abi.ptr(lin)))
        return lin

    def render_record(self, cam, width, height, spp, depth, sample_offset=0):
        """fp32 path-traced render that records each sample for the next
        adjoint of the same view: returns (rgb, linear)."""
        p = abi.RenderParams(width=width, height=height, spp=spp, max_depth=depth,
                             mode=abi.MODE_PATHTRACE, precision=abi.PREC_FP32,
                             sample_offset=sample_offset)
        rgb = np.zeros((height, width, 3), np.uint8)
        lin = np.zeros((height, width, 3), np.float32)
        self._check(self._L.mh_render_record(self._h, C.byref(cam), C.byref(p), abi.ptr(rgb), abi.ptr(lin)))
        return rgb, lin

    def adjoint(self, cam, width, height, spp, depth, d_image, sample_offset=0):
        """dLoss/dtheta for dLoss/d(pixel mean) d_image (h, w, 3): returns
        (per-material (n, 8) array [albedo rgb, roughness, metallic, emissive rgb],
        sun colour gradient (3,))."""
        d = np.ascontiguousarray(d_image, dtype=np.float32)
        assert d.shape == (height, width, 3)
        p = abi.RenderParams(width=width, height=height, spp=spp, max_depth=depth,
                             mode=abi.MODE_PATHTRACE, precision=abi.PREC_FP32,
                             sample_offset=sample_offset)
        n = self.n_materials
        g = np.zeros(n * 8 + 3, np.float32)
        self._check(self._L.mh_adjoint(self._h, C.byref(cam), C.byref(p), abi.ptr(d), abi.ptr(g)))
        return g[:n * 8].reshape(n, 8), g[n * 8:]

    @staticmethod
    def last_kernel_times():
        """{kernel: (total_ms, launches)} of the last render made with kernel_clock."""
        k = abi.KernelTimes()
        native.hip().mh_last_kernel_times(C.byref(k))
        return {n: (k.ms[i], int(k.launches[i])) for i, n in enumerate(abi.K_NAMES)}

    @staticmethod
    def last_timing():
        t = abi.Timing()
        native.hip().mh_last_timing(C.byref(t))
        return t

    @staticmethod
    def last_sample_stats():
        return last_sample_stats()


class Batch:
    """Batched renderer (mh_batch): many jobs over many GPU scenes in one
    persistent wavefront launch sequence."""

    def __init__(self):
        self._L = native.hip()
        h = C.c_void_p()
        rc = self._L.mh_batch_create(C.byref(h))
        if rc:
            raise MarmelError(rc, self._L.mh_last_error().decode())
        self._h = h

    def close(self):
        if self._h:
            self._L.mh_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, scenes, jobs, pool=0, grid_blocks=0, block=0, use_graph=True, autotune=False,
               kernel_clock=False, fetch=True, bvh_width=0, fixed_batches=False, engine=abi.BATCH_WAVEFRONT):
        """scenes: list of GpuScene; jobs: list of dicts (scene, cam, width, height, spp, depth,
        sample_offset). Returns [(rgb, linear)] per job (None when fetch is False)."""
        arr = (C.c_void_p * len(scenes))(*[s._h.value for s in scenes])
        js = (abi.Job * len(jobs))()
        npx = []
        for i, j in enumerate(jobs):
            js[i] = abi.Job(scene=j["scene"], width=j["width"], height=j["height"], spp=j["spp"],
                            max_depth=j["depth"], sample_offset=j.get("sample_offset", 0), cam=j["cam"])
            npx.append(max(j["width"], 0) * max(j["height"], 0))
        t = abi.Tune(pool=pool, grid_blocks=grid_blocks, block=block, use_graph=int(use_graph),
                     autotune=int(autotune), kernel_clock=int(kernel_clock), bvh_width=bvh_width,
                     fixed_batches=int(fixed_batches), engine=int(engine))
        tot = max(sum(npx), 1)
        rgb = np.zeros(tot * 3, np.uint8) if fetch else None
        lin = np.zeros(tot * 3, np.float32) if fetch else None
        rc = self._L.mh_render_batch(self._h, arr, len(scenes), js, len(jobs), C.byref(t),
                                     abi.ptr(rgb) if fetch else None, abi.ptr(lin) if fetch else None)
        if rc:
            raise MarmelError(rc, self._L.mh_last_error().decode())
        if not fetch:
            return None
        out, off = [], 0
        for j, n in zip(jobs, npx):
            shape = (j["height"], j["width"], 3)
            out.append((rgb[off * 3:(off + n) * 3].reshape(shape), lin[off * 3:(off + n) * 3].reshape(shape)))
            off += n
        return out

    def set_counting(self, on):
        """Count the traversal work of the following renders (mh_batch_set_counting)."""
        rc = self._L.mh_batch_set_counting(self._h, 1 if on else 0)
        if rc:
            raise MarmelError(rc, self._L.mh_last_error().decode())

    def last_counts(self):
        """Traversal counts of the last render made with counting on (dict)."""
        c = abi.TraversalCounts()
        rc = self._L.mh_batch_last_counts(self._h, C.byref(c))
        if rc:
            raise MarmelError(rc, self._L.mh_last_error().decode())
        return abi.struct_to_dict(c)

    def last_tune(self):
        t = abi.Tune()
        self._L.mh_batch_last_tune(self._h, C.byref(t))
        return {n: getattr(t, n) for n, _ in abi.Tune._fields_}


def flare_apply(rgb, params):
    """Drop-in form of reference flare_apply on a (h, w, 3) uint8 image, in
    place on the GPU; returns the code."""
    if rgb is None:
        return native.hip().mh_flare_apply(None, 1, 1, C.byref(params))
    assert rgb.dtype == np.uint8 and rgb.flags["C_CONTIGUOUS"] and rgb.ndim == 3
    return native.hip().mh_flare_apply(abi.ptr(rgb), rgb.shape[1], rgb.shape[0],
                                       C.byref(params) if params is not None else None)


def last_sample_stats():
    """(total primary samples, largest per-pixel count) of the last render on this thread."""
    tot, mx = C.c_ulonglong(), C.c_int()
    native.hip().mh_last_sample_stats(C.byref(tot), C.byref(mx))
    return tot.value, mx.value


def device_info():
    L = native.hip()
    name = C.create_string_buffer(256)
    cu = C.c_int()
    mem = C.c_size_t()
    rc = L.mh_device_info(name, 256, C.byref(cu), C.byref(mem))
    if rc:
        raise MarmelError(rc, L.mh_last_error().decode())
    return name.value.decode(), cu.value, mem.value
