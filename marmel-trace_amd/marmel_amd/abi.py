"""ctypes / numpy mirror of include/marmel_hip.h.

Every struct here has the exact layout of its C counterpart; the numpy dtypes
are used to build primitive / material arrays in bulk and hand them to the
C ABI without per-element Python work.
"""
import ctypes as C

import numpy as np

ABI_VERSION = 2

PRIM_SPHERE, PRIM_PLANE, PRIM_BOX, PRIM_TRIANGLE, PRIM_CYLINDER = 0, 1, 2, 3, 4
PRIM_SMOOTH, PRIM_UV = 1, 2
TEXTURE_NONE, TEXTURE_CHECKER, TEXTURE_UV_CHECKER = 0, 1, 7
MODE_PATHTRACE, MODE_WHITTED = 0, 1
PREC_FP32, PREC_FP64 = 0, 1


class Vec3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


def vec3(v):
    return Vec3(float(v[0]), float(v[1]), float(v[2]))


class Prim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("material", C.c_int32), ("flags", C.c_int32),
                ("reserved", C.c_int32), ("p0", Vec3), ("p1", Vec3), ("p2", Vec3),
                ("r0", C.c_double), ("r1", C.c_double)]


class PrimAttr(C.Structure):
    _fields_ = [(n, Vec3) for n in ("n0", "n1", "n2", "uv0", "uv1", "uv2")]


class Material(C.Structure):
    _fields_ = [("albedo", Vec3), ("specular", Vec3), ("shininess", C.c_double),
                ("reflectivity", C.c_double), ("transparency", C.c_double),
                ("ior", C.c_double), ("is_water", C.c_int32), ("beer_lambert", C.c_int32),
                ("pbr", C.c_int32), ("texture_kind", C.c_int32), ("absorption", Vec3),
                ("deep_color", Vec3), ("metallic", C.c_double), ("roughness", C.c_double),
                ("emissive", Vec3), ("texture_scale", C.c_double),
                ("texture_color_a", Vec3), ("texture_color_b", Vec3),
                ("atmosphere_glow", Vec3), ("bump_strength", C.c_double)]


class Sky(C.Structure):
    _fields_ = [("sun_dir", Vec3), ("sun_color", Vec3), ("horizon_color", Vec3),
                ("zenith_color", Vec3), ("gradient_gamma", C.c_double),
                ("sun_glow_exponent", C.c_double), ("sun_glow_strength", C.c_double),
                ("cloud_height", C.c_double), ("cloud_scale", C.c_double),
                ("cloud_coverage", C.c_double), ("cloud_softness", C.c_double),
                ("cloud_sharpness", C.c_double), ("cloud_octaves", C.c_int32),
                ("seed", C.c_uint32), ("sun_radius", C.c_double),
                ("cloud_thickness", C.c_double), ("star_intensity", C.c_double),
                ("nebula_intensity", C.c_double), ("galaxy_intensity", C.c_double),
                ("fog_density", C.c_double)]


class Fog(C.Structure):
    _fields_ = [("density", C.c_double), ("color", Vec3), ("height", C.c_double),
                ("height_falloff", C.c_double), ("inscatter_strength", C.c_double),
                ("sun_anisotropy", C.c_double), ("noise_scale", C.c_double),
                ("noise_amount", C.c_double), ("shadow_steps", C.c_int32),
                ("reserved", C.c_int32)]


class Ocean(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("seed", C.c_uint32), ("center", Vec3), ("direction", Vec3),
                ("amplitude", C.c_double), ("wavelength", C.c_double), ("steepness", C.c_double),
                ("chop", C.c_double), ("chop_wavelength", C.c_double)]


def fog_default():
    """Reference fog_default_params: disabled."""
    return Fog(density=0.0, color=Vec3(0.70, 0.75, 0.80), inscatter_strength=0.50,
               sun_anisotropy=0.70)


class Camera(C.Structure):
    _fields_ = [("position", Vec3), ("lower_left", Vec3), ("horizontal", Vec3),
                ("vertical", Vec3), ("aperture", C.c_double),
                ("focus_distance", C.c_double), ("dome_radius", C.c_double),
                ("anamorphic_squeeze", C.c_double)]


class BvhNode(C.Structure):
    _fields_ = [("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("left", C.c_int32),
                ("right", C.c_int32), ("first", C.c_int32), ("count", C.c_int32)]


class SceneSoa(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("n_prims", C.c_int32),
                ("prims", C.c_void_p), ("attrs", C.c_void_p),
                ("n_materials", C.c_int32), ("materials", C.c_void_p),
                ("sky", Sky), ("water_level", C.c_double), ("water_material", C.c_int32),
                ("n_nodes", C.c_int32), ("nodes", C.c_void_p),
                ("n_leaf_prims", C.c_int32), ("leaf_prims", C.c_void_p),
                ("n_planes", C.c_int32), ("planes", C.c_void_p), ("fog", Fog), ("ocean", Ocean)]


class BvhInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int32), ("n_leaf", C.c_int32), ("max_depth", C.c_int32),
                ("width4", C.c_int32), ("device_built", C.c_int32), ("reserved", C.c_int32),
                ("build_ms", C.c_double)]


class Ray(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("ox", "oy", "oz", "dx", "dy", "dz")]


class Hit(C.Structure):
    _fields_ = [("t", C.c_double), ("prim_index", C.c_int32), ("material", C.c_int32),
                ("front_face", C.c_int32), ("reserved", C.c_int32),
                ("nx", C.c_double), ("ny", C.c_double), ("nz", C.c_double),
                ("u", C.c_double), ("v", C.c_double)]


ENGINE_MEGAKERNEL, ENGINE_WAVEFRONT = 0, 1


class RenderParams(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("width", "height", "spp", "max_depth", "mode",
                                          "precision", "sample_offset", "engine", "pool",
                                          "grid_blocks", "use_graph", "kernel_clock", "bvh_width")]


class RenderExParams(C.Structure):
    """render_image_ex parameters (reference RenderParams)."""
    _fields_ = [("samples_per_pixel", C.c_int32), ("max_depth", C.c_int32), ("adaptive", C.c_int32),
                ("adaptive_max_spp", C.c_int32), ("adaptive_tau", C.c_double)]


class FlareParams(C.Structure):
    """Anamorphic streak flare parameters (reference FlareParams layout)."""
    _fields_ = [("enabled", C.c_int32), ("reserved", C.c_int32), ("intensity", C.c_double),
                ("threshold", C.c_double), ("streak_length", C.c_double), ("tint", Vec3)]


def flare_default(enabled=1):
    """The reference's defaults (src/flare.h:38-41, flare.c:10-19), enabled."""
    return FlareParams(enabled=enabled, intensity=0.6, threshold=0.85, streak_length=0.25,
                       tint=Vec3(0.15, 0.45, 1.0))


class Job(C.Structure):
    _fields_ = [("scene", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("max_depth", C.c_int32), ("sample_offset", C.c_int32), ("cam", Camera)]


class Tune(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("pool", "grid_blocks", "block", "use_graph", "autotune",
                                          "kernel_clock", "bvh_width", "fixed_batches", "engine")]


BATCH_WAVEFRONT, BATCH_MEGAKERNEL = 0, 1


K_GEN, K_EXTEND, K_SHADE, K_SHADOW, K_FINALIZE, K_OTHER = range(6)
K_NAMES = ("gen", "extend", "shade", "shadow", "finalize", "other")


class TraversalCounts(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("ext_queries", "ext_nodes", "ext_triangles", "ext_other",
                                        "shadow_queries", "shadow_nodes", "shadow_triangles", "shadow_other")]


class KernelTimes(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_int64 * 8)]


class Timing(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("kernel_ms", C.c_double),
                ("tonemap_ms", C.c_double), ("samples", C.c_int64), ("rays", C.c_int64)]


def _dtype_of(struct):
    """numpy structured dtype with the same layout as a ctypes Structure."""
    dt = np.dtype(struct)
    assert dt.itemsize == C.sizeof(struct)
    return dt


PRIM_DTYPE = _dtype_of(Prim)
ATTR_DTYPE = _dtype_of(PrimAttr)
MATERIAL_DTYPE = _dtype_of(Material)
NODE_DTYPE = _dtype_of(BvhNode)
RAY_DTYPE = _dtype_of(Ray)
HIT_DTYPE = _dtype_of(Hit)


def ptr(a):
    """c_void_p of a contiguous numpy array (None for None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return C.c_void_p(a.ctypes.data)


def struct_to_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        if isinstance(v, Vec3):
            v = (v.x, v.y, v.z)
        out[name] = v
    return out


def struct_from_dict(cls, d):
    s = cls()
    for name, typ in cls._fields_:
        if name not in d:
            continue
        v = d[name]
        if typ is Vec3:
            setattr(s, name, vec3(v))
        else:
            setattr(s, name, typ(v).value if hasattr(typ, "value") else v)
    return s
