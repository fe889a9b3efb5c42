"""Loaders for the in-tree native libraries of the package.

libmarmel_hip.so (GPU path) and libmarmel_host.so (C11 host) are built by
marmel-trace_amd/Makefile into marmel-trace_amd/lib/. Loading fails loudly
when they are missing: there is no Python or CPU fallback for the GPU path.
"""
import ctypes as C
import os

from . import abi

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# MARMEL_LIB_DIR selects another build of the same sources (tuning builds made
# with `make OUT=<dir> EXTRA=...`); the default is the in-tree build.
LIB_DIR = os.environ.get("MARMEL_LIB_DIR") or os.path.join(PKG_DIR, "lib")

_hip = None
_host = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name):
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run `make -C marmel-trace_amd` (or __graft_entry__.build())")
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


def host():
    """libmarmel_host.so with argtypes set."""
    global _host
    if _host is not None:
        return _host
    L = _load("libmarmel_host.so")
    P = C.c_void_p
    _sig(L, "mt_scene_new", P, [])
    _sig(L, "mt_scene_free", None, [P])
    _sig(L, "mt_sky_default", None, [C.POINTER(abi.Sky)])
    _sig(L, "mt_material_default", None, [C.POINTER(abi.Material)])
    _sig(L, "mt_material_preset", C.c_int, [C.c_char_p, C.POINTER(abi.Material)])
    _sig(L, "mt_scene_add_material", C.c_int, [P, C.POINTER(abi.Material)])
    _sig(L, "mt_scene_add_prims", C.c_int, [P, P, P, C.c_int])
    _sig(L, "mt_scene_set_sky", None, [P, C.POINTER(abi.Sky)])
    _sig(L, "mt_scene_set_water", None, [P, C.c_double, C.c_int])
    _sig(L, "mt_scene_set_fog", None, [P, C.POINTER(abi.Fog)])
    _sig(L, "mt_obj_xform_default", None, [C.POINTER(ObjXform)])
    _sig(L, "mt_scene_add_obj", C.c_long, [P, C.c_char_p, C.c_size_t, C.c_int, C.POINTER(ObjXform), C.c_int])
    _sig(L, "mt_scene_add_obj_file", C.c_long, [P, C.c_char_p, C.c_int, C.POINTER(ObjXform), C.c_int])
    _sig(L, "mt_ocean_default", None, [C.POINTER(OceanDesc)])
    _sig(L, "mt_scene_add_ocean", C.c_long, [P, C.POINTER(OceanDesc), C.c_int, C.c_int])
    _sig(L, "mt_scene_counts", C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int)])
    _sig(L, "mt_scene_build_bvh", C.c_int, [P, C.POINTER(BvhOptions), C.POINTER(BvhStats)])
    _sig(L, "mt_scene_soa", C.c_int, [P, C.POINTER(abi.SceneSoa)])
    _sig(L, "mt_scene_soa_no_bvh", C.c_int, [P, C.POINTER(abi.SceneSoa)])
    _sig(L, "mt_camera_create", None, [C.POINTER(C.c_double)] * 3 + [C.c_double, C.c_double,
                                                                      C.POINTER(abi.Camera)])
    _sig(L, "mt_scene_generate", C.c_int, [P, C.POINTER(GenParams), C.POINTER(GenInfo)])
    _sig(L, "mt_write_bmp", C.c_int, [C.c_char_p, P, C.c_int, C.c_int])
    _sig(L, "mt_write_ppm", C.c_int, [C.c_char_p, P, C.c_int, C.c_int])
    _sig(L, "mt_encode_bmp", C.c_long, [P, C.c_int, C.c_int, P, C.c_long])
    _sig(L, "mt_last_error", C.c_char_p, [])
    _host = L
    return L


def hip():
    """libmarmel_hip.so with argtypes set."""
    global _hip
    if _hip is not None:
        return _hip
    L = _load("libmarmel_hip.so")
    P = C.c_void_p
    _sig(L, "mh_device_info", C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int),
                                        C.POINTER(C.c_size_t)])
    _sig(L, "mh_scene_create", C.c_int, [C.POINTER(abi.SceneSoa), C.POINTER(P)])
    _sig(L, "mh_scene_create_gpu_bvh", C.c_int, [C.POINTER(abi.SceneSoa), C.c_int, C.c_int, C.POINTER(P)])
    _sig(L, "mh_scene_bvh_info", C.c_int, [P, C.POINTER(abi.BvhInfo)])
    _sig(L, "mh_scene_bvh_export", C.c_int, [P, P, P])
    _sig(L, "mh_scene_destroy", None, [P])
    _sig(L, "mh_scene_device_bytes", C.c_size_t, [P])
    _sig(L, "mh_trace_closest", C.c_int, [P, P, C.c_int, C.c_double, C.c_double, C.c_int, P])
    _sig(L, "mh_trace_any", C.c_int, [P, P, C.c_int, C.c_double, C.c_double, C.c_int, P])
    _sig(L, "mh_pathtrace_render", C.c_int, [P, C.POINTER(abi.Camera), C.c_int, C.c_int, C.c_int,
                                             C.c_int, P])
    _sig(L, "mh_render_image", C.c_int, [P, C.POINTER(abi.Camera), C.c_int, C.c_int, C.c_int,
                                         C.c_int, P])
    _sig(L, "mh_render", C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), P, P])
    _sig(L, "mh_render_image_ex", C.c_int, [P, C.POINTER(abi.Camera), C.c_int, C.c_int,
                                            C.POINTER(abi.RenderExParams), P])
    _sig(L, "mh_render_adaptive", C.c_int, [P, C.POINTER(abi.Camera), C.c_int, C.c_int,
                                            C.POINTER(abi.RenderExParams), C.c_int, P, P, P])
    _sig(L, "mh_last_sample_stats", C.c_int, [C.POINTER(C.c_ulonglong), C.POINTER(C.c_int)])
    _sig(L, "mh_flare_apply", C.c_int, [P, C.c_int, C.c_int, C.POINTER(abi.FlareParams)])
    _sig(L, "mh_scene_flare", C.c_int, [P, C.POINTER(abi.FlareParams)])
    _sig(L, "mh_framebuffer_download", C.c_int, [P, P, P])
    _sig(L, "mh_render_noalloc", C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), P, P])
    _sig(L, "mh_scene_set_materials", C.c_int, [P, P, C.c_int])
    _sig(L, "mh_scene_set_sky", C.c_int, [P, C.POINTER(abi.Sky)])
    _sig(L, "mh_scene_set_fog", C.c_int, [P, C.POINTER(abi.Fog)])
    _sig(L, "mh_adjoint", C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), P, P])
    _sig(L, "mh_render_record", C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), P, P])
    _sig(L, "mh_render_frozen", C.c_int, [P, C.POINTER(abi.Camera), C.POINTER(abi.RenderParams), P,
                                          C.c_int, P, P])
    _sig(L, "mh_batch_create", C.c_int, [C.POINTER(P)])
    _sig(L, "mh_batch_destroy", None, [P])
    _sig(L, "mh_render_batch", C.c_int, [P, P, C.c_int, C.POINTER(abi.Job), C.c_int,
                                         C.POINTER(abi.Tune), P, P])
    _sig(L, "mh_batch_last_tune", C.c_int, [P, C.POINTER(abi.Tune)])
    _sig(L, "mh_batch_set_counting", C.c_int, [P, C.c_int])
    _sig(L, "mh_batch_last_counts", C.c_int, [P, C.POINTER(abi.TraversalCounts)])
    _sig(L, "mh_last_timing", C.c_int, [C.POINTER(abi.Timing)])
    _sig(L, "mh_last_kernel_times", C.c_int, [C.POINTER(abi.KernelTimes)])
    _sig(L, "mh_last_seconds", C.c_double, [])
    _sig(L, "mh_last_error", C.c_char_p, [])
    _hip = L
    return L


class BvhOptions(C.Structure):
    _fields_ = [("max_leaf", C.c_int), ("bins", C.c_int), ("threads", C.c_int)]


class BvhStats(C.Structure):
    _fields_ = [("n_nodes", C.c_int), ("n_leaves", C.c_int), ("max_depth", C.c_int),
                ("sah_cost", C.c_double), ("build_seconds", C.c_double)]


class GenParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("seed", C.c_uint), ("grid", C.c_int),
                ("n_spheres", C.c_int), ("n_emissive", C.c_int), ("subdiv", C.c_int),
                ("count", C.c_int)]


class GenInfo(C.Structure):
    _fields_ = [("eye", C.c_double * 3), ("at", C.c_double * 3), ("up", C.c_double * 3),
                ("vfov", C.c_double), ("n_prims", C.c_int), ("n_triangles", C.c_int),
                ("n_spheres", C.c_int), ("n_emissive", C.c_int)]


class ObjXform(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("scale", C.c_double * 3), ("rotate_deg", C.c_double * 3),
                ("smooth", C.c_int32), ("auto_center", C.c_int32), ("auto_scale", C.c_double)]


class OceanDesc(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("size", C.c_double * 2), ("res_x", C.c_int32), ("res_z", C.c_int32),
                ("amplitude", C.c_double), ("wavelength", C.c_double), ("direction", C.c_double * 3),
                ("steepness", C.c_double), ("chop", C.c_double), ("chop_wavelength", C.c_double),
                ("depth", C.c_double), ("seed", C.c_uint32), ("reserved", C.c_int32)]


GEN_SPHERES, GEN_TERRAIN, GEN_ICOSPHERE, GEN_MATERIALS, GEN_FOREST = 1, 2, 3, 4, 5
