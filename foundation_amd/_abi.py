"""ctypes declarations of include/fnd_pt.h (the C ABI of libfnd_pt.so).

Field order and types follow the header one for one; the size checks in
``_check_sizes`` fail the import if the two ever drift apart.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
ABI_VERSION = 7  # FND_PT_ABI_VERSION of include/fnd_pt.h these declarations follow
LIB_PATH = os.path.join(_HERE, "libfnd_pt.so")

fnd_handle = C.c_uint64
INVALID_HANDLE = (1 << 64) - 1
NO_HIT = 0xFFFFFFFF

OK, E_INVALID_ARG, E_OOM_HOST, E_OOM_DEVICE, E_DEVICE, E_QUEUE_OVERFLOW, E_UNSUPPORTED, E_NOT_FOUND = range(8)

MAT_LAMBERT, MAT_GGX, MAT_EMISSIVE = 0, 1, 2
GEN_CORNELL_SOUP, GEN_DISPLACED_SPHERE, GEN_INSTANCED_SPHERES, GEN_FURNACE = 1, 2, 3, 4
TRACE_CLOSEST, TRACE_ANYHIT, TRACE_COUNT, ASYNC = 0, 1, 2, 0x100
JOB_USE_CACHE, JOB_TRAIN_CACHE = 1, 2
FB_FP64 = 1
PIPELINE_PATH_TRACE, PIPELINE_TRACE_CLOSEST, PIPELINE_TRACE_ANYHIT, PIPELINE_PRESENT, PIPELINE_CLEAR = 1, 2, 3, 4, 5
DESCRIPTOR_ACCELERATION_STRUCTURE, DESCRIPTOR_STORAGE_BUFFER, DESCRIPTOR_STORAGE_IMAGE = 1, 2, 3


class HostAllocator(C.Structure):
    _fields_ = [
        ("user", C.c_void_p),
        ("alloc", C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t)),
        ("realloc", C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t)),
        ("free", C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)),
    ]


LOG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_char_p)


class CreateDesc(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("flags", C.c_uint32),
        ("device_memory_budget", C.c_uint64),
        ("log", LOG_FN),
        ("log_user", C.c_void_p),
    ]


class Ray(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim", C.c_uint32), ("u", C.c_float), ("v", C.c_float)]


class Material(C.Structure):
    _fields_ = [
        ("type", C.c_uint32),
        ("albedo", C.c_float * 3),
        ("roughness", C.c_float),
        ("emission", C.c_float * 3),
    ]


class QuadLight(C.Structure):
    _fields_ = [
        ("corner", C.c_float * 3),
        ("edge0", C.c_float * 3),
        ("edge1", C.c_float * 3),
        ("radiance", C.c_float * 3),
    ]


class VertexStream(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("bytes", C.c_uint64),
        ("stride", C.c_uint32),
        ("position_offset", C.c_uint32),
        ("position_components", C.c_uint32),
        ("num_vertices", C.c_uint32),
    ]


class IndexStream(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("index_bits", C.c_uint32),
        ("num_indices", C.c_uint32),
    ]


class Arena(C.Structure):
    _fields_ = [
        ("memory", C.c_uint64),
        ("end", C.c_uint64),
        ("current", C.c_uint64),
    ]


class FrameArenaInfo(C.Structure):
    _fields_ = [
        ("size", C.c_uint64),
        ("used", C.c_uint64),
        ("batch_paths", C.c_uint32),
        ("caller_owned", C.c_uint32),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("positions", C.POINTER(C.c_float)),
        ("num_vertices", C.c_uint32),
        ("indices", C.POINTER(C.c_uint32)),
        ("num_triangles", C.c_uint32),
        ("material_ids", C.POINTER(C.c_uint32)),
        ("light_ids", C.POINTER(C.c_int32)),
    ]


class InstanceDesc(C.Structure):
    _fields_ = [
        ("mesh", C.c_uint32),
        ("flags", C.c_uint32),
        ("object_to_world", C.c_float * 12),
        ("world_to_object", C.c_float * 12),
    ]


class SceneDesc(C.Structure):
    _fields_ = [
        ("meshes", C.POINTER(MeshDesc)),
        ("num_meshes", C.c_uint32),
        ("instances", C.POINTER(InstanceDesc)),
        ("num_instances", C.c_uint32),
        ("materials", C.POINTER(Material)),
        ("num_materials", C.c_uint32),
        ("lights", C.POINTER(QuadLight)),
        ("num_lights", C.c_uint32),
        ("background", C.c_float * 3),
        ("build_flags", C.c_uint32),
    ]


class SceneInfo(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_uint64),
        ("num_prims", C.c_uint64),
        ("num_nodes", C.c_uint64),
        ("max_depth", C.c_uint32),
        ("stack_entries", C.c_uint32),
        ("device_bytes", C.c_uint64),
        ("build_ms", C.c_double),
        ("bounds_lo", C.c_float * 3),
        ("bounds_hi", C.c_float * 3),
    ]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("at", C.c_float * 3), ("up", C.c_float * 3), ("fov_y_deg", C.c_float)]


class RenderJob(C.Structure):
    _fields_ = [
        ("scene", fnd_handle),
        ("framebuffer", fnd_handle),
        ("camera", Camera),
        ("spp", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("rr_depth", C.c_uint32),
        ("flags", C.c_uint32),
        ("seed", C.c_uint64),
        ("sample_offset", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("last_kernel_ms", C.c_double),
        ("last_call_ms", C.c_double),
        ("trace_kernel_ms_total", C.c_double),
        ("trace_launches", C.c_uint64),
        ("rays_traced", C.c_uint64),
        ("node_visits", C.c_uint64),
        ("tri_tests", C.c_uint64),
        ("instance_visits", C.c_uint64),
        ("paths", C.c_uint64),
        ("render_rays", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("num_cus", C.c_uint32),
        ("last_grid", C.c_uint32),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("block_size", C.c_uint32),
        ("stack_lds", C.c_uint32),
        ("waves_per_cu", C.c_uint32),
        ("refill", C.c_uint32),
        ("queue_capacity", C.c_uint32),
        ("tile", C.c_uint32),
    ]


class GenParams(C.Structure):
    _fields_ = [
        ("kind", C.c_uint32),
        ("seed", C.c_uint64),
        ("soup_triangles", C.c_uint32),
        ("subdivision", C.c_uint32),
        ("instances", C.c_uint32),
        ("octaves", C.c_uint32),
        ("amplitude", C.c_float),
        ("albedo", C.c_float),
        ("emission", C.c_float),
        ("ggx_fraction_pct", C.c_uint32),
        ("lights", C.c_uint32),
    ]


class AutotuneResult(C.Structure):
    _fields_ = [
        ("tuning", Tuning),
        ("median", Tuning),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
        ("median_ms", C.c_double),
        ("candidates", C.c_uint32),
        ("renders", C.c_uint32),
        ("from_cache", C.c_uint32),
        ("reserved", C.c_uint32),
        ("key", C.c_char * 128),
    ]


AUTOTUNE_FORCE = 1


class CacheDesc(C.Structure):
    _fields_ = [
        ("enable", C.c_uint32),
        ("levels", C.c_uint32),
        ("features", C.c_uint32),
        ("log2_table", C.c_uint32),
        ("base_resolution", C.c_uint32),
        ("max_resolution", C.c_uint32),
        ("hidden", C.c_uint32),
        ("hidden_layers", C.c_uint32),
        ("learning_rate", C.c_float),
        ("train_paths", C.c_uint32),
        ("train_steps", C.c_uint32),
        ("train_depth", C.c_uint32),
        ("train_batch", C.c_uint32),
        ("glossy_min_roughness", C.c_float),
        ("glossy_any_bounce", C.c_uint32),
        ("seed", C.c_uint64),
        ("bounds_lo", C.c_float * 3),
        ("bounds_hi", C.c_float * 3),
        ("grid_gradient", C.c_uint32),
    ]


CACHE_GRAD_BF16, CACHE_GRAD_FP32 = 0, 1


class HostSceneView(C.Structure):
    _fields_ = [
        ("nodes", C.c_void_p),
        ("num_nodes", C.c_uint64),
        ("tris", C.c_void_p),
        ("num_tris", C.c_uint64),
        ("instances", C.c_void_p),
        ("num_instances", C.c_uint32),
        ("root", C.c_uint32),
        ("two_level", C.c_uint32),
        ("reserved", C.c_uint32),
        ("info", SceneInfo),
    ]


class Descriptor(C.Structure):
    _fields_ = [
        ("binding", C.c_uint32),
        ("type", C.c_uint32),
        ("resource", fnd_handle),
        ("offset", C.c_uint64),
        ("range", C.c_uint64),
        ("device_ptr", C.c_void_p),
    ]


class PathTraceConstants(C.Structure):
    _fields_ = [
        ("camera", Camera),
        ("max_depth", C.c_uint32),
        ("rr_depth", C.c_uint32),
        ("flags", C.c_uint32),
        ("sample_offset", C.c_uint32),
        ("seed", C.c_uint64),
    ]


def _check_sizes():
    assert C.sizeof(Ray) == 32 and C.sizeof(Hit) == 16
    assert C.sizeof(Material) == 32 and C.sizeof(QuadLight) == 48
    assert C.sizeof(InstanceDesc) == 104
    assert C.sizeof(Descriptor) == 40 and C.sizeof(PathTraceConstants) == 64


_check_sizes()

_P = C.c_void_p
_H = fnd_handle
_S = C.c_int
# name -> (restype, argtypes); every function declared in include/fnd_pt.h
PROTOTYPES = {
    "fnd_pt_create": (_S, [C.POINTER(CreateDesc), C.POINTER(HostAllocator), C.POINTER(_P)]),
    "fnd_pt_destroy": (None, [_P]),
    "fnd_pt_last_error": (C.c_char_p, [_P]),
    "fnd_pt_status_string": (C.c_char_p, [_S]),
    "fnd_pt_abi_version": (C.c_uint32, []),
    "fnd_pt_wait": (_S, [_P]),
    "fnd_pt_get_stats": (_S, [_P, C.POINTER(Stats)]),
    "fnd_pt_reset_stats": (_S, [_P]),
    "fnd_pt_set_tuning": (_S, [_P, C.POINTER(Tuning)]),
    "fnd_pt_get_tuning": (_S, [_P, C.POINTER(Tuning)]),
    "fnd_pt_scene_create": (_S, [_P, C.POINTER(SceneDesc), C.POINTER(_H)]),
    "fnd_pt_scene_destroy": (_S, [_P, _H]),
    "fnd_pt_scene_info": (_S, [_P, _H, C.POINTER(SceneInfo)]),
    "fnd_pt_scene_export": (_S, [_P, _H, _P, C.POINTER(C.c_uint64), _P, C.POINTER(C.c_uint64)]),
    "fnd_pt_buffer_create": (_S, [_P, C.c_uint64, C.POINTER(_H)]),
    "fnd_pt_buffer_destroy": (_S, [_P, _H]),
    "fnd_pt_buffer_write": (_S, [_P, _H, C.c_uint64, _P, C.c_uint64]),
    "fnd_pt_buffer_read": (_S, [_P, _H, C.c_uint64, _P, C.c_uint64]),
    "fnd_pt_trace": (_S, [_P, _H, _P, C.c_uint32, _P, C.c_uint32]),
    "fnd_pt_trace_buffers": (_S, [_P, _H, _H, _H, C.c_uint32, C.c_uint32]),
    "fnd_pt_trace_noalloc": (_S, [_P, _H, _P, _P, C.c_uint32, C.c_uint32]),
    "fnd_pt_framebuffer_create": (_S, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(_H)]),
    "fnd_pt_framebuffer_destroy": (_S, [_P, _H]),
    "fnd_pt_framebuffer_clear": (_S, [_P, _H]),
    "fnd_pt_framebuffer_read": (_S, [_P, _H, _P, C.POINTER(C.c_uint32)]),
    "fnd_pt_framebuffer_read_f64": (_S, [_P, _H, _P, C.POINTER(C.c_uint32)]),
    "fnd_pt_framebuffer_rel_mse": (_S, [_P, _H, _H, C.c_double, C.POINTER(C.c_double)]),
    "fnd_pt_framebuffer_read_rgba8": (_S, [_P, _H, C.c_float, _P, C.POINTER(C.c_uint32)]),
    "fnd_pt_framebuffer_present": (_S, [_P, _H, C.c_float, _P, C.c_uint64, C.POINTER(C.c_uint32)]),
    "fnd_pt_render": (_S, [_P, C.POINTER(RenderJob), C.c_uint32]),
    "fnd_camera_basis": (_S, [C.POINTER(Camera), C.c_uint32, C.c_uint32, _P]),
    "fnd_pt_camera_rays": (_S, [_P, C.POINTER(Camera), C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, _P]),
    "fnd_pt_rng_draws": (_S, [_P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "fnd_pt_cmd_create": (_S, [_P, C.POINTER(_P)]),
    "fnd_pt_cmd_destroy": (None, [_P]),
    "fnd_pt_cmd_begin": (_S, [_P]),
    "fnd_pt_cmd_set_pipeline": (_S, [_P, C.c_uint32]),
    "fnd_pt_cmd_bind_descriptors": (_S, [_P, C.POINTER(Descriptor), C.c_uint32]),
    "fnd_pt_cmd_push_constants": (_S, [_P, _P, C.c_uint32]),
    "fnd_pt_cmd_dispatch": (_S, [_P, C.c_uint32, C.c_uint32, C.c_uint32]),
    "fnd_pt_cmd_end": (_S, [_P]),
    "fnd_pt_cmd_size": (C.c_uint32, [_P]),
    "fnd_pt_fence_create": (_S, [_P, C.POINTER(_P)]),
    "fnd_pt_fence_destroy": (None, [_P]),
    "fnd_pt_submit": (_S, [_P, C.POINTER(_P), C.c_uint32, _P]),
    "fnd_pt_fence_wait": (_S, [_P, _P]),
    "fnd_arena_reset": (None, [C.POINTER(Arena), _P, C.c_uint64]),
    "fnd_arena_alloc": (_P, [C.POINTER(Arena), C.c_uint64]),
    "fnd_arena_alloc_aligned": (_P, [C.POINTER(Arena), C.c_uint64, C.c_uint64]),
    "fnd_arena_used": (C.c_uint64, [C.POINTER(Arena)]),
    "fnd_pt_queue_bytes": (_S, [C.c_uint32, C.POINTER(C.c_uint64)]),
    "fnd_pt_set_frame_arena": (_S, [_P, _P, C.c_uint64]),
    "fnd_pt_frame_arena_info": (_S, [_P, C.POINTER(FrameArenaInfo)]),
    "fnd_pt_cache_default_desc": (_S, [C.POINTER(CacheDesc)]),
    "fnd_pt_cache_configure": (_S, [_P, C.POINTER(CacheDesc)]),
    "fnd_pt_cache_param_count": (_S, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "fnd_pt_cache_set_params": (_S, [_P, _P, _P]),
    "fnd_pt_cache_get_params": (_S, [_P, _P, _P]),
    "fnd_pt_cache_infer": (_S, [_P, _P, C.c_uint32, _P]),
    "fnd_pt_cache_gradients": (_S, [_P, _P, _P, C.c_uint32, C.POINTER(C.c_float), _P, _P]),
    "fnd_pt_cache_train": (_S, [_P, _P, _P, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)]),
    "fnd_scene_encode": (_S, [C.POINTER(SceneDesc), C.POINTER(_P), C.c_char_p, C.c_size_t]),
    "fnd_mesh_from_streams": (_S, [C.POINTER(VertexStream), C.POINTER(IndexStream), _P, _P, C.c_char_p, C.c_size_t]),
    "fnd_obj_load": (_S, [C.c_char_p, C.POINTER(_P), C.c_char_p, C.c_size_t]),
    "fnd_obj_mesh_get": (_S, [_P, C.POINTER(MeshDesc)]),
    "fnd_obj_free": (None, [_P]),
    "fnd_host_scene_get": (_S, [_P, C.POINTER(HostSceneView)]),
    "fnd_host_scene_free": (None, [_P]),
    "fnd_gen_default_params": (_S, [C.c_uint32, C.POINTER(GenParams)]),
    "fnd_gen_scene": (_S, [C.POINTER(GenParams), C.POINTER(_P)]),
    "fnd_gen_scene_get": (_S, [_P, C.POINTER(SceneDesc), C.POINTER(Camera)]),
    "fnd_gen_scene_free": (None, [_P]),
    "fnd_gen_batch_params": (_S, [C.c_uint32, C.POINTER(GenParams), C.POINTER(C.c_uint32)]),
    "fnd_pt_autotune": (_S, [_P, C.POINTER(RenderJob), C.c_uint32, C.c_char_p, C.c_uint32,
                             C.POINTER(AutotuneResult)]),
}

_lib = None


def load():
    """Loads libfnd_pt.so from the package directory.  Raises if it is not built:
    there is no fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C foundation_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.fnd_pt_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{LIB_PATH} has ABI {lib.fnd_pt_abi_version()}, these bindings need {ABI_VERSION}: rebuild")
    _lib = lib
    return lib
