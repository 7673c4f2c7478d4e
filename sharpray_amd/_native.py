"""ctypes binding of libsharpray_hip.so (the C ABI declared in include/sharpray.h).

The library is built in-tree by sharpray_amd/csrc/build.sh. Loading fails loudly when it is
missing: there is no CPU fallback for the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SHARPRAY_LIB selects another build of the same library (e.g. to compare compile options).
LIB_PATH = os.environ.get("SHARPRAY_LIB") or os.path.join(_HERE, "libsharpray_hip.so")

SR_OK = 0
SR_RNG_REFERENCE = 0
SR_RNG_COUNTER = 1
SR_LIGHT_CDF = 0
SR_LIGHT_ALIAS = 1
SR_SHAPE_MESH, SR_SHAPE_SPHERE, SR_SHAPE_BOX = 0, 1, 2
SR_COUNTER_COUNT = 15
SR_BVH_BUILDER_HOST, SR_BVH_BUILDER_DEVICE = 0, 1
COUNTER_NAMES = [
    "Pixel", "Sample", "SampleBounce", "SampleHit", "SampleMiss", "SampleTerminate", "SampleFogScatter",
    "SampleFogEscape", "SamplePixelConverged", "SamplePixelMaxed", "SceneTrace", "SceneOcclude", "MeshIntersect",
    "BvhIntersectNode", "BvhIntersectShape",
]

# Counters.Type (counters.cs:24-79) in order: the layout of sr_counters_read_all.
COUNTER_TYPE_NAMES = [
    "Worker", "Block", "Pixel", "Sample", "SampleBounce", "SampleHit", "SampleMiss", "SampleTerminate",
    "SampleFogScatter", "SampleFogEscape", "SampleVarianceMax", "SamplePixelConverged", "SamplePixelMaxed",
    "SampleCountMin", "SampleCountMax", "SceneObject", "SceneBvhNodes", "SceneBvhDepth", "SceneTrace",
    "SceneOcclude", "MeshTriangle", "MeshIntersect", "MeshIntersectAny", "BvhIntersectNode", "BvhIntersectShape",
    "BvhOverlapNode", "BvhOverlapShape", "FilterSample", "DenoiseEarlyOut", "DenoiseWeightMax", "DenoiseMaxLum",
    "DenoiseMaxLumBoost", "DenoiseRejectMissing", "DenoiseRejectNormal", "DenoiseRejectDepth", "DenoiseRejectLum",
    "OverlayLine", "OverlayText", "RtPeakWorkingSet", "RtGcAllocatedBytes", "RtGcCurrentBytes", "RtGcGen0",
    "RtGcGen1", "RtGcGen2", "TimeSetup", "TimeObjLoad", "TimeMeshBvhBuild", "TimeSceneBvhBuild", "TimeRender",
    "TimeCompose", "TimeTotal",
]
SR_COUNTER_TYPE_COUNT = len(COUNTER_TYPE_NAMES)
# Reductions other than the sum (counters.cs:228-236), with the reference's initial ("unset") values.
COUNTER_TYPE_UNSET = {"SampleVarianceMax": -2 ** 63, "SampleCountMin": 2 ** 63 - 1, "SampleCountMax": -2 ** 63,
                      "DenoiseWeightMax": -2 ** 63, "DenoiseMaxLum": -2 ** 63, "DenoiseMaxLumBoost": -2 ** 63}

c_float_p = ctypes.POINTER(ctypes.c_float)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_uint32_p = ctypes.POINTER(ctypes.c_uint32)


class Transform_(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("rot", ctypes.c_float * 4), ("scale", ctypes.c_float * 3)]


class Material_(ctypes.Structure):
    _fields_ = [
        ("albedo", ctypes.c_float * 3), ("roughness", ctypes.c_float), ("metallic", ctypes.c_float),
        ("transparency", ctypes.c_float), ("ior", ctypes.c_float), ("radiance", ctypes.c_float * 3),
        ("color_texture", ctypes.c_int32), ("roughness_texture", ctypes.c_int32),
        ("metallic_texture", ctypes.c_int32), ("normal_texture", ctypes.c_int32),
    ]


class SkyProcedural_(ctypes.Structure):
    _fields_ = [
        ("radiance_top", ctypes.c_float * 3), ("radiance_middle", ctypes.c_float * 3),
        ("radiance_bottom", ctypes.c_float * 3), ("sun_dir", ctypes.c_float * 3),
        ("sun_radiance", ctypes.c_float * 3), ("sun_angle", ctypes.c_float),
    ]


class Fog_(ctypes.Structure):
    _fields_ = [("density", ctypes.c_float), ("color", ctypes.c_float * 3), ("anisotropy", ctypes.c_float),
                ("height_falloff", ctypes.c_float)]


class FogVolume_(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("sigma", ctypes.c_float), ("color", ctypes.c_float * 3),
                ("anisotropy", ctypes.c_float), ("bounds_min", ctypes.c_float * 3), ("bounds_max", ctypes.c_float * 3),
                ("res", ctypes.c_uint32 * 3), ("density", ctypes.POINTER(ctypes.c_uint16))]


class View_(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float * 3), ("rot", ctypes.c_float * 4), ("fov", ctypes.c_float),
                ("near_plane", ctypes.c_float)]


class RenderParams_(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("block_size", ctypes.c_uint32),
        ("min_samples", ctypes.c_uint32), ("max_samples", ctypes.c_uint32), ("variance_threshold", ctypes.c_float),
        ("bounces", ctypes.c_uint32), ("indirect_clamp", ctypes.c_float), ("rng_policy", ctypes.c_uint32),
        ("rng_seed", ctypes.c_uint32), ("tick_bands", ctypes.c_uint32), ("light_sampler", ctypes.c_uint32),
    ]


class DenoiseParams_(ctypes.Structure):
    _fields_ = [
        ("radius", ctypes.c_float), ("strength", ctypes.c_float), ("strength_max", ctypes.c_float),
        ("luminance_boost", ctypes.c_float), ("luminance_limit", ctypes.c_float), ("normal_limit", ctypes.c_float),
        ("depth_limit", ctypes.c_float),
    ]


class ComposeParams_(ctypes.Structure):
    _fields_ = [
        ("denoise", DenoiseParams_), ("tonemapper", ctypes.c_int32), ("exposure", ctypes.c_float),
        ("fog_radius", ctypes.c_float), ("fog_strength", ctypes.c_float),
    ]


class BvhStats_(ctypes.Structure):
    _fields_ = [
        ("node_count", ctypes.c_int32), ("depth_max", ctypes.c_int32), ("leaf_count", ctypes.c_int32),
        ("leaf_size_min", ctypes.c_int32), ("leaf_size_max", ctypes.c_int32), ("leaf_size_avg", ctypes.c_float),
        ("leaf_depth_avg", ctypes.c_float), ("sah_cost", ctypes.c_float),
    ]


c_uint8_p = ctypes.POINTER(ctypes.c_uint8)

# name -> (restype, argtypes); every entry point declared in include/sharpray.h.
SIGNATURES = {
    "sr_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "sr_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "sr_last_error": (ctypes.c_char_p, []),
    "sr_texture_create": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32,
                                         ctypes.c_float, ctypes.c_float, c_int32_p]),
    "sr_mesh_create": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_uint32, ctypes.c_uint32, c_int32_p]),
    "sr_scene_set_sky_texture": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, ctypes.c_float]),
    "sr_scene_set_sky_procedural": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SkyProcedural_)]),
    "sr_scene_set_fog": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Fog_)]),
    "sr_scene_set_fog_volume": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FogVolume_)]),
    "sr_scene_add_object": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Transform_), ctypes.POINTER(Material_),
                                           ctypes.c_int32, ctypes.c_int32, c_float_p, c_int32_p]),
    "sr_scene_build": (ctypes.c_int, [ctypes.c_void_p]),
    "sr_render_begin": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(View_), ctypes.POINTER(RenderParams_)]),
    "sr_render_tick": (ctypes.c_int, [ctypes.c_void_p, c_uint32_p, c_uint32_p]),
    "sr_render_read": (ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p, c_float_p, c_float_p, c_float_p,
                                      c_uint32_p, c_float_p]),
    "sr_render_device_aovs": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_void_p)] * 7),
    "sr_render_read_albedo": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "sr_denoise_bilateral": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DenoiseParams_), c_float_p]),
    "sr_denoise_bilateral_guided": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DenoiseParams_), ctypes.c_float,
                                                   c_float_p]),
    "sr_denoise_bilateral_mean": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DenoiseParams_), c_float_p]),
    "sr_denoise_bilateral_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(DenoiseParams_), ctypes.c_void_p,
                                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_uint32, ctypes.c_void_p]),
    "sr_compose": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ComposeParams_), c_uint8_p]),
    "sr_compose_device_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ComposeParams_)] + [ctypes.c_void_p] * 5
                                  + [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "sr_preview": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_float, c_uint8_p]),
    "sr_filter_fog": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, c_float_p]),
    "sr_image_encode_bmp": (ctypes.c_int, [c_uint8_p, ctypes.c_uint32, ctypes.c_uint32, c_uint8_p,
                                           ctypes.POINTER(ctypes.c_uint64)]),
    "sr_image_write_bmp": (ctypes.c_int, [ctypes.c_char_p, c_uint8_p, ctypes.c_uint32, ctypes.c_uint32]),
    "sr_trace_primary": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(View_), ctypes.c_uint32, ctypes.c_uint32,
                                        c_int32_p, c_int32_p, c_float_p]),
    "sr_bvh_export": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_float_p, c_int32_p, c_uint32_p, c_uint32_p]),
    "sr_counters_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "sr_counters_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "sr_counters_read_all": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.c_uint32]),
    "sr_counters_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "sr_bvh_stats_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(BvhStats_)]),
    "sr_profile_enable": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "sr_profile_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), c_uint32_p]),
    "sr_profile_busy": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "sr_set_slices": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "sr_set_samples_in_flight": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "sr_render_gpu_seconds": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "sr_adjoint_param_count": (ctypes.c_int, [ctypes.c_void_p, c_uint32_p]),
    "sr_adjoint_params_get": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "sr_adjoint_params_set": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "sr_adjoint_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(View_), ctypes.POINTER(RenderParams_),
                                          c_float_p]),
    "sr_adjoint_backward": (ctypes.c_int, [ctypes.c_void_p, c_float_p, c_float_p]),
    "sr_cnn_weight_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint64)]),
    "sr_cnn_load": (ctypes.c_int, [ctypes.c_void_p, c_float_p, ctypes.c_uint64]),
    "sr_cnn_denoise": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "sr_cnn_denoise_device": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 4 + [ctypes.c_uint32, ctypes.c_uint32,
                                                                                       ctypes.c_void_p]),
    "sr_cnn_layer_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_float_p, c_uint32_p, c_uint32_p, c_uint32_p]),
    "sr_cnn_benchmark": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "sr_cnn_train_begin": (ctypes.c_int, [ctypes.c_void_p]),
    "sr_cnn_train_add": (ctypes.c_int, [ctypes.c_void_p] + [c_float_p] * 5 + [ctypes.c_uint32, ctypes.c_uint32]),
    "sr_cnn_train_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32, c_float_p]),
    "sr_cnn_train_export": (ctypes.c_int, [ctypes.c_void_p, c_float_p]),
    "sr_cnn_train_gradient": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32, c_float_p, c_float_p]),
    "sr_cnn_flops": (ctypes.c_int, [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_double)]),
    "sr_triangles_build": (ctypes.c_int, [c_float_p, c_float_p, c_float_p, ctypes.c_uint32, ctypes.c_int32, c_float_p]),
    "sr_obj_load": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(Transform_),
                                   ctypes.POINTER(Material_), ctypes.c_int32, c_uint32_p]),
    "sr_hdr_load": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_float, c_float_p, c_uint32_p, c_uint32_p]),
    "sr_hdr_decode": (ctypes.c_int, [c_uint8_p, ctypes.c_uint64, ctypes.c_float, c_float_p, c_uint32_p, c_uint32_p]),
    "sr_tga_decode": (ctypes.c_int, [c_uint8_p, ctypes.c_uint64, c_uint8_p, c_uint32_p, c_uint32_p]),
    "sr_counters_read_fused": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "sr_counters_read_reference_only": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "sr_set_tail_segment": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint32]),
    "sr_set_bvh_builder": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "sr_scene_counts": (ctypes.c_int, [ctypes.c_void_p, c_uint32_p, c_uint32_p, c_uint32_p]),
    "sr_texture_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_float_p, c_uint32_p, c_uint32_p]),
    "sr_mesh_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, c_float_p, c_uint32_p]),
    "sr_scene_object_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(Transform_),
                                            ctypes.POINTER(Material_), c_int32_p, c_int32_p]),
}

_lib = None


class SharpRayError(RuntimeError):
    """A C ABI call returned a non-zero status."""


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SharpRayError(
                f"{LIB_PATH} is missing: build it with sharpray_amd/csrc/build.sh (there is no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status, what=""):
    if status != SR_OK:
        msg = lib().sr_last_error().decode(errors="replace")
        raise SharpRayError(f"{what}: status {status}: {msg}")
    return status


def fptr(a):
    return a.ctypes.data_as(c_float_p) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(c_int32_p) if a is not None else None


def uptr(a):
    return a.ctypes.data_as(c_uint32_p) if a is not None else None
