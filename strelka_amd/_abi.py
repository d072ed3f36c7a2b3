"""ctypes binding of libstrelka_amd.so (include/stk/stk.h, include/stk/stk_scene.h).

The structs below mirror include/stk/stk_types.h byte for byte; their sizes are
checked on import.  Loading fails loudly when the library has not been built:
there is no Python fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libstrelka_amd.so"

# ---- numpy record types (array buffers) ------------------------------------
VERTEX = np.dtype([("position", "<f4", 3), ("tangent", "<u4"), ("normal", "<u4"), ("uv", "<u4"),
                   ("pad0", "<f4"), ("pad1", "<f4")])
BVH_NODE = np.dtype([("minBounds", "<f4", 3), ("instId", "<i4"), ("maxBounds", "<f4", 3), ("nodeOffset", "<i4")])
INSTANCE = np.dtype([("objectToWorld", "<f4", 16), ("worldToObject", "<f4", 16), ("normalMatrix", "<f4", 16),
                     ("materialId", "<i4"), ("indexOffset", "<i4"), ("indexCount", "<i4"), ("lightId", "<i4")])
LIGHT = np.dtype([("points", "<f4", (4, 4)), ("color", "<f4", 4), ("normal", "<f4", 4), ("type", "<i4"),
                  ("pad", "<f4", 3)])
RAY = np.dtype([("o", "<f4", 4), ("d", "<f4", 4)])
HIT = np.dtype([("bary", "<f4", 2), ("t", "<f4"), ("instId", "<u4"), ("primId", "<u4")])
MAT_LAMBERT, MAT_GGX, MAT_MIRROR, MAT_GLASS = 0, 1, 2, 3  # StkMaterial.type
MATERIAL = np.dtype([("type", "<u4"), ("roughness", "<f4"), ("albedoTexture", "<u4"), ("roughnessTexture", "<u4"),
                     ("albedo", "<f4", 4), ("emission", "<f4", 3), ("ior", "<f4"), ("normalTexture", "<u4"),
                     ("emissionTexture", "<u4"), ("pad", "<u4", 2)])
TEX_REPEAT, TEX_CLAMP, TEX_MIRRORED_REPEAT, TEX_CLIP = 0, 1, 2, 3  # StkTexture.wrap
TEXTURE = np.dtype([("width", "<u4"), ("height", "<u4"), ("offset", "<u4"), ("wrap", "<u4")])
assert MATERIAL.itemsize == 64 and TEXTURE.itemsize == 16
TRIANGLE = np.dtype([("v0", "<f4", 3), ("leaf", "<u4"), ("e0", "<f4", 3), ("instId", "<u4"), ("e1", "<f4", 3),
                     ("primId", "<u4")])
PACKED_NODE = np.dtype([("x", "<f4", 4), ("y", "<f4", 4), ("z", "<f4", 4), ("child", "<i4", 2), ("pad", "<u4", 2)])
WIDE_NODE = np.dtype([("origin", "<f4", 3), ("exps", "u1", 3), ("imask", "u1"), ("childBase", "<u4"),
                      ("triBase", "<u4"), ("vmask", "<u4"), ("pad", "<u4"), ("qlo", "u1", (3, 8)),
                      ("qhi", "u1", (3, 8))])
BLAS_TRIANGLE = np.dtype([("p0", "<f4", 3), ("primId", "<u4"), ("p1", "<f4", 3), ("leaf", "<u4"), ("p2", "<f4", 3),
                          ("pad", "<u4")])
TLAS_LEAF = np.dtype([("lo", "<f4", 3), ("instId", "<u4"), ("hi", "<f4", 3), ("blasRoot", "<u4"), ("key", "<u4"),
                      ("pad", "<u4", 3)])

for _dt, _sz in ((VERTEX, 32), (BVH_NODE, 32), (INSTANCE, 208), (LIGHT, 112), (RAY, 32), (HIT, 20),
                 (MATERIAL, 64), (TRIANGLE, 48), (PACKED_NODE, 64), (WIDE_NODE, 80), (BLAS_TRIANGLE, 48),
                 (TLAS_LEAF, 48)):
    assert _dt.itemsize == _sz, (_dt, _sz)

INVALID = 0xFFFFFFFF


# ---- ctypes structs passed by pointer -------------------------------------
class PathTracerParam(C.Structure):
    _fields_ = [("viewToWorld", C.c_float * 16), ("clipToView", C.c_float * 16), ("viewToClip", C.c_float * 16),
                ("worldToView", C.c_float * 16), ("camPos", C.c_float * 4), ("dimension", C.c_int32 * 2),
                ("invDimension", C.c_float * 2), ("frameNumber", C.c_uint32), ("maxDepth", C.c_uint32),
                ("debug", C.c_uint32), ("numLights", C.c_uint32), ("len", C.c_int32), ("spp", C.c_int32),
                ("iteration", C.c_int32), ("stratifiedSamplingType", C.c_uint32), ("pad0", C.c_float * 4)]


class RenderParams(C.Structure):
    _fields_ = [("pt", PathTracerParam), ("rrStartDepth", C.c_uint32), ("refCompat", C.c_uint32),
                ("tonemapperType", C.c_uint32), ("enableAccumulation", C.c_uint32), ("skyColor", C.c_float * 4),
                ("enableTonemap", C.c_uint32), ("keepSamples", C.c_uint32), ("cacheMode", C.c_uint32),
                ("cacheQueryDepth", C.c_uint32), ("cacheTrainPaths", C.c_uint32),
                ("cacheMinRoughness", C.c_float), ("cacheTrainSteps", C.c_uint32), ("cacheTrainSuffix", C.c_uint32),
                ("cacheQueryDirect", C.c_uint32), ("pad", C.c_uint32 * 1)]


class CacheConfig(C.Structure):
    _fields_ = [("log2TableSize", C.c_uint32), ("baseResolution", C.c_uint32), ("maxResolution", C.c_uint32),
                ("seed", C.c_uint32), ("learningRate", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("epsilon", C.c_float), ("maxBatch", C.c_uint32), ("emaDecay", C.c_float), ("lrHalfLife", C.c_float),
                ("minLearningRate", C.c_float)]


class CacheInfo(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("scale", C.c_float), ("resolution", C.c_uint32 * 16),
                ("dense", C.c_uint32 * 16), ("log2TableSize", C.c_uint32), ("gridParams", C.c_uint32),
                ("mlpParams", C.c_uint32), ("step", C.c_uint32)]


CACHE_RECORD = np.dtype([("position", "<f4", 3), ("roughness", "<f4"), ("normal", "<f4", 3), ("payload", "<u4"),
                         ("direction", "<f4", 3), ("pad0", "<f4"), ("target", "<f4", 3), ("pad1", "<f4")])
assert CACHE_RECORD.itemsize == 64


class BvhOptions(C.Structure):
    _fields_ = [("maxDepth", C.c_uint32), ("numBins", C.c_uint32), ("traversalCost", C.c_float),
                ("intersectionCost", C.c_float), ("spatialSplitAlpha", C.c_float)]


class BvhHeader(C.Structure):
    _fields_ = [("lo", C.c_float * 3), ("root", C.c_int32), ("hi", C.c_float * 3), ("numTriangles", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [("header", BvhHeader), ("depth", C.c_uint32), ("wideDepth", C.c_uint32), ("numNodes", C.c_uint32),
                ("numWide", C.c_uint32), ("numEmitters", C.c_uint32), ("twoLevel", C.c_uint32),
                ("gpuBuilt", C.c_uint32), ("walkStack", C.c_uint32)]


class LaunchConfig(C.Structure):
    _fields_ = [("blockSize", C.c_uint32), ("tile", C.c_uint32), ("persistentBlocks", C.c_uint32),
                ("useGraph", C.c_uint32), ("pathsInFlight", C.c_uint32), ("binaryWalk", C.c_uint32),
                ("stripeThreads", C.c_uint32), ("passStreams", C.c_uint32), ("frameBlocks", C.c_uint32),
                ("frameChunk", C.c_uint32)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 64), ("computeUnits", C.c_uint32),
                ("wavefrontSize", C.c_uint32), ("globalMemBytes", C.c_uint64), ("l2Bytes", C.c_uint32),
                ("clockKhz", C.c_uint32), ("memClockKhz", C.c_uint32), ("memBusWidth", C.c_uint32)]


class BatchScene(C.Structure):
    _fields_ = [("vb", C.c_void_p), ("numVertices", C.c_uint32), ("ib", C.c_void_p), ("numIndices", C.c_uint32),
                ("inst", C.c_void_p), ("numInstances", C.c_uint32), ("lights", C.c_void_p),
                ("numLights", C.c_uint32), ("mats", C.c_void_p), ("numMaterials", C.c_uint32),
                ("camera", PathTracerParam), ("sky", C.c_float * 4)]


class BatchViewInfo(C.Structure):
    _fields_ = [("pixelBase", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32),
                ("numTriangles", C.c_uint32), ("bvhDepth", C.c_uint32), ("numScenes", C.c_uint32),
                ("pad", C.c_uint32 * 2)]


class ImageDesc(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("pitchBytes", C.c_uint32),
                ("format", C.c_uint32), ("onDevice", C.c_int32), ("pad", C.c_uint32)]


class SceneArrays(C.Structure):
    _fields_ = [("vb", C.c_void_p), ("numVertices", C.c_uint32), ("ib", C.c_void_p), ("numIndices", C.c_uint32),
                ("inst", C.c_void_p), ("numInstances", C.c_uint32), ("lights", C.c_void_p),
                ("numLights", C.c_uint32), ("mats", C.c_void_p), ("numMaterials", C.c_uint32),
                ("numTriangles", C.c_uint32), ("pad", C.c_uint32), ("camera", PathTracerParam),
                ("sky", C.c_float * 4)]


class CheckpointInfo(C.Structure):
    _fields_ = [("sections", C.c_uint32), ("iteration", C.c_uint32), ("pixels", C.c_uint64),
                ("userFrame", C.c_uint64), ("cacheStep", C.c_uint32), ("pad", C.c_uint32)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("xform", C.c_float * 16), ("color", C.c_float * 3), ("intensity", C.c_float),
                ("width", C.c_float), ("height", C.c_float), ("radius", C.c_float)]


assert C.sizeof(LightDesc) == 96
assert C.sizeof(PathTracerParam) == 336
assert C.sizeof(RenderParams) == 336 + 72
assert C.sizeof(CacheConfig) == 48
assert C.sizeof(BvhHeader) == 32

STATUS = {0: "STK_OK", 1: "STK_E_INVALID", 2: "STK_E_OOM", 3: "STK_E_DEVICE", 4: "STK_E_STATE", 5: "STK_E_TIMEOUT"}

# (name, restype, argtypes) for every entry point declared in include/stk/*.h
P = C.c_void_p
U32 = C.c_uint32
PROTOTYPES = [
    ("stk_version", C.c_char_p, []),
    ("stk_build_info", C.c_char_p, []),
    ("stk_create", C.c_int, [C.c_int, C.POINTER(P)]),
    ("stk_destroy", None, [P]),
    ("stk_last_error", C.c_char_p, [P]),
    ("stk_sync", C.c_int, [P]),
    ("stk_stream", P, [P]),
    ("stk_set_geometry", C.c_int, [P, P, U32, P, U32, P, U32]),
    ("stk_set_lights", C.c_int, [P, P, U32]),
    ("stk_set_textures", C.c_int, [P, P, C.c_uint32, P, C.c_uint64]),
    ("stk_set_materials", C.c_int, [P, P, U32]),
    ("stk_build_bvh", C.c_int, [P, C.POINTER(BvhOptions)]),
    ("stk_set_bvh", C.c_int, [P, P, U32]),
    ("stk_get_bvh", C.c_int, [P, P, U32, C.POINTER(U32)]),
    ("stk_get_bvh_info", C.c_int, [P, P]),
    ("stk_trace_closest", C.c_int, [P, P, U32, P, U32]),
    ("stk_trace_closest_noalloc", C.c_int, [P, P, U32, P, U32]),
    ("stk_trace_any", C.c_int, [P, P, U32, P, U32]),
    ("stk_trace_closest_device", C.c_int, [P, P, U32, P, U32]),
    ("stk_trace_closest_image_device", C.c_int, [P, P, U32, U32, P, U32]),
    ("stk_trace_any_device", C.c_int, [P, P, U32, P, U32]),
    ("stk_generate_camera_rays_device", C.c_int, [P, C.POINTER(PathTracerParam), U32, P]),
    ("stk_render", C.c_int, [P, C.POINTER(RenderParams), P, C.c_int]),
    ("stk_reset_accumulation", C.c_int, [P]),
    ("stk_get_samples", C.c_int, [P, P, C.c_size_t]),
    ("stk_get_accumulation", C.c_int, [P, P, C.c_size_t]),
    ("stk_get_counters", C.c_int, [P, P]),
    ("stk_render_image", C.c_int, [P, C.POINTER(RenderParams), C.POINTER(ImageDesc)]),
    ("stk_host_alloc", C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    ("stk_host_free", C.c_int, [P, P]),
    ("stk_get_ray_totals", C.c_int, [P, P]),
    ("stk_set_launch_config", C.c_int, [P, C.POINTER(LaunchConfig)]),
    ("stk_get_launch_config", C.c_int, [P, C.POINTER(LaunchConfig)]),
    ("stk_device_info", C.c_int, [P, C.POINTER(DeviceInfo)]),
    ("stk_batch_clear", C.c_int, [P]),
    ("stk_batch_add_scene", C.c_int, [P, C.POINTER(BatchScene), C.POINTER(C.c_uint32)]),
    ("stk_batch_commit", C.c_int, [P, C.POINTER(BvhOptions)]),
    ("stk_batch_info", C.c_int, [P, C.c_uint32, C.POINTER(BatchViewInfo)]),
    ("stk_batch_render", C.c_int, [P, C.POINTER(RenderParams), P, C.c_int]),
    ("stk_cache_create", C.c_int, [P, C.POINTER(CacheConfig)]),
    ("stk_cache_destroy", C.c_int, [P]),
    ("stk_cache_info", C.c_int, [P, C.POINTER(CacheInfo)]),
    ("stk_cache_infer", C.c_int, [P, P, U32, P]),
    ("stk_cache_train_step", C.c_int, [P, P, U32, C.POINTER(C.c_float)]),
    ("stk_cache_gradients", C.c_int, [P, P, U32, P, P, C.POINTER(C.c_float)]),
    ("stk_cache_get_params", C.c_int, [P, P, P]),
    ("stk_cache_set_params", C.c_int, [P, P, P]),
    ("stk_cache_last_records", C.c_int, [P, U32, P, U32, C.POINTER(U32)]),
    ("stk_device_alloc", C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    ("stk_device_free", C.c_int, [P, P]),
    ("stk_copy_to_device", C.c_int, [P, P, P, C.c_size_t]),
    ("stk_copy_to_host", C.c_int, [P, P, P, C.c_size_t]),
    ("stk_device_copy", C.c_int, [P, P, P, C.c_size_t]),
    ("stk_event_record", C.c_int, [P, C.c_int]),
    ("stk_event_elapsed", C.c_int, [P, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    ("stk_bvh_build_host", C.c_int, [P, U32, P, U32, P, U32, C.POINTER(BvhOptions), C.POINTER(P)]),
    ("stk_bvh_from_nodes", C.c_int, [P, U32, P, U32, P, U32, P, U32, C.POINTER(P)]),
    ("stk_bvh_free", None, [P]),
    ("stk_bvh_num_nodes", U32, [P]),
    ("stk_bvh_nodes", P, [P]),
    ("stk_bvh_num_packed", U32, [P]),
    ("stk_bvh_packed", P, [P]),
    ("stk_bvh_num_triangles", U32, [P]),
    ("stk_bvh_triangles", P, [P]),
    ("stk_bvh_header", BvhHeader, [P]),
    ("stk_bvh_depth", U32, [P]),
    ("stk_bvh_num_wide", U32, [P]),
    ("stk_bvh_wide_nodes", P, [P]),
    ("stk_bvh_wide_triangles", P, [P]),
    ("stk_bvh_wide_depth", U32, [P]),
    ("stk_bvh_num_emitters", U32, [P]),
    ("stk_build_two_level", C.c_int, [P, C.POINTER(BvhOptions)]),
    ("stk_two_level_world_triangles", C.c_int, [P, C.c_int]),
    ("stk_build_bvh_gpu", C.c_int, [P]),
    ("stk_build_bvh_gpu_method", C.c_int, [P, C.c_uint32]),
    ("stk_bvh_save_nodes", C.c_int, [P, U32, C.c_char_p]),
    ("stk_bvh_load_nodes", C.c_int, [C.c_char_p, P, U32, C.POINTER(C.c_uint32)]),
    ("stk_image_write_pfm", C.c_int, [C.c_char_p, P, U32, U32]),
    ("stk_image_write_exr", C.c_int, [C.c_char_p, P, U32, U32]),
    ("stk_image_read_pfm", C.c_int, [C.c_char_p, P, C.c_size_t, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("stk_tlas_build_host", C.c_int, [P, U32, P, U32, P, U32, C.POINTER(BvhOptions), C.POINTER(P)]),
    ("stk_tlas_free", None, [P]),
    ("stk_tlas_num_meshes", U32, [P]),
    ("stk_tlas_instance_mesh", U32, [P, U32]),
    ("stk_tlas_mesh_nodes", P, [P, U32, C.POINTER(C.c_uint32)]),
    ("stk_tlas_top_nodes", P, [P, C.POINTER(C.c_uint32)]),
    ("stk_tlas_top_wide", P, [P, C.POINTER(C.c_uint32)]),
    ("stk_tlas_top_leaves", P, [P, C.POINTER(C.c_uint32)]),
    ("stk_tlas_mesh_wide", P, [P, C.POINTER(C.c_uint32)]),
    ("stk_tlas_mesh_triangles", P, [P, C.POINTER(C.c_uint32)]),
    ("stk_tlas_instance_keys", P, [P]),
    ("stk_tlas_instance_inverses", P, [P]),
    ("stk_tlas_emitters", P, [P, C.POINTER(C.c_uint32)]),
    ("stk_tlas_depths", None, [P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    ("stk_scene_generate", C.c_int, [U32, U32, U32, C.POINTER(P)]),
    ("stk_scene_free", None, [P]),
    ("stk_scene_get", C.c_int, [P, C.POINTER(SceneArrays)]),
    ("stk_scene_jittered_rays", C.c_int, [P, C.c_uint64, P, U32]),
    ("stk_scene_create", C.c_int, [C.POINTER(P)]),
    ("stk_scene_create_mesh", C.c_int, [P, P, U32, P, U32, C.POINTER(U32)]),
    ("stk_scene_create_instance", C.c_int, [P, U32, C.c_int32, P, C.POINTER(U32)]),
    ("stk_scene_update_instance_transform", C.c_int, [P, U32, P]),
    ("stk_scene_add_material", C.c_int, [P, P, C.POINTER(U32)]),
    ("stk_scene_create_light", C.c_int, [P, C.POINTER(LightDesc), C.POINTER(U32)]),
    ("stk_scene_update_light", C.c_int, [P, U32, C.POINTER(LightDesc)]),
    ("stk_scene_light_instance", C.c_int, [P, U32, C.POINTER(U32)]),
    ("stk_scene_add_camera", C.c_int, [P, C.POINTER(PathTracerParam), C.POINTER(U32)]),
    ("stk_scene_update_camera", C.c_int, [P, U32, C.POINTER(PathTracerParam)]),
    ("stk_scene_camera_count", U32, [P]),
    ("stk_scene_get_camera", C.c_int, [P, U32, C.POINTER(PathTracerParam)]),
    ("stk_scene_set_sky", C.c_int, [P, C.c_float, C.c_float, C.c_float]),
    ("stk_camera_look_at", C.c_int, [C.c_float * 3, C.c_float * 3, C.c_float * 3, C.c_float, U32, U32, C.c_float,
                                     C.c_float, C.POINTER(PathTracerParam)]),
    ("stk_pack_normal", U32, [C.c_float, C.c_float, C.c_float]),
    ("stk_pack_uv", U32, [C.c_float, C.c_float]),
    ("stk_set_scene", C.c_int, [P, P, C.POINTER(BvhOptions)]),
    ("stk_checkpoint_save", C.c_int, [P, C.c_char_p, C.c_uint64]),
    ("stk_set_watchdog", C.c_int, [P, U32]),
    ("stk_abort", C.c_int, [P]),
    ("stk_checkpoint_load", C.c_int, [P, C.c_char_p, C.POINTER(CheckpointInfo)]),
]


def load(path: Path | None = None) -> C.CDLL:
    path = Path(path or LIB_PATH)
    if not path.exists():
        raise ImportError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                          "(there is no CPU fallback for the HIP library)")
    lib = C.CDLL(str(path))
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
