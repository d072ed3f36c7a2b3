"""strelka_amd - MI355X (gfx950) path-tracing hot path behind a C ABI.

Python mirror of the library's C entry points (include/stk/stk.h).  All work
runs in libstrelka_amd.so: the BVH builder and scene container in host C++, the
traversal / path-tracing / post-process kernels in HIP.  Importing this package
fails if the library has not been built; nothing here falls back to Python or
CPU code.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import (BLAS_TRIANGLE, BVH_NODE, CACHE_RECORD, HIT, INSTANCE, INVALID, LIGHT, MAT_GGX, MAT_GLASS,
                   MAT_LAMBERT, MAT_MIRROR, MATERIAL, TEX_CLAMP, TEX_CLIP, TEX_MIRRORED_REPEAT, TEX_REPEAT, PACKED_NODE, RAY, TEXTURE, TLAS_LEAF, TRIANGLE, WIDE_NODE,
                   VERTEX, BatchScene, BatchViewInfo, BvhHeader, BvhInfo, BvhOptions, CacheConfig, CacheInfo, DeviceInfo, ImageDesc, LaunchConfig,
                   PathTracerParam, RenderParams, SceneArrays)

__all__ = ["lib", "StkError", "SceneData", "Scene", "look_at", "pack_normal", "pack_uv", "HostBvh", "HostTwoLevel", "Context", "RenderParams", "PathTracerParam", "VERTEX",
           "BVH_NODE", "INSTANCE", "LIGHT", "RAY", "HIT", "MATERIAL", "TRIANGLE", "PACKED_NODE", "INVALID",
           "render_params", "CACHE_RECORD", "MAT_LAMBERT", "MAT_GGX", "MAT_MIRROR", "MAT_GLASS", "TEXTURE",
           "TEX_REPEAT", "TEX_CLAMP", "TEX_MIRRORED_REPEAT", "TEX_CLIP", "WIDE_NODE"]

_LIB = _abi.load()


def lib() -> C.CDLL:
    return _LIB


class StkError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{_abi.STATUS.get(status, status)}: {msg}")
        self.status = status


def _ptr(a: np.ndarray | None):
    return None if a is None else C.c_void_p(a.ctypes.data)


def _copy(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    buf = (C.c_char * (n * dtype.itemsize)).from_address(ptr)
    return np.frombuffer(bytes(buf), dtype=dtype).copy()


@dataclass
class SceneData:
    """Flat scene buffers (numpy copies) in the reference's GPU record layouts."""

    vb: np.ndarray
    ib: np.ndarray
    inst: np.ndarray
    lights: np.ndarray
    mats: np.ndarray
    camera: PathTracerParam
    sky: tuple
    _handle: C.c_void_p | None = None
    textures: np.ndarray | None = None  # TEXTURE records (the materials' textures)
    texels: np.ndarray | None = None    # RGBA32F texels, shape (n, 4)

    @property
    def num_triangles(self) -> int:
        return int(self.ib.size // 3)

    @classmethod
    def generate(cls, config: int, variant: int = 0, ref_icosphere: bool = False) -> "SceneData":
        h = C.c_void_p()
        st = _LIB.stk_scene_generate(config, variant, 1 if ref_icosphere else 0, C.byref(h))
        if st:
            raise StkError(st, f"stk_scene_generate(config={config}, variant={variant})")
        a = SceneArrays()
        _LIB.stk_scene_get(h, C.byref(a))
        cam = PathTracerParam()
        C.memmove(C.byref(cam), C.byref(a.camera), C.sizeof(PathTracerParam))
        sd = cls(vb=_copy(a.vb, a.numVertices, VERTEX), ib=_copy(a.ib, a.numIndices, np.dtype("<u4")),
                 inst=_copy(a.inst, a.numInstances, INSTANCE), lights=_copy(a.lights, a.numLights, LIGHT),
                 mats=_copy(a.mats, a.numMaterials, MATERIAL), camera=cam, sky=tuple(a.sky), _handle=h)
        return sd

    def jittered_rays(self, seed: int) -> np.ndarray:
        w, h = self.camera.dimension[0], self.camera.dimension[1]
        rays = np.zeros(w * h, dtype=RAY)
        st = _LIB.stk_scene_jittered_rays(self._handle, C.c_uint64(seed), _ptr(rays), rays.size)
        if st:
            raise StkError(st, "stk_scene_jittered_rays")
        return rays

    def __del__(self):
        if self._handle:
            _LIB.stk_scene_free(self._handle)
            self._handle = None


def look_at(eye, target, up, vfov_deg: float, width: int, height: int, znear: float = 0.01,
            zfar: float = 100.0) -> PathTracerParam:
    """stk_camera_look_at: oka::Camera's view and reverse-Z projection
    matrices for a look-at camera (src/scene/camera.cpp:61-131)."""
    f3 = C.c_float * 3
    cam = PathTracerParam()
    st = _LIB.stk_camera_look_at(f3(*eye), f3(*target), f3(*up), vfov_deg, width, height, znear, zfar,
                                 C.byref(cam))
    if st:
        raise StkError(st, "stk_camera_look_at")
    return cam


def pack_normal(x: float, y: float, z: float) -> int:
    """The vertex record's 10:10:10 normal packing (src/hdStrelka/RenderPass.cpp:50-57)."""
    return int(_LIB.stk_pack_normal(x, y, z))


def pack_uv(u: float, v: float) -> int:
    """The vertex record's 16:16 UV packing (src/hdStrelka/RenderPass.cpp:59-64)."""
    return int(_LIB.stk_pack_uv(u, v))


def _mat16(m) -> np.ndarray:
    """A 4x4 transform as the column-major float[16] the C ABI takes: a (4, 4)
    array is read as a matrix (rows, columns); a flat 16-array is taken as
    already column-major."""
    a = np.asarray(m, dtype=np.float32)
    if a.shape == (4, 4):
        a = a.T
    a = np.ascontiguousarray(a.reshape(16), dtype=np.float32)
    return a


class Scene:
    """The scene container of include/stk/stk_scene.h, built through oka::Scene's
    calls (include/scene/scene.h:231-389): meshes, instances, materials,
    lights and cameras.  Every call raises StkError when the library refuses
    it; ids count from 0 in creation order."""

    def __init__(self):
        self._handle = C.c_void_p()
        st = _LIB.stk_scene_create(C.byref(self._handle))
        if st:
            raise StkError(st, "stk_scene_create")

    def _call(self, fn, what, *args):
        st = fn(self._handle, *args)
        if st:
            raise StkError(st, what)

    def create_mesh(self, vb: np.ndarray, ib) -> int:
        vb = np.ascontiguousarray(vb, dtype=VERTEX)
        ib = np.ascontiguousarray(ib, dtype=np.uint32)
        out = C.c_uint32()
        self._call(_LIB.stk_scene_create_mesh, "stk_scene_create_mesh", _ptr(vb), vb.size, _ptr(ib), ib.size,
                   C.byref(out))
        return out.value

    def create_instance(self, mesh_id: int, material_id: int, transform) -> int:
        m = _mat16(transform)
        out = C.c_uint32()
        self._call(_LIB.stk_scene_create_instance, "stk_scene_create_instance", mesh_id, material_id, _ptr(m),
                   C.byref(out))
        return out.value

    def update_instance_transform(self, inst_id: int, transform):
        m = _mat16(transform)
        self._call(_LIB.stk_scene_update_instance_transform, "stk_scene_update_instance_transform", inst_id, _ptr(m))

    def add_material(self, mat) -> int:
        rec = np.ascontiguousarray(np.asarray(mat, dtype=MATERIAL).reshape(1))
        out = C.c_uint32()
        self._call(_LIB.stk_scene_add_material, "stk_scene_add_material", _ptr(rec), C.byref(out))
        return out.value

    @staticmethod
    def light_desc(type: int = 0, xform=None, color=(1.0, 1.0, 1.0), intensity: float = 1.0, width: float = 1.0,
                   height: float = 1.0, radius: float = 1.0) -> _abi.LightDesc:
        d = _abi.LightDesc()
        d.type = type
        m = _mat16(np.eye(4, dtype=np.float32) if xform is None else xform)
        for i in range(16):
            d.xform[i] = float(m[i])
        for i in range(3):
            d.color[i] = color[i]
        d.intensity, d.width, d.height, d.radius = intensity, width, height, radius
        return d

    def create_light(self, desc: _abi.LightDesc) -> int:
        out = C.c_uint32()
        self._call(_LIB.stk_scene_create_light, "stk_scene_create_light", C.byref(desc), C.byref(out))
        return out.value

    def update_light(self, light_id: int, desc: _abi.LightDesc):
        self._call(_LIB.stk_scene_update_light, "stk_scene_update_light", light_id, C.byref(desc))

    def light_instance(self, light_id: int) -> int:
        out = C.c_uint32()
        self._call(_LIB.stk_scene_light_instance, "stk_scene_light_instance", light_id, C.byref(out))
        return out.value

    def add_camera(self, camera: PathTracerParam) -> int:
        out = C.c_uint32()
        self._call(_LIB.stk_scene_add_camera, "stk_scene_add_camera", C.byref(camera), C.byref(out))
        return out.value

    def update_camera(self, camera_id: int, camera: PathTracerParam):
        self._call(_LIB.stk_scene_update_camera, "stk_scene_update_camera", camera_id, C.byref(camera))

    def camera_count(self) -> int:
        return int(_LIB.stk_scene_camera_count(self._handle))

    def camera(self, camera_id: int) -> PathTracerParam:
        cam = PathTracerParam()
        self._call(_LIB.stk_scene_get_camera, "stk_scene_get_camera", camera_id, C.byref(cam))
        return cam

    def set_sky(self, r: float, g: float, b: float):
        self._call(_LIB.stk_scene_set_sky, "stk_scene_set_sky", r, g, b)

    def data(self) -> "SceneData":
        """Copies of the container's arrays (stk_scene_get) as a SceneData."""
        a = SceneArrays()
        self._call(_LIB.stk_scene_get, "stk_scene_get", C.byref(a))
        cam = PathTracerParam()
        C.memmove(C.byref(cam), C.byref(a.camera), C.sizeof(PathTracerParam))
        return SceneData(vb=_copy(a.vb, a.numVertices, VERTEX), ib=_copy(a.ib, a.numIndices, np.dtype("<u4")),
                         inst=_copy(a.inst, a.numInstances, INSTANCE), lights=_copy(a.lights, a.numLights, LIGHT),
                         mats=_copy(a.mats, a.numMaterials, MATERIAL), camera=cam, sky=tuple(a.sky))

    def __del__(self):
        if getattr(self, "_handle", None):
            _LIB.stk_scene_free(self._handle)
            self._handle = None


class HostBvh:
    """BVH built or packed on the host (no device needed)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        n = _LIB.stk_bvh_num_nodes(handle)
        self.nodes = _copy(_LIB.stk_bvh_nodes(handle), n, BVH_NODE)
        self.packed = _copy(_LIB.stk_bvh_packed(handle), _LIB.stk_bvh_num_packed(handle), PACKED_NODE)
        self.tris = _copy(_LIB.stk_bvh_triangles(handle), _LIB.stk_bvh_num_triangles(handle), TRIANGLE)
        self.header = _LIB.stk_bvh_header(handle)
        self.depth = int(_LIB.stk_bvh_depth(handle))
        self.wide = _copy(_LIB.stk_bvh_wide_nodes(handle), _LIB.stk_bvh_num_wide(handle), WIDE_NODE)
        self.wide_tris = _copy(_LIB.stk_bvh_wide_triangles(handle), self.tris.size, TRIANGLE)
        self.wide_depth = int(_LIB.stk_bvh_wide_depth(handle))
        self.num_emitters = int(_LIB.stk_bvh_num_emitters(handle))

    @classmethod
    def build(cls, scene: SceneData, max_depth: int = 32, bins: int = 32, spatial_split_alpha: float = 0.0) -> "HostBvh":
        opts = BvhOptions(max_depth, bins, 1.0, 1.0, spatial_split_alpha)
        h = C.c_void_p()
        st = _LIB.stk_bvh_build_host(_ptr(scene.vb), scene.vb.size, _ptr(scene.ib), scene.ib.size, _ptr(scene.inst),
                                     scene.inst.size, C.byref(opts), C.byref(h))
        if st:
            raise StkError(st, "stk_bvh_build_host")
        return cls(h)

    @classmethod
    def from_nodes(cls, scene: SceneData, nodes: np.ndarray) -> "HostBvh":
        nodes = np.ascontiguousarray(nodes, dtype=BVH_NODE)
        h = C.c_void_p()
        st = _LIB.stk_bvh_from_nodes(_ptr(scene.vb), scene.vb.size, _ptr(scene.ib), scene.ib.size, _ptr(scene.inst),
                                     scene.inst.size, _ptr(nodes), nodes.size, C.byref(h))
        if st:
            raise StkError(st, "stk_bvh_from_nodes")
        return cls(h)

    def __del__(self):
        if getattr(self, "_h", None):
            _LIB.stk_bvh_free(self._h)
            self._h = None


class HostTwoLevel:
    """Two-level structure built on the host (no device needed): a binary and a
    compressed 8-wide tree per distinct mesh in object space, and over the
    instances' world boxes (the build stk_build_two_level runs)."""

    def __init__(self, handle: C.c_void_p, num_instances: int):
        self._h = handle
        n = C.c_uint32()

        def arr(fn, dtype, *args):
            p = fn(handle, *args, C.byref(n))
            return _copy(p, n.value, dtype)

        self.num_meshes = int(_LIB.stk_tlas_num_meshes(handle))
        self.instance_mesh = np.array([_LIB.stk_tlas_instance_mesh(handle, i) for i in range(num_instances)],
                                      dtype=np.uint32)
        self.mesh_nodes = [arr(_LIB.stk_tlas_mesh_nodes, BVH_NODE, m) for m in range(self.num_meshes)]
        self.top_nodes = arr(_LIB.stk_tlas_top_nodes, BVH_NODE)
        self.top_wide = arr(_LIB.stk_tlas_top_wide, WIDE_NODE)
        self.top_leaves = arr(_LIB.stk_tlas_top_leaves, TLAS_LEAF)
        self.mesh_wide = arr(_LIB.stk_tlas_mesh_wide, WIDE_NODE)
        self.mesh_tris = arr(_LIB.stk_tlas_mesh_triangles, BLAS_TRIANGLE)
        self.emitters = arr(_LIB.stk_tlas_emitters, TRIANGLE)
        self.instance_keys = _copy(_LIB.stk_tlas_instance_keys(handle), num_instances, np.dtype(np.uint32)) \
            if num_instances else np.zeros(0, np.uint32)
        self.instance_inverses = _copy(_LIB.stk_tlas_instance_inverses(handle), 16 * num_instances,
                                       np.dtype(np.float32)).reshape(num_instances, 16) \
            if num_instances else np.zeros((0, 16), np.float32)
        top, mesh = C.c_uint32(), C.c_uint32()
        _LIB.stk_tlas_depths(handle, C.byref(top), C.byref(mesh))
        self.top_depth, self.mesh_depth = int(top.value), int(mesh.value)

    @classmethod
    def build(cls, scene: SceneData, max_depth: int = 32, bins: int = 32) -> "HostTwoLevel":
        opts = BvhOptions(max_depth, bins, 1.0, 1.0, 0.0)
        h = C.c_void_p()
        st = _LIB.stk_tlas_build_host(_ptr(scene.vb), scene.vb.size, _ptr(scene.ib), scene.ib.size,
                                      _ptr(scene.inst), scene.inst.size, C.byref(opts), C.byref(h))
        if st:
            raise StkError(st, "stk_tlas_build_host")
        return cls(h, scene.inst.size)

    def __del__(self):
        if getattr(self, "_h", None):
            _LIB.stk_tlas_free(self._h)
            self._h = None


def save_bvh(path, nodes: np.ndarray) -> None:
    """stk_bvh_save_nodes: the skip-link nodes to a BVH file."""
    nodes = np.ascontiguousarray(nodes, dtype=BVH_NODE)
    st = _LIB.stk_bvh_save_nodes(_ptr(nodes), nodes.size, os.fsencode(path))
    if st:
        raise StkError(st, f"stk_bvh_save_nodes({path})")


def load_bvh(path) -> np.ndarray:
    """stk_bvh_load_nodes: the skip-link nodes of a BVH file."""
    n = C.c_uint32()
    st = _LIB.stk_bvh_load_nodes(os.fsencode(path), None, 0, C.byref(n))
    if st:
        raise StkError(st, f"stk_bvh_load_nodes({path})")
    nodes = np.zeros(n.value, dtype=BVH_NODE)
    st = _LIB.stk_bvh_load_nodes(os.fsencode(path), _ptr(nodes), nodes.size, C.byref(n))
    if st:
        raise StkError(st, f"stk_bvh_load_nodes({path})")
    return nodes


def write_pfm(path, rgba: np.ndarray) -> None:
    """stk_image_write_pfm: an (h, w, 4) float32 image, top row first."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[0], rgba.shape[1]
    assert rgba.shape == (h, w, 4)
    st = _LIB.stk_image_write_pfm(os.fsencode(path), _ptr(rgba), w, h)
    if st:
        raise StkError(st, f"stk_image_write_pfm({path})")


def write_exr(path, rgba: np.ndarray) -> None:
    """stk_image_write_exr: an (h, w, 4) float32 image, top row first, as an
    uncompressed OpenEXR file of float B, G, R channels."""
    rgba = np.ascontiguousarray(rgba, dtype=np.float32)
    h, w = rgba.shape[0], rgba.shape[1]
    assert rgba.shape == (h, w, 4)
    st = _LIB.stk_image_write_exr(os.fsencode(path), _ptr(rgba), w, h)
    if st:
        raise StkError(st, f"stk_image_write_exr({path})")


def read_pfm(path) -> np.ndarray:
    """stk_image_read_pfm: an (h, w, 4) float32 image, top row first."""
    w, h = C.c_uint32(), C.c_uint32()
    st = _LIB.stk_image_read_pfm(os.fsencode(path), None, 0, C.byref(w), C.byref(h))
    if st:
        raise StkError(st, f"stk_image_read_pfm({path})")
    out = np.zeros((h.value, w.value, 4), np.float32)
    st = _LIB.stk_image_read_pfm(os.fsencode(path), _ptr(out), out.size, C.byref(w), C.byref(h))
    if st:
        raise StkError(st, f"stk_image_read_pfm({path})")
    return out


def rmse(a: np.ndarray, b: np.ndarray, tonemapped: bool = True) -> float:
    """Root-mean-square difference of two images' RGB channels; with
    tonemapped, both are first clamped to [0, 1] (the display range of an
    ACES-tonemapped frame)."""
    x = np.asarray(a, np.float64)[..., :3]
    y = np.asarray(b, np.float64)[..., :3]
    if tonemapped:
        x, y = np.clip(x, 0.0, 1.0), np.clip(y, 0.0, 1.0)
    return float(np.sqrt(np.mean((x - y) ** 2)))


def render_params(camera: PathTracerParam, spp: int = 1, max_depth: int = 8, frame: int = 0,
                  sampling: int = 1, sky=(0.0, 0.0, 0.0), tonemap: int = 1, accumulate: bool = True,
                  keep_samples: bool = False, rr_start: int = 3, enable_tonemap: bool = True,
                  cache_mode: int = 0, cache_query_depth: int = 2, cache_train_paths: int = 65536,
                  cache_min_roughness: float = 0.0, cache_train_steps: int = 1,
                  ref_compat: bool = True, cache_train_suffix: int = 0,
                  cache_query_direct: bool = False) -> RenderParams:
    """StkRenderParams for a render call.  ref_compat: reproduce the
    reference estimator exactly (True) or run the corrected one (False: the
    light sample's solid-angle pdf, only non-emitters occlude, BSDF-sampled
    light hits weighted by MIS, Russian roulette with p <= 1)."""
    p = RenderParams()
    C.memmove(C.byref(p.pt), C.byref(camera), C.sizeof(PathTracerParam))
    p.pt.spp = spp
    p.pt.maxDepth = max_depth
    p.pt.frameNumber = frame
    p.pt.stratifiedSamplingType = sampling
    p.rrStartDepth = rr_start
    p.refCompat = 1 if ref_compat else 0
    p.tonemapperType = tonemap
    p.enableAccumulation = 1 if accumulate else 0
    p.enableTonemap = 1 if enable_tonemap else 0
    p.keepSamples = 1 if keep_samples else 0
    p.cacheMode = cache_mode
    p.cacheQueryDepth = cache_query_depth
    p.cacheTrainPaths = cache_train_paths
    p.cacheMinRoughness = cache_min_roughness
    p.cacheTrainSteps = cache_train_steps
    p.cacheTrainSuffix = cache_train_suffix
    p.cacheQueryDirect = 1 if cache_query_direct else 0
    for i in range(3):
        p.skyColor[i] = float(sky[i])
    p.skyColor[3] = 0.0
    return p


class Context:
    """One device context (stk_ctx): owns device buffers and one HIP stream."""

    def __init__(self, device: int = 0):
        self._h = None  # set before stk_create so that __del__ of a failed create is a no-op
        h = C.c_void_p()
        st = _LIB.stk_create(device, C.byref(h))
        if st:
            raise StkError(st, f"stk_create(device={device})")
        self._h = h
        self._batch_views = []

    @property
    def handle(self):
        return self._h

    def _check(self, st: int, what: str):
        if st:
            raise StkError(st, f"{what}: {_LIB.stk_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            _LIB.stk_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()

    # -- scene -----------------------------------------------------------------
    def set_scene(self, s: SceneData):
        self._check(_LIB.stk_set_geometry(self._h, _ptr(s.vb), s.vb.size, _ptr(s.ib), s.ib.size, _ptr(s.inst),
                                          s.inst.size), "stk_set_geometry")
        self._check(_LIB.stk_set_lights(self._h, _ptr(s.lights), s.lights.size), "stk_set_lights")
        self.set_materials(s.mats)
        if s.textures is not None:
            self.set_textures(s.textures, s.texels)

    def set_materials(self, mats: np.ndarray):
        """stk_set_materials: MATERIAL records (an unknown type or a negative
        index of refraction is refused)."""
        mats = np.ascontiguousarray(mats, dtype=_abi.MATERIAL)
        self._check(_LIB.stk_set_materials(self._h, _ptr(mats), mats.size), "stk_set_materials")

    def load_scene(self, s, max_depth: int = 32, bins: int = 32, spatial_split_alpha: float = 0.0):
        """stk_set_scene: the whole container (a Scene, or a generated
        SceneData) goes to the device and the flat BVH is built over it."""
        opts = BvhOptions(max_depth, bins, 1.0, 1.0, spatial_split_alpha)
        self._check(_LIB.stk_set_scene(self._h, s._handle, C.byref(opts)), "stk_set_scene")

    def set_textures(self, textures: np.ndarray, texels: np.ndarray):
        """stk_set_textures: TEXTURE records over RGBA32F texels (n x 4)."""
        textures = np.ascontiguousarray(textures, dtype=_abi.TEXTURE)
        texels = np.ascontiguousarray(texels, dtype=np.float32).reshape(-1, 4)
        self._check(_LIB.stk_set_textures(self._h, _ptr(textures), textures.size, _ptr(texels), texels.shape[0]),
                    "stk_set_textures")

    def build_bvh(self, max_depth: int = 32, bins: int = 32, spatial_split_alpha: float = 0.0):
        """stk_build_bvh: the flat tree by binned SAH on the host;
        spatial_split_alpha > 0 allows spatial splits (1e-5 is the usual
        threshold: children overlapping by more than that fraction of the
        root's area)."""
        opts = BvhOptions(max_depth, bins, 1.0, 1.0, spatial_split_alpha)
        self._check(_LIB.stk_build_bvh(self._h, C.byref(opts)), "stk_build_bvh")

    def build_bvh_gpu(self, method: str = "lbvh"):
        """stk_build_bvh_gpu_method: the flat tree built on the device, as a
        linear BVH ("lbvh") or by locally-ordered clustering ("ploc")."""
        m = {"lbvh": 0, "ploc": 1}[method]
        self._check(_LIB.stk_build_bvh_gpu_method(self._h, m), "stk_build_bvh_gpu_method")

    def build_two_level(self, max_depth: int = 32, bins: int = 32):
        """stk_build_two_level: per-mesh trees under a tree of instances."""
        opts = BvhOptions(max_depth, bins, 1.0, 1.0, 0.0)
        self._check(_LIB.stk_build_two_level(self._h, C.byref(opts)), "stk_build_two_level")

    def two_level_world_triangles(self, enable: bool = True):
        """stk_two_level_world_triangles: keep (or free) every instance's world
        triangles in device memory for the two-level walks."""
        self._check(_LIB.stk_two_level_world_triangles(self._h, 1 if enable else 0),
                    "stk_two_level_world_triangles")

    def set_bvh(self, nodes: np.ndarray):
        nodes = np.ascontiguousarray(nodes, dtype=BVH_NODE)
        self._check(_LIB.stk_set_bvh(self._h, _ptr(nodes), nodes.size), "stk_set_bvh")

    def get_bvh(self) -> np.ndarray:
        n = C.c_uint32()
        self._check(_LIB.stk_get_bvh(self._h, None, 0, C.byref(n)), "stk_get_bvh")
        out = np.zeros(n.value, dtype=BVH_NODE)
        self._check(_LIB.stk_get_bvh(self._h, _ptr(out), out.size, C.byref(n)), "stk_get_bvh")
        return out

    def bvh_info(self) -> BvhInfo:
        """stk_get_bvh_info: depths and sizes of what the context traces, and
        the stack entries of the closest-hit kernel its tree takes."""
        info = BvhInfo()
        self._check(_LIB.stk_get_bvh_info(self._h, C.byref(info)), "stk_get_bvh_info")
        return info

    # -- traversal (host buffers) ----------------------------------------------
    def trace_closest(self, rays: np.ndarray, reference_order: bool = False, binary: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(rays.size, dtype=HIT)
        flags = (1 if reference_order else 0) | (2 if binary else 0)
        self._check(_LIB.stk_trace_closest(self._h, _ptr(rays), rays.size, _ptr(hits), flags), "stk_trace_closest")
        return hits

    def trace_closest_noalloc(self, rays: np.ndarray, reference_order: bool = False,
                              binary: bool = False) -> np.ndarray:
        """stk_trace_closest_noalloc: the same hits, staged through the
        library's static device buffer (at most 1,048,576 rays per call)."""
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(rays.size, dtype=HIT)
        flags = (1 if reference_order else 0) | (2 if binary else 0)
        self._check(_LIB.stk_trace_closest_noalloc(self._h, _ptr(rays), rays.size, _ptr(hits), flags),
                    "stk_trace_closest_noalloc")
        return hits

    def trace_any(self, rays: np.ndarray, reference_order: bool = False) -> np.ndarray:
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(rays.size, dtype=HIT)
        self._check(_LIB.stk_trace_any(self._h, _ptr(rays), rays.size, _ptr(hits), 1 if reference_order else 0),
                    "stk_trace_any")
        return hits

    # -- device buffers, events ------------------------------------------------
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(_LIB.stk_device_alloc(self._h, nbytes, C.byref(p)), "stk_device_alloc")
        return p.value

    def free(self, ptr: int):
        self._check(_LIB.stk_device_free(self._h, C.c_void_p(ptr)), "stk_device_free")

    def upload(self, dst: int, a: np.ndarray):
        self._check(_LIB.stk_copy_to_device(self._h, C.c_void_p(dst), _ptr(a), a.nbytes), "stk_copy_to_device")

    def download(self, src: int, out: np.ndarray):
        self._check(_LIB.stk_copy_to_host(self._h, _ptr(out), C.c_void_p(src), out.nbytes), "stk_copy_to_host")
        return out

    def copy_device(self, dst: int, src: int, nbytes: int):
        self._check(_LIB.stk_device_copy(self._h, C.c_void_p(dst), C.c_void_p(src), nbytes), "stk_device_copy")

    def trace_closest_device(self, d_rays: int, n: int, d_hits: int, reference_order: bool = False,
                             binary: bool = False):
        flags = (1 if reference_order else 0) | (2 if binary else 0)
        self._check(_LIB.stk_trace_closest_device(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits), flags),
                    "stk_trace_closest_device")

    def trace_closest_image_device(self, d_rays: int, width: int, height: int, d_hits: int, binary: bool = False):
        """stk_trace_closest_image_device: an image's rays (row-major) traced
        in 8x8 tiles; the hits of trace_closest_device."""
        self._check(_LIB.stk_trace_closest_image_device(self._h, C.c_void_p(d_rays), width, height,
                                                        C.c_void_p(d_hits), 2 if binary else 0),
                    "stk_trace_closest_image_device")

    def trace_any_device(self, d_rays: int, n: int, d_hits: int, reference_order: bool = False):
        self._check(_LIB.stk_trace_any_device(self._h, C.c_void_p(d_rays), n, C.c_void_p(d_hits),
                                              1 if reference_order else 0), "stk_trace_any_device")

    def camera_rays_device(self, camera: PathTracerParam, sample: int, d_rays: int):
        self._check(_LIB.stk_generate_camera_rays_device(self._h, C.byref(camera), sample, C.c_void_p(d_rays)),
                    "stk_generate_camera_rays_device")

    def record(self, slot: int):
        self._check(_LIB.stk_event_record(self._h, slot), "stk_event_record")

    def elapsed_ms(self, a: int, b: int) -> float:
        ms = C.c_float()
        self._check(_LIB.stk_event_elapsed(self._h, a, b, C.byref(ms)), "stk_event_elapsed")
        return float(ms.value)

    def sync(self):
        self._check(_LIB.stk_sync(self._h), "stk_sync")

    # -- rendering ---------------------------------------------------------------
    def render(self, p: RenderParams, out: np.ndarray | None = None) -> np.ndarray:
        w, h = p.pt.dimension[0], p.pt.dimension[1]
        if out is None:
            out = np.zeros((h, w, 4), dtype=np.float32)
        self._check(_LIB.stk_render(self._h, C.byref(p), _ptr(out), 0), "stk_render")
        return out

    def render_image(self, p: RenderParams, rgba8: bool = False, pitch: int = 0) -> np.ndarray:
        """stk_render_image into a host array: RGBA32F or RGBA8, rows `pitch`
        bytes apart (0 = packed); returns the (h, w, 4) pixel view."""
        w, h = p.pt.dimension[0], p.pt.dimension[1]
        px = 4 if rgba8 else 16
        row = pitch or w * px
        buf = np.zeros(h * row, np.uint8)
        d = ImageDesc(buf.ctypes.data, w, h, pitch, 1 if rgba8 else 0, 0)
        self._check(_LIB.stk_render_image(self._h, C.byref(p), C.byref(d)), "stk_render_image")
        rows = buf.reshape(h, row)[:, :w * px]
        return (rows.copy().view(np.uint8) if rgba8 else rows.copy().view(np.float32)).reshape(h, w, 4)

    def host_alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(_LIB.stk_host_alloc(self._h, nbytes, C.byref(p)), "stk_host_alloc")
        return p.value

    def host_free(self, ptr: int):
        self._check(_LIB.stk_host_free(self._h, C.c_void_p(ptr)), "stk_host_free")

    def render_device(self, p: RenderParams, d_out: int):
        self._check(_LIB.stk_render(self._h, C.byref(p), C.c_void_p(d_out), 1), "stk_render")

    def set_watchdog(self, milliseconds: int):
        """stk_set_watchdog: frames running longer than this are abandoned
        (the render call raises StkError with STK_E_TIMEOUT); 0 disarms."""
        self._check(_LIB.stk_set_watchdog(self._h, milliseconds), "stk_set_watchdog")

    def abort(self):
        """stk_abort: abandon the frame in flight (or the next one)."""
        self._check(_LIB.stk_abort(self._h), "stk_abort")

    def checkpoint_save(self, path, user_frame: int = 0):
        """stk_checkpoint_save: the accumulated image, its iteration count and,
        with a cache, the cache's weights, Adam moments, averages and step."""
        self._check(_LIB.stk_checkpoint_save(self._h, os.fsencode(str(path)), C.c_uint64(user_frame)),
                    "stk_checkpoint_save")

    def checkpoint_load(self, path) -> dict:
        """stk_checkpoint_load: restores a checkpoint into this context (same
        scene; same cache configuration for the cache section)."""
        info = _abi.CheckpointInfo()
        self._check(_LIB.stk_checkpoint_load(self._h, os.fsencode(str(path)), C.byref(info)), "stk_checkpoint_load")
        return {"accumulation": bool(info.sections & 1), "cache": bool(info.sections & 2),
                "iteration": info.iteration, "pixels": info.pixels, "user_frame": info.userFrame,
                "cache_step": info.cacheStep}

    def reset_accumulation(self):
        self._check(_LIB.stk_reset_accumulation(self._h), "stk_reset_accumulation")

    def samples(self, n: int) -> np.ndarray:
        out = np.zeros(n, dtype=np.float32)
        self._check(_LIB.stk_get_samples(self._h, _ptr(out), n), "stk_get_samples")
        return out

    def accumulation(self, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), dtype=np.float32)
        self._check(_LIB.stk_get_accumulation(self._h, _ptr(out), out.size), "stk_get_accumulation")
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros(8, np.uint32)
        self._check(_LIB.stk_get_counters(self._h, _ptr(out)), "stk_get_counters")
        return out

    def ray_totals(self) -> tuple[int, int]:
        """(closest-hit rays, shadow rays) traced by the last render call."""
        out = np.zeros(2, np.uint64)
        self._check(_LIB.stk_get_ray_totals(self._h, _ptr(out)), "stk_get_ray_totals")
        return int(out[0]), int(out[1])

    # -- launch shape --------------------------------------------------------------
    def set_launch_config(self, block_size: int = 256, tile: int = 8, persistent_blocks: int = 0,
                          graph: bool = False, paths_in_flight: int = 0, binary_walk: bool = False,
                          stripe_threads: int = 0, pass_streams: int = 1, frame_blocks: int = 0,
                          frame_chunk: int = 0):
        lc = LaunchConfig(block_size, tile, persistent_blocks, 1 if graph else 0, paths_in_flight,
                          1 if binary_walk else 0, stripe_threads, pass_streams, frame_blocks, frame_chunk)
        self._check(_LIB.stk_set_launch_config(self._h, C.byref(lc)), "stk_set_launch_config")

    def launch_config(self) -> dict:
        lc = LaunchConfig()
        self._check(_LIB.stk_get_launch_config(self._h, C.byref(lc)), "stk_get_launch_config")
        return {"block_size": lc.blockSize, "tile": lc.tile, "persistent_blocks": lc.persistentBlocks,
                "graph": bool(lc.useGraph), "paths_in_flight": lc.pathsInFlight,
                "binary_walk": bool(lc.binaryWalk), "stripe_threads": lc.stripeThreads,
                "pass_streams": lc.passStreams, "frame_blocks": lc.frameBlocks,
                "frame_chunk": lc.frameChunk}

    def device_info(self) -> dict:
        d = DeviceInfo()
        self._check(_LIB.stk_device_info(self._h, C.byref(d)), "stk_device_info")
        return {"name": d.name.decode(errors="replace"), "arch": d.arch.decode(errors="replace"),
                "compute_units": d.computeUnits, "wavefront_size": d.wavefrontSize,
                "global_mem_bytes": d.globalMemBytes, "l2_bytes": d.l2Bytes, "clock_khz": d.clockKhz,
                "mem_clock_khz": d.memClockKhz, "mem_bus_width": d.memBusWidth}

    # -- batched scenes -------------------------------------------------------------
    def batch_clear(self):
        self._check(_LIB.stk_batch_clear(self._h), "stk_batch_clear")
        self._batch_views = []

    def batch_add(self, s: SceneData, camera: PathTracerParam | None = None, sky=None) -> int:
        cam = camera if camera is not None else s.camera
        b = BatchScene(_ptr(s.vb), s.vb.size, _ptr(s.ib), s.ib.size, _ptr(s.inst), s.inst.size, _ptr(s.lights),
                       s.lights.size, _ptr(s.mats), s.mats.size)
        C.memmove(C.byref(b.camera), C.byref(cam), C.sizeof(PathTracerParam))
        sk = tuple(sky) if sky is not None else tuple(s.sky)
        for i in range(4):
            b.sky[i] = float(sk[i]) if i < len(sk) else 0.0
        idx = C.c_uint32()
        self._check(_LIB.stk_batch_add_scene(self._h, C.byref(b), C.byref(idx)), "stk_batch_add_scene")
        return int(idx.value)

    def batch_commit(self, max_depth: int = 32, bins: int = 32, spatial_split_alpha: float = 0.0):
        opts = BvhOptions(max_depth, bins, 1.0, 1.0, spatial_split_alpha)
        self._check(_LIB.stk_batch_commit(self._h, C.byref(opts)), "stk_batch_commit")
        self._batch_views = []
        i = 0
        while True:
            info = BatchViewInfo()
            self._check(_LIB.stk_batch_info(self._h, i, C.byref(info)), "stk_batch_info")
            self._batch_views.append(info)
            i += 1
            if i >= info.numScenes:
                break

    def batch_views(self) -> list:
        return [dict(pixel_base=v.pixelBase, width=v.width, height=v.height, triangles=v.numTriangles,
                     depth=v.bvhDepth) for v in self._batch_views]

    def batch_pixels(self) -> int:
        return sum(v.width * v.height for v in self._batch_views)

    def batch_render(self, p: RenderParams, out: np.ndarray | None = None) -> list:
        """Renders every committed view; returns one (h, w, 4) array per view
        (views of the flat output buffer)."""
        n = self.batch_pixels()
        if out is None:
            out = np.zeros((n, 4), dtype=np.float32)
        self._check(_LIB.stk_batch_render(self._h, C.byref(p), _ptr(out), 0), "stk_batch_render")
        return [out[v.pixelBase:v.pixelBase + v.width * v.height].reshape(v.height, v.width, 4)
                for v in self._batch_views]

    def batch_render_device(self, p: RenderParams, d_out: int):
        self._check(_LIB.stk_batch_render(self._h, C.byref(p), C.c_void_p(d_out), 1), "stk_batch_render")

    # -- radiance cache ----------------------------------------------------------
    def cache_create(self, log2_table: int = 19, base_res: int = 16, max_res: int = 2048, seed: int = 1,
                     lr: float = 1e-2, beta1: float = 0.9, beta2: float = 0.99, eps: float = 1e-15,
                     max_batch: int = 1 << 18, ema: float = 0.99, lr_half_life: float = 0.0,
                     min_lr: float = 0.0):
        cfg = CacheConfig(log2_table, base_res, max_res, seed, lr, beta1, beta2, eps, max_batch, ema, lr_half_life,
                          min_lr)
        self._check(_LIB.stk_cache_create(self._h, C.byref(cfg)), "stk_cache_create")

    def cache_info(self) -> CacheInfo:
        info = CacheInfo()
        self._check(_LIB.stk_cache_info(self._h, C.byref(info)), "stk_cache_info")
        return info

    def _upload_records(self, recs: np.ndarray) -> int:
        recs = np.ascontiguousarray(recs, dtype=CACHE_RECORD)
        d = self.alloc(max(1, recs.nbytes))
        self.upload(d, recs)
        return d

    def cache_infer(self, recs: np.ndarray) -> np.ndarray:
        d = self._upload_records(recs)
        o = self.alloc(max(1, recs.size * 16))
        try:
            self._check(_LIB.stk_cache_infer(self._h, C.c_void_p(d), recs.size, C.c_void_p(o)), "stk_cache_infer")
            self.sync()
            return self.download(o, np.zeros((recs.size, 4), np.float32))
        finally:
            self.free(d)
            self.free(o)

    def cache_train_step(self, recs: np.ndarray) -> float:
        d = self._upload_records(recs)
        loss = C.c_float()
        try:
            self._check(_LIB.stk_cache_train_step(self._h, C.c_void_p(d), recs.size, C.byref(loss)),
                        "stk_cache_train_step")
        finally:
            self.free(d)
        return float(loss.value)

    def cache_gradients(self, recs: np.ndarray):
        info = self.cache_info()
        g = np.zeros(info.gridParams, np.float32)
        m = np.zeros(info.mlpParams, np.float32)
        loss = C.c_float()
        d = self._upload_records(recs)
        try:
            self._check(_LIB.stk_cache_gradients(self._h, C.c_void_p(d), recs.size, _ptr(g), _ptr(m), C.byref(loss)),
                        "stk_cache_gradients")
        finally:
            self.free(d)
        return g, m, float(loss.value)

    def cache_get_params(self):
        info = self.cache_info()
        g = np.zeros(info.gridParams, np.float32)
        m = np.zeros(info.mlpParams, np.float32)
        self._check(_LIB.stk_cache_get_params(self._h, _ptr(g), _ptr(m)), "stk_cache_get_params")
        return g, m

    def cache_last_records(self, which: str = "train") -> np.ndarray:
        """stk_cache_last_records: the last cached render's query ("query") or
        training ("train") records."""
        w = {"query": 0, "train": 1}[which]
        n = C.c_uint32()
        self._check(_LIB.stk_cache_last_records(self._h, w, None, 0, C.byref(n)), "stk_cache_last_records")
        out = np.zeros(n.value, CACHE_RECORD)
        if n.value:
            self._check(_LIB.stk_cache_last_records(self._h, w, _ptr(out), out.size, C.byref(n)),
                        "stk_cache_last_records")
        return out[:n.value]

    def cache_set_params(self, grid: np.ndarray | None, mlp: np.ndarray | None):
        grid = None if grid is None else np.ascontiguousarray(grid, np.float32)
        mlp = None if mlp is None else np.ascontiguousarray(mlp, np.float32)
        self._check(_LIB.stk_cache_set_params(self._h, _ptr(grid), _ptr(mlp)), "stk_cache_set_params")
