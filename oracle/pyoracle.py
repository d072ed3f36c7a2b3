"""ctypes binding of oracle/build/liboracle.so (TEST INFRASTRUCTURE ONLY)."""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "build" / "liboracle.so"

# record layouts of include/stk/stk_types.h (kept here so the oracle does not
# import the product package)
VERTEX = np.dtype([("position", "<f4", 3), ("tangent", "<u4"), ("normal", "<u4"), ("uv", "<u4"),
                   ("pad0", "<f4"), ("pad1", "<f4")])
BVH_NODE = np.dtype([("minBounds", "<f4", 3), ("instId", "<i4"), ("maxBounds", "<f4", 3), ("nodeOffset", "<i4")])
RAY = np.dtype([("o", "<f4", 4), ("d", "<f4", 4)])
HIT = np.dtype([("bary", "<f4", 2), ("t", "<f4"), ("instId", "<u4"), ("primId", "<u4")])

MODE_REF, MODE_PRUNED, MODE_BRUTE, MODE_ANY, MODE_WIDE = 0, 1, 2, 3, 4
MODE_TWO_LEVEL, MODE_ANY_TWO_LEVEL, MODE_ANY_BRUTE = 5, 6, 7


class OrcScene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("nnodes", C.c_uint32), ("inst", C.c_void_p), ("ninst", C.c_uint32),
                ("vb", C.c_void_p), ("nv", C.c_uint32), ("ib", C.c_void_p), ("ni", C.c_uint32),
                ("lights", C.c_void_p), ("nlights", C.c_uint32), ("mats", C.c_void_p), ("nmats", C.c_uint32),
                ("textures", C.c_void_p), ("ntextures", C.c_uint32), ("texels", C.c_void_p),
                ("wide", C.c_void_p), ("nwide", C.c_uint32), ("wideTris", C.c_void_p), ("nwideTris", C.c_uint32),
                ("tlas", C.c_void_p), ("ntlas", C.c_uint32), ("tlasLeaves", C.c_void_p), ("blas", C.c_void_p),
                ("blasTris", C.c_void_p), ("instKey", C.c_void_p), ("instW", C.c_void_p)]


class OracleStats(C.Structure):
    _fields_ = [("inner", C.c_uint64), ("leaf", C.c_uint64), ("rays", C.c_uint64), ("hits", C.c_uint64)]


def build_if_missing() -> Path:
    if not LIB_PATH.exists():
        from . import build as _b
        _b.build()
    return LIB_PATH


_lib = None
_native = None


def load_native() -> C.CDLL:
    """The -march=native timing build (compiled on first use on this host)."""
    global _native
    if _native is None:
        from . import build as _b
        path = _b.build_native()
        _native = _bind(C.CDLL(str(path)))
    return _native


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        build_if_missing()
        _lib = _bind(C.CDLL(str(LIB_PATH)))
    return _lib


def _bind(lib: C.CDLL) -> C.CDLL:
    P, U32, F = C.c_void_p, C.c_uint32, C.c_float
    FP = C.POINTER(C.c_float)
    sigs = {
        "orc_intersect_ray_box": (C.c_int, [P, FP, FP, FP, FP]),
        "orc_ray_triangle": (C.c_int, [FP, FP, FP, FP, FP, FP, FP]),
        "orc_transform_point": (None, [FP, FP, FP]),
        "orc_get_triangle": (None, [P, U32, U32, FP, FP, FP]),
        "orc_closest_hit": (C.c_int, [P, P, P, P]),
        "orc_closest_hit_pruned": (C.c_int, [P, P, P, P]),
        "orc_closest_hit_brute": (C.c_int, [P, P, P]),
        "orc_any_hit": (C.c_int, [P, P, P, P]),
        "orc_trace_batch": (None, [P, P, U32, P, C.c_int, C.c_int, P]),
        "orc_expand_two_level": (U32, [P, U32, P, P, P, P, P, P, U32]),
        "orc_jenkins": (U32, [U32]),
        "orc_init_rng": (U32, [U32, U32, U32, U32]),
        "orc_xorshift": (U32, [C.POINTER(U32)]),
        "orc_rand": (F, [C.POINTER(U32)]),
        "orc_permute": (U32, [U32, U32, U32]),
        "orc_randfloat": (F, [U32, U32]),
        "orc_camera_ray": (None, [P, U32, U32, C.POINTER(U32), U32, P]),
        "orc_pixel_jitter": (None, [P, U32, C.POINTER(U32), FP, FP]),
        "orc_offset_ray": (None, [FP, FP, FP]),
        "orc_sample_rect_light": (None, [P, FP, FP, FP, FP, FP, FP, FP, FP, FP]),
        "orc_bsdf_sample": (None, [P, FP, FP, FP, FP, P]),
        "orc_bsdf_eval": (None, [P, FP, FP, FP, FP, FP, FP, FP]),
        "orc_pt_sample": (None, [P, P, U32, U32, U32, FP]),
        "orc_pt_frame": (None, [P, P, U32, U32, U32, U32, FP, C.c_int]),
        "orc_texture_lookup": (None, [P, FP, C.c_float, C.c_float, FP]),
        "orc_bumped_normal": (None, [FP, FP, FP, FP]),
        "orc_reduce": (None, [FP, U32, U32, FP]),
        "orc_accumulate": (None, [FP, FP, U32, U32, FP]),
        "orc_tonemap": (None, [FP, U32, U32, FP]),
        "orc_fmath": (None, [U32, FP, U32, FP]),
        "orc_pt_sample_fp64": (None, [P, P, U32, U32, U32, C.POINTER(C.c_double)]),
        "orc_pt_frame_fp64": (None, [P, P, U32, U32, U32, U32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.c_int]),
        "orc_occluded": (C.c_int, [P, P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def vptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


def texture_lookup(texture: np.ndarray, texels: np.ndarray, s: float, t: float) -> np.ndarray:
    """orc_texture_lookup: rgb of one StkTexture record (16-B structured
    scalar) over RGBA32F texels at texture coordinates (s, t)."""
    rec = np.ascontiguousarray(np.asarray(texture).reshape(1))
    tx = np.ascontiguousarray(texels, dtype=np.float32)
    out = np.zeros(3, np.float32)
    load().orc_texture_lookup(rec.ctypes.data, fptr(tx), C.c_float(s), C.c_float(t), fptr(out))
    return out


def bumped_normal(normal, tangent, texel) -> np.ndarray:
    """orc_bumped_normal: the normal map's shading normal for a world normal,
    a world tangent and a map texel (rgb)."""
    n, t, x = (np.ascontiguousarray(v, dtype=np.float32)[:3] for v in (normal, tangent, texel))
    out = np.zeros(3, np.float32)
    load().orc_bumped_normal(fptr(n), fptr(t), fptr(x), fptr(out))
    return out


def fmath(op: int, x: np.ndarray) -> np.ndarray:
    """The shared binary32 transcendentals: op 0 sin, 1 cos, 2 acos, 3 x^5."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros_like(x)
    load().orc_fmath(op, fptr(x), x.size, fptr(out))
    return out


def default_threads() -> int:
    n = os.cpu_count() or 1
    return max(1, min(n, 64))


class Oracle:
    """Holds the scene arrays and the skip-link BVH the oracle walks."""

    def __init__(self, vb, ib, inst, nodes, lights=None, mats=None, wide=None, wide_tris=None, textures=None,
                 texels=None, native: bool = False):
        self.lib = load_native() if native else load()
        self.vb = np.ascontiguousarray(vb)
        self.ib = np.ascontiguousarray(ib, dtype=np.uint32)
        self.inst = np.ascontiguousarray(inst)
        self.nodes = np.ascontiguousarray(nodes)
        self.lights = None if lights is None else np.ascontiguousarray(lights)
        self.mats = None if mats is None else np.ascontiguousarray(mats)
        s = OrcScene()
        s.nodes, s.nnodes = self.nodes.ctypes.data, self.nodes.size
        s.inst, s.ninst = self.inst.ctypes.data, self.inst.size
        s.vb, s.nv = self.vb.ctypes.data, self.vb.size
        s.ib, s.ni = self.ib.ctypes.data, self.ib.size
        if self.lights is not None and self.lights.size:
            s.lights, s.nlights = self.lights.ctypes.data, self.lights.size
        if self.mats is not None and self.mats.size:
            s.mats, s.nmats = self.mats.ctypes.data, self.mats.size
        # material textures (StkTexture records, RGBA32F texels), kept alive here
        self.textures = None if textures is None else np.ascontiguousarray(textures)
        self.texels = None if texels is None else np.ascontiguousarray(texels, dtype=np.float32)
        if self.textures is not None and self.textures.size:
            assert self.textures.dtype.itemsize == 16
            s.textures, s.ntextures, s.texels = self.textures.ctypes.data, self.textures.size, self.texels.ctypes.data
        # the wide layout is an input (records of 80 B and 48 B), kept alive here
        self.wide = None if wide is None else np.ascontiguousarray(wide)
        self.wide_tris = None if wide_tris is None else np.ascontiguousarray(wide_tris)
        if self.wide is not None and self.wide.size:
            assert self.wide.dtype.itemsize == 80 and self.wide_tris.dtype.itemsize == 48
            s.wide, s.nwide = self.wide.ctypes.data, self.wide.size
            s.wideTris, s.nwideTris = self.wide_tris.ctypes.data, self.wide_tris.size
        self.scene = s

    @classmethod
    def from_scene(cls, scene, nodes, wide=None, wide_tris=None, native: bool = False):
        return cls(scene.vb, scene.ib, scene.inst, nodes, scene.lights, scene.mats, wide, wide_tris,
                   getattr(scene, "textures", None), getattr(scene, "texels", None), native)

    @classmethod
    def from_two_level(cls, scene, tl):
        """The two-level structure `tl` (arrays: top_nodes, mesh_nodes,
        instance_mesh, top_wide, top_leaves, mesh_wide, mesh_tris,
        instance_keys, instance_inverses) expanded into one skip-link tree for
        the reference walks, with the two-level arrays kept for
        MODE_TWO_LEVEL / MODE_ANY_TWO_LEVEL."""
        nodes = expand_two_level(tl, scene.inst)
        o = cls(scene.vb, scene.ib, scene.inst, nodes, scene.lights, scene.mats)
        o.tl_arrays = [np.ascontiguousarray(a) for a in (tl.top_wide, tl.top_leaves, tl.mesh_wide, tl.mesh_tris,
                                                         tl.instance_keys, tl.instance_inverses)]
        tw, tlv, mw, mt, ik, iw = o.tl_arrays
        assert tw.dtype.itemsize == 80 and tlv.dtype.itemsize == 48 and mt.dtype.itemsize == 48
        s = o.scene
        s.tlas, s.ntlas = tw.ctypes.data, tw.size
        s.tlasLeaves, s.blas, s.blasTris = tlv.ctypes.data, mw.ctypes.data, mt.ctypes.data
        s.instKey = ik.astype(np.uint32).ctypes.data if ik.dtype != np.uint32 else ik.ctypes.data
        s.instW = iw.ctypes.data
        assert iw.dtype == np.float32 and ik.dtype == np.uint32
        return o

    def trace(self, rays: np.ndarray, mode: int = MODE_REF, threads: int | None = None):
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(rays.size, dtype=HIT)
        st = OracleStats()
        self.lib.orc_trace_batch(C.byref(self.scene), vptr(rays), rays.size, vptr(hits), mode,
                                 threads or default_threads(), C.byref(st))
        return hits, st

    def get_triangle(self, inst: int, prim: int):
        v0, e0, e1 = (np.zeros(3, np.float32) for _ in range(3))
        self.lib.orc_get_triangle(C.byref(self.scene), inst, prim, fptr(v0), fptr(e0), fptr(e1))
        return v0, e0, e1

    def pt_frame(self, params, x0: int, y0: int, w: int, h: int, threads: int | None = None) -> np.ndarray:
        spp = params.pt.spp
        out = np.zeros(w * h * spp * 3, dtype=np.float32)
        self.lib.orc_pt_frame(C.byref(self.scene), C.byref(params), x0, y0, w, h, fptr(out),
                              threads or default_threads())
        return out

    def pt_frame_fp64(self, params, x0: int, y0: int, w: int, h: int, threads: int | None = None):
        """Per-pixel mean and variance of the mean (h x w x 3, float64) of the
        binary64 twin's params.pt.spp samples per pixel.  The twin models the
        materials' constant parameters only: a scene whose materials read
        textures is refused."""
        if self.mats is not None and self.mats.size and any(
                np.any(self.mats[f] != 0) for f in ("albedoTexture", "roughnessTexture", "normalTexture",
                                                   "emissionTexture")):
            raise ValueError("the binary64 twin does not model material textures")
        n = params.pt.spp
        s1 = np.zeros(w * h * 3, np.float64)
        s2 = np.zeros(w * h * 3, np.float64)
        dp = C.POINTER(C.c_double)
        self.lib.orc_pt_frame_fp64(C.byref(self.scene), C.byref(params), x0, y0, w, h, s1.ctypes.data_as(dp),
                                   s2.ctypes.data_as(dp), threads or default_threads())
        mean = s1 / n
        var = np.maximum(s2 / n - mean * mean, 0.0) * n / max(n - 1, 1)
        return mean.reshape(h, w, 3), (var / n).reshape(h, w, 3)

    def pt_sample(self, params, px: int, py: int, s: int) -> np.ndarray:
        out = np.zeros(3, dtype=np.float32)
        self.lib.orc_pt_sample(C.byref(self.scene), C.byref(params), px, py, s, fptr(out))
        return out


def expand_two_level(tl, inst) -> np.ndarray:
    """orc_expand_two_level on a two-level structure's binary trees."""
    lib = load()
    top = np.ascontiguousarray(tl.top_nodes, dtype=BVH_NODE)
    counts = np.array([m.size for m in tl.mesh_nodes], dtype=np.uint32)
    starts = np.zeros(counts.size, dtype=np.uint32)
    if counts.size:
        starts[1:] = np.cumsum(counts)[:-1]
    mesh = np.ascontiguousarray(np.concatenate(tl.mesh_nodes) if counts.size else np.zeros(0, BVH_NODE),
                                dtype=BVH_NODE)
    inst_mesh = np.ascontiguousarray(tl.instance_mesh, dtype=np.uint32)
    inst = np.ascontiguousarray(inst)
    total = 0
    for i in range(inst_mesh.size):
        if inst_mesh[i] != 0xFFFFFFFF:
            total += int(counts[inst_mesh[i]])
    cap = total + top.size
    out = np.zeros(cap, dtype=BVH_NODE)
    n = lib.orc_expand_two_level(vptr(top), top.size, vptr(mesh), vptr(starts), vptr(counts), vptr(inst_mesh),
                                 vptr(inst), vptr(out), cap)
    assert n > 0, "expansion failed"
    return out[:n]

