"""Scene building through the C ABI (stk_scene_create_mesh / _instance /
_light / _add_material / _add_camera, include/stk/stk_scene.h), the calls of
the reference's oka::Scene (include/scene/scene.h:231-389,
src/scene/scene.cpp:20-91, 418-511).

A generated workload rebuilt through these calls from nothing but its own
arrays (and its light's size) must give byte-identical arrays: the same global
vertex and index buffers, the same instance records (transforms, inverses and
normal matrices the library derives), light records, materials and camera.
The GPU half (tests/test_gpu_scene_api.py) renders such a scene and the
generated one and requires identical frames."""
import ctypes as C

import numpy as np
import pytest

import strelka_amd as sa

# generated workloads whose rect lights are 1 x 1 (the light instance's
# transform is then the light frame itself)
REBUILDABLE = [1, 6, 7, 8, 9]


def rebuild(gen: sa.SceneData) -> sa.Scene:
    """The workload's content re-entered through the scene-building calls."""
    s = sa.Scene()
    for i, m in enumerate(gen.mats):
        assert s.add_material(m) == i
    meshes = {}
    for rec in gen.inst:
        if rec["lightId"] >= 0:
            continue
        key = (int(rec["indexOffset"]), int(rec["indexCount"]))
        if key not in meshes:
            idx = gen.ib[key[0]:key[0] + key[1]]
            lo, hi = int(idx.min()), int(idx.max())
            meshes[key] = s.create_mesh(gen.vb[lo:hi + 1], idx - lo)
        xf = np.asarray(rec["objectToWorld"], dtype=np.float32).reshape(16)
        s.create_instance(meshes[key], int(rec["materialId"]), xf)
    for li, rec in enumerate(gen.inst[gen.inst["lightId"] >= 0]):
        light = gen.lights[rec["lightId"]]
        assert light["type"] == 0
        xf = np.asarray(rec["objectToWorld"], dtype=np.float32).reshape(16)
        d = sa.Scene.light_desc(0, xf, color=(1.0, 1.0, 1.0), intensity=float(light["color"][0]), width=1.0,
                                height=1.0)
        assert s.create_light(d) == li
    assert s.add_camera(gen.camera) == 0
    s.set_sky(*gen.sky[:3])
    return s


def same_bytes(a: np.ndarray, b: np.ndarray) -> bool:
    return a.dtype == b.dtype and a.shape == b.shape and a.tobytes() == b.tobytes()


@pytest.mark.parametrize("config", REBUILDABLE)
def test_rebuilt_workload_has_identical_arrays(config):
    gen = sa.SceneData.generate(config)
    got = rebuild(gen).data()
    for name in ("vb", "ib", "inst", "lights", "mats"):
        assert same_bytes(getattr(got, name), getattr(gen, name)), (config, name)
    assert bytes(got.camera) == bytes(gen.camera)
    assert got.sky == gen.sky


def test_lit_floor_from_its_description():
    """The lit floor (a 40 x 40 Lambert floor of albedo 0.5 under a 1 x 1 rect
    light of intensity 10 at y = 2 facing down, a camera at (0, 1, 0) looking
    down) written from its description, not from generated arrays."""
    gen = sa.SceneData.generate(9)
    s = sa.Scene()
    m = np.zeros(1, dtype=sa.MATERIAL)
    m["type"] = sa.MAT_LAMBERT
    m["roughness"] = 1.0
    m["albedo"] = (0.5, 0.5, 0.5, 1.0)
    mat = s.add_material(m)
    vb = np.zeros(4, dtype=sa.VERTEX)
    vb["position"][:, 0] = (-20.0, 20.0, 20.0, -20.0)
    vb["position"][:, 2] = (-20.0, -20.0, 20.0, 20.0)
    vb["normal"] = sa.pack_normal(0.0, 1.0, 0.0)
    vb["tangent"] = sa.pack_normal(1.0, 0.0, 0.0)
    mesh = s.create_mesh(vb, [0, 2, 1, 0, 3, 2])
    s.create_instance(mesh, mat, np.eye(4, dtype=np.float32))
    # translate(0, 2, 0) * rotateX(-pi/2): local -z (the light's emitting side) maps to -y
    c, sn = np.float32(np.cos(np.float32(-np.pi / 2))), np.float32(np.sin(np.float32(-np.pi / 2)))
    xf = np.array([[1, 0, 0, 0], [0, c, -sn, 2], [0, sn, c, 0], [0, 0, 0, 1]], dtype=np.float32)
    s.create_light(sa.Scene.light_desc(0, xf, intensity=10.0, width=1.0, height=1.0))
    s.add_camera(sa.look_at((0, 1, 0), (0, 0, 0), (0, 0, -1), 60.0, 64, 64, 0.01, 100.0))
    got = s.data()
    for name in ("vb", "ib", "mats"):
        assert same_bytes(getattr(got, name), getattr(gen, name)), name
    assert bytes(got.camera) == bytes(gen.camera)
    # matrices through numpy's float32 cos/sin: equal to the generator's to the last bit or one ulp
    np.testing.assert_allclose(got.inst["objectToWorld"], gen.inst["objectToWorld"], rtol=0, atol=1e-7)
    np.testing.assert_allclose(got.lights["points"], gen.lights["points"], rtol=0, atol=1e-6)
    assert np.array_equal(got.lights["color"], gen.lights["color"])


def test_updates_change_the_records():
    gen = sa.SceneData.generate(9)
    s = rebuild(gen)
    li = s.light_instance(0)
    assert li == 1
    # move the light up by one and double its size: corners and instance transform follow
    xf = np.asarray(gen.inst[li]["objectToWorld"], dtype=np.float32).reshape(16).copy()
    xf[13] += 1.0
    s.update_light(0, sa.Scene.light_desc(0, xf, intensity=40.0, width=2.0, height=2.0))
    d = s.data()
    pts = d.lights["points"][0]
    assert np.allclose(pts[:, 1], 3.0)
    assert np.isclose(np.ptp(pts[:, 0]), 2.0) and np.isclose(np.ptp(pts[:, 2]), 2.0)
    assert np.allclose(d.lights["color"][0][:3], 10.0)   # 40 / (2 * 2)
    assert np.isclose(d.inst[li]["objectToWorld"][13], 3.0)
    # an instance transform update reaches the inverse as well
    t = np.eye(4, dtype=np.float32)
    t[0, 3] = 5.0
    s.update_instance_transform(0, t)
    d = s.data()
    assert np.isclose(d.inst[0]["objectToWorld"][12], 5.0)
    assert np.isclose(d.inst[0]["worldToObject"][12], -5.0)
    # a light cannot change type
    with pytest.raises(sa.StkError):
        s.update_light(0, sa.Scene.light_desc(1, xf, radius=1.0))


def test_cameras():
    s = sa.Scene()
    assert s.camera_count() == 0
    a = sa.look_at((0, 0, 5), (0, 0, 0), (0, 1, 0), 45.0, 32, 16)
    b = sa.look_at((1, 2, 3), (0, 0, 0), (0, 1, 0), 30.0, 8, 8)
    assert s.add_camera(a) == 0 and s.add_camera(b) == 1
    assert bytes(s.camera(1)) == bytes(b)
    s.update_camera(0, b)
    assert bytes(s.camera(0)) == bytes(b)
    assert bytes(s.data().camera) == bytes(b)   # camera 0 is the one the arrays report
    with pytest.raises(sa.StkError):
        s.camera(2)
    with pytest.raises(sa.StkError):
        sa.look_at((0, 0, 0), (0, 0, 1), (0, 1, 0), 0.0, 8, 8)


def test_invalid_calls_are_refused():
    s = sa.Scene()
    vb = np.zeros(3, dtype=sa.VERTEX)
    with pytest.raises(sa.StkError):
        s.create_mesh(vb, [0, 1, 3])                 # index past the vertices
    with pytest.raises(sa.StkError):
        s.create_mesh(vb, [0, 1])                    # not whole triangles
    mesh = s.create_mesh(vb, [0, 1, 2])
    with pytest.raises(sa.StkError):
        s.create_instance(mesh + 1, -1, np.eye(4))   # unknown mesh
    with pytest.raises(sa.StkError):
        s.create_instance(mesh, -2, np.eye(4))       # material below -1
    bad = np.eye(4, dtype=np.float32)
    bad[0, 0] = np.nan
    with pytest.raises(sa.StkError):
        s.create_instance(mesh, -1, bad)
    with pytest.raises(sa.StkError):
        s.create_light(sa.Scene.light_desc(3))       # unknown light type
    with pytest.raises(sa.StkError):
        s.create_light(sa.Scene.light_desc(0, width=0.0))
    with pytest.raises(sa.StkError):
        s.update_instance_transform(7, np.eye(4))
    with pytest.raises(sa.StkError):
        s.set_sky(float("inf"), 0, 0)
    # nothing refused changed the container
    d = s.data()
    assert d.vb.size == 3 and d.ib.size == 3 and d.inst.size == 0 and d.lights.size == 0


def test_packing_matches_the_vertex_records():
    gen = sa.SceneData.generate(9)
    assert int(gen.vb["normal"][0]) == sa.pack_normal(0.0, 1.0, 0.0)
    assert sa.pack_normal(-1.0, -1.0, -1.0) == 0
    assert sa.pack_uv(-10.0, -10.0) == 0
    # the reference's scale literals 511.99999f and 16383.99999f round to 512 and 16384 in binary32, so the
    # top of each range packs to 512 (of a 10-bit field) and 16384 (of a 16-bit field)
    assert sa.pack_normal(1.0, 1.0, 1.0) == 512 | (512 << 10) | (512 << 20)
    assert sa.pack_uv(10.0, 10.0) == (16384 | (16384 << 16))
    lib = sa.lib()
    assert lib.stk_scene_camera_count(None) == 0
    assert lib.stk_scene_create(None) != 0
    assert lib.stk_scene_create_mesh(None, None, 0, None, 0, C.byref(C.c_uint32())) != 0


def test_add_material_refuses_unknown_records():
    """The container refuses a material of an unknown type or a negative
    index of refraction, and keeps the specular and emissive ones."""
    s = sa.Scene()
    m = np.zeros(1, dtype=sa.MATERIAL)
    for bad in (dict(type=9), dict(type=sa.MAT_GLASS, ior=-0.5)):
        r = m.copy()
        for k, v in bad.items():
            r[k] = v
        with pytest.raises(sa.StkError):
            s.add_material(r)
    for kind in (sa.MAT_MIRROR, sa.MAT_GLASS):
        r = m.copy()
        r["type"] = kind
        r["ior"] = 1.4
        r["emission"] = (1.0, 2.0, 3.0)
        assert s.add_material(r) == kind - sa.MAT_MIRROR
    got = s.data().mats
    assert got["type"].tolist() == [sa.MAT_MIRROR, sa.MAT_GLASS] and got["emission"][1].tolist() == [1.0, 2.0, 3.0]
