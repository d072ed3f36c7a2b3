"""The C-ABI library loads, exports every declared entry point and keeps the
record layouts of include/stk/stk_types.h (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

import strelka_amd as sa
from strelka_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include" / "stk").glob("*.h"):
        text = h.read_text()
        for m in re.finditer(r"STK_API\s+[\w\s\*]+?\b(stk_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported_and_bound():
    names = declared_symbols()
    assert len(names) >= 40
    lib = C.CDLL(str(_abi.LIB_PATH))
    bound = {n for n, _, _ in _abi.PROTOTYPES}
    for n in sorted(names):
        assert hasattr(lib, n), f"{n} declared in include/stk but not exported"
        assert n in bound, f"{n} has no ctypes prototype"


def test_record_sizes_match_reference_layouts():
    assert sa.VERTEX.itemsize == 32          # shaders/raytracing.h:17-25
    assert sa.BVH_NODE.itemsize == 32        # shaders/raytracing.h:27-33
    assert sa.INSTANCE.itemsize == 208       # shaders/instanceconstants.h:19-28
    assert sa.LIGHT.itemsize == 112          # shaders/lights.h:10-19
    assert sa.RAY.itemsize == 32             # shaders/ray.h:12-16
    assert sa.HIT.itemsize == 20             # shaders/raytracing.h:90-96
    assert C.sizeof(sa.PathTracerParam) == 336  # shaders/pathtracerparam.h:20-41
    assert sa.BVH_NODE.fields["nodeOffset"][1] == 28
    assert sa.INSTANCE.fields["lightId"][1] == 204
    assert sa.LIGHT.fields["type"][1] == 96


def test_version_and_device_probe():
    assert b"gfx950" in sa.lib().stk_version()
    h = C.c_void_p()
    st = sa.lib().stk_create(0, C.byref(h))
    # STK_OK on a GPU box, STK_E_DEVICE where no device is present
    assert st in (0, 3)
    if st == 0:
        sa.lib().stk_destroy(h)


def test_null_arguments_are_rejected_without_a_device():
    lib = sa.lib()
    assert lib.stk_create(0, None) == 1
    assert lib.stk_sync(None) == 1
    assert lib.stk_set_geometry(None, None, 0, None, 0, None, 0) == 1
    assert lib.stk_trace_closest(None, None, 0, None, 0) == 1
    assert lib.stk_trace_any(None, None, 0, None, 0) == 1
    assert lib.stk_last_error(None) == b"null context"
    assert lib.stk_scene_generate(99, 0, 0, C.byref(C.c_void_p())) == 1
    assert lib.stk_bvh_build_host(None, 0, None, 0, None, 0, None, None) == 1
    assert lib.stk_checkpoint_save(None, b"/tmp/x.stkckpt", 0) == 1
    assert lib.stk_checkpoint_load(None, b"/tmp/x.stkckpt", None) == 1
    assert lib.stk_set_scene(None, None, None) == 1


def test_scene_generators_are_deterministic():
    a = sa.SceneData.generate(1)
    b = sa.SceneData.generate(1)
    assert a.vb.tobytes() == b.vb.tobytes() and a.ib.tobytes() == b.ib.tobytes()
    # config 1: 1,280 sphere triangles + 12 box triangles
    assert int(a.inst["indexCount"].sum()) // 3 == 1292
    c2 = sa.SceneData.generate(2)
    assert int(c2.inst["indexCount"].sum()) // 3 == 524288 + 2000 * 320 + 2
    assert c2.camera.dimension[0] == 1920 and c2.camera.dimension[1] == 1080
    r1 = a.jittered_rays(7)
    r2 = a.jittered_rays(7)
    assert r1.tobytes() == r2.tobytes()
    d = r1["d"][:, :3]
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)


def test_reference_icosphere_faces():
    fixed = sa.SceneData.generate(1)
    ref = sa.SceneData.generate(1, ref_icosphere=True)
    # same triangle count (1,280) but different connectivity: the reference's
    # last column emits (10,2,1) and (10,11,3) (src/scene/scene.cpp:275-304)
    assert fixed.inst[0]["indexCount"] == ref.inst[0]["indexCount"] == 3840
    assert fixed.ib.tobytes() != ref.ib.tobytes()


HEADER_STRUCTS = {
    "StkPathTracerParam": "PathTracerParam", "StkRenderParams": "RenderParams", "StkCacheConfig": "CacheConfig",
    "StkCacheInfo": "CacheInfo", "StkBvhOptions": "BvhOptions", "StkBvhHeader": "BvhHeader",
    "StkLaunchConfig": "LaunchConfig", "StkDeviceInfo": "DeviceInfo", "StkBatchScene": "BatchScene",
    "StkBatchViewInfo": "BatchViewInfo", "StkImageDesc": "ImageDesc", "StkBvhInfo": "BvhInfo",
    "StkLightDesc": "LightDesc", "StkSceneArrays": "SceneArrays", "StkCheckpointInfo": "CheckpointInfo",
}


def test_ctypes_records_match_the_header(tmp_path):
    """Every ctypes record has the size and field offsets the C compiler gives
    the header's struct, so no entry point writes past a caller's record."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "stk/stk.h"', '#include "stk/stk_scene.h"',
             "int main(void) {"]
    for c_name, py_name in HEADER_STRUCTS.items():
        lines.append(f'printf("{c_name} %zu\\n", sizeof({c_name}));')
        for f, *_ in getattr(_abi, py_name)._fields_:
            lines.append(f'printf("{c_name}.{f} %zu\\n", offsetof({c_name}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    import subprocess
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    for c_name, py_name in HEADER_STRUCTS.items():
        cls = getattr(_abi, py_name)
        assert int(got[c_name]) == C.sizeof(cls), c_name
        for f, *_ in cls._fields_:
            assert int(got[f"{c_name}.{f}"]) == getattr(cls, f).offset, (c_name, f)


def test_oracle_scene_record_matches_its_header(tmp_path):
    """oracle/pyoracle.OrcScene has the size and field offsets of oracle.h's
    OrcScene (the oracle reads the scene through it)."""
    import subprocess

    from oracle import pyoracle
    fields = [f for f, *_ in pyoracle.OrcScene._fields_]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "oracle.h"', "int main(void) {",
             'printf("size %zu\\n", sizeof(OrcScene));']
    lines += [f'printf("{f} %zu\\n", offsetof(OrcScene, {f}));' for f in fields]
    lines.append("return 0; }")
    src = tmp_path / "orc.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "orc"
    subprocess.run(["gcc", "-I", str(ROOT / "oracle"), "-I", str(ROOT / "include"), str(src), "-o", str(exe)],
                   check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                       check=True).stdout.splitlines())
    assert int(got["size"]) == C.sizeof(pyoracle.OrcScene)
    for f in fields:
        assert int(got[f]) == getattr(pyoracle.OrcScene, f).offset, f
