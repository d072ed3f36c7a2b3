"""stk_set_scene (include/stk/stk.h) on scenes built through the scene calls
of include/stk/stk_scene.h: the same BVH and the same frames, byte for byte,
as the generated workload uploaded array by array."""
import numpy as np
import pytest

import strelka_amd as sa
from test_scene_api import rebuild

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("config", [6, 7])
def test_scene_built_through_the_api_renders_bit_identical(config):
    print(f"case: stk_set_scene on config {config} rebuilt through the scene API", flush=True)
    gen = sa.SceneData.generate(config)
    built = rebuild(gen)
    frames = []
    nodes = []
    for load in ("arrays", "scene"):
        ctx = sa.Context(0)
        if load == "arrays":
            ctx.set_scene(gen)
            ctx.build_bvh()
        else:
            ctx.load_scene(built)
        nodes.append(ctx.get_bvh())
        for f in range(3):
            img = ctx.render(sa.render_params(gen.camera, spp=2, max_depth=8, frame=f, sampling=1, sky=gen.sky))
        frames.append(img)
        ctx.close()
    assert nodes[0].tobytes() == nodes[1].tobytes()
    assert frames[0].tobytes() == frames[1].tobytes()
    assert np.isfinite(frames[0]).all() and frames[0][..., :3].max() > 0


def test_generated_scene_loads_through_stk_set_scene():
    print("case: stk_set_scene on generated config 9", flush=True)
    gen = sa.SceneData.generate(9)
    ctx = sa.Context(0)
    ctx.load_scene(gen)
    info = ctx.bvh_info()
    assert info.numNodes == 2 * (gen.ib.size // 3) - 1
    ctx.close()


def test_scene_with_unknown_material_is_refused():
    print("case: stk_set_scene refuses a material id past the table", flush=True)
    s = sa.Scene()
    vb = np.zeros(3, dtype=sa.VERTEX)
    vb["position"] = [(0, 0, 0), (1, 0, 0), (0, 1, 0)]
    mesh = s.create_mesh(vb, [0, 1, 2])
    s.create_instance(mesh, 3, np.eye(4))   # no material 3
    ctx = sa.Context(0)
    with pytest.raises(sa.StkError):
        ctx.load_scene(s)
    ctx.close()
