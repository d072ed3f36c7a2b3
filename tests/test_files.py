"""On-disk formats: skip-link BVH files (round trip into the packed and wide
layouts stk_set_bvh derives) and PFM images (checked against an independent
reading of the format: header, little-endian float32 RGB, bottom row first)."""
import numpy as np
import pytest

import strelka_amd as sa


def test_bvh_file_round_trip(tmp_path):
    s = sa.SceneData.generate(1)
    b = sa.HostBvh.build(s)
    path = tmp_path / "scene.stkbvh"
    sa.save_bvh(path, b.nodes)
    raw = path.read_bytes()
    assert raw[:8] == b"STKBVH01"
    assert int.from_bytes(raw[8:12], "little") == b.nodes.size and len(raw) == 16 + 32 * b.nodes.size
    nodes = sa.load_bvh(path)
    assert nodes.tobytes() == b.nodes.tobytes()
    again = sa.HostBvh.from_nodes(s, nodes)
    assert again.wide.tobytes() == b.wide.tobytes()
    assert again.packed.tobytes() == b.packed.tobytes()


def test_bvh_file_rejects_bad_input(tmp_path):
    bad = tmp_path / "bad.stkbvh"
    bad.write_bytes(b"NOTABVH!" + bytes(8))
    with pytest.raises(sa.StkError):
        sa.load_bvh(bad)
    s = sa.SceneData.generate(1)
    good = tmp_path / "good.stkbvh"
    sa.save_bvh(good, sa.HostBvh.build(s).nodes)
    truncated = tmp_path / "short.stkbvh"
    truncated.write_bytes(good.read_bytes()[:-5])
    with pytest.raises(sa.StkError):
        sa.load_bvh(truncated)
    with pytest.raises(sa.StkError):
        sa.load_bvh(tmp_path / "missing.stkbvh")


def parse_pfm(raw: bytes) -> np.ndarray:
    """Independent PFM reader: three whitespace-separated header tokens after
    'PF', one separator byte, then rows bottom to top."""
    parts = raw.split(b"\n", 3)
    assert parts[0] == b"PF"
    w, h = (int(v) for v in parts[1].split())
    scale = float(parts[2])
    assert scale < 0  # little-endian
    data = np.frombuffer(parts[3], dtype="<f4").reshape(h, w, 3)
    return data[::-1]


def test_pfm_write_read(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.uniform(-2, 5, (37, 53, 4)).astype(np.float32)
    img[0, 0, :3] = [np.inf, -0.0, 1e-30]
    path = tmp_path / "frame.pfm"
    sa.write_pfm(path, img)
    raw = path.read_bytes()
    assert raw.startswith(b"PF\n53 37\n-1\n")
    ref = parse_pfm(raw)
    assert ref.tobytes() == np.ascontiguousarray(img[..., :3]).tobytes()
    back = sa.read_pfm(path)
    assert back[..., :3].tobytes() == img[..., :3].tobytes()
    assert np.all(back[..., 3] == 1.0)


def test_rmse_tool():
    a = np.zeros((4, 4, 4), np.float32)
    b = np.full((4, 4, 4), 0.5, np.float32)
    assert sa.rmse(a, b) == pytest.approx(0.5)
    b[0, 0, :3] = 7.0  # clamped to 1 in the display range
    expected = np.sqrt((15 * 3 * 0.25 + 3 * 1.0) / 48)
    assert sa.rmse(a, b) == pytest.approx(expected)
    assert sa.rmse(a, b, tonemapped=False) > expected
