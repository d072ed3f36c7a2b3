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


def parse_exr(raw: bytes):
    """Independent reading of an uncompressed single-part scanline OpenEXR
    file: magic and version, the attribute list up to its empty name, the
    scanline offset table, then per scanline {y, byte count, channel rows in
    the channel list's order}."""
    import struct
    assert struct.unpack_from("<I", raw, 0)[0] == 20000630
    version = struct.unpack_from("<I", raw, 4)[0]
    assert version & 0xFF == 2 and not version & 0x200  # scanline, single part
    pos, attrs = 8, {}
    while raw[pos] != 0:
        end = raw.index(b"\0", pos)
        name = raw[pos:end].decode()
        tend = raw.index(b"\0", end + 1)
        typ = raw[end + 1:tend].decode()
        size = struct.unpack_from("<i", raw, tend + 1)[0]
        attrs[name] = (typ, raw[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    chans, cl = [], attrs["channels"][1]
    p = 0
    while cl[p] != 0:
        end = cl.index(b"\0", p)
        ptype, _lin, xs, ys = struct.unpack_from("<iB3xii", cl, end + 1)
        chans.append((cl[p:end].decode(), ptype, xs, ys))
        p = end + 1 + 16
    assert attrs["compression"] == ("compression", b"\0")
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", raw, pos)
    out = {c[0]: np.zeros((h, w), np.float32) for c in chans}
    for off in offsets:
        y, nbytes = struct.unpack_from("<ii", raw, off)
        assert nbytes == 4 * w * len(chans)
        data = np.frombuffer(raw, "<f4", w * len(chans), off + 8).reshape(len(chans), w)
        for k, c in enumerate(chans):
            out[c[0]][y - y0] = data[k]
    return attrs, chans, out


def test_exr_write(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.uniform(-1, 8, (29, 41, 4)).astype(np.float32)
    img[3, 7, :3] = [np.inf, -0.0, 1e-40]
    path = tmp_path / "frame.exr"
    sa.write_exr(path, img)
    attrs, chans, planes = parse_exr(path.read_bytes())
    assert [c[0] for c in chans] == ["B", "G", "R"] and all(c[1:] == (2, 1, 1) for c in chans)
    for k in ("displayWindow", "lineOrder", "pixelAspectRatio", "screenWindowCenter", "screenWindowWidth"):
        assert k in attrs
    for k, name in enumerate("RGB"):
        assert planes[name].tobytes() == np.ascontiguousarray(img[..., k]).tobytes()
    with pytest.raises(sa.StkError):
        sa.write_exr(tmp_path / "missing_dir" / "x.exr", img)
