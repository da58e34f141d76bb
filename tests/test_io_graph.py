"""TIFF frame I/O and the JSON graph loader (CPU): round trips, byte orders,
sample formats, strips and tiles, and the graph checks that run before any
device is touched (json-sample.json's read -> ir task -> write chain)."""
import json
import os
import struct

import numpy as np
import pytest

import uir


def test_tiff_roundtrip_and_append(tmp_path):
    p = str(tmp_path / "a.tif")
    rng = np.random.default_rng(1)
    pages = [rng.standard_normal((7, 13)).astype(np.float32), rng.standard_normal((5, 4)).astype(np.float32)]
    uir.tiff_write(p, pages[0])
    uir.tiff_write(p, pages[1], append=True)
    assert uir.tiff_info(p) == (2, 13, 7)
    assert np.array_equal(uir.tiff_read(p, 0), pages[0])
    assert np.array_equal(uir.tiff_read(p, 1), pages[1])
    with pytest.raises(uir.UirError):
        uir.tiff_read(p, 2)


def _tiff(byteorder, w, h, fmt, bits, data_bytes, tiles=None, rows_per_strip=None, strips=None):
    """A hand-built one-page TIFF (independent of the library's writer)."""
    e = "<" if byteorder == "II" else ">"
    entries = []
    blocks = tiles if tiles else (strips if strips else [data_bytes])
    data_off = 8
    offs = []
    blob = b""
    for blk in blocks:
        offs.append(data_off + len(blob))
        blob += blk
    ifd_off = data_off + len(blob)
    ifd_off += ifd_off & 1
    extra_off = ifd_off + 2 + 12 * 16 + 4
    extra = b""

    def ent(tag, typ, vals):
        nonlocal extra
        fmtc = "H" if typ == 3 else "I"
        if len(vals) * (2 if typ == 3 else 4) <= 4:
            v = struct.pack(e + fmtc * len(vals), *vals).ljust(4, b"\0")
        else:
            v = struct.pack(e + "I", extra_off + len(extra))
            extra += struct.pack(e + fmtc * len(vals), *vals)
        entries.append(struct.pack(e + "HHI", tag, typ, len(vals)) + v)

    ent(256, 4, [w])
    ent(257, 4, [h])
    ent(258, 3, [bits])
    ent(259, 3, [1])
    ent(262, 3, [1])
    if tiles:
        ent(322, 3, [tiles_w])
        ent(323, 3, [tiles_h])
        ent(324, 4, offs)
        ent(325, 4, [len(b) for b in blocks])
    else:
        ent(273, 4, offs)
        ent(278, 4, [h if rows_per_strip is None else rows_per_strip])
        ent(279, 4, [len(b) for b in blocks])
    ent(277, 3, [1])
    ent(339, 3, [fmt])
    entries.sort(key=lambda x: struct.unpack(e + "H", x[:2])[0])
    ifd = struct.pack(e + "H", len(entries)) + b"".join(entries) + struct.pack(e + "I", 0)
    ifd = ifd.ljust(2 + 12 * 16 + 4, b"\0")
    head = (b"II" if byteorder == "II" else b"MM") + struct.pack(e + "HI", 42, ifd_off)
    body = blob + (b"\0" if (data_off + len(blob)) & 1 else b"")
    return head + body + ifd + extra


tiles_w = tiles_h = 16


@pytest.mark.parametrize("order", ["II", "MM"])
@pytest.mark.parametrize("dtype,fmt", [(np.uint8, 1), (np.uint16, 1), (np.int16, 2), (np.float32, 3),
                                        (np.float64, 3), (np.int32, 2)])
def test_tiff_reads_sample_formats_and_byte_orders(tmp_path, order, dtype, fmt):
    rng = np.random.default_rng(2)
    img = (rng.random((6, 9)) * 100).astype(dtype)
    raw = img.astype(img.dtype.newbyteorder("<" if order == "II" else ">")).tobytes()
    p = tmp_path / "x.tif"
    p.write_bytes(_tiff(order, 9, 6, fmt, img.dtype.itemsize * 8, raw))
    got = uir.tiff_read(str(p))
    assert np.array_equal(got, img.astype(np.float32))


def test_tiff_reads_tiles_and_strips(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.standard_normal((20, 35)).astype(np.float32)
    tiles = []
    for ty in range(0, 20, 16):
        for tx in range(0, 35, 16):
            t = np.zeros((16, 16), np.float32)
            part = img[ty:ty + 16, tx:tx + 16]
            t[:part.shape[0], :part.shape[1]] = part
            tiles.append(t.astype("<f4").tobytes())
    p = tmp_path / "t.tif"
    p.write_bytes(_tiff("II", 35, 20, 3, 32, None, tiles=tiles))
    assert np.array_equal(uir.tiff_read(str(p)), img)
    # seven strips of 3 rows (the last one short)
    s = tmp_path / "s.tif"
    strips = [img[y:y + 3].astype(">f4").tobytes() for y in range(0, 20, 3)]
    s.write_bytes(_tiff("MM", 35, 20, 3, 32, None, rows_per_strip=3, strips=strips))
    assert np.array_equal(uir.tiff_read(str(s)), img)
    # a strip table shorter than the image needs
    bad = tmp_path / "b.tif"
    bad.write_bytes(_tiff("II", 35, 20, 3, 32, img.astype("<f4").tobytes(), rows_per_strip=3))
    with pytest.raises(uir.UirError):
        uir.tiff_read(str(bad))


def test_tiff_strip_rows_edge_values(tmp_path):
    img = np.arange(35 * 20, dtype=np.float32).reshape(20, 35)
    # RowsPerStrip 0 is invalid (TIFF 6.0: at least 1): an error, not a crash
    z = tmp_path / "z.tif"
    z.write_bytes(_tiff("II", 35, 20, 3, 32, img.astype("<f4").tobytes(), rows_per_strip=0))
    with pytest.raises(uir.UirError) as e:
        uir.tiff_read(str(z))
    assert "RowsPerStrip" in str(e.value)
    # a huge RowsPerStrip means one strip (no 32-bit overflow of the row count)
    h = tmp_path / "h.tif"
    h.write_bytes(_tiff("II", 35, 20, 3, 32, img.astype("<f4").tobytes(), rows_per_strip=0xFFFFFFF0))
    assert np.array_equal(uir.tiff_read(str(h)), img)


def test_tiff_rejects_garbage(tmp_path):
    p = tmp_path / "bad.tif"
    p.write_bytes(b"not a tiff at all")
    with pytest.raises(uir.UirError):
        uir.tiff_info(str(p))
    with pytest.raises(uir.UirError):
        uir.tiff_info(str(tmp_path / "missing.tif"))


SAMPLE = {
    "nodes": [
        {"plugin": "read", "name": "read", "properties": {"path": "sino.tif"}},
        {"plugin": "write", "name": "write", "properties": {"filename": "out/out%03i.tif"}},
        {"plugin": "asdpocs", "name": "asdpocs", "package": "ir", "properties": {
            "projector": {"plugin": "parallel-projector", "name": "parallel-projector", "package": "ir",
                          "properties": {"model": "joseph", "is_forward": False, "step": 0.034906585}},
            "df_minimizer": {"plugin": "sart", "package": "ir", "name": "sart_minimizer",
                             "properties": {"relaxation_factor": 0.25, "num-iterations": 2}},
            "num-iterations": 2}},
    ],
    "edges": [{"from": {"name": "read"}, "to": {"name": "asdpocs"}},
              {"from": {"name": "asdpocs"}, "to": {"name": "write"}}],
}


@pytest.mark.parametrize("mutate,status", [
    (lambda g: g.update(nodes="x"), 1),
    (lambda g: g["edges"].append({"from": {"name": "read"}, "to": {"name": "write"}}), 4),   # not a chain
    (lambda g: g["edges"].pop(), 4),                                                         # no write
    (lambda g: g["nodes"][2].update(plugin="fbp"), 6),                                       # unknown plugin
    (lambda g: g["nodes"][2].update(package="filters"), 6),
    (lambda g: g["nodes"][2]["properties"].update({"num-iterations": -3}), 1),
    (lambda g: g["nodes"][2]["properties"]["df_minimizer"]["properties"].update(relaxation_factor=3.0), 1),
    (lambda g: g["nodes"][1]["properties"].update(filename="out/%s.tif"), 1),                 # bad format
    (lambda g: g["nodes"][0]["properties"].update(path="/nonexistent/*.tif"), 6),
])
def test_graph_errors_before_the_device(tmp_path, mutate, status):
    g = json.loads(json.dumps(SAMPLE))
    uir.tiff_write(str(tmp_path / "sino.tif"), np.zeros((4, 5), np.float32))
    mutate(g)
    with pytest.raises(uir.UirError) as e:
        uir.json_run(g, base_dir=str(tmp_path))
    assert e.value.status == status, str(e.value)


def test_graph_rejects_malformed_json():
    for text in ('{"nodes": [}', '{"nodes": [], "edges": []} x', '{"a": "\\q"}', '[1, 2', '{"a": tru}'):
        with pytest.raises(uir.UirError) as e:
            uir.json_run(text)
        assert e.value.status == 1
        assert "JSON" in str(e.value)
