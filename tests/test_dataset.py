"""The dataset harness on the CPU: PNG and EXR files, both folder layouts of
the SiSter dataset (raw capture folders as python/extract_dataset.py:9-96
reads them, and the exported layout of README.md:40-61), the re-layout, and
depth -> disparity.  The dataset itself is not available offline: every
fixture here is synthetic, written in the dataset's layout.  EXR decoding
against the OpenEXR library is parity unpinned (no OpenEXR here): the reader
is checked on files assembled byte by byte from the published file layout."""
import struct
import zlib

import numpy as np
import pytest

from sister_amd import dataset, imageio
from tests import pngutil


def test_png_reads_every_row_filter(tmp_path):
    rng = np.random.default_rng(1)
    g = rng.integers(0, 256, size=(9, 13)).astype(np.uint8)
    c = rng.integers(0, 256, size=(7, 11, 3)).astype(np.uint8)
    for f in range(5):
        pngutil.write_png(str(tmp_path / "g.png"), g, (f, (f + 1) % 5, (f + 3) % 5))
        assert np.array_equal(imageio.read_png(tmp_path / "g.png"), g), f
        pngutil.write_png(str(tmp_path / "c.png"), c, (f, (f + 2) % 5))
        assert np.array_equal(imageio.read_png(tmp_path / "c.png"), c[..., ::-1]), f   # returned as BGR


def test_png_round_trip_8_and_16_bit(tmp_path):
    rng = np.random.default_rng(2)
    a16 = rng.integers(0, 65536, size=(5, 8)).astype(np.uint16)
    imageio.write_png(tmp_path / "a.png", a16)
    assert np.array_equal(imageio.read_png(tmp_path / "a.png"), a16)
    assert np.array_equal(pngutil.read_png16(str(tmp_path / "a.png")), a16)     # an independent reader
    assert np.array_equal(imageio.read_png(tmp_path / "a.png", anydepth=False), (a16 >> 8).astype(np.uint8))
    bgr = rng.integers(0, 256, size=(4, 6, 3)).astype(np.uint8)
    imageio.write_png(tmp_path / "b.png", bgr)
    assert np.array_equal(imageio.read_png(tmp_path / "b.png"), bgr)
    assert np.array_equal(pngutil.read_png_rgb8(str(tmp_path / "b.png")), bgr[..., ::-1])


def _exr_bytes(W, H, name, ptype, rows, compression=0, lines_per_chunk=1, zip_chunk=None):
    """An EXR file assembled field by field from the published layout."""
    def attr(n, t, v):
        return n.encode() + b"\0" + t.encode() + b"\0" + struct.pack("<i", len(v)) + v
    chl = name.encode() + b"\0" + struct.pack("<i", ptype) + b"\0\0\0\0" + struct.pack("<ii", 1, 1) + b"\0"
    hdr = (attr("channels", "chlist", chl) + attr("compression", "compression", bytes([compression])) +
           attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1)) +
           attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1)) +
           attr("lineOrder", "lineOrder", b"\0") + attr("pixelAspectRatio", "float", struct.pack("<f", 1)) +
           attr("screenWindowCenter", "v2f", struct.pack("<2f", 0, 0)) +
           attr("screenWindowWidth", "float", struct.pack("<f", 1)) + b"\0")
    chunks = []
    for y in range(0, H, lines_per_chunk):
        raw = b"".join(rows[y + i] for i in range(min(lines_per_chunk, H - y)))
        body = zip_chunk(raw) if zip_chunk else raw
        if len(body) >= len(raw):
            body = raw                    # the format stores a chunk raw when deflate does not shrink it
        chunks.append(struct.pack("<ii", y, len(body)) + body)
    base = 8 + len(hdr) + 8 * len(chunks)
    offs = []
    for ch in chunks:
        offs.append(base)
        base += len(ch)
    return struct.pack("<ii", 20000630, 2) + hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks)


def _zip_chunk(raw):
    # OpenEXR ZIP: even bytes then odd bytes, each byte minus its predecessor
    # plus 128 (mod 256), deflated
    b = list(raw[0::2]) + list(raw[1::2])
    out = [b[0]] + [(b[i] - b[i - 1] + 128) % 256 for i in range(1, len(b))]
    return zlib.compress(bytes(out))


def test_exr_reader_on_assembled_files(tmp_path):
    rng = np.random.default_rng(3)
    H, W = 37, 48
    # a depth map like the dataset's: a plane with a raised part (deflate
    # shrinks it, so the compressed chunks are stored compressed) ...
    plane = np.full((H, W), 0.1, np.float32)
    plane[10:25, 20:40] = 0.08
    # ... and noise (deflate does not: chunks stored raw)
    noise = rng.uniform(0.05, 0.12, size=(H, W)).astype(np.float32)
    for depth in (plane, noise):
        rows = [depth[y].astype("<f4").tobytes() for y in range(H)]
        for comp, lpc, z in ((0, 1, None), (2, 1, _zip_chunk), (3, 16, _zip_chunk)):
            p = tmp_path / f"d{comp}.exr"
            p.write_bytes(_exr_bytes(W, H, "Y", 2, rows, comp, lpc, z))
            got = imageio.read_exr(p)
            assert got.dtype == np.float32 and np.array_equal(got, depth), comp
    assert len(_zip_chunk(b"".join(plane[y].tobytes() for y in range(16)))) < 16 * W * 4
    half = depth.astype(np.float16)
    p = tmp_path / "h.exr"
    p.write_bytes(_exr_bytes(W, H, "Z", 1, [half[y].astype("<f2").tobytes() for y in range(H)]))
    assert np.array_equal(imageio.read_exr(p), half.astype(np.float32))
    p.write_bytes(_exr_bytes(W, H, "Y", 2, rows, compression=4))    # PIZ: refused
    with pytest.raises(ValueError):
        imageio.read_exr(p)


def test_exr_writer_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    d = rng.uniform(0.01, 1.0, size=(40, 23)).astype(np.float32)
    for comp in ("NONE", "ZIPS", "ZIP"):
        imageio.write_exr(tmp_path / "w.exr", {"Y": d}, comp)
        assert np.array_equal(imageio.read_exr(tmp_path / "w.exr"), d), comp
    rgb = {k: rng.uniform(0, 1, size=(6, 5)).astype(np.float32) for k in "RGB"}
    imageio.write_exr(tmp_path / "c.exr", rgb)
    got = imageio.read_exr(tmp_path / "c.exr")
    assert np.array_equal(got, np.stack([rgb["B"], rgb["G"], rgb["R"]], axis=2))


def _raw_layout(tmp_path, rng):
    """Raw capture folders: objects with underscores in their names, three
    levels, two baselines, a prediction, and ground-truth files."""
    scenes, gts = tmp_path / "scenes", tmp_path / "gt"
    gts.mkdir()
    views = {}
    for obj in ("hexa_nut", "washer"):
        for level in ("0", "2"):
            for base in ("010", "025"):
                f = scenes / obj / f"rig_{level}_{base}"
                (f / "output").mkdir(parents=True)
                for v in dataset.VIEWS:
                    img = rng.integers(0, 256, size=(8, 12, 3)).astype(np.uint8)
                    imageio.write_png(f / f"00000_{v}.png", img)
                    views[(obj, level, base, v)] = img
                imageio.write_png(f / "output" / "FULL_mccnn_refine.png", np.full((8, 12), 513, np.uint16))
            imageio.write_exr(gts / f"{obj}_{level}.exr", {"Y": np.full((8, 12), 0.05 * (1 + int(level)), np.float32)})
    (scenes / "no_gt_object" / "rig_1_010").mkdir(parents=True)
    return scenes, gts, views


def test_raw_layout_and_export(tmp_path):
    rng = np.random.default_rng(5)
    scenes, gts, views = _raw_layout(tmp_path, rng)
    ds = dataset.SisterDataset(scenes, gts)
    assert ds.object_names == {"hexa_nut", "washer"}                  # objects with ground truth
    assert sorted(ds.gt_levels("hexa_nut")) == ["0", "2"]
    assert sorted(ds.object_baselines("washer", "2")) == ["010", "025"]
    assert np.array_equal(ds.load_image("hexa_nut", "0", "025", "left"), views[("hexa_nut", "0", "025", "left")])
    assert np.all(ds.load_prediction("washer", "2", "010") == 513)
    assert np.allclose(ds.load_gt("washer", "2"), 0.15)
    out = dataset.export(scenes, gts, tmp_path / "exported")
    ex = dataset.ExportedDataset(out)
    assert ex.objects() == ["hexa_nut", "washer"]
    assert ex.distances("hexa_nut") == ["10cm", "1cm"]                # levels 2 and 0
    assert ex.baselines("hexa_nut", "1cm") == ["010mm", "025mm"]
    assert len(ex.rigs()) == 8
    got = ex.load_views("washer", "10cm", "025mm")
    for v, img in zip(dataset.VIEWS, got):
        assert np.array_equal(img, views[("washer", "2", "025", v)]), v
    assert np.allclose(ex.load_gt("hexa_nut", "10cm"), 0.15)


def test_depth_to_disparity():
    z = np.array([[0.1, 0.05], [0.0, np.inf]], np.float32)
    d = dataset.depth_to_disparity(z, "025mm", focal_px=1000.0)
    assert np.allclose(d, [[250.0, 500.0], [0.0, 0.0]])
    assert dataset.baseline_mm("010mm") == 10.0 and dataset.baseline_mm("100") == 100.0
