"""Image files of the SiSter dataset without OpenCV: PNG and OpenEXR.

PNG (ISO/IEC 15948): 8- or 16-bit gray, gray+alpha, RGB or RGBA,
non-interlaced, every row filter (None, Sub, Up, Average, Paeth); returned the
way cv2.imread returns them (colour as B, G, R; alpha dropped; 16 bits kept
for IMREAD_ANYDEPTH, scaled to 8 bits otherwise is not done here).  Written:
8-bit gray / BGR and 16-bit gray.

OpenEXR (the published file layout: magic, attribute header, chunk offset
table, scanline chunks): single-part scanline files, channels of type UINT,
HALF or FLOAT, compression NONE, ZIPS (one line per chunk) or ZIP (16 lines
per chunk: zlib over bytes split into even and odd halves and delta-coded).
The depth ground truth of the dataset is one float32 channel in metres
(README.md:57-61); read_exr returns it as an H x W float32 array, three
channels R, G, B as H x W x 3 in B, G, R order (as cv2.imread with
IMREAD_ANYCOLOR | IMREAD_ANYDEPTH).  Other compressions (PIZ, PXR24, B44,
DWA) and tiled or multi-part files raise ValueError.
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

PNG_SIG = b"\x89PNG\r\n\x1a\n"


# ---------------------------------------------------------------------------
# PNG
# ---------------------------------------------------------------------------
def _png_chunks(data):
    if data[:8] != PNG_SIG:
        raise ValueError("not a PNG file")
    p = 8
    while p + 8 <= len(data):
        n = struct.unpack(">I", data[p:p + 4])[0]
        yield data[p + 4:p + 8], data[p + 8:p + 8 + n]
        p += 12 + n


def _unfilter(raw, H, stride, bpp):
    """Undo the five PNG row filters; bpp = bytes per complete pixel."""
    out = np.zeros((H, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    pos = 0
    for y in range(H):
        f = raw[pos]
        line = np.frombuffer(raw, np.uint8, stride, pos + 1).astype(np.int32)
        pos += stride + 1
        if f == 0:
            cur = line
        elif f == 1:
            cur = line.copy()
            for x in range(bpp, stride, bpp):
                cur[x:x + bpp] = (cur[x:x + bpp] + cur[x - bpp:x]) & 255
        elif f == 2:
            cur = (line + prev) & 255
        elif f == 3:
            cur = line.copy()
            for x in range(stride):
                a = cur[x - bpp] if x >= bpp else 0
                cur[x] = (cur[x] + ((a + prev[x]) >> 1)) & 255
        elif f == 4:
            cur = line.copy()
            for x in range(stride):
                a = cur[x - bpp] if x >= bpp else 0
                b = prev[x]
                c = prev[x - bpp] if x >= bpp else 0
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                cur[x] = (cur[x] + pred) & 255
        else:
            raise ValueError(f"PNG row filter {f} is not defined")
        out[y] = cur
        prev = cur
    return out


def read_png(path, anydepth=True):
    """PNG -> uint8 / uint16 array: H x W for gray, H x W x 3 (B, G, R) for
    colour.  With anydepth=False 16-bit samples are reduced to their high
    byte (as an 8-bit load does)."""
    data = open(path, "rb").read()
    idat = []
    hdr = None
    for t, body in _png_chunks(data):
        if t == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body[:13])
        elif t == b"IDAT":
            idat.append(body)
        elif t == b"IEND":
            break
    if hdr is None:
        raise ValueError("PNG without IHDR")
    W, H, depth, ctype, comp, filt, inter = hdr
    if comp != 0 or filt != 0 or inter != 0 or depth not in (8, 16):
        raise ValueError("only non-interlaced 8- and 16-bit PNG is read")
    nch = {0: 1, 2: 3, 4: 2, 6: 4}.get(ctype)
    if nch is None:
        raise ValueError(f"PNG colour type {ctype} is not read (palette images)")
    bps = depth // 8
    stride = W * nch * bps
    rows = _unfilter(zlib.decompress(b"".join(idat)), H, stride, nch * bps)
    if bps == 2:
        px = (rows[:, 0::2].astype(np.uint16) << 8) | rows[:, 1::2]
    else:
        px = rows
    px = px.reshape(H, W, nch)
    if nch in (2, 4):
        px = px[..., :nch - 1]                     # alpha dropped
    if px.shape[2] == 3:
        px = px[..., ::-1]                         # RGB -> BGR
    else:
        px = px[..., 0]
    if bps == 2 and not anydepth:
        px = (px >> 8).astype(np.uint8)
    return np.ascontiguousarray(px)


def _png_chunk(t, d):
    c = t + d
    return struct.pack(">I", len(d)) + c + struct.pack(">I", zlib.crc32(c) & 0xFFFFFFFF)


def write_png(path, img):
    """uint8 H x W (gray) or H x W x 3 (B, G, R), or uint16 H x W (gray)."""
    img = np.asarray(img)
    if img.dtype == np.uint16 and img.ndim == 2:
        H, W = img.shape
        rows = img.astype(">u2").tobytes()
        depth, ctype, stride = 16, 0, 2 * W
    elif img.dtype == np.uint8 and img.ndim in (2, 3):
        H, W = img.shape[:2]
        if img.ndim == 3:
            if img.shape[2] != 3:
                raise ValueError("colour images are B, G, R")
            img = img[..., ::-1]
        rows = np.ascontiguousarray(img).tobytes()
        depth, ctype = 8, (0 if img.ndim == 2 else 2)
        stride = W * (1 if img.ndim == 2 else 3)
    else:
        raise ValueError("write_png takes uint8 gray / BGR or uint16 gray")
    raw = b"".join(b"\x00" + rows[y * stride:(y + 1) * stride] for y in range(H))
    ihdr = struct.pack(">IIBBBBB", W, H, depth, ctype, 0, 0, 0)
    with open(path, "wb") as f:
        f.write(PNG_SIG + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(raw, 6)) +
                _png_chunk(b"IEND", b""))


# ---------------------------------------------------------------------------
# OpenEXR
# ---------------------------------------------------------------------------
EXR_MAGIC = 20000630
_EXR_COMP = {0: ("NONE", 1), 2: ("ZIPS", 1), 3: ("ZIP", 16)}
_EXR_TYPES = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}


def _zip_decode(buf, n):
    """OpenEXR ZIP/ZIPS block: zlib, then undo the byte delta, then merge the
    two halves (even bytes first, odd bytes second)."""
    t = np.frombuffer(zlib.decompress(buf), np.uint8)
    if t.size != n:
        raise ValueError("EXR chunk decompressed to the wrong size")
    d = (np.cumsum(t.astype(np.int64) - np.concatenate([[0], np.full(n - 1, 128)])) & 255).astype(np.uint8) \
        if n else t
    out = np.empty(n, np.uint8)
    h = (n + 1) // 2
    out[0::2] = d[:h]
    out[1::2] = d[h:]
    return out.tobytes()


def _zip_encode(raw):
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int64)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 255
    return zlib.compress(d.astype(np.uint8).tobytes(), 6)


def _read_attr_header(data, p):
    attrs = {}
    while True:
        if data[p] == 0:
            return attrs, p + 1
        e = data.index(b"\x00", p)
        name = data[p:e].decode()
        e2 = data.index(b"\x00", e + 1)
        typ = data[e + 1:e2].decode()
        size = struct.unpack("<i", data[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, data[e2 + 5:e2 + 5 + size])
        p = e2 + 5 + size


def _channels(val):
    chans = []
    p = 0
    while val[p] != 0:
        e = val.index(b"\x00", p)
        name = val[p:e].decode()
        ptype, _plin, xs, ys = struct.unpack("<iB3xii", val[e + 1:e + 17])
        if xs != 1 or ys != 1:
            raise ValueError("subsampled EXR channels are not read")
        chans.append((name, ptype))
        p = e + 17
    return sorted(chans)


def read_exr_channels(path):
    """Every channel of a single-part scanline EXR: {name: H x W array}."""
    data = open(path, "rb").read()
    magic, version = struct.unpack("<ii", data[:8])
    if magic != EXR_MAGIC:
        raise ValueError("not an OpenEXR file")
    if version & 0x200 or version & 0x1000 or version & 0x800:
        raise ValueError("tiled, deep or multi-part EXR files are not read")
    attrs, p = _read_attr_header(data, 8)
    chans = _channels(attrs["channels"][1])
    comp = attrs["compression"][1][0]
    if comp not in _EXR_COMP:
        raise ValueError(f"EXR compression {comp} is not read (NONE, ZIPS and ZIP are)")
    lpc = _EXR_COMP[comp][1]
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    nchunk = (H + lpc - 1) // lpc
    offsets = struct.unpack(f"<{nchunk}Q", data[p:p + 8 * nchunk])
    out = {n: np.zeros((H, W), _EXR_TYPES[t].newbyteorder("=")) for n, t in chans}
    line_bytes = sum(W * _EXR_TYPES[t].itemsize for _, t in chans)
    for off in offsets:
        y, size = struct.unpack("<ii", data[off:off + 8])
        r0 = y - y0
        nl = min(lpc, H - r0)
        block = data[off + 8:off + 8 + size]
        n = nl * line_bytes
        if comp != 0 and size < n:
            block = _zip_decode(block, n)
        elif len(block) != n:
            raise ValueError("EXR chunk of the wrong size")
        q = 0
        for li in range(nl):
            for name, t in chans:
                dt = _EXR_TYPES[t]
                out[name][r0 + li] = np.frombuffer(block, dt, W, q)
                q += W * dt.itemsize
    return out


def read_exr(path):
    """The EXR as cv2.imread(path, IMREAD_ANYCOLOR | IMREAD_ANYDEPTH) returns
    it: one channel -> H x W float32; R, G, B -> H x W x 3 (B, G, R)."""
    ch = read_exr_channels(path)
    if len(ch) == 1:
        return next(iter(ch.values())).astype(np.float32)
    if {"R", "G", "B"} <= set(ch):
        return np.stack([ch["B"], ch["G"], ch["R"]], axis=2).astype(np.float32)
    raise ValueError(f"EXR channels {sorted(ch)}: expected one channel or R, G, B")


def write_exr(path, channels, compression="ZIP"):
    """channels: {name: H x W float32 / float16 / uint32}; scanline, single part."""
    codes = {v[0]: k for k, v in _EXR_COMP.items()}
    comp = codes[compression]
    lpc = _EXR_COMP[comp][1]
    names = sorted(channels)
    H, W = channels[names[0]].shape
    tcode = {np.dtype(np.float32): 2, np.dtype(np.float16): 1, np.dtype(np.uint32): 0}
    chl = b""
    for n in names:
        chl += n.encode() + b"\x00" + struct.pack("<iB3xii", tcode[channels[n].dtype], 0, 1, 1)
    chl += b"\x00"

    def attr(name, typ, val):
        return name.encode() + b"\x00" + typ.encode() + b"\x00" + struct.pack("<i", len(val)) + val

    hdr = (attr("channels", "chlist", chl) + attr("compression", "compression", bytes([comp])) +
           attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, W - 1, H - 1)) +
           attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, W - 1, H - 1)) +
           attr("lineOrder", "lineOrder", b"\x00") + attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)) +
           attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)) +
           attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\x00")
    chunks = []
    for y in range(0, H, lpc):
        raw = b"".join(channels[n][y + i].astype(channels[n].dtype.newbyteorder("<")).tobytes()
                       for i in range(min(lpc, H - y)) for n in names)
        body = raw if comp == 0 else _zip_encode(raw)
        if comp != 0 and len(body) >= len(raw):
            body = raw                                 # stored as is when it does not shrink
        chunks.append(struct.pack("<ii", y, len(body)) + body)
    start = 8 + len(hdr) + 8 * len(chunks)
    offs, pos = [], start
    for c in chunks:
        offs.append(pos)
        pos += len(c)
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", EXR_MAGIC, 2) + hdr + struct.pack(f"<{len(offs)}Q", *offs) + b"".join(chunks))
