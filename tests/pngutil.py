"""Tiny PNG writer/reader (zlib) used by the CLI tests."""
import struct
import zlib

import numpy as np


def _chunk(t, d):
    c = t + d
    return struct.pack(">I", len(d)) + c + struct.pack(">I", zlib.crc32(c) & 0xFFFFFFFF)


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)


def write_png(path, img, filters=(0,)):
    """img: HxW (gray) or HxWx3 (RGB) uint8; row y uses filters[y % len]."""
    img = np.asarray(img, np.uint8)
    H, W = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    ctype = 0 if ch == 1 else 2
    rows = img.reshape(H, W * ch).astype(np.int32)
    raw = bytearray()
    for y in range(H):
        f = filters[y % len(filters)]
        cur = rows[y]
        up = rows[y - 1] if y else np.zeros_like(cur)
        out = []
        for x in range(W * ch):
            a = int(cur[x - ch]) if x >= ch else 0
            b = int(up[x])
            c = int(up[x - ch]) if x >= ch else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][f]
            out.append((int(cur[x]) - pred) & 255)
        raw.append(f)
        raw.extend(out)
    ihdr = struct.pack(">IIBBBBB", W, H, 8, ctype, 0, 0, 0)
    data = b"\x89PNG\r\n\x1a\n" + _chunk(b"IHDR", ihdr) + _chunk(b"IDAT", zlib.compress(bytes(raw))) + _chunk(b"IEND", b"")
    open(path, "wb").write(data)


def read_png16(path):
    """Gray 8/16-bit, filter type 0 only (what compute_disp writes)."""
    d = open(path, "rb").read()
    p = 8
    idat = b""
    while p < len(d):
        n = struct.unpack(">I", d[p:p + 4])[0]
        t = d[p + 4:p + 8]
        body = d[p + 8:p + 8 + n]
        if t == b"IHDR":
            W, H, depth = struct.unpack(">IIB", body[:9])
        elif t == b"IDAT":
            idat += body
        p += 12 + n
    raw = zlib.decompress(idat)
    bps = depth // 8
    out = np.zeros((H, W), np.uint16)
    stride = W * bps + 1
    for y in range(H):
        row = np.frombuffer(raw[y * stride + 1:(y + 1) * stride], np.uint8)
        out[y] = (row[0::2].astype(np.uint16) << 8 | row[1::2]) if bps == 2 else row
    return out


def read_png_rgb8(path):
    """8-bit RGB (colour type 2), filter type 0 only (what compute_disp writes)."""
    d = open(path, "rb").read()
    p = 8
    idat = b""
    while p < len(d):
        n = struct.unpack(">I", d[p:p + 4])[0]
        t = d[p + 4:p + 8]
        body = d[p + 8:p + 8 + n]
        if t == b"IHDR":
            W, H, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 2
        elif t == b"IDAT":
            idat += body
        p += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(H, 3 * W + 1)
    assert np.all(raw[:, 0] == 0)
    return raw[:, 1:].reshape(H, W, 3).copy()
