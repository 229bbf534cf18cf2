"""PNG fixture decoding and the reference's PNG export transform
(src/export_png.zig:153-360), so a rendered surface can be compared with the
reference's spec/files fixtures pixel by pixel instead of by file hash."""
import struct
import zlib

import numpy as np

from z2d_amd import abi


def png_chunks(data):
    """(type, payload, crc_ok) for each chunk of a PNG file's bytes."""
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    out = []
    i = 8
    while i < len(data):
        (n,) = struct.unpack(">I", data[i:i + 4])
        t = data[i + 4:i + 8]
        d = data[i + 8:i + 8 + n]
        (crc,) = struct.unpack(">I", data[i + 8 + n:i + 12 + n])
        out.append((t, d, zlib.crc32(t + d) & 0xFFFFFFFF == crc))
        i += 12 + n
    assert i == len(data)
    return out


def png_stream(data):
    """The decompressed IDAT stream (filter bytes + scanlines) of a PNG
    file's bytes or path."""
    if isinstance(data, str):
        data = open(data, "rb").read()
    return zlib.decompress(b"".join(d for t, d, _ in png_chunks(data) if t == b"IDAT"))


def decode_png(path):
    """Returns (color_type, bit_depth, width, height, rows) where rows is a
    list of unfiltered scanline byte strings."""
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    i = 8
    idat = b""
    ihdr = None
    while i < len(data):
        (n,) = struct.unpack(">I", data[i:i + 4])
        t = data[i + 4:i + 8]
        d = data[i + 8:i + 8 + n]
        if t == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", d)
        elif t == b"IDAT":
            idat += d
        i += 12 + n
    w, h, depth, ctype = ihdr[0], ihdr[1], ihdr[2], ihdr[3]
    raw = zlib.decompress(idat)
    chans = {0: 1, 2: 3, 6: 4}[ctype]
    row_bytes = (w * chans * depth + 7) // 8
    bpp = max(1, chans * depth // 8)
    rows = []
    prev = bytearray(row_bytes)
    pos = 0
    for _ in range(h):
        f = raw[pos]
        line = bytearray(raw[pos + 1:pos + 1 + row_bytes])
        pos += 1 + row_bytes
        for k in range(row_bytes):
            a = line[k - bpp] if k >= bpp else 0
            b = prev[k]
            c = prev[k - bpp] if k >= bpp else 0
            if f == 1:
                line[k] = (line[k] + a) & 0xFF
            elif f == 2:
                line[k] = (line[k] + b) & 0xFF
            elif f == 3:
                line[k] = (line[k] + (a + b) // 2) & 0xFF
            elif f == 4:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pr = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
                line[k] = (line[k] + pr) & 0xFF
        rows.append(bytes(line))
        prev = line
    return ctype, depth, w, h, rows


def png_pixels(path):
    """Decoded fixture as a numpy array: (h, w, 3|4) for color, (h, w) for
    gray at native depth (MSB-first unpacking)."""
    ctype, depth, w, h, rows = decode_png(path)
    if ctype in (2, 6):
        ch = 3 if ctype == 2 else 4
        return np.frombuffer(b"".join(rows), np.uint8).reshape(h, w, ch)
    out = np.zeros((h, w), np.uint8)
    per = 8 // depth
    mask = (1 << depth) - 1
    for y, r in enumerate(rows):
        for x in range(w):
            byte = r[x // per]
            sh = 8 - depth - (x % per) * depth
            out[y, x] = (byte >> sh) & mask
    return out


def surface_channels(fmt, w, h, data):
    """Unpacks a surface buffer to (h, w, 4) r,g,b,a (alpha formats at native
    depth in the `a` channel)."""
    n = w * h
    if abi.FORMAT_BITS[fmt] == 32:
        px = np.frombuffer(bytes(data[: n * 4]), np.uint8).reshape(h, w, 4)
        if fmt == abi.RGBA:
            return px.copy()
        if fmt == abi.ARGB:
            return px[..., [2, 1, 0, 3]].copy()
        if fmt == abi.RGB:
            o = px.copy()
            o[..., 3] = 255
            return o
        o = px[..., [2, 1, 0, 3]].copy()
        o[..., 3] = 255
        return o
    out = np.zeros((h, w, 4), np.uint8)
    bits = abi.FORMAT_BITS[fmt]
    if bits == 8:
        out[..., 3] = np.frombuffer(bytes(data[:n]), np.uint8).reshape(h, w)
        return out
    b = np.frombuffer(bytes(data), np.uint8)
    idx = np.arange(n, dtype=np.int64) * bits
    vals = (b[idx >> 3] >> (idx & 7).astype(np.uint8)) & ((1 << bits) - 1)
    out[..., 3] = vals.reshape(h, w)
    return out


def export_surface(fmt, w, h, data):
    """What writeToPNGFile would store for this surface (linear profile):
    rgb/xrgb -> RGB bytes; rgba/argb -> demultiplied RGBA; alpha -> gray at
    native depth."""
    ch = surface_channels(fmt, w, h, data).astype(np.uint32)
    if fmt in (abi.RGB, abi.XRGB):
        return ch[..., :3].astype(np.uint8)
    if fmt in (abi.RGBA, abi.ARGB):
        a = ch[..., 3]
        o = np.zeros_like(ch)
        den = np.maximum(a, 1)
        for k in range(3):
            o[..., k] = np.where(a == 0, 0, ch[..., k] * 255 // den)
        o[..., 3] = a
        return (o & 0xFF).astype(np.uint8)
    return ch[..., 3].astype(np.uint8)


def surface_from_fixture(path):
    """The surface buffer a fixture was exported from, recovered from the
    file: color type 2 -> rgb, 6 -> rgba (re-premultiplied: c*255//a is
    strictly increasing in c for a <= 255, so the premultiplied value is the
    unique c that exports to the stored byte), gray -> alpha8/4/2/1 at the
    file's depth.  Returns (format, width, height, bytes)."""
    ctype, depth, w, h, rows = decode_png(path)
    n = w * h
    if ctype == 2:
        px = np.frombuffer(b"".join(rows), np.uint8).reshape(h, w, 3)
        out = np.zeros((h, w, 4), np.uint8)
        out[..., :3] = px
        return abi.RGB, w, h, out.tobytes()
    if ctype == 6:
        px = np.frombuffer(b"".join(rows), np.uint8).reshape(h, w, 4).astype(np.int64)
        a = px[..., 3]
        out = np.zeros((h, w, 4), np.uint8)
        out[..., 3] = a
        c = np.arange(256, dtype=np.int64)
        for av in np.unique(a):
            if av == 0:
                continue
            exp = c * 255 // av          # what each premultiplied c exports to
            inv = np.full(256, -1, np.int64)
            inv[exp[exp < 256]] = c[exp < 256]
            sel = a == av
            for k in range(3):
                v = inv[px[..., k][sel]]
                assert (v >= 0).all(), "stored byte no premultiplied value exports to"
                out[..., k][sel] = v
        return abi.RGBA, w, h, out.tobytes()
    vals = png_pixels(path).reshape(-1).astype(np.uint32)
    if depth == 8:
        return abi.ALPHA8, w, h, vals.astype(np.uint8).tobytes()
    fmt = {4: abi.ALPHA4, 2: abi.ALPHA2, 1: abi.ALPHA1}[depth]
    buf = np.zeros((n * depth + 7) // 8, np.uint8)
    idx = np.arange(n, dtype=np.int64) * depth
    np.bitwise_or.at(buf, idx >> 3, (vals << (idx & 7).astype(np.uint32)).astype(np.uint8))
    return fmt, w, h, buf.tobytes()
