"""Device PNG export (z2d_gpu_png_encode) against the reference's files and
the oracle.

- Every spec fixture is turned back into the surface it was exported from;
  the device's file must decode to the fixture's scanline stream byte for
  byte, carry the fixture's IHDR and gAMA chunks, and have valid CRCs (the
  zlib decoder checks the Adler-32).
- 072 (sRGB profile) is exported from the oracle's render of the spec case:
  its stream must equal the fixture's within 1 LSB (the device pow uses the
  device's expf/logf).
- Random surfaces of every format and ragged sizes (1x1, rows not a whole
  byte, streams that end exactly on a deflate segment or one byte past it,
  rows longer than the 32 KiB window) against the oracle's stream; noise
  (stored blocks), flat content (matches) and mixed content (blocks with
  codes built per segment).
- argb/rgba and xrgb/rgb surfaces of the same picture give identical files
  (the reference's own test, src/export_png.zig:448-491).
"""
import ctypes as C
import glob
import os
import struct

import numpy as np
import pytest

from z2d_amd import abi, gpu
from gpu_util import gpu_backend
from oracle_backend import OracleBackend, png_scanlines
from pngfix import png_chunks, png_stream, surface_from_fixture

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")
FIXTURES = sorted(p for p in glob.glob(os.path.join(GOLD, "*.png"))
                  if not os.path.basename(p).startswith("072_"))


def _ctx():
    return gpu_backend().ctx


def _desc(fmt, w, h, data):
    buf = np.frombuffer(bytearray(data), np.uint8)
    d = abi.SurfaceDesc(fmt, w, h, 0, buf.ctypes.data, buf.nbytes)
    d._keep = buf
    return d


def encode(fmt, w, h, data, profile=abi.PNG_PROFILE_NONE):
    return gpu.png_encode(_ctx(), _desc(fmt, w, h, data), profile)


def check_file(png, fmt, w, h, profile):
    chunks = png_chunks(png)
    assert all(ok for _, _, ok in chunks), "bad chunk CRC"
    types = [t for t, _, _ in chunks]
    head = [b"IHDR"] + ([b"gAMA"] if profile != abi.PNG_PROFILE_NONE else [])
    assert types[:len(head)] == head and types[-1] == b"IEND"
    assert set(types[len(head):-1]) == {b"IDAT"}
    ihdr = struct.unpack(">IIBBBBB", chunks[0][1])
    depth = {abi.ALPHA4: 4, abi.ALPHA2: 2, abi.ALPHA1: 1}.get(fmt, 8)
    ctype = 6 if fmt in (abi.ARGB, abi.RGBA) else 2 if fmt in (abi.XRGB, abi.RGB) else 0
    assert ihdr == (w, h, depth, ctype, 0, 0, 0)
    idat = [d for t, d, _ in chunks if t == b"IDAT"]
    assert all(len(d) == 16384 for d in idat[:-1]) and 0 < len(idat[-1]) <= 16384


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_fixture_roundtrip(path):
    fmt, w, h, data = surface_from_fixture(path)
    prof = abi.PNG_PROFILE_LINEAR if os.path.basename(path).startswith("071_") \
        else abi.PNG_PROFILE_NONE
    png = encode(fmt, w, h, data, prof)
    print(os.path.basename(path), fmt, w, h, len(png), flush=True)
    check_file(png, fmt, w, h, prof)
    ref = png_chunks(open(path, "rb").read())
    ours = png_chunks(png)
    assert ours[0][1] == ref[0][1]  # IHDR payload
    ref_g = [d for t, d, _ in ref if t == b"gAMA"]
    our_g = [d for t, d, _ in ours if t == b"gAMA"]
    assert our_g == ref_g
    assert png_stream(png) == png_stream(path)


def test_srgb_fixture_072():
    from spec_comp_scenes import COMP_SCENES
    fn = dict((n, f) for n, f, _ in COMP_SCENES)["072_gamma_srgb"]
    out = fn(OracleBackend(), None)
    png = encode(out.format, out.width, out.height, out.data().tobytes(), abi.PNG_PROFILE_SRGB)
    check_file(png, out.format, out.width, out.height, abi.PNG_PROFILE_SRGB)
    ref_path = os.path.join(GOLD, "072_gamma_srgb.png")
    ref = png_chunks(open(ref_path, "rb").read())
    ours = png_chunks(png)
    assert ours[0][1] == ref[0][1]
    assert [d for t, d, _ in ours if t == b"gAMA"] == [d for t, d, _ in ref if t == b"gAMA"]
    got = np.frombuffer(png_stream(png), np.uint8).astype(np.int32)
    want = np.frombuffer(png_stream(ref_path), np.uint8).astype(np.int32)
    d = np.abs(got - want)
    print("072 stream bytes differing:", int((d > 0).sum()), flush=True)
    assert d.max() <= 1


def _surface_bytes(fmt, w, h):
    bits = abi.FORMAT_BITS[fmt]
    return w * h * 4 if bits == 32 else (w * h * bits + 7) // 8


def random_surface(fmt, w, h, kind, seed):
    rng = np.random.default_rng(seed)
    n = _surface_bytes(fmt, w, h)
    if kind == "noise":
        data = rng.integers(0, 256, n, dtype=np.uint8)
    elif kind == "flat":
        data = np.full(n, rng.integers(0, 256), np.uint8)
    else:  # blocks: runs of pixels and repeated rows, some noise
        if abi.FORMAT_BITS[fmt] == 32:
            px = rng.integers(0, 2**32, max(1, w // 7 + 1), dtype=np.uint64).astype(np.uint32)
            row = np.repeat(px, 7)[:w]
            rows = np.tile(row, (h, 1))
            noisy = rng.random(h) < 0.3
            rows[noisy] = rng.integers(0, 2**32, (int(noisy.sum()), w), dtype=np.uint64)
            data = rows.astype(np.uint32).view(np.uint8).reshape(-1)
        else:
            data = np.repeat(rng.integers(0, 256, n // 13 + 1, dtype=np.uint8), 13)[:n]
    data = data.copy()
    if fmt in (abi.ARGB, abi.RGBA) and kind != "flat":
        px = data.reshape(-1, 4)
        a = px[:, 3].astype(np.uint32)
        for k in range(3):  # keep the surface premultiplied-valid (c <= a)
            px[:, k] = (px[:, k].astype(np.uint32) * a // 255).astype(np.uint8)
    return data.tobytes()


SIZES = [(1, 1), (7, 3), (33, 17), (300, 7), (257, 129), (9000, 3)]


@pytest.mark.parametrize("fmt", range(8), ids=abi.FORMAT_NAMES)
@pytest.mark.parametrize("kind", ["noise", "flat", "blocks"])
def test_random_surfaces(fmt, kind):
    for i, (w, h) in enumerate(SIZES):
        data = random_surface(fmt, w, h, kind, 1000 * fmt + i)
        print(abi.FORMAT_NAMES[fmt], kind, w, h, flush=True)
        png = encode(fmt, w, h, data)
        check_file(png, fmt, w, h, abi.PNG_PROFILE_NONE)
        assert png_stream(png) == png_scanlines(fmt, w, h, data)


SEG = 4096  # stream bytes per deflate segment (PNG_SEG)


def test_segment_boundaries():
    # alpha8 streams that end exactly on a segment (rows of SEG and 4 SEG
    # bytes with their filter byte), one byte past it and one byte short;
    # rgb rows wider than the 32 KiB window
    for w, h in [(SEG - 1, 2), (SEG - 1, 1), (SEG, 2), (SEG - 2, 2), (4 * SEG - 1, 3),
                 (4 * SEG, 1), (12000, 4)]:
        for kind in ("flat", "blocks", "noise"):
            fmt = abi.ALPHA8 if w != 12000 else abi.RGB
            data = random_surface(fmt, w, h, kind, w + h)
            png = encode(fmt, w, h, data)
            check_file(png, fmt, w, h, abi.PNG_PROFILE_NONE)
            assert png_stream(png) == png_scanlines(fmt, w, h, data), (w, h, kind)


@pytest.mark.parametrize("fmt", [abi.ARGB, abi.XRGB, abi.RGB, abi.RGBA])
def test_srgb_profile_random(fmt):
    data = random_surface(fmt, 333, 77, "noise", 7 + fmt)
    png = encode(fmt, 333, 77, data, abi.PNG_PROFILE_SRGB)
    check_file(png, fmt, 333, 77, abi.PNG_PROFILE_SRGB)
    got = np.frombuffer(png_stream(png), np.uint8).astype(np.int32)
    want = np.frombuffer(png_scanlines(fmt, 333, 77, data, abi.PNG_PROFILE_SRGB),
                         np.uint8).astype(np.int32)
    d = np.abs(got - want)
    print("srgb bytes differing:", int((d > 0).sum()), flush=True)
    assert d.max() <= 1


def test_compression_flat_and_noise():
    w, h = 1024, 512
    flat = random_surface(abi.RGBA, w, h, "flat", 3)
    png = encode(abi.RGBA, w, h, flat)
    assert len(png) < w * h * 4 // 100
    noise = random_surface(abi.RGB, w, h, "noise", 4)
    png = encode(abi.RGB, w, h, noise)
    raw = h * (1 + 3 * w)
    assert len(png) <= gpu.png_bound(abi.RGB, w, h) and len(png) < raw * 1.01


def test_same_picture_same_file():
    # the reference's "RGB/ARGB formats all export to same image"
    w, h = 300, 300
    rgba = random_surface(abi.RGBA, w, h, "blocks", 11)
    px = np.frombuffer(rgba, np.uint8).reshape(-1, 4)
    argb = px[:, [2, 1, 0, 3]].copy().tobytes()
    assert encode(abi.RGBA, w, h, rgba) == encode(abi.ARGB, w, h, argb)
    rgb = px.copy()
    rgb[:, 3] = 0
    xrgb = px[:, [2, 1, 0, 3]].copy()
    xrgb[:, 3] = 0
    assert encode(abi.RGB, w, h, rgb.tobytes()) == encode(abi.XRGB, w, h, xrgb.tobytes())


def test_device_surface_and_errors():
    ctx = _ctx()
    w, h = 97, 41
    data = random_surface(abi.RGBA, w, h, "blocks", 5)
    s = gpu.DeviceSurface(ctx, abi.RGBA, w, h)
    s.upload(np.frombuffer(data, np.uint8))
    png = gpu.png_encode(ctx, s.desc(), abi.PNG_PROFILE_NONE)
    assert png == encode(abi.RGBA, w, h, data)
    L = gpu.lib()
    n = C.c_uint64()
    assert L.z2d_gpu_png_bound(abi.RGBA, 0, 5, 0, C.byref(n)) == 3  # InvalidWidth
    assert L.z2d_gpu_png_bound(abi.RGBA, 5, 0, 0, C.byref(n)) == 4  # InvalidHeight
    assert L.z2d_gpu_png_bound(abi.RGBA, 5, 5, 3, C.byref(n)) == -1
    small = C.create_string_buffer(16)
    d = s.desc()
    assert L.z2d_gpu_png_encode(ctx.h, C.byref(d), 0, small, 16, C.byref(n)) == -1
    assert n.value == len(png)
    s.close()
