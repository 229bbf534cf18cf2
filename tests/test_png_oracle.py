"""The PNG export's scanline stream (src/export_png.zig:153-360) on the
oracle, pinned against the reference's own files: every fixture in
spec/files is turned back into the surface it was exported from, and the
oracle's stream for that surface must equal the fixture's decompressed IDAT
byte for byte (filter bytes, RGB byte order, demultiplied RGBA, MSB-first
packed gray with padded row ends).  072 (sRGB profile) is covered by the
rendered-scene check in test_oracle_comp_fixtures.py, since its gamma step is
not invertible."""
import glob
import os

import pytest

from z2d_amd import abi
from oracle_backend import png_scanlines
from pngfix import png_chunks, png_stream, surface_from_fixture

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")
FIXTURES = sorted(p for p in glob.glob(os.path.join(GOLD, "*.png"))
                  if not os.path.basename(p).startswith("072_"))
PROFILE = {"071": abi.PNG_PROFILE_LINEAR}


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_stream_of_recovered_surface(path):
    fmt, w, h, data = surface_from_fixture(path)
    prof = PROFILE.get(os.path.basename(path)[:3], abi.PNG_PROFILE_NONE)
    assert png_scanlines(fmt, w, h, data, prof) == png_stream(path)


def test_fixture_chunks_are_the_reference_layout():
    for path in FIXTURES:
        types = [t for t, _, ok in png_chunks(open(path, "rb").read()) if ok]
        assert types[0] == b"IHDR" and types[-1] == b"IEND"
        assert set(types[1:-1]) <= {b"gAMA", b"IDAT"}


def test_packed_rows_pad_to_a_byte():
    # alpha1, 3x2: rows are 1 byte each, pixels MSB-first
    import numpy as np
    vals = np.array([1, 0, 1, 0, 1, 1], np.uint32)
    buf = np.zeros(1, np.uint8)
    for i, v in enumerate(vals):
        buf[i >> 3] |= np.uint8(v << (i & 7))
    s = png_scanlines(abi.ALPHA1, 3, 2, buf.tobytes())
    assert s == bytes([0, 0b10100000, 0, 0b01100000])
