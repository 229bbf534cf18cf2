"""Text (src/text.zig, src/Font.zig, src/internal/Glyph.zig) on the CPU:
the library's font parser and layout against the reference's own unit-test
expectations (Font.loadBuffer, Glyph.init, Glyph.Outline.init), and spec
case 074 rendered from the library's text paths by the oracle, pixel for
pixel against the fixture (its scanline stream byte for byte)."""
import os

import numpy as np
import pytest

from z2d_amd import abi, gpu
from oracle_backend import OracleBackend, png_scanlines
from pngfix import export_surface, png_pixels, png_stream
from spec_text import FONTS, font, render

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")
AA = [(abi.AA_NONE, "_pixelated"), (abi.AA_SUPERSAMPLE_4X, "_smooth"),
      (abi.AA_MULTISAMPLE_4X, "_smooth")]


def test_font_meta():  # src/Font.zig:342-390 (Inter-Regular.subset)
    info = font("Inter-Regular.subset.ttf").info()
    assert info == dict(units_per_em=2048, advance_width_max=5492, number_of_hmetrics=292,
                        lsb_is_at_x_zero=True, loca_long=False, cmap_full=False)


def test_glyph_init():  # src/internal/Glyph.zig:877-926, :1010-1043
    f = font("Inter-Regular.subset.ttf")
    assert f.glyph(ord("a")) == dict(index=56, advance=1150, lsb=90, offset=3256)
    assert f.glyph(ord(" ")) == dict(index=252, advance=576, lsb=0, offset=None)
    assert f.glyph(0x10FFFD) == dict(index=0, advance=1344, lsb=328, offset=236)
    m = font("AnnotationMono-Regular.ttf")
    g = m.glyph(ord("u"))  # past the long hmetrics
    assert (g["index"], g["advance"], g["lsb"]) == (676, 500, 85)
    g = m.glyph(0x1E95)  # the last glyph
    assert (g["index"], g["advance"], g["lsb"]) == (765, 500, 64)


def _closed(tags):  # isClosedNodeSet (src/internal/path_nodes.zig:22-37)
    closed = False
    for i, t in enumerate(tags):
        if t == abi.MOVE_TO:
            if not closed and i != 0:
                return False
        elif t == abi.CLOSE_PATH:
            closed = True
        else:
            closed = False
    return closed


@pytest.mark.parametrize("name,cp,curves", [("Inter-Regular.subset.ttf", ord("a"), True),
                                             ("Inter-Regular.subset.ttf", ord("á"), True),
                                             ("Inter-Regular.subset.ttf", 0x10FFFD, False),
                                             ("Inter-SemiBold.subset.ttf", ord("i"), False)])
def test_outline(name, cp, curves):  # src/internal/Glyph.zig:937-1008
    rc, nodes = font(name).text_path(chr(cp), 0.0, 0.0, abi.text_opts(2048.0))
    assert rc == 0
    tags = list(nodes["tag"])
    assert _closed(tags)
    if name.startswith("Inter-Regular"):
        assert (abi.CURVE_TO in tags) == curves


def test_invalid_inputs():
    with pytest.raises(gpu.GpuError):
        gpu.Font(b"not a font at all")
    f = font()
    rc, _ = f.text_path(b"ab\xff", 0.0, 0.0, abi.text_opts())
    assert rc == -6  # Z2D_ERR_INVALID_TEXT (Utf8InvalidStartByte)
    rc, _ = f.text_path(b"\xc3", 0.0, 0.0, abi.text_opts())
    assert rc == -6  # truncated sequence
    rc, nodes = f.text_path("   ", 0.0, 0.0, abi.text_opts())
    assert rc == 0 and len(nodes) == 0  # whitespace only: an empty path


@pytest.mark.parametrize("aa,suffix", AA, ids=["none", "ssaa", "msaa"])
def test_074_text_oracle(aa, suffix):
    s = render(OracleBackend(), aa)
    path = os.path.join(GOLD, "074_text" + suffix + ".png")
    got = export_surface(s.format, s.width, s.height, s.data())
    want = png_pixels(path)
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"{len(diff)} mismatches, first at {diff[:3].tolist()}"
    assert png_scanlines(s.format, s.width, s.height, s.data()) == png_stream(path)
