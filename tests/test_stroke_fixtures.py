"""The host stroke tessellator (a restatement of src/internal/tess/
stroke_plotter.zig, dashed_plotter.zig, Face.zig, Pen.zig, Slope.zig,
Dasher.zig) against the reference's stroke fixtures: polygons rasterized by
the pinned oracle must reproduce spec/files/NNN_*.png pixel for pixel."""
import os

import numpy as np
import pytest

from z2d_amd import abi
from oracle_backend import png_scanlines
from pngfix import export_surface, png_pixels, png_stream
from spec_scenes import AA_SUFFIX
from spec_stroke_scenes import STROKE_SCENES
from stroke_check import StrokeCheckBackend

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")


def stroke_fixture(name, aa):
    suffix = AA_SUFFIX[abi.AA_SUPERSAMPLE_4X if aa == abi.AA_MULTISAMPLE_4X else aa]
    return os.path.join(GOLD, name + suffix + ".png")


@pytest.mark.parametrize("aa", [abi.AA_NONE, abi.AA_SUPERSAMPLE_4X, abi.AA_MULTISAMPLE_4X],
                         ids=["none", "ssaa", "msaa"])
@pytest.mark.parametrize("builder", STROKE_SCENES, ids=lambda f: f.__name__)
def test_stroke_matches_fixture(builder, aa):
    scene = builder(aa)
    out = scene.run(StrokeCheckBackend())
    got = export_surface(out.format, out.width, out.height, out.data())
    want = png_pixels(stroke_fixture(scene.name, aa))
    assert got.shape == want.shape
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"{len(diff)} mismatches, first at {diff[:3].tolist()}"
    # the export's scanline stream, byte for byte (filter bytes, packing, demultiply)
    assert png_scanlines(out.format, out.width, out.height, out.data()) == png_stream(stroke_fixture(scene.name, aa))
