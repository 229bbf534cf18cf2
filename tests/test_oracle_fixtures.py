"""Pins the CPU oracle against the reference's own acceptance fixtures
(spec/files/*.png copied to tests/golden/spec/): each spec case's inputs are
restated in spec_scenes.py, rendered by the oracle under the three AA modes
the reference's pathTestRun uses (spec/main_spec.zig:677-770), exported with
the reference's PNG transform and compared pixel-for-pixel."""
import os

import numpy as np
import pytest

from z2d_amd import abi
from oracle_backend import OracleBackend, png_scanlines
from pngfix import export_surface, png_pixels, png_stream
from spec_scenes import AA_SUFFIX, PATH_SCENES

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")


def fixture_path(name, aa):
    p = os.path.join(GOLD, name + AA_SUFFIX[aa] + ".png")
    if not os.path.exists(p) and aa == abi.AA_MULTISAMPLE_4X:
        p = os.path.join(GOLD, name + "_smooth.png")
    return p


@pytest.mark.parametrize("aa", [abi.AA_NONE, abi.AA_SUPERSAMPLE_4X, abi.AA_MULTISAMPLE_4X],
                         ids=["none", "ssaa", "msaa"])
@pytest.mark.parametrize("builder", PATH_SCENES, ids=lambda f: f.__name__)
def test_oracle_matches_fixture(builder, aa):
    scene = builder(aa)
    out = scene.run(OracleBackend())
    got = export_surface(out.format, out.width, out.height, out.data())
    want = png_pixels(fixture_path(scene.name, aa))
    assert got.shape == want.shape
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"{len(diff)} mismatches, first at {diff[:3].tolist()}"
    # the export's scanline stream, byte for byte (filter bytes, packing, demultiply)
    assert png_scanlines(out.format, out.width, out.height, out.data()) == png_stream(fixture_path(scene.name, aa))
