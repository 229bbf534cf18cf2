"""GPU parity on the reference's spec scenes: the HIP library must produce
the same bytes as the oracle (and therefore the reference's fixture PNGs)
under all three AA modes."""
import os

import numpy as np
import pytest

from z2d_amd import abi
from gpu_util import diff_report, gpu_backend
from oracle_backend import OracleBackend
from pngfix import export_surface, png_pixels
from spec_scenes import PATH_SCENES
from test_oracle_fixtures import fixture_path

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("aa", [abi.AA_NONE, abi.AA_SUPERSAMPLE_4X, abi.AA_MULTISAMPLE_4X],
                         ids=["none", "ssaa", "msaa"])
@pytest.mark.parametrize("builder", PATH_SCENES, ids=lambda f: f.__name__)
def test_gpu_spec_scene(builder, aa):
    scene = builder(aa)
    want = scene.run(OracleBackend())
    got = scene.run(gpu_backend())
    msg = diff_report(got.data(), want.data())
    assert not msg, msg
    exp = export_surface(got.format, got.width, got.height, got.data())
    assert np.array_equal(exp, png_pixels(fixture_path(scene.name, aa)))
