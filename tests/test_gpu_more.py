"""The transformed-stroke, gradient-stroke, curved-gradient-fill,
out-of-bounds and double-close cases (spec 043-045, 059, 060, 075-078) on
the GPU: against the reference's fixtures where the CPU path matches them,
and against the CPU path (host tessellator + oracle) everywhere, bit-exact."""
import numpy as np
import pytest

from z2d_amd import abi
from gpu_util import diff_report, gpu_backend
from pngfix import export_surface, png_pixels
from spec_more_scenes import MORE_SCENES, NEAR_SCENES
from stroke_check import StrokeCheckBackend
from test_stroke_fixtures import stroke_fixture

pytestmark = pytest.mark.gpu
AA = [abi.AA_NONE, abi.AA_SUPERSAMPLE_4X, abi.AA_MULTISAMPLE_4X]


@pytest.mark.parametrize("aa", AA, ids=["none", "ssaa", "msaa"])
@pytest.mark.parametrize("builder", MORE_SCENES, ids=lambda f: f.__name__)
def test_gpu_more_fixture(builder, aa):
    scene = builder(aa)
    out = scene.run(gpu_backend())
    got = export_surface(out.format, out.width, out.height, out.data())
    want = png_pixels(stroke_fixture(scene.name, aa))
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"{len(diff)} mismatches, first at {diff[:3].tolist()}"


@pytest.mark.parametrize("aa", AA, ids=["none", "ssaa", "msaa"])
@pytest.mark.parametrize("builder", MORE_SCENES + NEAR_SCENES, ids=lambda f: f.__name__)
def test_gpu_more_equals_cpu(builder, aa):
    scene = builder(aa)
    got = scene.run(gpu_backend())
    want = scene.run(StrokeCheckBackend())
    msg = diff_report(np.asarray(got.data()), np.asarray(want.data()))
    assert not msg, msg
