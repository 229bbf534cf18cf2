"""painter.stroke on the GPU (host tessellation, device raster) against the
reference's stroke fixtures and against the host tessellator rasterized by
the oracle; plus random strokes (widths, joins, caps, dashes, curves, every
AA mode) against the same CPU path, bit-exact."""
import numpy as np
import pytest

from z2d_amd import abi, scenes
from gpu_util import diff_report, gpu_backend
from pngfix import export_surface, png_pixels
from spec_stroke_scenes import STROKE_SCENES
from stroke_check import StrokeCheckBackend
from test_stroke_fixtures import stroke_fixture
from zscene import Context, Surface

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("aa", [abi.AA_NONE, abi.AA_SUPERSAMPLE_4X, abi.AA_MULTISAMPLE_4X],
                         ids=["none", "ssaa", "msaa"])
@pytest.mark.parametrize("builder", STROKE_SCENES, ids=lambda f: f.__name__)
def test_gpu_stroke_fixture(builder, aa):
    scene = builder(aa)
    out = scene.run(gpu_backend())
    got = export_surface(out.format, out.width, out.height, out.data())
    want = png_pixels(stroke_fixture(scene.name, aa))
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"{len(diff)} mismatches, first at {diff[:3].tolist()}"


def _random_stroke_scene(seed):
    rng = scenes.SplitMix64(0x57E0 + seed)
    fmt = [abi.RGBA, abi.RGB, abi.ALPHA8, abi.ALPHA4, abi.ALPHA1][rng.randint(0, 4)]
    w, h = rng.randint(40, 200), rng.randint(40, 200)
    sfc = Surface(fmt, w, h)
    c = Context(sfc)
    c.aa = [abi.AA_NONE, abi.AA_DEFAULT, abi.AA_SUPERSAMPLE_4X][rng.randint(0, 2)]
    for _ in range(rng.randint(1, 4)):
        c.reset_path()
        c.line_width = rng.uniform(0.5, 12.0)
        c.join = rng.randint(0, 2)
        c.cap = rng.randint(0, 2)
        c.miter_limit = rng.uniform(1.0, 12.0)
        c.dashes = ([rng.uniform(1, 20) for _ in range(rng.randint(2, 6))]
                    if rng.uniform() < 0.35 else [])
        c.dash_offset = rng.uniform(-10, 10) if c.dashes else 0.0
        c.set_source_pixel(scenes.rgba_from_clamped(rng.uniform(), rng.uniform(), rng.uniform(),
                                                    rng.uniform(0.3, 1.0)))
        for _ in range(rng.randint(1, 3)):
            c.move_to(rng.uniform(-10, w + 10), rng.uniform(-10, h + 10))
            for _ in range(rng.randint(1, 8)):
                if rng.uniform() < 0.7:
                    c.line_to(rng.uniform(-10, w + 10), rng.uniform(-10, h + 10))
                else:
                    c.curve_to(*[rng.uniform(-10, w + 10) if i % 2 == 0 else
                                 rng.uniform(-10, h + 10) for i in range(6)])
            if rng.uniform() < 0.4:
                c.close_path()
        c.stroke()
    return sfc, c


@pytest.mark.parametrize("seed", list(range(120)))
def test_gpu_stroke_random(seed):
    sfc, c = _random_stroke_scene(seed)
    want = sfc.copy()
    got = sfc.copy()
    for kind, pat, nodes, opts in c.ops:
        assert StrokeCheckBackend().stroke(want, pat, nodes, opts) == 0
        assert gpu_backend().stroke(got, pat, nodes, opts) == 0
    msg = diff_report(got.data(), want.data())
    assert not msg, f"seed {seed}: {msg}"
