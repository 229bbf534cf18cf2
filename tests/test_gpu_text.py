"""Text on the device: spec case 074 through z2d_gpu_text_show and through a
batch (z2d_gpu_batch_text), pixel for pixel against the fixture; and many
strings (sizes, rotations, scales, solid and gradient paints, four fonts'
worth of glyphs, accents) on several surfaces in one batch, bit-exact against
the oracle rendering the same text paths."""
import math
import os

import numpy as np
import pytest

from z2d_amd import abi, gpu, scenes
from gpu_util import diff_report, gpu_backend
from oracle_backend import OracleBackend
from pngfix import export_surface, png_pixels
from spec_text import H, TEXT, W, font, lines

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")
AA = [(abi.AA_NONE, "_pixelated"), (abi.AA_SUPERSAMPLE_4X, "_smooth"),
      (abi.AA_MULTISAMPLE_4X, "_smooth")]


def _check_fixture(fmt, data, suffix):
    got = export_surface(fmt, W, H, data)
    want = png_pixels(os.path.join(GOLD, "074_text" + suffix + ".png"))
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"{len(diff)} mismatches, first at {diff[:3].tolist()}"


@pytest.mark.parametrize("aa,suffix", AA, ids=["none", "ssaa", "msaa"])
def test_074_text_show(aa, suffix):
    from zscene import Surface
    ctx = gpu_backend().ctx
    f = font()
    s = Surface(abi.RGB, W, H)
    d = s.desc()
    for pat, x, y, opts, _keep in lines(aa):
        rc = gpu.text_show(ctx, d, pat, f, TEXT, x, y, opts)
        assert rc == 0, rc
    _check_fixture(abi.RGB, s.data(), suffix)


@pytest.mark.parametrize("aa,suffix", AA, ids=["none", "ssaa", "msaa"])
def test_074_text_batch(aa, suffix):
    ctx = gpu_backend().ctx
    f = font()
    surf = gpu.DeviceSurface(ctx, abi.RGB, W, H)
    b = gpu.Batch(ctx)
    ls = lines(aa)
    for pat, x, y, opts, _keep in ls:
        assert b.text(surf, pat, f, TEXT, x, y, opts) == 0
    b.prepare()
    b.render(0)
    b.wait()
    out = np.zeros(W * H * 4, np.uint8)
    surf.download(out)
    _check_fixture(abi.RGB, out, suffix)
    b.close()
    surf.close()


WORDS = ["The", "quick", "brown", "fox", "jumps", "over", "the", "lázy", "dög", "Ærøskøbing",
         "façade", "naïve", "0123456789", "{[()]}", "@#%&*", "Wavy", "AVATAR", "fi fl ff"]


def test_many_strings_one_batch():
    from zscene import Surface
    ctx = gpu_backend().ctx
    fonts = [font("Inter-Regular.ttf"), font("AnnotationMono-Regular.ttf"),
             font("Inter-Regular.subset.ttf")]
    rng = scenes.SplitMix64(0x7E47)
    sizes = [(512, 384, abi.RGBA), (640, 200, abi.RGB), (300, 300, abi.ALPHA8)]
    surfs = [gpu.DeviceSurface(ctx, fmt, w, h) for w, h, fmt in sizes]
    refs = [Surface(fmt, w, h) for w, h, fmt in sizes]
    keep = []
    b = gpu.Batch(ctx)
    n = 0
    for k in range(120):
        si = k % len(sizes)
        w, h, fmt = sizes[si]
        f = fonts[int(rng.uniform() * len(fonts)) % len(fonts)]
        text = " ".join(WORDS[int(rng.uniform() * len(WORDS)) % len(WORDS)]
                        for _ in range(1 + int(rng.uniform() * 4)))
        size = rng.uniform(6.0, 48.0)
        ang = rng.uniform(-0.6, 0.6) if rng.uniform() < 0.4 else 0.0
        sx = rng.uniform(0.6, 1.6) if rng.uniform() < 0.3 else 1.0
        c, s_ = math.cos(ang), math.sin(ang)
        t = (c * sx, -s_, s_ * sx, c, rng.uniform(0, w * 0.3), rng.uniform(0, h * 0.5))
        aa = [abi.AA_NONE, abi.AA_DEFAULT, abi.AA_SUPERSAMPLE_4X][k % 3]
        opts = abi.text_opts(size, transformation=t, fill=abi.fill_opts(aa, k % 2))
        if rng.uniform() < 0.3:
            g = abi.Gradient(abi.GRADIENT_LINEAR, [0, 0, w, h],
                             [(0.0, abi.color(abi.COLOR_LINEAR_RGB, 1, 0.2, 0, 1.0)),
                              (1.0, abi.color(abi.COLOR_LINEAR_RGB, 0, 0.3, 1, 0.8))])
            keep.append(g)
            pat = g.pattern()
        else:
            pat = abi.solid_pattern(abi.rgba(int(rng.uniform() * 255), int(rng.uniform() * 255),
                                             int(rng.uniform() * 255), 255))
        x, y = rng.uniform(-10, w * 0.6), rng.uniform(0, h * 0.9)
        assert b.text(surfs[si], pat, f, text, x, y, opts) == 0
        rc, nodes = f.text_path(text, x, y, opts)
        assert rc == 0
        if len(nodes):
            assert OracleBackend().fill(refs[si], pat, nodes, opts.fill) == 0
        n += 1
    b.prepare()
    b.render(0)
    b.wait()
    for (w, h, fmt), ds, ref in zip(sizes, surfs, refs):
        out = np.zeros(ref.nbytes, np.uint8)
        ds.download(out)
        msg = diff_report(out, ref.data())
        assert not msg, f"{abi.FORMAT_NAMES[fmt]} {w}x{h}: {msg}"
    b.close()
    for ds in surfs:
        ds.close()
