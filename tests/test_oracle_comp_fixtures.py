"""Pins the oracle's compositor, gradient (linear/radial/conic; linear-RGB,
sRGB and HSL interpolation), dither (Bayer, blue noise) and gamma paths
against the reference's fixtures spec/files/061-072.  Every case matches
bit for bit, including the sRGB / HSL / conic ones that go through the
oracle's restatements of std.math.pow and atan2."""
import math
import os

import numpy as np
import pytest

from z2d_amd import abi
from oracle_backend import OracleBackend, lib, png_scanlines
from pngfix import export_surface, png_pixels, png_stream
from spec_comp_scenes import COMP_SCENES

GOLD = os.path.join(os.path.dirname(__file__), "golden", "spec")
AA = [(abi.AA_NONE, "_pixelated"), (abi.AA_SUPERSAMPLE_4X, "_smooth"),
      (abi.AA_MULTISAMPLE_4X, "_smooth_multisample")]


PROFILES = {"071_gamma_linear": abi.PNG_PROFILE_LINEAR, "072_gamma_srgb": abi.PNG_PROFILE_SRGB}


def srgb_export(img):
    """export_png's sRGB profile step: pow(c/255, 1/2.2), round(255*x)."""
    f32 = np.float32
    ig = float(f32(1.0) / f32(2.2))
    lut = np.array([int(math.floor(float(f32(255.0) * f32(lib().zo_powf(float(f32(v) / f32(255.0)),
                                                                          ig))) + 0.5))
                    for v in range(256)], dtype=np.uint8)
    return lut[img]


def cases():
    out = []
    for name, fn, aa_dep in COMP_SCENES:
        if aa_dep:
            for aa, suffix in AA:
                out.append((name, fn, aa, suffix))
        else:
            out.append((name, fn, None, ""))
    return out


def fixture(name, suffix):
    p = os.path.join(GOLD, name + suffix + ".png")
    if not os.path.exists(p) and suffix == "_smooth_multisample":
        p = os.path.join(GOLD, name + "_smooth.png")
    return p


@pytest.mark.parametrize("name,fn,aa,suffix", cases(), ids=lambda v: v if isinstance(v, str) else "")
def test_oracle_comp_fixture(name, fn, aa, suffix):
    out = fn(OracleBackend(), aa)
    got = export_surface(out.format, out.width, out.height, out.data())
    if name == "072_gamma_srgb":
        got = srgb_export(got)
    want = png_pixels(fixture(name, suffix))
    assert got.shape == want.shape
    d = np.abs(got.astype(np.int32) - want.astype(np.int32))
    assert d.max(initial=0) == 0, f"max diff {d.max()} at {np.argwhere(d > 0)[:3].tolist()}"
    # the export's scanline stream, byte for byte, with the profile the spec
    # case exports with (spec/071_gamma_linear.zig:29, 072_gamma_srgb.zig:40)
    profile = PROFILES.get(name, abi.PNG_PROFILE_NONE)
    assert png_scanlines(out.format, out.width, out.height, out.data(), profile) == \
        png_stream(fixture(name, suffix))
