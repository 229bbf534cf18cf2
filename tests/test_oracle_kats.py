"""The oracle against the reference's own known-answer tests, read from its
Zig unit tests into tests/golden/kats.json (by tests/golden/make_kats.py):

  compositor runPixel over format pairs     src/compositor.zig:2452-3076
  all 28 operators, integer and float       src/compositor.zig:3078-3860
  T.fromPixel between every format pair     src/pixel.zig:845-1168
  Pixel.fromColor                           src/pixel.zig:1371-1415
  Dither.getPixel (incl. scale 2)           src/Dither.zig:164-266
  Linear/Radial/Conic.getOffset             src/gradient.zig:1200-1766
  Linear/Radial/Conic.getPixel (incl. HSL)  src/gradient.zig:1252-1779
  Stop.List.search (incl. hard stops)       src/gradient.zig:963-1146
  Polygon.inBox                             src/internal/tess/Polygon.zig:452-784

Every vector is exact (the reference uses expectEqual / expectEqualDeep),
except Stop.List.search offsets, compared within floatEps(f32) as there."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from z2d_amd import abi
from oracle_backend import lib

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "kats.json")) as _f:
    KATS = json.load(_f)

FMT = {n: i for i, n in enumerate(abi.FORMAT_NAMES)}
SPACE = {"linear_rgb": abi.COLOR_LINEAR_RGB, "srgb": abi.COLOR_SRGB, "hsl": abi.COLOR_HSL}
PREC = {"integer": abi.PRECISION_INTEGER, "float": abi.PRECISION_FLOAT}
DITHER = {"none": abi.DITHER_NONE, "bayer": abi.DITHER_BAYER,
          "blue_noise": abi.DITHER_BLUE_NOISE}
GTYPE = {"linear": abi.GRADIENT_LINEAR, "radial": abi.GRADIENT_RADIAL,
         "conic": abi.GRADIENT_CONIC}
METHOD = {"linear_rgb": abi.INTERP_LINEAR_RGB, "srgb": abi.INTERP_SRGB, "hsl": abi.INTERP_HSL}
POLAR = {"shorter": abi.POLAR_SHORTER, "longer": abi.POLAR_LONGER,
         "increasing": abi.POLAR_INCREASING, "decreasing": abi.POLAR_DECREASING}


def _setup():
    L = lib()
    L.zo_pixel_convert.argtypes = [C.c_uint32, abi.Pixel]
    L.zo_pixel_convert.restype = abi.Pixel
    L.zo_pixel_multiply.argtypes = [abi.Pixel]
    L.zo_pixel_multiply.restype = abi.Pixel
    L.zo_pixel_from_color.argtypes = [C.POINTER(abi.Color)]
    L.zo_pixel_from_color.restype = abi.Pixel
    L.zo_stop_search.argtypes = [C.POINTER(abi.GradientDesc), C.c_float, C.POINTER(abi.Color),
                                 C.POINTER(abi.Color), C.POINTER(C.c_float)]
    L.zo_stop_search.restype = None
    return L


L = _setup()


def px(spec):
    fmt, r, g, b, a = spec
    return abi.Pixel(FMT[fmt], r, g, b, a)


def px_tuple(p):
    """Comparable form: channels the format does not have read as 0."""
    f = p.format
    if f in (abi.ARGB, abi.RGBA):
        return (f, p.r, p.g, p.b, p.a)
    if f in (abi.XRGB, abi.RGB):
        return (f, p.r, p.g, p.b, 0)
    return (f, 0, 0, 0, p.a)


def chain(c):
    p = px(c["base"])
    for op in c["ops"]:
        if op[0] == "convert":
            p = L.zo_pixel_convert(FMT[op[1]], p)
        elif op[0] == "multiply":
            p = L.zo_pixel_multiply(p)
        elif op[0] == "set_a":
            p.a = op[1]
        else:
            raise ValueError(op)
    return p


def color(spec):
    space, v = spec
    return abi.color(SPACE[space], *v)


def _ids(cases, key="name"):
    return [f"{i}-{c.get(key, '')}" for i, c in enumerate(cases)]


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", KATS["pixel_from_pixel"]["cases"],
                         ids=_ids(KATS["pixel_from_pixel"]["cases"]))
def test_kat_from_pixel(case):
    """src/pixel.zig:845-1168: @TypeOf(a).fromPixel(b) == a for each pair."""
    a, b = px(case["a"]), px(case["b"])
    got = L.zo_pixel_convert(a.format, b)
    assert px_tuple(got) == px_tuple(a)


@pytest.mark.parametrize("case", KATS["pixel_from_color"]["cases"],
                         ids=_ids(KATS["pixel_from_color"]["cases"]))
def test_kat_pixel_from_color(case):
    c = color(case["color"])
    assert px_tuple(L.zo_pixel_from_color(C.byref(c))) == px_tuple(px(case["expected"]))


@pytest.mark.parametrize("case", KATS["compositor_pairs"]["cases"],
                         ids=_ids(KATS["compositor_pairs"]["cases"], "op"))
def test_kat_compositor_format_pairs(case):
    """runPixel(.integer, dst, src, op) over the format cross-product."""
    dst, src = chain(case["dst"]), chain(case["src"])
    want = chain(case["expected"])
    got = L.zo_run_pixel(PREC[case["precision"]], dst, src, abi.OP[case["op"]])
    assert px_tuple(got) == px_tuple(want), (px_tuple(dst), px_tuple(src))


@pytest.mark.parametrize("case", KATS["compositor_ops"]["cases"],
                         ids=_ids(KATS["compositor_ops"]["cases"]))
def test_kat_compositor_all_operators(case):
    """runPixel(precision, Pixel.fromColor(bg), Pixel.fromColor(fg), op)."""
    bg, fg = color(case["bg"]), color(case["fg"])
    dst = L.zo_pixel_from_color(C.byref(bg))
    src = L.zo_pixel_from_color(C.byref(fg))
    got = L.zo_run_pixel(PREC[case["precision"]], dst, src, abi.OP[case["op"]])
    assert px_tuple(got) == px_tuple(px(case["expected"]))


def _gradient(spec, matrix=abi.IDENTITY):
    stops = [(off, color(c)) for off, c in spec.get("stops", [])]
    return abi.Gradient(GTYPE[spec["type"]], spec["geom"], stops,
                        METHOD[spec.get("method", "linear_rgb")],
                        POLAR[spec.get("polar", "shorter")], matrix)


@pytest.mark.parametrize("case", KATS["dither"]["cases"], ids=_ids(KATS["dither"]["cases"]))
def test_kat_dither(case):
    d = abi.DitherDesc()
    d.type = DITHER[case["type"]]
    d.scale = case["scale"]
    keep = None
    s = case["source"]
    if "pixel" in s:
        d.source = abi.DITHER_SRC_PIXEL
        d.pixel = px(s["pixel"])
    elif "color" in s:
        d.source = abi.DITHER_SRC_COLOR
        d.color = color(s["color"])
    else:
        d.source = abi.DITHER_SRC_GRADIENT
        keep = _gradient(s["gradient"])
        d.gradient = C.pointer(keep.desc)
    got = lib().zo_dither_pixel(C.byref(d), case["x"], case["y"])
    assert px_tuple(got) == px_tuple(px(case["expected"]))


def _inverse_of_scale(m):
    """Transformation.inverse of the diagonal matrices these tests use: the
    gradient stores the inverse of the matrix given to setTransformation."""
    ax, by, cx, dy, tx, ty = m
    assert by == 0 and cx == 0 and tx == 0 and ty == 0
    return (1.0 / ax, 0.0, 0.0, 1.0 / dy, 0.0, 0.0)


@pytest.mark.parametrize("case", KATS["gradient_offset"]["cases"],
                         ids=_ids(KATS["gradient_offset"]["cases"]))
def test_kat_gradient_offset(case):
    g = _gradient(case, _inverse_of_scale(case["matrix_given"]))
    got = lib().zo_gradient_offset(C.byref(g.desc), case["x"], case["y"])
    assert np.float32(got) == np.float32(case["expected"])


@pytest.mark.parametrize("case", KATS["gradient_pixel"]["cases"],
                         ids=_ids(KATS["gradient_pixel"]["cases"], "x"))
def test_kat_gradient_pixel(case):
    g = _gradient(case["gradient"])
    got = lib().zo_gradient_pixel(C.byref(g.desc), case["x"], case["y"])
    assert px_tuple(got) == px_tuple(px(case["expected"]))


@pytest.mark.parametrize("case", KATS["stop_search"]["cases"],
                         ids=_ids(KATS["stop_search"]["cases"], "t"))
def test_kat_stop_search(case):
    g = _gradient({"type": "linear", "geom": [0, 0, 1, 1], "stops": case["stops"]})
    c0, c1, t = abi.Color(), abi.Color(), C.c_float()
    L.zo_stop_search(C.byref(g.desc), case["t"], C.byref(c0), C.byref(c1), C.byref(t))
    for got, want in ((c0, case["c0"]), (c1, case["c1"])):
        w = color(want)
        assert got.space == w.space
        assert [np.float32(v) for v in got.v] == [np.float32(v) for v in w.v]
    assert abs(np.float32(t.value) - np.float32(case["offset"])) <= case["tol"]


@pytest.mark.parametrize("case", KATS["in_box"]["cases"], ids=_ids(KATS["in_box"]["cases"]))
def test_kat_in_box(case):
    ext = (C.c_double * 4)(*case["extents"])
    got = lib().zo_in_box(ext, case["scale"], case["width_arg"], case["height_arg"])
    assert bool(got) == case["expected"]


def test_kat_counts():
    """Every table the generator reads is present and non-trivial."""
    want = {"compositor_pairs": 76, "compositor_ops": 102, "pixel_from_pixel": 64,
            "pixel_from_color": 6, "dither": 6, "gradient_offset": 12, "gradient_pixel": 17,
            "stop_search": 18, "in_box": 24}
    for k, n in want.items():
        assert len(KATS[k]["cases"]) == n, k
