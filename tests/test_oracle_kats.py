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
  color interpolation, every space and the
    four HSL polar modes                    src/color.zig:788-1779
  fill_plotter.plot, degenerate nodes       src/internal/tess/fill_plotter.zig:121-209
  isClosedNodeSet (oracle and the product's
    host check through the C ABI)           src/internal/path_nodes.zig:63-137
  multisample.addSpan                       src/internal/raster/multisample.zig:281-428

The SparseCoverageBuffer tests (src/internal/raster/sparse_coverage.zig:
214-361) check the run-length layout of the reference's coverage buffer
with all-zero values; the oracle keeps coverage as a dense per-pixel array,
so what they pin here is the addSpan result per pixel, checked below.

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
    L.zo_interpolate_encode.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(abi.Color),
                                        C.POINTER(abi.Color), C.c_float]
    L.zo_interpolate_encode.restype = abi.Pixel
    L.zo_msaa_add_span.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32]
    L.zo_msaa_add_span.restype = None
    L.zo_is_closed_node_set.argtypes = [C.c_void_p, C.c_uint64]
    L.zo_is_closed_node_set.restype = C.c_int
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
    """[space, [v0..v3]]; a component ["powf", x, y] is std.math.pow(f32, x, y)
    (the oracle's restatement of it)."""
    space, v = spec
    vals = [float(L.zo_powf(np.float32(c[1]), np.float32(c[2]))) if isinstance(c, list) else c
            for c in v]
    return abi.color(SPACE[space], *vals)


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


@pytest.mark.parametrize("case", KATS["interpolate"]["cases"],
                         ids=_ids(KATS["interpolate"]["cases"]))
def test_kat_interpolate_encode(case):
    """interpolateEncode(a, b, t) == LinearRGB.fromColor(expected).encodeRGBA()."""
    a, b, e = color(case["a"]), color(case["b"]), color(case["expected"])
    got = L.zo_interpolate_encode(METHOD[case["method"]], POLAR[case["polar"]], C.byref(a),
                                  C.byref(b), case["t"])
    assert px_tuple(got) == px_tuple(L.zo_pixel_from_color(C.byref(e)))


def _nodes(spec):
    tags = {"move_to": abi.MOVE_TO, "line_to": abi.LINE_TO, "close_path": abi.CLOSE_PATH}
    return abi.nodes_array([(tags[n[0]], *n[1:]) for n in spec])


@pytest.mark.parametrize("case", KATS["fill_plotter"]["cases"],
                         ids=_ids(KATS["fill_plotter"]["cases"], "n_edges"))
def test_kat_fill_plotter(case):
    from oracle_backend import plot_fill
    edges, _ = plot_fill(_nodes(case["nodes"]), case["scale"], case["tolerance"])
    assert len(edges) == case["n_edges"]
    got = [[float(e[k]) for k in ("y0", "y1", "x_start", "x_inc")] for e in edges]
    assert got == case["edges"]


@pytest.mark.parametrize("case", KATS["closed_node_set"]["cases"],
                         ids=_ids(KATS["closed_node_set"]["cases"], "expected"))
def test_kat_closed_node_set(case):
    nodes = _nodes(case["nodes"]) if case["nodes"] else np.zeros(0, abi.NODE_DTYPE)
    assert bool(L.zo_is_closed_node_set(nodes.ctypes.data, len(nodes))) == case["expected"]


@pytest.mark.parametrize("case", KATS["closed_node_set"]["cases"],
                         ids=_ids(KATS["closed_node_set"]["cases"], "expected"))
def test_kat_closed_node_set_product(case):
    """The library's host-side painter.fill checks (no device needed): an
    unclosed set is PathNotClosed before the context is looked at; a closed
    one passes that check and stops at the missing context; no nodes is the
    reference's silent no-op (src/painter.zig:73-82)."""
    from z2d_amd import gpu
    nodes = _nodes(case["nodes"]) if case["nodes"] else np.zeros(1, abi.NODE_DTYPE)
    buf = np.zeros(16 * 16 * 4, np.uint8)
    d = abi.SurfaceDesc(abi.RGBA, 16, 16, 0, buf.ctypes.data, buf.nbytes)
    pat = abi.solid_pattern(abi.rgba(1, 2, 3, 255))
    opts = abi.fill_opts()
    rc = gpu.lib().z2d_gpu_fill(None, C.byref(d), C.byref(pat), nodes.ctypes.data,
                                len(case["nodes"]), C.byref(opts))
    if not case["nodes"]:
        assert rc == abi.OK
    else:
        assert rc == (abi.ERR_NO_CONTEXT if case["expected"] else abi.ERR_PATH_NOT_CLOSED)


def _cov_runs(cov, n):
    """Maximal runs of equal coverage over [0, n): [(start, value, length)]."""
    runs, i = [], 0
    while i < n:
        j = i
        while j < n and cov[j] == cov[i]:
            j += 1
        runs.append((i, int(cov[i]), j - i))
        i = j
    return runs


def test_kat_msaa_add_span():
    """multisample.zig:281-314: four nested spans of a 100-px cross-section."""
    cov = np.zeros(1024 + 1, np.uint8)
    L.zo_msaa_add_span(cov.ctypes.data, 200, 400)
    assert (cov[50], cov[149]) == (4, 4) and cov[50:150].tolist() == [4] * 100
    L.zo_msaa_add_span(cov.ctypes.data, 201, 398)
    assert (cov[50], cov[51:149].tolist(), cov[149]) == (7, [8] * 98, 7)
    L.zo_msaa_add_span(cov.ctypes.data, 202, 396)
    assert (cov[50], cov[51:149].tolist(), cov[149]) == (9, [12] * 98, 9)
    L.zo_msaa_add_span(cov.ctypes.data, 203, 394)
    assert (cov[50], cov[51:149].tolist(), cov[149]) == (10, [16] * 98, 10)
    # four runs over the touched range: zeros, 10, 16 x 98, 10
    assert len(_cov_runs(cov, 150)) == 4


@pytest.mark.parametrize("cap", [255, 256, 65535, 65536, 131072])
def test_kat_msaa_add_span_capacity(cap):
    """multisample.zig:316-428: spans over the whole capacity, each later one
    one sample further in at both ends."""
    cov = np.zeros(cap + 1, np.uint8)
    full = cap * 4
    L.zo_msaa_add_span(cov.ctypes.data, 0, full)
    assert _cov_runs(cov, cap) == [(0, 4, cap)]
    for k, (edge, mid) in enumerate(((7, 8), (9, 12), (10, 16)), start=1):
        L.zo_msaa_add_span(cov.ctypes.data, k, full - 2 * k)
        assert _cov_runs(cov, cap) == [(0, edge, 1), (1, mid, cap - 2), (cap - 1, edge, 1)]


def test_kat_counts():
    """Every table the generator reads is present and non-trivial."""
    want = {"compositor_pairs": 76, "compositor_ops": 102, "pixel_from_pixel": 64,
            "pixel_from_color": 6, "dither": 6, "gradient_offset": 12, "gradient_pixel": 17,
            "stop_search": 18, "in_box": 24, "interpolate": 69, "fill_plotter": 3,
            "closed_node_set": 6}
    for k, n in want.items():
        assert len(KATS[k]["cases"]) == n, k
