"""ctypes mirror of include/z2d_types.h: the POD descriptors that cross the
backend's C ABI.  Enum values keep the reference's Zig declaration order
(src/pixel.zig:47, src/compositor.zig:46-158, src/options.zig:23-91)."""
import ctypes as C

import numpy as np

# pixel.Format (src/pixel.zig:47-56)
ARGB, XRGB, RGB, RGBA, ALPHA8, ALPHA4, ALPHA2, ALPHA1 = range(8)
FORMAT_NAMES = ["argb", "xrgb", "rgb", "rgba", "alpha8", "alpha4", "alpha2", "alpha1"]
FORMAT_BITS = [32, 32, 32, 32, 8, 4, 2, 1]

# compositor.Operator (src/compositor.zig:46-158)
OPERATORS = [
    "clear", "src", "dst", "src_over", "dst_over", "src_in", "dst_in", "src_out",
    "dst_out", "src_atop", "dst_atop", "xor", "plus", "multiply", "screen",
    "overlay", "darken", "lighten", "color_dodge", "color_burn", "hard_light",
    "soft_light", "difference", "exclusion", "hue", "saturation", "color",
    "luminosity",
]
OP = {name: i for i, name in enumerate(OPERATORS)}
FLOAT_ONLY_OPS = {"color_dodge", "color_burn", "soft_light", "hue", "saturation",
                  "color", "luminosity"}
UNBOUNDED_OPS = {"src_in", "dst_in", "src_out", "dst_atop"}

PRECISION_INTEGER, PRECISION_FLOAT = 0, 1
NON_ZERO, EVEN_ODD = 0, 1
JOIN_MITER, JOIN_ROUND, JOIN_BEVEL = 0, 1, 2
CAP_BUTT, CAP_ROUND, CAP_SQUARE = 0, 1, 2
AA_NONE, AA_DEFAULT, AA_MULTISAMPLE_4X, AA_SUPERSAMPLE_4X = 0, 1, 2, 3
MOVE_TO, LINE_TO, CURVE_TO, CLOSE_PATH = 0, 1, 2, 3
GRADIENT_LINEAR, GRADIENT_RADIAL, GRADIENT_CONIC = 0, 1, 2
INTERP_LINEAR_RGB, INTERP_SRGB, INTERP_HSL = 0, 1, 2
POLAR_SHORTER, POLAR_LONGER, POLAR_INCREASING, POLAR_DECREASING = 0, 1, 2, 3
COLOR_LINEAR_RGB, COLOR_SRGB, COLOR_HSL = 0, 1, 2
DITHER_NONE, DITHER_BAYER, DITHER_BLUE_NOISE = 0, 1, 2
DITHER_SRC_PIXEL, DITHER_SRC_COLOR, DITHER_SRC_GRADIENT = 0, 1, 2
PATTERN_OPAQUE, PATTERN_GRADIENT, PATTERN_DITHER = 0, 1, 2
# WriteToPNGFileOptions.color_profile (src/export_png.zig:33-41): none = null
PNG_PROFILE_NONE, PNG_PROFILE_LINEAR, PNG_PROFILE_SRGB = 0, 1, 2
PARAM_NONE, PARAM_DITHER, PARAM_GRADIENT, PARAM_PIXEL, PARAM_SURFACE = range(5)

# status codes (include/z2d_types.h)
OK = 0
ERR_PATH_NOT_CLOSED = 1
ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED = 2
ERR_INVALID_WIDTH = 3
ERR_INVALID_HEIGHT = 4
ERR_INVALID_STATE = 5
ERR_OUT_OF_MEMORY = 6
ERR_INVALID_MATRIX = 7
ERR_INVALID_ARGUMENT = -1
ERR_DEVICE = -2
ERR_BUDGET_EXCEEDED = -3
ERR_NO_CONTEXT = -4
ERR_UNSUPPORTED = -5


class SurfaceDesc(C.Structure):
    _fields_ = [("format", C.c_uint32), ("width", C.c_int32), ("height", C.c_int32),
                ("reserved", C.c_uint32), ("data", C.c_void_p), ("len", C.c_uint64)]


class Pixel(C.Structure):
    _fields_ = [("format", C.c_uint32), ("r", C.c_uint8), ("g", C.c_uint8),
                ("b", C.c_uint8), ("a", C.c_uint8)]


class Color(C.Structure):
    _fields_ = [("space", C.c_uint32), ("v", C.c_float * 4)]


class Stop(C.Structure):
    _fields_ = [("offset", C.c_float), ("color", Color)]


class GradientDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("method", C.c_uint32), ("polar", C.c_uint32),
                ("n_stops", C.c_uint32), ("geom", C.c_double * 6),
                ("matrix", C.c_double * 6), ("stops", C.POINTER(Stop))]


class DitherDesc(C.Structure):
    _fields_ = [("type", C.c_uint32), ("source", C.c_uint32), ("pixel", Pixel),
                ("color", Color), ("scale", C.c_uint32),
                ("gradient", C.POINTER(GradientDesc))]


class PatternDesc(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pixel", Pixel),
                ("gradient", C.POINTER(GradientDesc)), ("dither", DitherDesc)]


class PathNode(C.Structure):
    _fields_ = [("tag", C.c_uint32), ("reserved", C.c_uint32), ("v", C.c_double * 6)]


class Edge(C.Structure):
    _fields_ = [("y0", C.c_double), ("y1", C.c_double), ("x_start", C.c_double),
                ("x_inc", C.c_double)]


class FillOpts(C.Structure):
    _fields_ = [("anti_aliasing_mode", C.c_uint32), ("fill_rule", C.c_uint32),
                ("op", C.c_uint32), ("precision", C.c_uint32), ("tolerance", C.c_double)]


class StrokeOpts(C.Structure):
    _fields_ = [("anti_aliasing_mode", C.c_uint32), ("line_cap_mode", C.c_uint32),
                ("line_join_mode", C.c_uint32), ("op", C.c_uint32),
                ("precision", C.c_uint32), ("hairline", C.c_uint32),
                ("dashes", C.POINTER(C.c_double)), ("n_dashes", C.c_uint64),
                ("dash_offset", C.c_double), ("line_width", C.c_double),
                ("miter_limit", C.c_double), ("tolerance", C.c_double),
                ("transformation", C.c_double * 6)]


class CompParam(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pixel", Pixel),
                ("gradient", C.POINTER(GradientDesc)), ("dither", DitherDesc),
                ("surface", C.POINTER(SurfaceDesc))]


class CompOp(C.Structure):
    _fields_ = [("op", C.c_uint32), ("dst", CompParam), ("src", CompParam)]


class FillItem(C.Structure):
    _fields_ = [("pattern", C.POINTER(PatternDesc)), ("nodes", C.c_void_p),
                ("n_nodes", C.c_uint64), ("opts", FillOpts)]


DRAW_FILL, DRAW_STROKE = 0, 1


class DrawItem(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("pattern", C.POINTER(PatternDesc)), ("nodes", C.c_void_p),
                ("n_nodes", C.c_uint64), ("fill", FillOpts), ("stroke", C.POINTER(StrokeOpts))]


class RequestDesc(C.Structure):
    _fields_ = [("format", C.c_uint32), ("width", C.c_int32), ("height", C.c_int32),
                ("clear", Pixel), ("clip_mask", C.c_void_p), ("fills", C.POINTER(FillItem)),
                ("n_fills", C.c_uint64), ("draws", C.POINTER(DrawItem)), ("n_draws", C.c_uint64),
                ("encode", C.c_uint32), ("reserved", C.c_uint32),
                ("out_written", C.POINTER(C.c_uint64))]


IDENTITY = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)

NODE_DTYPE = np.dtype([("tag", "<u4"), ("reserved", "<u4"), ("v", "<f8", (6,))])
EDGE_DTYPE = np.dtype([("y0", "<f8"), ("y1", "<f8"), ("x_start", "<f8"), ("x_inc", "<f8")])
assert NODE_DTYPE.itemsize == C.sizeof(PathNode) == 56
assert EDGE_DTYPE.itemsize == C.sizeof(Edge) == 32


def surface_nbytes(fmt, w, h):
    bits = FORMAT_BITS[fmt]
    if bits == 32:
        return w * h * 4
    return (w * h * bits + 7) // 8


def pixel(fmt, r=0, g=0, b=0, a=0):
    return Pixel(fmt, r, g, b, a)


def rgba(r, g, b, a=255):
    return Pixel(RGBA, r, g, b, a)


def solid_pattern(px):
    p = PatternDesc()
    p.kind = PATTERN_OPAQUE
    p.pixel = px
    return p


def color(space, v0, v1, v2, v3=1.0):
    c = Color()
    c.space = space
    c.v[:] = (v0, v1, v2, v3)
    return c


class Gradient:
    """Owns a GradientDesc and its stop array (keeps ctypes memory alive)."""

    def __init__(self, gtype, geom, stops, method=INTERP_LINEAR_RGB, polar=POLAR_SHORTER,
                 matrix=IDENTITY):
        self.stop_arr = (Stop * max(1, len(stops)))()
        for i, (off, col) in enumerate(stops):
            self.stop_arr[i].offset = off
            self.stop_arr[i].color = col
        self.desc = GradientDesc()
        self.desc.type = gtype
        self.desc.method = method
        self.desc.polar = polar
        self.desc.n_stops = len(stops)
        g = list(geom) + [0.0] * (6 - len(geom))
        self.desc.geom[:] = g
        self.desc.matrix[:] = matrix
        self.desc.stops = C.cast(self.stop_arr, C.POINTER(Stop))

    def pattern(self):
        p = PatternDesc()
        p.kind = PATTERN_GRADIENT
        p.gradient = C.pointer(self.desc)
        return p


def dither_pattern(dtype, scale, px=None, col=None, gradient=None):
    p = PatternDesc()
    p.kind = PATTERN_DITHER
    p.dither.type = dtype
    p.dither.scale = scale
    if px is not None:
        p.dither.source = DITHER_SRC_PIXEL
        p.dither.pixel = px
    elif col is not None:
        p.dither.source = DITHER_SRC_COLOR
        p.dither.color = col
    else:
        p.dither.source = DITHER_SRC_GRADIENT
        p.dither.gradient = C.pointer(gradient.desc)
    return p


def nodes_array(nodes):
    """nodes: list of (tag, v...) tuples -> numpy structured array."""
    arr = np.zeros(len(nodes), dtype=NODE_DTYPE)
    arr["tag"] = [nd[0] for nd in nodes]
    arr["v"] = [tuple(nd[1:]) + (0.0,) * (7 - len(nd)) for nd in nodes]
    return arr


def fill_opts(aa=AA_DEFAULT, rule=NON_ZERO, op="src_over", precision=PRECISION_INTEGER,
              tolerance=0.1):
    o = FillOpts()
    o.anti_aliasing_mode = aa
    o.fill_rule = rule
    o.op = OP[op] if isinstance(op, str) else op
    o.precision = precision
    o.tolerance = tolerance
    return o


class TextOpts(C.Structure):
    """text.ShowTextOptions: transformation {ax, by, cx, dy, tx, ty}, size, fill options."""
    _fields_ = [("transformation", C.c_double * 6), ("size", C.c_double), ("fill", FillOpts)]


def text_opts(size=16.0, transformation=(1.0, 0.0, 0.0, 1.0, 0.0, 0.0), fill=None):
    o = TextOpts()
    for i, v in enumerate(transformation):
        o.transformation[i] = v
    o.size = size
    o.fill = fill if fill is not None else fill_opts()
    return o
