"""Test-scene builder: a small Python restatement of the reference's managed
Context/Path front end (src/Context.zig, src/Path.zig,
src/Transformation.zig) that produces the same PathNode arrays, patterns and
options the painter receives.  Scenes recorded with it replay on any backend
(the oracle or the HIP library) through `Scene.run(backend)`."""
import ctypes as C
import math
from fractions import Fraction

import numpy as np

from z2d_amd import abi

I24_MIN, I24_MAX = -(1 << 23), (1 << 23) - 1


def cf(x):
    """Comptime-float arithmetic helper: exact rational -> nearest f64 (Zig
    evaluates comptime float expressions at f128 and coerces once)."""
    return float(Fraction(x))


class Transformation:
    """src/Transformation.zig (identity, mul, translate, scale, rotate,
    inverse, userToDevice)."""

    def __init__(self, ax=1.0, by=0.0, cx=0.0, dy=1.0, tx=0.0, ty=0.0):
        self.ax, self.by, self.cx, self.dy, self.tx, self.ty = ax, by, cx, dy, tx, ty

    def tuple(self):
        return (self.ax, self.by, self.cx, self.dy, self.tx, self.ty)

    def is_identity(self):
        return self.tuple() == abi.IDENTITY

    def mul(self, b):
        a = self
        return Transformation(
            ax=a.ax * b.ax + a.by * b.cx,
            by=a.ax * b.by + a.by * b.dy,
            cx=a.cx * b.ax + a.dy * b.cx,
            dy=a.cx * b.by + a.dy * b.dy,
            tx=a.ax * b.tx + a.by * b.ty + a.tx,
            ty=a.cx * b.tx + a.dy * b.ty + a.ty,
        )

    def translate(self, tx, ty):
        return self.mul(Transformation(tx=float(tx), ty=float(ty)))

    def scale(self, sx, sy):
        return self.mul(Transformation(ax=float(sx), dy=float(sy)))

    def rotate(self, angle):
        s, c = math.sin(angle), math.cos(angle)
        return self.mul(Transformation(ax=c, by=-s, cx=s, dy=c))

    def inverse(self):
        a = self
        if a.by == 0 and a.cx == 0:
            if a.ax == 0 or a.dy == 0:
                raise ValueError("InvalidMatrix")
            if a.ax != 1 or a.dy != 1:
                return Transformation(1 / a.ax, 0.0, 0.0, 1 / a.dy, -a.tx / a.ax, -a.ty / a.dy)
            return Transformation(1.0, 0.0, 0.0, 1.0, -a.tx, -a.ty)
        det = a.ax * a.dy - a.by * a.cx
        if det == 0:
            raise ValueError("InvalidMatrix")
        adj = (a.dy, -a.by, -a.cx, a.ax, a.by * a.ty - a.dy * a.tx, a.cx * a.tx - a.ax * a.ty)
        inv = 1 / det
        ax, by, cx, dy, tx, ty = adj
        return Transformation(ax * inv, by * inv, cx * inv, dy * inv, tx * inv, ty * inv)

    def user_to_device(self, x, y):
        ox = self.ax * x + self.by * y
        oy = self.cx * x + self.dy * y
        return ox + self.tx, oy + self.ty

    def device_to_user_distance(self, x, y):
        """deviceToUserDistance: the inverse's userToDeviceDistance."""
        inv = self.inverse()
        return inv.ax * x + inv.by * y, inv.cx * x + inv.dy * y


def clamp_i24(v):
    return min(max(float(v), float(I24_MIN)), float(I24_MAX))


_ARC_TABLE = [(math.pi / 1.0, 0.0185185185185185036127),
              (math.pi / 2.0, 0.000272567143730179811158),
              (math.pi / 3.0, 2.38647043651461047433e-05),
              (math.pi / 4.0, 4.2455377443222443279e-06),
              (math.pi / 5.0, 1.11281001494389081528e-06),
              (math.pi / 6.0, 3.72662000942734705475e-07),
              (math.pi / 7.0, 1.47783685574284411325e-07),
              (math.pi / 8.0, 6.63240432022601149057e-08),
              (math.pi / 9.0, 3.2715520137536980553e-08),
              (math.pi / 10.0, 1.73863223499021216974e-08),
              (math.pi / 11.0, 9.81410988043554039085e-09)]


def _arc_max_angle(tolerance):
    """arc_max_angle_for_tolerance_normalized (src/internal/arc.zig)."""
    for angle, err in _ARC_TABLE:
        if err < tolerance:
            return angle
    angle = 0.0
    for i in range(len(_ARC_TABLE), 1000):
        angle = math.pi / float(i)
        err = 2.0 / 27.0 * math.pow(math.sin(angle / 4), 6) / math.pow(math.cos(angle / 4), 2)
        if err <= tolerance:
            break
    return angle


def _circle_major_axis(m, radius):
    """transformed_circle_major_axis (src/internal/arc.zig)."""
    eps = 0.00390625
    det = m.ax * m.dy - m.by * m.cx
    if abs(det * det - 1.0) < eps:
        if abs(m.by) < eps and abs(m.cx) < eps:
            return radius
        if abs(m.ax) < eps and abs(m.dy) < eps:
            return radius
    i = m.ax * m.ax + m.by * m.by
    j = m.cx * m.cx + m.dy * m.dy
    f = 0.5 * (i + j)
    g = 0.5 * (i - j)
    h = m.ax * m.cx + m.by * m.dy
    return radius * math.sqrt(f + math.hypot(g, h))


def _arc_segment(path, xc, yc, radius, a, b):
    rsa, rca = radius * math.sin(a), radius * math.cos(a)
    rsb, rcb = radius * math.sin(b), radius * math.cos(b)
    h = 4.0 / 3.0 * math.tan((b - a) / 4.0)
    path.curve_to(xc + rca - h * rsa, yc + rsa + h * rca, xc + rcb + h * rsb,
                  yc + rsb - h * rcb, xc + rcb, yc + rsb)


def _arc_in_direction(path, xc, yc, radius, amin, amax, forward, ctm, tolerance):
    """arc_in_direction (src/internal/arc.zig)."""
    if not (amax * amax >= 0.0) or not (amin * amin >= 0.0):
        return
    full = 2 * math.pi * 65536.0
    if amax - amin > full:
        amax = math.fmod(amax - amin, 2 * math.pi)
        amin = math.fmod(amin, 2 * math.pi)
        amax += amin + full
    if amax - amin > math.pi:
        amid = amin + (amax - amin) / 2.0
        halves = ((amin, amid), (amid, amax)) if forward else ((amid, amax), (amin, amid))
        for lo, hi in halves:
            _arc_in_direction(path, xc, yc, radius, lo, hi, forward, ctm, tolerance)
    elif amax != amin:
        major = _circle_major_axis(ctm, radius)
        segments = int(math.ceil(abs(amax - amin) / _arc_max_angle(tolerance / major)))
        step = (amax - amin) / float(segments)
        segments -= 1
        if not forward:
            amin, amax = amax, amin
            step = -step
        path._arc_line_to(xc + radius * math.cos(amin), yc + radius * math.sin(amin))
        for _ in range(max(0, segments)):
            _arc_segment(path, xc, yc, radius, amin, amin + step)
            amin += step
        _arc_segment(path, xc, yc, radius, amin, amax)
    else:
        path._arc_line_to(xc + radius * math.cos(amin), yc + radius * math.sin(amin))


class Path:
    """src/Path.zig:116-262, :447-477, :570-572."""

    def __init__(self):
        self.nodes = []
        self.initial = None
        self.current = None
        self.transformation = Transformation()

    def reset(self):
        self.nodes = []
        self.initial = None
        self.current = None

    def _pt(self, x, y):
        return self.transformation.user_to_device(clamp_i24(x), clamp_i24(y))

    def move_to(self, x, y):
        p = self._pt(x, y)
        if self.nodes and self.nodes[-1][0] == abi.MOVE_TO and self.nodes[-1][1:3] == p:
            return
        self.nodes.append((abi.MOVE_TO, p[0], p[1]))
        self.initial = p
        self.current = p

    def line_to(self, x, y):
        if self.current is None:
            return self.move_to(x, y)
        p = self._pt(x, y)
        self.nodes.append((abi.LINE_TO, p[0], p[1]))
        self.current = p

    def rel_line_to(self, dx, dy):
        """Path.relLineTo (src/Path.zig:187-192)."""
        if self.current is None:
            raise ValueError("NoCurrentPoint")
        ux, uy = self.transformation.inverse().user_to_device(*self.current)
        self.line_to(ux + dx, uy + dy)

    def curve_to(self, x1, y1, x2, y2, x3, y3):
        if self.current is None:
            raise ValueError("NoCurrentPoint")
        p1, p2, p3 = self._pt(x1, y1), self._pt(x2, y2), self._pt(x3, y3)
        self.nodes.append((abi.CURVE_TO, p1[0], p1[1], p2[0], p2[1], p3[0], p3[1]))
        self.current = p3

    # --- arcs: Path.arc / arcNegative (src/Path.zig:345-443) over
    # arc_in_direction (src/internal/arc.zig), in user space; points go
    # through the path's transformation like any line_to / curve_to.
    def arc(self, xc, yc, radius, angle1, angle2, tolerance=0.1):
        a2 = float(angle2)
        while a2 < angle1:
            a2 += math.pi * 2
        _arc_in_direction(self, float(xc), float(yc), float(radius), float(angle1), a2, True,
                          self.transformation, max(tolerance, 0.001))

    def arc_negative(self, xc, yc, radius, angle1, angle2, tolerance=0.1):
        a2 = float(angle2)
        while a2 > angle1:
            a2 -= math.pi * 2
        _arc_in_direction(self, float(xc), float(yc), float(radius), a2, float(angle1), False,
                          self.transformation, max(tolerance, 0.001))

    def _arc_line_to(self, x, y):
        # no-op when the current point (device space) equals the target as
        # given (user space), as the reference compares them
        if self.current is not None and self.current[0] == x and self.current[1] == y:
            return
        self.line_to(x, y)

    def close(self):
        if self.current is None:
            return
        self.nodes.append((abi.CLOSE_PATH,))
        self.nodes.append((abi.MOVE_TO, self.initial[0], self.initial[1]))

    def node_array(self):
        return abi.nodes_array(self.nodes)


class Surface:
    """A host surface buffer (src/surface.zig:97-184 init/initPixel)."""

    def __init__(self, fmt, width, height, init_px=None):
        self.format, self.width, self.height = fmt, width, height
        self.buf = np.zeros(abi.surface_nbytes(fmt, width, height) + 8, dtype=np.uint8)
        self.nbytes = abi.surface_nbytes(fmt, width, height)
        if init_px is not None:
            self.paint(init_px)

    def desc(self):
        d = abi.SurfaceDesc()
        d.format = self.format
        d.width = self.width
        d.height = self.height
        d.data = self.buf.ctypes.data
        d.len = self.nbytes
        return d

    def paint(self, px):
        """ImageSurface.initBuffer/paintPixel: every pixel = T.fromPixel(px)."""
        raw = pixel_bytes(self.format, px)
        bits = abi.FORMAT_BITS[self.format]
        if bits == 32:
            self.buf[: self.nbytes] = np.tile(np.frombuffer(bytes(raw), np.uint8),
                                              self.width * self.height)
        elif bits == 8:
            self.buf[: self.nbytes] = raw[0]
        else:
            v = raw[0]
            packed = 0
            sh = 0
            while sh <= 8 - bits:
                packed |= v << sh
                sh += bits
            self.buf[: self.nbytes] = packed if v != 0 else 0

    def copy(self):
        s = Surface.__new__(Surface)
        s.format, s.width, s.height, s.nbytes = self.format, self.width, self.height, self.nbytes
        s.buf = self.buf.copy()
        return s

    def data(self):
        return self.buf[: self.nbytes]


def _scale_alpha(v, from_bits, to_bits):
    """Alpha.scaleIntoIntAlpha (src/pixel.zig:582-630)."""
    if from_bits == to_bits or v == 0:
        return v
    if from_bits == 1:
        return v * ((1 << to_bits) - 1)
    if from_bits > to_bits:
        return v >> (from_bits - to_bits)
    d = to_bits - from_bits
    if d in (2, 4):
        return (v << d) + v
    return (v << 6) | (v << 4) | (v << 2) | v


def pixel_rgba(px):
    """pixel.RGBA.fromPixel (src/pixel.zig:399-432)."""
    f = px.format
    if f in (abi.RGBA, abi.ARGB):
        return px.r, px.g, px.b, px.a
    if f in (abi.RGB, abi.XRGB):
        return px.r, px.g, px.b, 255
    return 0, 0, 0, _scale_alpha(px.a, abi.FORMAT_BITS[f], 8)


def pixel_bytes(fmt, px):
    """Bytes of T.fromPixel(px) for a surface of format `fmt`."""
    r, g, b, a = pixel_rgba(px)
    if fmt == abi.RGBA:
        return [r, g, b, a]
    if fmt == abi.ARGB:
        return [b, g, r, a]
    if fmt == abi.RGB:
        return [r, g, b, 0]
    if fmt == abi.XRGB:
        return [b, g, r, 0]
    if px.format in (abi.RGB, abi.XRGB):
        return [(1 << abi.FORMAT_BITS[fmt]) - 1 if abi.FORMAT_BITS[fmt] < 8 else 255]
    bits = abi.FORMAT_BITS[fmt]
    src_bits = abi.FORMAT_BITS[px.format] if px.format >= abi.ALPHA8 else 8
    return [_scale_alpha(px.a if px.format >= abi.ALPHA8 else a, src_bits, bits)]


def color_rgba_encode(r, g, b, a=1.0):
    """Pixel.fromColor for a linear rgb(a) color (src/pixel.zig:118-121):
    round(255*c) per channel (f32) then RGBA.multiply."""
    f32 = np.float32
    vals = [int(math.floor(float(f32(255.0) * f32(v)) + 0.5)) for v in (r, g, b, a)]
    aa = vals[3]
    return abi.rgba(vals[0] * aa // 255, vals[1] * aa // 255, vals[2] * aa // 255, aa)


class Context:
    """Recorder for the managed Context (src/Context.zig) subset the hot-path
    scenes use.  Each fill/stroke/composite is recorded with the options in
    effect at the call."""

    def __init__(self, surface):
        self.surface = surface
        self.path = Path()
        self.pattern = abi.solid_pattern(abi.rgba(0, 0, 0, 255))
        self.keep = []
        self.aa = abi.AA_DEFAULT
        self.fill_rule = abi.NON_ZERO
        self.op = abi.OP["src_over"]
        self.precision = abi.PRECISION_INTEGER
        self.tolerance = 0.1
        self.dither = abi.DITHER_NONE
        self.line_width = 2.0
        self.cap = abi.CAP_BUTT
        self.join = abi.JOIN_MITER
        self.miter_limit = 10.0
        self.dashes = []
        self.dash_offset = 0.0
        self.transformation = Transformation()
        self.ops = []

    # --- setters --------------------------------------------------------
    def set_source_pixel(self, px):
        self.pattern = abi.solid_pattern(px)

    def set_source_gradient(self, grad):
        self.keep.append(grad)
        self.pattern = grad.pattern()

    def set_transformation(self, t):
        self.transformation = t
        self.path.transformation = t

    def translate(self, tx, ty):
        self.set_transformation(self.transformation.translate(tx, ty))

    def scale(self, sx, sy):
        self.set_transformation(self.transformation.scale(sx, sy))

    def rotate(self, a):
        self.set_transformation(self.transformation.rotate(a))

    def set_identity(self):
        self.set_transformation(Transformation())

    # --- path -----------------------------------------------------------
    def move_to(self, x, y):
        self.path.move_to(x, y)

    def line_to(self, x, y):
        self.path.line_to(x, y)

    def rel_line_to(self, dx, dy):
        self.path.rel_line_to(dx, dy)

    def curve_to(self, *a):
        self.path.curve_to(*a)

    def arc(self, xc, yc, radius, a1, a2):
        self.path.arc(xc, yc, radius, a1, a2, self.tolerance)

    def arc_negative(self, xc, yc, radius, a1, a2):
        self.path.arc_negative(xc, yc, radius, a1, a2, self.tolerance)

    def close_path(self):
        self.path.close()

    def reset_path(self):
        self.path.reset()

    # --- drawing ----------------------------------------------------------
    def _wrapped_pattern(self):
        """Context.wrapDither (src/Context.zig:679-699)."""
        if self.dither == abi.DITHER_NONE:
            return self.pattern
        p = self.pattern
        scale = {abi.ALPHA1: 1, abi.ALPHA2: 2, abi.ALPHA4: 4}.get(self.surface.format, 8)
        d = abi.PatternDesc()
        d.kind = abi.PATTERN_DITHER
        d.dither.type = self.dither
        d.dither.scale = scale
        if p.kind == abi.PATTERN_OPAQUE:
            d.dither.source = abi.DITHER_SRC_PIXEL
            d.dither.pixel = p.pixel
        elif p.kind == abi.PATTERN_GRADIENT:
            d.dither.source = abi.DITHER_SRC_GRADIENT
            d.dither.gradient = p.gradient
        else:
            return p
        return d

    def fill(self):
        opts = abi.fill_opts(self.aa, self.fill_rule, self.op, self.precision, self.tolerance)
        self.ops.append(("fill", self._wrapped_pattern(), self.path.node_array(), opts))

    def stroke(self):
        o = abi.StrokeOpts()
        o.anti_aliasing_mode = self.aa
        o.line_cap_mode = self.cap
        o.line_join_mode = self.join
        o.op = self.op
        o.precision = self.precision
        o.hairline = 0
        dashes = (C.c_double * max(1, len(self.dashes)))(*self.dashes)
        self.keep.append(dashes)
        o.dashes = C.cast(dashes, C.POINTER(C.c_double))
        o.n_dashes = len(self.dashes)
        o.dash_offset = self.dash_offset
        o.line_width = self.line_width
        o.miter_limit = self.miter_limit
        o.tolerance = self.tolerance
        o.transformation[:] = self.transformation.tuple()
        self.ops.append(("stroke", self._wrapped_pattern(), self.path.node_array(), o))


class Scene:
    """A recorded scene: a target surface plus the painter calls made on it."""

    def __init__(self, name, surface, ops, keep=(), extra_surfaces=()):
        self.name = name
        self.surface = surface
        self.ops = ops
        self.keep = list(keep)
        self.extra = list(extra_surfaces)

    def run(self, backend):
        """Replays the calls on a fresh copy of the target; returns it."""
        sfc = self.surface.copy()
        for op in self.ops:
            kind = op[0]
            if kind == "fill":
                _, pat, nodes, opts = op
                rc = backend.fill(sfc, pat, nodes, opts)
            elif kind == "stroke":
                _, pat, nodes, opts = op
                rc = backend.stroke(sfc, pat, nodes, opts)
            elif kind == "composite":
                _, src_fn, dst_x, dst_y, ops, precision = op
                rc = backend.composite(sfc, dst_x, dst_y, ops, precision)
            elif kind == "call":
                rc = op[1](sfc, backend)
            else:
                raise ValueError(kind)
            if rc != 0:
                raise RuntimeError(f"{self.name}: {kind} returned {rc}")
        return sfc
