"""Inputs of the reference's gradient, dither, operator and gamma acceptance
cases (spec/061-072), restated from the spec sources.  Unlike the fill-only
cases these drive the compositor directly (SurfaceCompositor.run,
Surface.composite) on scratch surfaces, so each case is a function that
replays its calls on a backend (oracle or HIP library) and returns the
target surface."""
import ctypes as C
import math

from z2d_amd import abi
from zscene import Context, Surface, Transformation, color_rgba_encode

RGB1 = lambda r, g, b: abi.color(abi.COLOR_LINEAR_RGB, r, g, b, 1.0)  # noqa: E731
RGBA1 = lambda r, g, b, a: abi.color(abi.COLOR_LINEAR_RGB, r, g, b, a)  # noqa: E731
SRGB1 = lambda r, g, b: abi.color(abi.COLOR_SRGB, r, g, b, 1.0)  # noqa: E731
HSL1 = lambda h, s, l: abi.color(abi.COLOR_HSL, h, s, l, 1.0)  # noqa: E731
HSLA1 = lambda h, s, l, a: abi.color(abi.COLOR_HSL, h, s, l, a)  # noqa: E731


def comp_op(op, src=None, dst=None):
    o = abi.CompOp()
    o.op = abi.OP[op] if isinstance(op, str) else op
    for side, v in (("src", src), ("dst", dst)):
        p = getattr(o, side)
        if v is None:
            p.kind = abi.PARAM_NONE
        elif isinstance(v, abi.SurfaceDesc):
            p.kind = abi.PARAM_SURFACE
            p.surface = C.pointer(v)
        elif isinstance(v, abi.Gradient):
            p.kind = abi.PARAM_GRADIENT
            p.gradient = C.pointer(v.desc)
        elif isinstance(v, abi.DitherDesc):
            p.kind = abi.PARAM_DITHER
            p.dither = v
        else:
            p.kind = abi.PARAM_PIXEL
            p.pixel = v
    return o


def run_ops(backend, dst, x, y, ops, precision=abi.PRECISION_INTEGER):
    rc = backend.composite(dst, x, y, ops, precision)
    assert rc == 0, rc


def surface_composite(backend, dst, src, op, x, y):
    """Surface.composite (src/surface.zig:225-241)."""
    sd = src.desc()
    run_ops(backend, dst, x, y, [comp_op(op, src=sd)])


def run_fills(backend, ctx):
    for kind, pat, nodes, opts in ctx.ops:
        assert kind == "fill"
        rc = backend.fill(ctx.surface, pat, nodes, opts)
        assert rc == 0, rc
    ctx.ops = []


def c061(backend, aa=None):  # spec/061_linear_gradient.zig
    dst = Surface(abi.RGB, 200, 400)
    scratch = Surface(abi.RGB, 100, 100)
    for sx, sy, x0, y0, x1, y1 in [(0, 0, 0, 49, 99, 49), (100, 0, 99, 49, 0, 49),
                                   (0, 100, 49, 0, 49, 99), (100, 100, 49, 99, 49, 0),
                                   (0, 200, 0, 0, 99, 99), (100, 200, 99, 0, 0, 99),
                                   (0, 300, 0, 99, 99, 0), (100, 300, 99, 99, 0, 0)]:
        g = abi.Gradient(abi.GRADIENT_LINEAR, [x0, y0, x1, y1],
                         [(0.0, RGB1(1, 0, 0)), (0.5, RGB1(0, 1, 0)), (1.0, RGB1(0, 0, 1))])
        run_ops(backend, scratch, 0, 0, [comp_op("src_over", src=g)])
        surface_composite(backend, dst, scratch, "src_over", sx, sy)
    return dst


def c062(backend, aa=None):  # spec/062_hsl_gradient.zig
    dst = Surface(abi.RGB, 100, 200)
    calls = [(100, 100, 0, 0, 0, 49, 99, 49, HSL1(300, 1, 0.5), HSL1(60, 1, 0.5)),
             (50, 100, 0, 100, 0, 49, 49, 49, HSL1(0, 1, 0.5), HSL1(0, 0, 0.5)),
             (50, 100, 50, 100, 49, 49, 0, 49, HSL1(180, 1, 0.5), HSL1(180, 0, 0.5)),
             (100, 50, 0, 100, 49, 0, 49, 49, HSLA1(0, 0, 1, 1), HSLA1(0, 0, 0.5, 0)),
             (100, 50, 0, 150, 49, 49, 49, 0, HSLA1(0, 0, 0, 1), HSLA1(0, 0, 0.5, 0))]
    for w, h, sx, sy, x0, y0, x1, y1, c0, c1 in calls:
        scratch = Surface(abi.RGBA, w, h)
        g = abi.Gradient(abi.GRADIENT_LINEAR, [x0, y0, x1, y1], [(0.0, c0), (1.0, c1)],
                         abi.INTERP_HSL, abi.POLAR_SHORTER)
        run_ops(backend, scratch, 0, 0, [comp_op("src_over", src=g)])
        surface_composite(backend, dst, scratch, "src_over", sx, sy)
    return dst


def _radial_draw(backend, dst, sx, sy, ix, iy, ir, ox, oy, orr):
    scratch = Surface(abi.RGBA, 100, 100)
    g = abi.Gradient(abi.GRADIENT_RADIAL, [ix, iy, ir, ox, oy, orr],
                     [(0.0, RGB1(1, 0, 0)), (0.5, RGB1(0, 1, 0)), (1.0, RGB1(0, 0, 1))])
    run_ops(backend, scratch, 0, 0, [comp_op("src_over", src=g)])
    surface_composite(backend, dst, scratch, "src_over", sx, sy)


def c063(backend, aa=None):  # spec/063_radial_gradient.zig
    dst = Surface(abi.RGBA, 300, 500)
    for y in range(3):
        for x in range(3):
            _radial_draw(backend, dst, 100 * x, 100 * y, 49 + 25 * (float(x) - 1),
                         49 + 25 * (float(y) - 1), 5, 49, 49, 50)
    _radial_draw(backend, dst, 0, 300, 49, 49, 0, 49, 49, 50)
    _radial_draw(backend, dst, 100, 300, 49, 49, 50, 49, 49, 0)
    _radial_draw(backend, dst, 200, 300, 49, 49, 50, 49, 49, 50)
    _radial_draw(backend, dst, 0, 400, 49, 49, 0, 49, 49, 0)
    _radial_draw(backend, dst, 100, 400, 10, 49, 0, 49, 49, 25)
    _radial_draw(backend, dst, 200, 400, 49, 49, 49, 49, 49, 50)
    return dst


def _rect(ctx, w, h):
    ctx.move_to(0, 0)
    ctx.line_to(w, 0)
    ctx.line_to(w, h)
    ctx.line_to(0, h)
    ctx.close_path()


def _set_gradient(ctx, g):
    """Context.setSource for a gradient: locks the inverse CTM into a
    snapshot of the gradient (src/Context.zig:115-121)."""
    snap = abi.Gradient(g.desc.type, list(g.desc.geom), [(g.stop_arr[i].offset, g.stop_arr[i].color)
                                                         for i in range(g.desc.n_stops)],
                        g.desc.method, g.desc.polar, ctx.transformation.inverse().tuple())
    ctx.set_source_gradient(snap)


def c064(backend, aa):  # spec/064_radial_source.zig
    sfc = Surface(abi.RGB, 100, 100)
    ctx = Context(sfc)
    ctx.aa = aa
    g = abi.Gradient(abi.GRADIENT_RADIAL, [49, 49, 0, 49, 49, 50],
                     [(0.0, RGB1(1, 0, 0)), (0.5, RGB1(0, 1, 0)), (1.0, RGB1(0, 0, 1))])
    _set_gradient(ctx, g)
    _rect(ctx, 100, 100)
    ctx.fill()
    run_fills(backend, ctx)
    return sfc


def c065(backend, aa=None):  # spec/065_conic_gradient.zig
    dst = Surface(abi.RGBA, 200, 200)
    for y in range(2):
        for x in range(2):
            scratch = Surface(abi.RGBA, 100, 100)
            g = abi.Gradient(abi.GRADIENT_CONIC, [49, 49, math.pi / 2.0 * float(y * 2 + x)],
                             [(0.0, HSL1(0, 1, 0.5)), (1.0, HSL1(360, 1, 0.5))],
                             abi.INTERP_HSL, abi.POLAR_INCREASING)
            run_ops(backend, scratch, 0, 0, [comp_op("src_over", src=g)])
            surface_composite(backend, dst, scratch, "src_over", 100 * x, 100 * y)
    return dst


def c067(backend, aa):  # spec/067_gradient_transforms.zig
    sfc = Surface(abi.RGB, 200, 200)
    ctx = Context(sfc)
    ctx.scale(2, 2)
    ctx.aa = aa
    lin = abi.Gradient(abi.GRADIENT_LINEAR, [0, 0, 50, 50],
                       [(0.0, RGB1(1, 0, 0)), (0.5, RGB1(0, 1, 0)), (1.0, RGB1(0, 0, 1))])
    _set_gradient(ctx, lin)
    _rect(ctx, 50, 50)
    ctx.fill()
    rad = abi.Gradient(abi.GRADIENT_RADIAL, [25, 50, 0, 25, 50, 50],
                       [(0.0, RGB1(1, 0, 0)), (0.5, RGB1(0, 1, 0)), (1.0, RGB1(0, 0, 1))])
    ctx.set_identity()
    skew = Transformation(by=0.5)
    ctx.set_transformation(ctx.transformation.mul(skew))
    ctx.translate(100, 0)
    _set_gradient(ctx, rad)
    ctx.set_identity()
    ctx.translate(100, 0)
    ctx.reset_path()
    _rect(ctx, 100, 100)
    ctx.fill()
    con = abi.Gradient(abi.GRADIENT_CONIC, [50, 50, 0],
                       [(0.0, HSL1(0, 1, 0.5)), (1.0, HSL1(360, 1, 0.5))],
                       abi.INTERP_HSL, abi.POLAR_INCREASING)
    ctx.set_identity()
    ctx.scale(2, 1)
    ctx.translate(0, 100)
    _set_gradient(ctx, con)
    ctx.reset_path()
    _rect(ctx, 100, 100)
    ctx.fill()
    run_fills(backend, ctx)
    return sfc


def c068(backend, aa=None):  # spec/068_gradient_deband.zig
    dst = Surface(abi.RGB, 500, 500, init_px=abi.Pixel(abi.RGB, 255, 255, 255, 0))
    for sy, gray, bits, dt in [(0, False, 8, abi.DITHER_NONE), (100, False, 8, abi.DITHER_BAYER),
                               (200, True, 8, abi.DITHER_NONE), (300, True, 4, abi.DITHER_NONE),
                               (400, True, 4, abi.DITHER_BAYER)]:
        fmt = (abi.ALPHA4 if bits == 4 else abi.ALPHA8) if gray else abi.RGB
        scratch = Surface(fmt, 500, 100)
        if gray:
            stops = [(0.0, RGBA1(1, 1, 1, 0)), (1.0, RGBA1(0, 0, 0, 1))]
        else:
            stops = [(0.0, RGB1(27.0 / 255.0, 93.0 / 255.0, 124.0 / 255.0)),
                     (1.0, RGB1(38.0 / 255.0, 32.0 / 255.0, 16.0 / 255.0))]
        g = abi.Gradient(abi.GRADIENT_LINEAR, [0, 50, 500, 50], stops)
        d = abi.DitherDesc()
        d.type = dt
        d.source = abi.DITHER_SRC_GRADIENT
        d.gradient = C.pointer(g.desc)
        d.scale = bits
        run_ops(backend, scratch, 0, 0, [comp_op("src_over", src=d)])
        surface_composite(backend, dst, scratch, "src_over", 0, sy)
    return dst


def c069(backend, aa):  # spec/069_gradient_dither_context.zig
    sfc = Surface(abi.ALPHA4, 400, 150)
    ctx = Context(sfc)
    ctx.aa = aa
    g = abi.Gradient(abi.GRADIENT_LINEAR, [0, 25, 400, 25],
                     [(0.0, RGBA1(0, 0, 0, 0)), (1.0, RGBA1(1, 1, 1, 1))])
    _set_gradient(ctx, g)
    _rect(ctx, 400, 50)
    ctx.fill()
    ctx.reset_path()
    ctx.translate(0, 50)
    _set_gradient(ctx, g)
    ctx.dither = abi.DITHER_BAYER
    _rect(ctx, 400, 50)
    ctx.fill()
    ctx.reset_path()
    ctx.set_identity()
    ctx.translate(0, 100)
    _set_gradient(ctx, g)
    ctx.dither = abi.DITHER_BLUE_NOISE
    _rect(ctx, 400, 50)
    ctx.fill()
    run_fills(backend, ctx)
    return sfc


def c070(backend, aa):  # spec/070_compositor_ops.zig
    main = Surface(abi.RGB, 460, 3090, init_px=abi.Pixel(abi.RGB, 255, 255, 255, 0))
    for i, op in enumerate(abi.OPERATORS):
        for j, prec, transparent in [(0, 0, False), (1, 0, True), (2, 1, False), (3, 1, True)]:
            if transparent:
                bg = color_rgba_encode(0.69, 0.23, 0.21, 0.9)
                fg = color_rgba_encode(0.56, 0.50, 0.89, 0.8)
            else:
                bg = color_rgba_encode(0.69, 0.23, 0.21)
                fg = color_rgba_encode(0.56, 0.50, 0.89)
            scratch = Surface(abi.RGBA, 100, 100)
            ctx = Context(scratch)
            ctx.aa = aa
            ctx.precision = prec
            ctx.set_source_pixel(bg)
            _rect(ctx, 75, 75)
            ctx.fill()
            ctx.set_source_pixel(fg)
            ctx.op = abi.OP[op]
            ctx.reset_path()
            ctx.translate(25, 25)
            _rect(ctx, 75, 75)
            ctx.fill()
            run_fills(backend, ctx)
            surface_composite(backend, main, scratch, "src_over", 10 + 110 * j, 10 + 110 * i)
    return main


def _gamma_case(backend, aa=None):  # spec/071_gamma_linear.zig, 072_gamma_srgb.zig
    w, h = 400, 300
    sfc = Surface(abi.RGB, w, h)
    ctx = Context(sfc)
    ctx.aa = abi.AA_NONE
    for pos, profile in ((0.0, "linear"), (1.0, "srgb")):
        if profile == "linear":
            px = color_rgba_encode(0.3, 0.3, 0.3)
        else:
            px = None
        if px is None:
            px = _srgb_pixel(0.3, 0.3, 0.3)
        ctx.set_source_pixel(px)
        ctx.translate(0, pos * h / 2)
        _rect(ctx, w, h / 2)
        ctx.set_identity()
        ctx.fill()
        ctx.reset_path()
    off = 30.0
    gw = w - off * 2
    gh = (h - off * 2) / 4
    for pos, stop_profile, method in ((0, "linear", abi.INTERP_LINEAR_RGB),
                                      (1, "srgb", abi.INTERP_LINEAR_RGB),
                                      (2, "linear", abi.INTERP_SRGB),
                                      (3, "srgb", abi.INTERP_SRGB)):
        mk = RGB1 if stop_profile == "linear" else SRGB1
        g = abi.Gradient(abi.GRADIENT_LINEAR, [0, gh / 2, gw, gh / 2],
                         [(0.0, mk(0.8, 0, 0)), (0.5, mk(0, 0.8, 0)), (1.0, mk(0, 0, 0.8))],
                         method)
        ctx.translate(off, off + gh * pos)
        _set_gradient(ctx, g)
        _rect(ctx, gw, gh)
        ctx.set_identity()
        ctx.fill()
        ctx.reset_path()
        ctx.set_source_pixel(abi.Pixel(abi.RGB, 0, 0, 0, 0))
    run_fills(backend, ctx)
    return sfc


def _srgb_pixel(r, g, b):
    """Pixel.fromColor(.srgb): LinearRGB.fromColor (pow 2.2) then encodeRGBA.
    Uses the oracle's std.math.pow restatement."""
    import numpy as np

    from oracle_backend import lib
    f32 = np.float32
    vals = []
    for v in (r, g, b):
        lin = lib().zo_powf(float(f32(v)), float(f32(2.2)))
        vals.append(int(math.floor(float(f32(255.0) * f32(lin)) + 0.5)))
    return abi.rgba(vals[0], vals[1], vals[2], 255)


c071 = _gamma_case
c072 = _gamma_case

COMP_SCENES = [("061_linear_gradient", c061, False), ("062_hsl_gradient", c062, False),
               ("063_radial_gradient", c063, False), ("064_radial_source", c064, True),
               ("065_conic_gradient", c065, False), ("067_gradient_transforms", c067, True),
               ("068_gradient_deband", c068, False), ("069_gradient_dither_context", c069, True),
               ("070_compositor_ops", c070, True), ("071_gamma_linear", c071, False),
               ("072_gamma_srgb", c072, False)]
