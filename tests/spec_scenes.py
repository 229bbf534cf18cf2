"""Inputs of the reference's acceptance ("spec") cases that exercise the fill
and composite hot path, restated from spec/NNN_*.zig.  Each builder returns a
zscene.Scene; the expected outputs are the reference's own fixture PNGs
(copied under tests/golden/spec/ by tests/golden/make_golden.py)."""
from fractions import Fraction

from z2d_amd import abi
from zscene import Context, Scene, Surface, color_rgba_encode

WHITE_RGB = abi.Pixel(abi.RGB, 255, 255, 255, 0)
GREEN_RGB = abi.Pixel(abi.RGB, 0, 255, 0, 0)


def _ctx(fmt, w, h, aa, init_px=None):
    sfc = Surface(fmt, w, h, init_px)
    ctx = Context(sfc)
    ctx.aa = aa
    return sfc, ctx


def _scene(name, sfc, ctx):
    return Scene(name, sfc, ctx.ops, keep=ctx.keep)


def s003(aa):  # spec/003_fill_triangle.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    m = 10
    c.move_to(0 + m, 0 + m)
    c.line_to(w - m - 1, 0 + m)
    c.line_to(w // 2 - 1, h - m - 1)
    c.close_path()
    c.fill()
    return _scene("003_fill_triangle", sfc, c)


def s004(aa):  # spec/004_fill_square.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    m = 50
    c.move_to(0 + m, 0 + m)
    c.line_to(w - m - 1, 0 + m)
    c.line_to(w - m - 1, h - m - 1)
    c.line_to(0 + m, h - m - 1)
    c.close_path()
    c.fill()
    return _scene("004_fill_square", sfc, c)


def s005(aa):  # spec/005_fill_trapezoid.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    mt, mb, my = 89, 50, 100
    c.move_to(0 + mt, 0 + my)
    c.line_to(w - mt - 1, 0 + my)
    c.line_to(w - mb - 1, h - my - 1)
    c.line_to(0 + mb, h - my - 1)
    c.close_path()
    c.fill()
    return _scene("005_fill_trapezoid", sfc, c)


def _star(c, w, h, margin, xs, ys, xo=0, yo=0):
    c.move_to(xo + w // 2, yo + 0 + margin)
    c.line_to(xo + w - margin * xs - 1, yo + h - margin - 1)
    c.line_to(xo + 0 + margin, yo + 0 + margin * ys)
    c.line_to(xo + w - margin - 1, yo + 0 + margin * ys)
    c.line_to(xo + 0 + margin * xs, yo + h - margin - 1)
    c.close_path()


def s006(aa):  # spec/006_fill_star_even_odd.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    c.fill_rule = abi.EVEN_ODD
    _star(c, w, h, 20, 3, 5)
    c.fill()
    return _scene("006_fill_star_even_odd", sfc, c)


def s016(aa):  # spec/016_fill_star_non_zero.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    c.fill_rule = abi.NON_ZERO
    _star(c, w, h, 20, 3, 5)
    c.fill()
    return _scene("016_fill_star_non_zero", sfc, c)


def s007(aa):  # spec/007_fill_bezier.zig
    sfc, c = _ctx(abi.RGB, 300, 300, aa)
    c.set_source_pixel(WHITE_RGB)
    c.move_to(19, 249)
    c.curve_to(89, 49, 209, 49, 279, 249)
    c.close_path()
    c.fill()
    return _scene("007_fill_bezier", sfc, c)


def s013(aa):  # spec/013_fill_combined.zig
    w, h = 600, 400
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    scw, sch = w // 3, h // 2
    m = 10
    c.move_to(0 + m, 0 + m)
    c.line_to(scw - m - 1, 0 + m)
    c.line_to(scw // 2 - 1, sch - m - 1)
    c.close_path()
    m = 50
    xo = scw
    c.move_to(xo + m, 0 + m)
    c.line_to(xo + scw - m - 1, 0 + m)
    c.line_to(xo + scw - m - 1, sch - m - 1)
    c.line_to(xo + m, sch - m - 1)
    c.close_path()
    tmt, tmb, tmy = 59, 33, 66
    xo = scw * 2
    c.move_to(xo + tmt, 0 + tmy)
    c.line_to(xo + scw - tmt - 1, 0 + tmy)
    c.line_to(xo + scw - tmb - 1, sch - tmy - 1)
    c.line_to(xo + tmb, sch - tmy - 1)
    c.close_path()
    m, xs, ys = 13, 3, 5
    xo = w // 6
    yo = sch
    c.move_to(xo + scw // 2, yo + m)
    c.line_to(xo + scw - m * xs - 1, yo + sch - m - 1)
    c.line_to(xo + m, yo + m * ys)
    c.line_to(xo + scw - m - 1, yo + m * ys)
    c.line_to(xo + m * xs, yo + sch - m - 1)
    c.close_path()
    xo += scw
    c.move_to(xo + 12, yo + 166)
    c.curve_to(xo + 59, yo + 32, xo + 139, yo + 32, xo + 186, yo + 166)
    c.close_path()
    c.fill()
    return _scene("013_fill_combined", sfc, c)


def s024(aa):  # spec/024_fill_triangle_direct_cross_format.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(abi.rgba(0xFF, 0x8A, 0xA5, 0xFF))
    m = 10
    c.move_to(0 + m, 0 + m)
    c.line_to(w - m - 1, 0 + m)
    c.line_to(w // 2 - 1, h - m - 1)
    c.close_path()
    c.fill()
    return _scene("024_fill_triangle_direct_cross_format", sfc, c)


def s025(aa):  # spec/025_fill_diamond_clipped.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    c.move_to(w // 2, 0 - h // 10)
    c.line_to(w + w // 10, h // 2)
    c.line_to(w // 2, h + h // 10)
    c.line_to(0 - w // 10, h // 2)
    c.close_path()
    c.fill()
    return _scene("025_fill_diamond_clipped", sfc, c)


def s026(aa):  # spec/026_fill_triangle_full.zig
    w = h = 300
    sfc, c = _ctx(abi.RGB, w, h, aa)
    c.set_source_pixel(WHITE_RGB)
    c.move_to(0, 0)
    c.line_to(w, 0)
    c.line_to(w // 2, h)
    c.close_path()
    c.fill()
    return _scene("026_fill_triangle_full", sfc, c)


def s028(aa):  # spec/028_fill_bezier_tolerance.zig
    sfc, c = _ctx(abi.RGB, 900, 300, aa)
    c.set_source_pixel(WHITE_RGB)
    c.move_to(19, 224)
    c.curve_to(89, 49, 209, 49, 279, 224)
    c.close_path()
    c.fill()
    c.tolerance = 3.0
    c.reset_path()
    c.move_to(319, 224)
    c.curve_to(389, 49, 509, 49, 579, 224)
    c.close_path()
    c.fill()
    c.tolerance = 10.0
    c.reset_path()
    c.move_to(619, 224)
    c.curve_to(689, 49, 809, 49, 879, 224)
    c.close_path()
    c.fill()
    return _scene("028_fill_bezier_tolerance", sfc, c)


def s031(aa):  # spec/031_fill_quad_bezier.zig
    sfc, c = _ctx(abi.RGB, 300, 300, aa)
    c.set_source_pixel(WHITE_RGB)
    c.move_to(20, 130)
    c.curve_to(20, 130, 20, 20, 130, 20)
    c.close_path()
    c.move_to(170, 20)
    c.curve_to(280, 20, 280, 20, 280, 130)
    c.close_path()
    c.move_to(280, 170)
    c.curve_to(280, 280, 170, 280, 170, 280)
    c.close_path()
    c.move_to(130, 280)
    c.curve_to(20, 280, 20, 280, 20, 170)
    c.close_path()
    c.fill()
    return _scene("031_fill_quad_bezier", sfc, c)


def s032(aa):  # spec/032_fill_arc.zig (comptime f128 arithmetic -> f64)
    sfc, c = _ctx(abi.RGB, 300, 300, aa)
    c.set_source_pixel(WHITE_RGB)
    ratio = Fraction("0.552284749831")
    cy = 150

    def f(x):
        return float(Fraction(x))

    cx, r = 100, 50
    p12 = r * ratio
    c.move_to(cx, cy - r)
    c.curve_to(f(cx + p12), cy - r, cx + r, f(cy - p12), cx + r, cy)
    c.curve_to(cx + r, f(cy + p12), f(cx + p12), cy + r, cx, cy + r)
    c.curve_to(f(cx - p12), cy + r, cx - r, f(cy + p12), cx - r, cy)
    c.curve_to(cx - r, f(cy - p12), f(cx - p12), cy - r, cx, cy - r)
    c.close_path()
    c.fill()
    cx, rM, rm = 200, 50, 25
    pM, pm = rM * ratio, rm * ratio
    c.reset_path()
    c.move_to(cx, cy - rM)
    c.curve_to(f(cx + pm), cy - rM, cx + rm, f(cy - pM), cx + rm, cy)
    c.curve_to(cx + rm, f(cy + pM), f(cx + pm), cy + rM, cx, cy + rM)
    c.curve_to(f(cx - pm), cy + rM, cx - rm, f(cy + pM), cx - rm, cy)
    c.curve_to(cx - rm, f(cy - pM), f(cx - pm), cy - rM, cx, cy - rM)
    c.close_path()
    c.fill()
    return _scene("032_fill_arc", sfc, c)


def s033(aa):  # spec/033_fill_zig_mark.zig
    sfc, c = _ctx(abi.RGBA, 153, 140, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xF7, 0xA4, 0x1D, 0))
    polys = [
        [(46, 22), (28, 44), (19, 30)],
        [(46, 22), (33, 33), (28, 44), (22, 44), (22, 95), (31, 95), (20, 100), (12, 117),
         (0, 117), (0, 22)],
        [(31, 95), (12, 117), (4, 106)],
        [(56, 22), (62, 36), (37, 44)],
        [(56, 22), (111, 22), (111, 44), (37, 44), (56, 32)],
        [(116, 95), (97, 117), (90, 104)],
        [(116, 95), (100, 104), (97, 117), (42, 117), (42, 95)],
        [(150, 0), (52, 117), (3, 140), (101, 22)],
        [(141, 22), (140, 40), (122, 45)],
        [(153, 22), (153, 117), (106, 117), (120, 105), (125, 95), (131, 95), (131, 45),
         (122, 45), (132, 36), (141, 22)],
        [(125, 95), (130, 110), (106, 117)],
    ]
    for poly in polys:
        c.move_to(*poly[0])
        for p in poly[1:]:
            c.line_to(*p)
        c.close_path()
    c.fill()
    return _scene("033_fill_zig_mark", sfc, c)


def _tri(c, w=300, h=300, m=10):
    c.move_to(0 + m, 0 + m)
    c.line_to(w - m - 1, 0 + m)
    c.line_to(w // 2 - 1, h - m - 1)
    c.close_path()


def s046(aa):  # spec/046_fill_triangle_alpha.zig
    sfc, c = _ctx(abi.RGB, 300, 300, aa, init_px=WHITE_RGB)
    c.set_source_pixel(abi.Pixel(abi.ALPHA8, 0, 0, 0, 255))
    _tri(c)
    c.fill()
    return _scene("046_fill_triangle_alpha", sfc, c)


def s047(aa):  # spec/047_fill_triangle_alpha_gray.zig
    sfc, c = _ctx(abi.ALPHA8, 300, 300, aa)
    c.set_source_pixel(abi.Pixel(abi.ALPHA8, 0, 0, 0, 119))
    _tri(c)
    c.fill()
    return _scene("047_fill_triangle_alpha_gray", sfc, c)


def s048(aa):  # spec/048_fill_triangle_static.zig (painter.fill directly)
    sfc, c = _ctx(abi.RGB, 300, 300, aa)
    c.set_source_pixel(WHITE_RGB)
    _tri(c)
    c.fill()
    return _scene("048_fill_triangle_static", sfc, c)


def s049(aa):  # spec/049_fill_triangle_alpha4_gray.zig
    sfc, c = _ctx(abi.ALPHA4, 300, 300, aa)
    c.set_source_pixel(abi.Pixel(abi.ALPHA4, 0, 0, 0, 7))
    _tri(c)
    c.fill()
    return _scene("049_fill_triangle_alpha4_gray", sfc, c)


def s050(aa):  # spec/050_fill_triangle_alpha2_gray.zig
    sfc, c = _ctx(abi.ALPHA2, 300, 300, aa)
    c.set_source_pixel(abi.Pixel(abi.ALPHA2, 0, 0, 0, 2))
    _tri(c)
    c.fill()
    return _scene("050_fill_triangle_alpha2_gray", sfc, c)


def s051(aa):  # spec/051_fill_triangle_alpha1_gray.zig
    sfc, c = _ctx(abi.ALPHA1, 300, 300, aa)
    c.set_source_pixel(abi.Pixel(abi.ALPHA1, 0, 0, 0, 1))
    _tri(c)
    c.fill()
    return _scene("051_fill_triangle_alpha1_gray", sfc, c)


def s052(aa):  # spec/052_fill_triangle_alpha4_gray_scaledown.zig
    sfc, c = _ctx(abi.ALPHA4, 300, 300, aa)
    c.set_source_pixel(abi.Pixel(abi.ALPHA8, 0, 0, 0, 119))
    _tri(c)
    c.fill()
    return _scene("052_fill_triangle_alpha4_gray_scaledown", sfc, c)


def s053(aa):  # spec/053_fill_triangle_alpha8_gray_scaleup.zig
    sfc, c = _ctx(abi.ALPHA8, 300, 300, aa)
    c.set_source_pixel(abi.Pixel(abi.ALPHA4, 0, 0, 0, 7))
    _tri(c)
    c.fill()
    return _scene("053_fill_triangle_alpha8_gray_scaleup", sfc, c)


def s079(aa):  # spec/079_fill_degenerate_lineto.zig
    sfc, c = _ctx(abi.RGB, 100, 200, aa)
    c.set_source_pixel(WHITE_RGB)
    c.move_to(180, 180)
    c.line_to(190, 190)
    c.close_path()
    c.fill()
    c.reset_path()
    c.move_to(10, 10); c.line_to(20, 10); c.line_to(20, 20); c.line_to(10, 20); c.close_path()
    c.move_to(50, 10); c.line_to(70, 20); c.close_path()
    c.fill()
    c.reset_path()
    c.move_to(10, 50); c.line_to(30, 70); c.close_path()
    c.move_to(50, 50); c.line_to(60, 50); c.line_to(60, 60); c.line_to(50, 60); c.close_path()
    c.fill()
    c.reset_path()
    c.move_to(-30, 80); c.line_to(-10, 100); c.close_path()
    c.move_to(50, 80); c.line_to(60, 80); c.line_to(60, 90); c.line_to(50, 90); c.close_path()
    c.fill()
    c.reset_path()
    c.move_to(10, 110); c.line_to(30, 130); c.close_path()
    c.move_to(150, 110); c.line_to(160, 110); c.line_to(160, 120); c.line_to(150, 120)
    c.close_path()
    c.fill()
    c.reset_path()
    c.move_to(10, 140); c.line_to(20, 140); c.line_to(20, 150); c.line_to(10, 150)
    c.close_path()
    c.move_to(50, 140); c.line_to(70, 150); c.close_path()
    c.fill()
    return _scene("079_fill_degenerate_lineto", sfc, c)


PATH_SCENES = [s003, s004, s005, s006, s007, s013, s016, s024, s025, s026, s028, s031, s032,
               s033, s046, s047, s048, s049, s050, s051, s052, s053, s079]

AA_SUFFIX = {abi.AA_NONE: "_pixelated", abi.AA_SUPERSAMPLE_4X: "_smooth",
             abi.AA_MULTISAMPLE_4X: "_smooth_multisample"}
