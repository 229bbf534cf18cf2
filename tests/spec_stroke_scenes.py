"""Inputs of the reference's stroke acceptance cases (spec/NNN_*.zig),
restated as recorded Context calls.  The expected outputs are the
reference's own fixture PNGs (tests/golden/spec/); stroke fixtures exist for
the aliased and supersampled modes, and the multisample mode is compared
with the supersampled fixture (the two agree for 8-bit surfaces:
floor(255k/16) = 16k - 1 for k = 1..15)."""
from z2d_amd import abi
from spec_scenes import _ctx, _scene  # noqa: F401


def s008(aa):  # spec/008_stroke_triangle.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    margin = 10
    c.move_to(0 + margin, 0 + margin)
    c.line_to(width - margin - 1, 0 + margin)
    c.line_to(width / 2 - 1, height - margin - 1)
    c.close_path()
    c.stroke()
    return _scene("008_stroke_triangle", sfc, c)


def s009(aa):  # spec/009_stroke_square.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    margin = 50
    c.move_to(0 + margin, 0 + margin)
    c.line_to(width - margin - 1, 0 + margin)
    c.line_to(width - margin - 1, height - margin - 1)
    c.line_to(0 + margin, height - margin - 1)
    c.close_path()
    c.stroke()
    return _scene("009_stroke_square", sfc, c)


def s010(aa):  # spec/010_stroke_trapezoid.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    margin_top = 89
    margin_bottom = 50
    margin_y = 100
    c.move_to(0 + margin_top, 0 + margin_y)
    c.line_to(width - margin_top - 1, 0 + margin_y)
    c.line_to(width - margin_bottom - 1, height - margin_y - 1)
    c.line_to(0 + margin_bottom, height - margin_y - 1)
    c.close_path()
    c.stroke()
    return _scene("010_stroke_trapezoid", sfc, c)


def s011(aa):  # spec/011_stroke_star.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(6)
    margin = 20
    x_scale = 3
    y_scale = 5
    c.move_to(width / 2, 0 + margin)
    c.line_to(width - margin * x_scale - 1, height - margin - 1)
    c.line_to(0 + margin, 0 + margin * y_scale)
    c.line_to(width - margin - 1, 0 + margin * y_scale)
    c.line_to(0 + margin * x_scale, height - margin - 1)
    c.close_path()
    c.stroke()
    return _scene("011_stroke_star", sfc, c)


def s012(aa):  # spec/012_stroke_bezier.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.move_to(19, 149)
    c.curve_to(89, 0, 209, 0, 279, 149)
    c.stroke()
    c.line_width = float(6)
    c.reset_path()
    c.move_to(19, 199)
    c.curve_to(89, 24, 209, 24, 279, 199)
    c.stroke()
    c.line_width = float(10)
    c.reset_path()
    c.move_to(19, 249)
    c.curve_to(89, 49, 209, 49, 279, 249)
    c.stroke()
    return _scene("012_stroke_bezier", sfc, c)


def s014(aa):  # spec/014_stroke_lines.zig
    width = 800
    height = 400
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    sub_canvas_width = width / 4
    sub_canvas_height = height / 2
    margin = 10
    x_offset = 0
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = 0
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    x_offset = sub_canvas_width * 2
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    x_offset = sub_canvas_width * 2
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    y_offset = 0
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    x_offset = sub_canvas_width * 3
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.stroke()
    return _scene("014_stroke_lines", sfc, c)


def s015(aa):  # spec/015_stroke_miter.zig
    width = 240
    height = 260
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(6)
    c.move_to(10, 10)
    c.line_to(100, 20)
    c.line_to(110, 120)
    c.line_to(10, 110)
    c.line_to(20, 30)
    c.line_to(90, 40)
    c.line_to(95, 100)
    c.line_to(30, 95)
    c.line_to(30, 50)
    c.line_to(80, 50)
    c.line_to(75, 85)
    c.line_to(45, 80)
    c.line_to(50, 60)
    c.line_to(65, 70)
    x_offset = 120
    c.move_to(x_offset + 110, 10)
    c.line_to(x_offset + 20, 20)
    c.line_to(x_offset + 10, 120)
    c.line_to(x_offset + 110, 110)
    c.line_to(x_offset + 100, 30)
    c.line_to(x_offset + 30, 40)
    c.line_to(x_offset + 25, 100)
    c.line_to(x_offset + 90, 95)
    c.line_to(x_offset + 90, 50)
    c.line_to(x_offset + 40, 50)
    c.line_to(x_offset + 45, 85)
    c.line_to(x_offset + 75, 80)
    c.line_to(x_offset + 70, 60)
    c.line_to(x_offset + 55, 70)
    y_offset = 130
    c.move_to(10, y_offset + 120)
    c.line_to(100, y_offset + 110)
    c.line_to(110, y_offset + 10)
    c.line_to(10, y_offset + 20)
    c.line_to(20, y_offset + 100)
    c.line_to(90, y_offset + 90)
    c.line_to(95, y_offset + 30)
    c.line_to(30, y_offset + 35)
    c.line_to(30, y_offset + 80)
    c.line_to(80, y_offset + 80)
    c.line_to(75, y_offset + 45)
    c.line_to(45, y_offset + 50)
    c.line_to(50, y_offset + 70)
    c.line_to(65, y_offset + 60)
    c.move_to(x_offset + 110, y_offset + 120)
    c.line_to(x_offset + 20, y_offset + 110)
    c.line_to(x_offset + 10, y_offset + 10)
    c.line_to(x_offset + 110, y_offset + 20)
    c.line_to(x_offset + 100, y_offset + 100)
    c.line_to(x_offset + 30, y_offset + 90)
    c.line_to(x_offset + 25, y_offset + 30)
    c.line_to(x_offset + 90, y_offset + 35)
    c.line_to(x_offset + 90, y_offset + 80)
    c.line_to(x_offset + 40, y_offset + 80)
    c.line_to(x_offset + 45, y_offset + 45)
    c.line_to(x_offset + 75, y_offset + 50)
    c.line_to(x_offset + 70, y_offset + 70)
    c.line_to(x_offset + 55, y_offset + 60)
    c.stroke()
    return _scene("015_stroke_miter", sfc, c)


def s017(aa):  # spec/017_stroke_star_round.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(10)
    c.join = abi.JOIN_ROUND
    margin = 20
    x_scale = 3
    y_scale = 5
    c.move_to(width / 2, 0 + margin)
    c.line_to(width - margin * x_scale - 1, height - margin - 1)
    c.line_to(0 + margin, 0 + margin * y_scale)
    c.line_to(width - margin - 1, 0 + margin * y_scale)
    c.line_to(0 + margin * x_scale, height - margin - 1)
    c.close_path()
    c.stroke()
    return _scene("017_stroke_star_round", sfc, c)


def s018(aa):  # spec/018_stroke_square_spiral_round.zig
    width = 240
    height = 260
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(10)
    c.join = abi.JOIN_ROUND
    c.move_to(10, 10)
    c.line_to(100, 20)
    c.line_to(110, 120)
    c.line_to(10, 110)
    c.line_to(20, 30)
    c.line_to(90, 40)
    c.line_to(95, 100)
    c.line_to(30, 95)
    c.line_to(30, 50)
    c.line_to(80, 50)
    c.line_to(75, 85)
    c.line_to(45, 80)
    c.line_to(50, 60)
    c.line_to(65, 70)
    x_offset = 120
    c.move_to(x_offset + 110, 10)
    c.line_to(x_offset + 20, 20)
    c.line_to(x_offset + 10, 120)
    c.line_to(x_offset + 110, 110)
    c.line_to(x_offset + 100, 30)
    c.line_to(x_offset + 30, 40)
    c.line_to(x_offset + 25, 100)
    c.line_to(x_offset + 90, 95)
    c.line_to(x_offset + 90, 50)
    c.line_to(x_offset + 40, 50)
    c.line_to(x_offset + 45, 85)
    c.line_to(x_offset + 75, 80)
    c.line_to(x_offset + 70, 60)
    c.line_to(x_offset + 55, 70)
    y_offset = 130
    c.move_to(10, y_offset + 120)
    c.line_to(100, y_offset + 110)
    c.line_to(110, y_offset + 10)
    c.line_to(10, y_offset + 20)
    c.line_to(20, y_offset + 100)
    c.line_to(90, y_offset + 90)
    c.line_to(95, y_offset + 30)
    c.line_to(30, y_offset + 35)
    c.line_to(30, y_offset + 80)
    c.line_to(80, y_offset + 80)
    c.line_to(75, y_offset + 45)
    c.line_to(45, y_offset + 50)
    c.line_to(50, y_offset + 70)
    c.line_to(65, y_offset + 60)
    c.move_to(x_offset + 110, y_offset + 120)
    c.line_to(x_offset + 20, y_offset + 110)
    c.line_to(x_offset + 10, y_offset + 10)
    c.line_to(x_offset + 110, y_offset + 20)
    c.line_to(x_offset + 100, y_offset + 100)
    c.line_to(x_offset + 30, y_offset + 90)
    c.line_to(x_offset + 25, y_offset + 30)
    c.line_to(x_offset + 90, y_offset + 35)
    c.line_to(x_offset + 90, y_offset + 80)
    c.line_to(x_offset + 40, y_offset + 80)
    c.line_to(x_offset + 45, y_offset + 45)
    c.line_to(x_offset + 75, y_offset + 50)
    c.line_to(x_offset + 70, y_offset + 70)
    c.line_to(x_offset + 55, y_offset + 60)
    c.stroke()
    return _scene("018_stroke_square_spiral_round", sfc, c)


def s019(aa):  # spec/019_stroke_bevel_miterlimit.zig
    width = 375
    height = 560
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(5)
    c.join = abi.JOIN_BEVEL
    c.move_to(10, 90)
    c.line_to(80, 60)
    c.line_to(150, 90)
    c.move_to(170, 90)
    c.line_to(205, 60)
    c.line_to(240, 90)
    c.move_to(260, 90)
    c.line_to(280, 60)
    c.line_to(300, 90)
    c.move_to(320, 90)
    c.line_to(327.5, 60)
    c.line_to(335, 90)
    c.move_to(355, 90)
    c.line_to(360, 60)
    c.line_to(365, 90)
    c.stroke()
    c.join = abi.JOIN_MITER
    c.miter_limit = float(4)
    c.reset_path()
    c.move_to(10, 190)
    c.line_to(80, 160)
    c.line_to(150, 190)
    c.move_to(170, 190)
    c.line_to(205, 160)
    c.line_to(240, 190)
    c.move_to(260, 190)
    c.line_to(280, 160)
    c.line_to(300, 190)
    c.move_to(320, 190)
    c.line_to(327.5, 160)
    c.line_to(335, 190)
    c.move_to(355, 190)
    c.line_to(360, 160)
    c.line_to(365, 190)
    c.stroke()
    c.miter_limit = float(1)
    c.reset_path()
    c.move_to(10, 290)
    c.line_to(80, 260)
    c.line_to(150, 290)
    c.move_to(170, 290)
    c.line_to(205, 260)
    c.line_to(240, 290)
    c.move_to(260, 290)
    c.line_to(280, 260)
    c.line_to(300, 290)
    c.move_to(320, 290)
    c.line_to(327.5, 260)
    c.line_to(335, 290)
    c.move_to(355, 290)
    c.line_to(360, 260)
    c.line_to(365, 290)
    c.stroke()
    c.miter_limit = float(6)
    c.reset_path()
    c.move_to(10, 390)
    c.line_to(80, 360)
    c.line_to(150, 390)
    c.move_to(170, 390)
    c.line_to(205, 360)
    c.line_to(240, 390)
    c.move_to(260, 390)
    c.line_to(280, 360)
    c.line_to(300, 390)
    c.move_to(320, 390)
    c.line_to(327.5, 360)
    c.line_to(335, 390)
    c.move_to(355, 390)
    c.line_to(360, 360)
    c.line_to(365, 390)
    c.stroke()
    c.miter_limit = float(10)
    c.reset_path()
    c.move_to(10, 490)
    c.line_to(80, 460)
    c.line_to(150, 490)
    c.move_to(170, 490)
    c.line_to(205, 460)
    c.line_to(240, 490)
    c.move_to(260, 490)
    c.line_to(280, 460)
    c.line_to(300, 490)
    c.move_to(320, 490)
    c.line_to(327.5, 460)
    c.line_to(335, 490)
    c.move_to(355, 490)
    c.line_to(360, 460)
    c.line_to(365, 490)
    c.stroke()
    return _scene("019_stroke_bevel_miterlimit", sfc, c)


def s020(aa):  # spec/020_stroke_lines_round_caps.zig
    width = 800
    height = 600
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.cap = abi.CAP_ROUND
    c.join = abi.JOIN_ROUND
    c.line_width = float(20)
    sub_canvas_width = width / 4
    sub_canvas_height = height / 3
    margin = 30
    x_offset = 0
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = 0
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    x_offset = sub_canvas_width * 2
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    x_offset = sub_canvas_width * 2
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    y_offset = 0
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    x_offset = sub_canvas_width * 3
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    x_offset = 0
    y_offset = sub_canvas_height * 2
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2 - 1)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2 - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 2
    sub_canvas_height = sub_canvas_height / 2
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = sub_canvas_width * 3
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 2
    y_offset = y_offset + sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    c.stroke()
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xF3, 0x00, 0x00, 0))
    c.line_width = float(1)
    c.stroke()
    return _scene("020_stroke_lines_round_caps", sfc, c)


def s021(aa):  # spec/021_stroke_lines_square_caps.zig
    width = 800
    height = 600
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.cap = abi.CAP_SQUARE
    c.line_width = float(20)
    sub_canvas_width = width / 4
    sub_canvas_height = height / 3
    margin = 30
    x_offset = 0
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = 0
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    x_offset = sub_canvas_width * 2
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    x_offset = sub_canvas_width * 2
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    y_offset = 0
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    x_offset = sub_canvas_width * 3
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    x_offset = 0
    y_offset = sub_canvas_height * 2
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2 - 1)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2 - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 2
    sub_canvas_height = sub_canvas_height / 2
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = sub_canvas_width * 3
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 2
    y_offset = y_offset + sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    c.stroke()
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xF3, 0x00, 0x00, 0))
    c.line_width = float(1)
    c.stroke()
    return _scene("021_stroke_lines_square_caps", sfc, c)


def s022(aa):  # spec/022_stroke_lines_butt_caps.zig
    width = 800
    height = 600
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.cap = abi.CAP_BUTT
    c.line_width = float(20)
    sub_canvas_width = width / 4
    sub_canvas_height = height / 3
    margin = 30
    x_offset = 0
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = 0
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    x_offset = sub_canvas_width * 2
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    x_offset = sub_canvas_width * 2
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    y_offset = 0
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    x_offset = sub_canvas_width * 3
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    x_offset = 0
    y_offset = sub_canvas_height * 2
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2 - 1)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2 - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 2
    sub_canvas_height = sub_canvas_height / 2
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = sub_canvas_width * 3
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 2
    y_offset = y_offset + sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2 - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    c.stroke()
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xF3, 0x00, 0x00, 0))
    c.line_width = float(1)
    c.stroke()
    return _scene("022_stroke_lines_butt_caps", sfc, c)


def s027(aa):  # spec/027_stroke_bezier_tolerance.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(5)
    c.move_to(19, 149)
    c.curve_to(89, 0, 209, 0, 279, 149)
    c.stroke()
    c.tolerance = float(3)
    c.reset_path()
    c.move_to(19, 199)
    c.curve_to(89, 24, 209, 24, 279, 199)
    c.stroke()
    c.tolerance = float(10)
    c.reset_path()
    c.move_to(19, 249)
    c.curve_to(89, 49, 209, 49, 279, 249)
    c.stroke()
    return _scene("027_bezier_tolerance", sfc, c)


def s029(aa):  # spec/029_stroke_lines_round_caps_tolerance.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.cap = abi.CAP_ROUND
    c.join = abi.JOIN_ROUND
    c.line_width = float(30)
    c.move_to(30, 50)
    c.line_to(270, 50)
    c.stroke()
    c.tolerance = float(3)
    c.reset_path()
    c.move_to(30, 150)
    c.line_to(270, 150)
    c.stroke()
    c.tolerance = float(10)
    c.reset_path()
    c.move_to(30, 250)
    c.line_to(270, 250)
    c.stroke()
    return _scene("029_stroke_lines_round_caps_tolerance", sfc, c)


def s030(aa):  # spec/030_stroke_star_round_tolerance.zig
    width = 900
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(30)
    c.join = abi.JOIN_ROUND
    margin = 20
    sub_canvas_width = 300
    x_scale = 3
    y_scale = 5
    c.move_to(sub_canvas_width / 2, 0 + margin)
    c.line_to(sub_canvas_width - margin * x_scale - 1, height - margin - 1)
    c.line_to(0 + margin, 0 + margin * y_scale)
    c.line_to(sub_canvas_width - margin - 1, 0 + margin * y_scale)
    c.line_to(0 + margin * x_scale, height - margin - 1)
    c.close_path()
    c.stroke()
    c.reset_path()
    c.tolerance = float(3)
    x_offset = 300
    c.move_to(x_offset + sub_canvas_width / 2, 0 + margin)
    c.line_to(x_offset + sub_canvas_width - margin * x_scale - 1, height - margin - 1)
    c.line_to(x_offset + 0 + margin, 0 + margin * y_scale)
    c.line_to(x_offset + sub_canvas_width - margin - 1, 0 + margin * y_scale)
    c.line_to(x_offset + 0 + margin * x_scale, height - margin - 1)
    c.close_path()
    c.stroke()
    c.reset_path()
    c.tolerance = float(10)
    x_offset = 600
    c.move_to(x_offset + sub_canvas_width / 2, 0 + margin)
    c.line_to(x_offset + sub_canvas_width - margin * x_scale - 1, height - margin - 1)
    c.line_to(x_offset + 0 + margin, 0 + margin * y_scale)
    c.line_to(x_offset + sub_canvas_width - margin - 1, 0 + margin * y_scale)
    c.line_to(x_offset + 0 + margin * x_scale, height - margin - 1)
    c.close_path()
    c.stroke()
    return _scene("030_stroke_star_round_tolerance", sfc, c)


def s034(aa):  # spec/034_stroke_cross.zig
    width = 250
    height = 250
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(5)
    c.move_to(100, 50)
    c.rel_line_to(0, 50)
    c.rel_line_to(-50, 0)
    c.rel_line_to(0, 50)
    c.rel_line_to(50, 0)
    c.rel_line_to(0, 50)
    c.rel_line_to(50, 0)
    c.rel_line_to(0, -50)
    c.rel_line_to(50, 0)
    c.rel_line_to(0, -50)
    c.rel_line_to(-50, 0)
    c.rel_line_to(0, -50)
    c.close_path()
    c.stroke()
    return _scene("034_stroke_cross", sfc, c)


def s036(aa):  # spec/036_stroke_colinear.zig
    width = 100
    height = 240
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(4)
    c.move_to(40, 50)
    c.line_to(35, 60)
    c.line_to(30, 70)
    c.line_to(10, 50)
    c.move_to(10, 10)
    c.line_to(20, 20)
    c.line_to(30, 30)
    c.line_to(40, 10)
    c.move_to(90, 50)
    c.line_to(85, 60)
    c.line_to(80, 70)
    c.line_to(60, 50)
    c.close_path()
    c.move_to(60, 10)
    c.line_to(70, 20)
    c.line_to(80, 30)
    c.line_to(90, 10)
    c.close_path()
    c.move_to(10, 90)
    c.line_to(25, 100)
    c.line_to(40, 110)
    c.move_to(10, 150)
    c.line_to(25, 140)
    c.line_to(40, 130)
    c.move_to(90, 90)
    c.line_to(75, 100)
    c.line_to(60, 110)
    c.move_to(90, 150)
    c.line_to(75, 140)
    c.line_to(60, 130)
    c.move_to(10, 170)
    c.line_to(30, 190)
    c.line_to(20, 180)
    c.line_to(40, 170)
    c.move_to(90, 170)
    c.line_to(70, 190)
    c.line_to(80, 180)
    c.line_to(60, 170)
    c.move_to(40, 210)
    c.line_to(30, 230)
    c.line_to(20, 220)
    c.line_to(10, 210)
    c.move_to(60, 210)
    c.line_to(70, 230)
    c.line_to(80, 220)
    c.line_to(90, 210)
    c.stroke()
    return _scene("036_stroke_colinear", sfc, c)


def s037(aa):  # spec/037_stroke_join_overlap.zig
    width = 200
    height = 200
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(5)
    c.miter_limit = float(4)
    c.move_to(10, 10)
    c.line_to(50, 90)
    c.line_to(35, 50)
    c.move_to(190, 10)
    c.line_to(150, 90)
    c.line_to(165, 50)
    c.move_to(10, 190)
    c.line_to(50, 110)
    c.line_to(35, 150)
    c.move_to(190, 190)
    c.line_to(150, 110)
    c.line_to(165, 150)
    c.stroke()
    return _scene("037_stroke_join_overlap", sfc, c)


def s038(aa):  # spec/038_stroke_zero_length.zig
    width = 60
    height = 60
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(10)
    c.cap = abi.CAP_ROUND
    c.move_to(10, 10)
    c.close_path()
    c.move_to(30, 30)
    c.line_to(30, 30)
    c.line_to(30, 30)
    c.close_path()
    c.move_to(50, 50)
    c.stroke()
    return _scene("038_stroke_zero_length", sfc, c)


def s039(aa):  # spec/039_stroke_paint_extent_dontclip.zig
    width = 50
    height = 80
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(5)
    c.move_to(40, 50)
    c.line_to(35, 60)
    c.line_to(30, 70)
    c.line_to(10, 50)
    c.stroke()
    return _scene("039_stroke_paint_extent_dontclip", sfc, c)


def s040(aa):  # spec/040_stroke_corner_symmetrical.zig
    width = 18
    height = 36
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.move_to(0, 0)
    c.line_to(width, height)
    c.stroke()
    return _scene("040_stroke_corner_symmetrical", sfc, c)


def s041(aa):  # spec/041_stroke_noop_lineto.zig
    width = 18
    height = 36
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.move_to(9, 0)
    c.line_to(9, 9)
    c.line_to(0, 18)
    c.line_to(0, 18)
    c.stroke()
    return _scene("041_stroke_noop_lineto", sfc, c)


def s054(aa):  # spec/054_stroke_lines_dashed.zig
    width = 800
    height = 400
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.dashes = [float(v) for v in ( 10, 5 ,)]
    sub_canvas_width = width / 4
    sub_canvas_height = height / 2
    margin = 10
    x_offset = 0
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    c.move_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    x_offset = 0
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + margin)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + margin, y_offset + margin)
    x_offset = sub_canvas_width * 2
    y_offset = 0
    c.move_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    x_offset = sub_canvas_width * 2
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    x_offset = sub_canvas_width * 3
    y_offset = 0
    c.move_to(x_offset + sub_canvas_width / 2, y_offset + sub_canvas_height - margin - 1)
    c.line_to(x_offset + sub_canvas_width / 2, y_offset + margin)
    x_offset = sub_canvas_width * 3
    y_offset = sub_canvas_height
    c.move_to(x_offset + sub_canvas_width - margin - 1, y_offset + sub_canvas_height / 2)
    c.line_to(x_offset + margin, y_offset + sub_canvas_height / 2)
    c.stroke()
    return _scene("054_stroke_lines_dashed", sfc, c)


def s055(aa):  # spec/055_stroke_miter_dashed.zig
    width = 240
    height = 260
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.dashes = [float(v) for v in ( 15, 5 ,)]
    c.move_to(10, 10)
    c.line_to(100, 20)
    c.line_to(110, 120)
    c.line_to(10, 110)
    c.line_to(20, 30)
    c.line_to(90, 40)
    c.line_to(95, 100)
    c.line_to(30, 95)
    c.line_to(30, 50)
    c.line_to(80, 50)
    c.line_to(75, 85)
    c.line_to(45, 80)
    c.line_to(50, 60)
    c.line_to(65, 70)
    x_offset = 120
    c.move_to(x_offset + 110, 10)
    c.line_to(x_offset + 20, 20)
    c.line_to(x_offset + 10, 120)
    c.line_to(x_offset + 110, 110)
    c.line_to(x_offset + 100, 30)
    c.line_to(x_offset + 30, 40)
    c.line_to(x_offset + 25, 100)
    c.line_to(x_offset + 90, 95)
    c.line_to(x_offset + 90, 50)
    c.line_to(x_offset + 40, 50)
    c.line_to(x_offset + 45, 85)
    c.line_to(x_offset + 75, 80)
    c.line_to(x_offset + 70, 60)
    c.line_to(x_offset + 55, 70)
    y_offset = 130
    c.move_to(10, y_offset + 120)
    c.line_to(100, y_offset + 110)
    c.line_to(110, y_offset + 10)
    c.line_to(10, y_offset + 20)
    c.line_to(20, y_offset + 100)
    c.line_to(90, y_offset + 90)
    c.line_to(95, y_offset + 30)
    c.line_to(30, y_offset + 35)
    c.line_to(30, y_offset + 80)
    c.line_to(80, y_offset + 80)
    c.line_to(75, y_offset + 45)
    c.line_to(45, y_offset + 50)
    c.line_to(50, y_offset + 70)
    c.line_to(65, y_offset + 60)
    c.move_to(x_offset + 110, y_offset + 120)
    c.line_to(x_offset + 20, y_offset + 110)
    c.line_to(x_offset + 10, y_offset + 10)
    c.line_to(x_offset + 110, y_offset + 20)
    c.line_to(x_offset + 100, y_offset + 100)
    c.line_to(x_offset + 30, y_offset + 90)
    c.line_to(x_offset + 25, y_offset + 30)
    c.line_to(x_offset + 90, y_offset + 35)
    c.line_to(x_offset + 90, y_offset + 80)
    c.line_to(x_offset + 40, y_offset + 80)
    c.line_to(x_offset + 45, y_offset + 45)
    c.line_to(x_offset + 75, y_offset + 50)
    c.line_to(x_offset + 70, y_offset + 70)
    c.line_to(x_offset + 55, y_offset + 60)
    c.stroke()
    return _scene("055_stroke_miter_dashed", sfc, c)


def s056(aa):  # spec/056_stroke_star_dashed.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.line_width = float(6)
    c.dashes = [float(v) for v in ( 12, 3 ,)]
    margin = 20
    x_scale = 3
    y_scale = 5
    c.move_to(width / 2, 0 + margin)
    c.line_to(width - margin * x_scale - 1, height - margin - 1)
    c.line_to(0 + margin, 0 + margin * y_scale)
    c.line_to(width - margin - 1, 0 + margin * y_scale)
    c.line_to(0 + margin * x_scale, height - margin - 1)
    c.close_path()
    c.stroke()
    return _scene("056_stroke_star_dashed", sfc, c)


def s057(aa):  # spec/057_stroke_bezier_dashed.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    c.dashes = [float(v) for v in ( 6, 2 ,)]
    c.move_to(19, 149)
    c.curve_to(89, 0, 209, 0, 279, 149)
    c.stroke()
    c.line_width = float(6)
    c.dashes = [float(v) for v in ( 18, 6 ,)]
    c.reset_path()
    c.move_to(19, 199)
    c.curve_to(89, 24, 209, 24, 279, 199)
    c.stroke()
    c.line_width = float(10)
    c.dashes = [float(v) for v in ( 30, 10 ,)]
    c.reset_path()
    c.move_to(19, 249)
    c.curve_to(89, 49, 209, 49, 279, 249)
    c.stroke()
    return _scene("057_stroke_bezier_dashed", sfc, c)


def s073(aa):  # spec/073_stroke_sameclose.zig
    width = 300
    height = 300
    sfc, c = _ctx(abi.RGB, width, height, aa)
    c.set_source_pixel(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    margin = 50
    c.move_to(0 + margin, 0 + margin)
    c.line_to(width - margin - 1, 0 + margin)
    c.line_to(width - margin - 1, height - margin - 1)
    c.line_to(0 + margin, height - margin - 1)
    c.line_to(0 + margin, 0 + margin)
    c.close_path()
    c.stroke()
    return _scene("073_stroke_sameclose", sfc, c)


STROKE_SCENES = [s008, s009, s010, s011, s012, s014, s015, s017, s018, s019, s020, s021, s022, s027, s029, s030, s034, s036, s037, s038, s039, s040, s041, s054, s055, s056, s057, s073]
