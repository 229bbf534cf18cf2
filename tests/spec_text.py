"""Spec case 074 (spec/074_text.zig), restated: three lines of Inter Regular
at size 27 on a 900x100 RGB surface (black), white with text.show at (10, 0),
white through Context.showText at (10, 30), and a linear-RGB gradient
(10, 60) -> (900, 90) under scale(1.5, 1) at (10, 60).  The text path of
each line comes from the library's layout (z2d_text_path, the nodes
text.show hands painter.fill); `fill` renders it on any backend."""
import os

from z2d_amd import abi, gpu

FONTS = os.path.join(os.path.dirname(__file__), "golden", "fonts")
TEXT = "The quick brown fox jumps over the lázy dog"
W, H = 900, 100


def font(name="Inter-Regular.ttf"):
    with open(os.path.join(FONTS, name), "rb") as fh:
        return gpu.Font(fh.read())


def _rgb1(r, g, b):
    return abi.color(abi.COLOR_LINEAR_RGB, r, g, b, 1.0)


def lines(aa):
    """(pattern, x, y, text_opts, keep) for the three lines of the case."""
    white = abi.solid_pattern(abi.Pixel(abi.RGB, 0xFF, 0xFF, 0xFF, 0))
    grad = abi.Gradient(abi.GRADIENT_LINEAR, [10, 60, 900, 90],
                        [(0.0, _rgb1(1, 0, 0)), (0.5, _rgb1(0, 1, 0)), (1.0, _rgb1(0, 0, 1))])
    fo = abi.fill_opts(aa)
    return [
        (white, 10.0, 0.0, abi.text_opts(27.0, fill=fo), None),
        (white, 10.0, 30.0, abi.text_opts(27.0, fill=fo), None),
        (grad.pattern(), 10.0, 60.0,
         abi.text_opts(27.0, transformation=(1.5, 0.0, 0.0, 1.0, 0.0, 0.0), fill=fo), grad),
    ]


def render(backend, aa, f=None):
    """Fills the three text paths with backend.fill on a fresh surface."""
    from zscene import Surface
    f = f or font()
    s = Surface(abi.RGB, W, H)
    for pat, x, y, opts, _keep in lines(aa):
        rc, nodes = f.text_path(TEXT, x, y, opts)
        assert rc == 0, rc
        rc = backend.fill(s, pat, nodes, opts.fill)
        assert rc == 0, f"{backend.name} fill rc={rc}"
    return s
