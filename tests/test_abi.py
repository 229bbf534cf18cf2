"""CPU-side checks of the C ABI: the library loads, exports every function
include/z2d_gpu.h declares, and the host-side argument checks return the
reference's error codes (src/painter.zig:73-82, :221-234) without a device."""
import ctypes as C
import os
import re

import numpy as np

from z2d_amd import abi, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "z2d_gpu.h")).read()
    return sorted(set(re.findall(r"\b(z2d_(?:gpu|stroke|font|text)_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = gpu.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert set(declared_functions()) == set(gpu.exported_symbols())


def test_struct_sizes_match_header():
    assert C.sizeof(abi.PathNode) == 56
    assert C.sizeof(abi.Edge) == 32
    assert C.sizeof(abi.Pixel) == 8
    assert C.sizeof(abi.SurfaceDesc) == 32


def _fill(pattern, nodes, opts=None):
    buf = np.zeros(16, np.uint8)
    d = abi.SurfaceDesc(abi.RGBA, 2, 2, 0, buf.ctypes.data, 16)
    opts = opts or abi.fill_opts()
    arr = abi.nodes_array(nodes) if nodes else np.zeros(0, abi.NODE_DTYPE)
    return gpu.lib().z2d_gpu_fill(None, C.byref(d), C.byref(pattern), arr.ctypes.data, len(arr),
                                  C.byref(opts))


def test_not_premultiplied_is_rejected_first():
    # painter.zig "fill, non pre-multiplied pixel" (src/painter.zig:369-383)
    p = abi.solid_pattern(abi.rgba(0xFF, 0xFF, 0xFF, 0xAA))
    assert _fill(p, []) == abi.ERR_PIXEL_SOURCE_NOT_PREMULTIPLIED


def test_zero_nodes_is_a_noop():
    assert _fill(abi.solid_pattern(abi.rgba(1, 1, 1, 255)), []) == abi.OK


def test_unclosed_path_is_rejected():
    nodes = [(abi.MOVE_TO, 0, 0), (abi.LINE_TO, 1, 1), (abi.LINE_TO, 0, 1)]
    assert _fill(abi.solid_pattern(abi.rgba(1, 1, 1, 255)), nodes) == abi.ERR_PATH_NOT_CLOSED


def test_valid_fill_needs_a_context():
    nodes = [(abi.MOVE_TO, 0, 0), (abi.LINE_TO, 1, 1), (abi.LINE_TO, 0, 1), (abi.CLOSE_PATH,)]
    assert _fill(abi.solid_pattern(abi.rgba(1, 1, 1, 255)), nodes) == abi.ERR_NO_CONTEXT


def test_div255_identity():
    """The renderer's imul8: (x * 0x8081) >> 23 == x // 255 for every x below
    2^16 (products of two 8-bit channels), and fails beyond that range."""
    import numpy as np
    x = np.arange(1 << 16, dtype=np.uint64)
    assert np.array_equal((x * 0x8081) >> 23, x // 255)
    assert (66299 * 0x8081) >> 23 != 66299 // 255


def test_node_walk_errors_before_the_device():
    """The node walk's own errors come from the host before any device work:
    a line before the first move is InvalidState (fill_plotter.zig:52-71);
    a non-finite coordinate on a node the walk uses (where the reference
    panics in Polygon.inBox, src/internal/tess/Polygon.zig:146-178, or
    recurses without end in Spline.decompose) is an invalid argument; the
    coordinates of a final move, which the walk ignores, are not checked."""
    red = abi.solid_pattern(abi.rgba(1, 1, 1, 255))
    inf, nan = float("inf"), float("nan")
    tri = [(abi.LINE_TO, 1, 1), (abi.LINE_TO, 0, 1), (abi.CLOSE_PATH,)]
    assert _fill(red, [(abi.MOVE_TO, 0, 0)] + tri) == abi.ERR_NO_CONTEXT
    assert _fill(red, [(abi.LINE_TO, 0, 0)] + tri) == abi.ERR_INVALID_STATE
    assert _fill(red, [(abi.MOVE_TO, nan, 0)] + tri) == abi.ERR_INVALID_ARGUMENT
    assert _fill(red, [(abi.MOVE_TO, 0, 0), (abi.LINE_TO, 1, inf)] + tri[1:]) == \
        abi.ERR_INVALID_ARGUMENT
    assert _fill(red, [(abi.MOVE_TO, 0, 0), (abi.CURVE_TO, 1, 1, -inf, 2, 3, 3)] + tri) == \
        abi.ERR_INVALID_ARGUMENT
    # a final move is ignored by the walk, whatever its coordinates
    assert _fill(red, [(abi.MOVE_TO, 0, 0)] + tri + [(abi.MOVE_TO, nan, nan)]) == \
        abi.ERR_NO_CONTEXT
