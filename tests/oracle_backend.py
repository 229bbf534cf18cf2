"""ctypes loader for the CPU oracle (oracle/liboracle.so).  Test
infrastructure only: the product never imports this module."""
import ctypes as C
import os

import numpy as np

from z2d_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle` or build()")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.zo_fill.argtypes = [P(abi.SurfaceDesc), P(abi.PatternDesc), C.c_void_p, C.c_uint64,
                              P(abi.FillOpts)]
        L.zo_fill.restype = C.c_int32
        L.zo_fill_polygon.argtypes = [P(abi.SurfaceDesc), P(abi.PatternDesc), C.c_void_p,
                                      C.c_uint64, P(C.c_double), C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.c_uint32]
        L.zo_fill_polygon.restype = C.c_int32
        L.zo_plot_fill.argtypes = [C.c_void_p, C.c_uint64, C.c_double, C.c_double, C.c_void_p,
                                   C.c_uint64, P(C.c_double)]
        L.zo_plot_fill.restype = C.c_int64
        L.zo_composite.argtypes = [P(abi.SurfaceDesc), C.c_int32, C.c_int32, P(abi.CompOp),
                                   C.c_uint64, C.c_uint32]
        L.zo_composite.restype = None
        L.zo_run_pixel.argtypes = [C.c_uint32, abi.Pixel, abi.Pixel, C.c_uint32]
        L.zo_run_pixel.restype = abi.Pixel
        L.zo_gradient_offset.argtypes = [P(abi.GradientDesc), C.c_int32, C.c_int32]
        L.zo_gradient_offset.restype = C.c_float
        L.zo_gradient_pixel.argtypes = [P(abi.GradientDesc), C.c_int32, C.c_int32]
        L.zo_gradient_pixel.restype = abi.Pixel
        L.zo_dither_pixel.argtypes = [P(abi.DitherDesc), C.c_int32, C.c_int32]
        L.zo_dither_pixel.restype = abi.Pixel
        L.zo_sample_pattern_int.argtypes = [P(abi.PatternDesc), C.c_int32, C.c_int32,
                                            P(C.c_uint16)]
        L.zo_powf.argtypes = [C.c_float, C.c_float]
        L.zo_powf.restype = C.c_float
        L.zo_in_box.argtypes = [P(C.c_double), C.c_double, C.c_int32, C.c_int32]
        L.zo_in_box.restype = C.c_int
        L.zo_count_coverage.argtypes = [C.c_void_p, C.c_uint64, P(C.c_double), C.c_uint32,
                                        C.c_int32, C.c_int32]
        L.zo_count_coverage.restype = C.c_uint64
        L.zo_png_scanlines.argtypes = [P(abi.SurfaceDesc), C.c_uint32, C.c_void_p]
        L.zo_png_scanlines.restype = C.c_uint64
        _lib = L
    return _lib


class OracleBackend:
    """Replays scene calls on the oracle (same signature as the GPU
    backend's painter calls)."""

    name = "oracle"

    def fill(self, sfc, pattern, nodes, opts):
        d = sfc.desc()
        return lib().zo_fill(C.byref(d), C.byref(pattern), nodes.ctypes.data, len(nodes),
                             C.byref(opts))

    def stroke(self, sfc, pattern, nodes, opts):
        raise NotImplementedError("stroke tessellation is not part of the oracle yet")

    def composite(self, sfc, dst_x, dst_y, ops, precision):
        d = sfc.desc()
        arr = (abi.CompOp * len(ops))(*ops)
        lib().zo_composite(C.byref(d), dst_x, dst_y, arr, len(ops), precision)
        return 0


def plot_fill(nodes, scale, tolerance, cap=1 << 20):
    edges = np.zeros(cap, dtype=abi.EDGE_DTYPE)
    ext = (C.c_double * 4)()
    n = lib().zo_plot_fill(nodes.ctypes.data, len(nodes), scale, tolerance,
                           edges.ctypes.data, cap, ext)
    if n < 0:
        raise RuntimeError(f"plot_fill error {-n}")
    return edges[:n].copy(), tuple(ext)


def png_scanlines(fmt, width, height, data, profile=abi.PNG_PROFILE_NONE):
    """The oracle's PNG scanline stream (zo_png_scanlines) for a surface
    buffer: the bytes the reference hands its zlib compressor."""
    buf = np.ascontiguousarray(np.frombuffer(bytes(data), np.uint8))
    d = abi.SurfaceDesc(fmt, width, height, 0, buf.ctypes.data, buf.nbytes)
    n = lib().zo_png_scanlines(C.byref(d), profile, None)
    out = np.zeros(n, np.uint8)
    lib().zo_png_scanlines(C.byref(d), profile, out.ctypes.data)
    return out.tobytes()
