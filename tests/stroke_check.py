"""A CPU backend for the stroke fixtures: the library's host stroke
tessellator (z2d_stroke_polygon, no device work) feeding the oracle's
polygon rasterizer, with every other call on the oracle.  It checks the
tessellator against the reference's PNGs without a GPU."""
import ctypes as C

from z2d_amd import abi, gpu
from oracle_backend import OracleBackend, lib as oracle_lib


class StrokeCheckBackend(OracleBackend):
    name = "host-tessellator+oracle"

    def stroke(self, sfc, pattern, nodes, opts):
        rc, edges, ext, aa = gpu.stroke_polygon(nodes, opts, sfc.format)
        if rc:
            return rc
        d = sfc.desc()
        e = (C.c_double * 4)(*ext)
        return oracle_lib().zo_fill_polygon(C.byref(d), C.byref(pattern), edges.ctypes.data,
                                            len(edges), e, aa, abi.NON_ZERO, opts.op,
                                            opts.precision)
