"""ctypes binding of libz2d_gpu.so (include/z2d_gpu.h).

Loading fails loudly when the library is missing: there is no CPU fallback
on the product path."""
import ctypes as C
import os

import numpy as np

from . import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libz2d_gpu.so")

_lib = None

P = C.POINTER
_SIGS = {
    "z2d_gpu_create": ([C.c_int32, P(C.c_void_p)], C.c_int32),
    "z2d_gpu_destroy": ([C.c_void_p], None),
    "z2d_gpu_last_error": ([C.c_void_p], C.c_char_p),
    "z2d_gpu_sync": ([C.c_void_p], C.c_int32),
    "z2d_gpu_fill": ([C.c_void_p, P(abi.SurfaceDesc), P(abi.PatternDesc), C.c_void_p,
                      C.c_uint64, P(abi.FillOpts)], C.c_int32),
    "z2d_gpu_stroke": ([C.c_void_p, P(abi.SurfaceDesc), P(abi.PatternDesc), C.c_void_p,
                        C.c_uint64, P(abi.StrokeOpts)], C.c_int32),
    "z2d_gpu_fill_polygon": ([C.c_void_p, P(abi.SurfaceDesc), P(abi.PatternDesc), C.c_void_p,
                              C.c_uint64, P(C.c_double), C.c_uint32, C.c_uint32, C.c_uint32,
                              C.c_uint32], C.c_int32),
    "z2d_gpu_composite": ([C.c_void_p, P(abi.SurfaceDesc), C.c_int32, C.c_int32, P(abi.CompOp),
                           C.c_uint64, C.c_uint32, P(abi.SurfaceDesc)], C.c_int32),
    "z2d_gpu_surface_create": ([C.c_void_p, C.c_uint32, C.c_int32, C.c_int32, P(C.c_void_p)],
                               C.c_int32),
    "z2d_gpu_surface_upload": ([C.c_void_p, C.c_void_p, C.c_uint64], C.c_int32),
    "z2d_gpu_surface_download": ([C.c_void_p, C.c_void_p, C.c_uint64], C.c_int32),
    "z2d_gpu_surface_desc": ([C.c_void_p, P(abi.SurfaceDesc)], C.c_int32),
    "z2d_gpu_surface_destroy": ([C.c_void_p], None),
    "z2d_gpu_batch_create": ([C.c_void_p, P(C.c_void_p)], C.c_int32),
    "z2d_gpu_batch_fill": ([C.c_void_p, C.c_void_p, P(abi.PatternDesc), C.c_void_p, C.c_uint64,
                            P(abi.FillOpts)], C.c_int32),
    "z2d_gpu_batch_stroke": ([C.c_void_p, C.c_void_p, P(abi.PatternDesc), C.c_void_p,
                              C.c_uint64, P(abi.StrokeOpts)], C.c_int32),
    "z2d_gpu_batch_draw_n": ([C.c_void_p, P(C.c_void_p), P(abi.DrawItem), C.c_uint64, C.c_uint32,
                              P(C.c_int32)], C.c_int32),
    "z2d_gpu_batch_fill_polygon": ([C.c_void_p, C.c_void_p, P(abi.PatternDesc), C.c_void_p,
                                    C.c_uint64, P(C.c_double), C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_uint32], C.c_int32),
    "z2d_gpu_batch_prepare": ([C.c_void_p], C.c_int32),
    "z2d_gpu_batch_render": ([C.c_void_p, C.c_uint32], C.c_int32),
    "z2d_gpu_batch_wait": ([C.c_void_p], C.c_int32),
    "z2d_gpu_batch_timing": ([C.c_void_p, P(C.c_double), P(C.c_double)], C.c_int32),
    "z2d_gpu_batch_use_graph": ([C.c_void_p, C.c_uint32], C.c_int32),
    "z2d_gpu_batch_set_band_height": ([C.c_void_p, C.c_uint32], C.c_int32),
    "z2d_gpu_batch_set_queue_order": ([C.c_void_p, C.c_uint32], C.c_int32),
    "z2d_gpu_batch_autotune": ([C.c_void_p, C.c_uint32, C.c_uint32, P(C.c_uint32),
                                P(C.c_uint32), P(C.c_double)], C.c_int32),
    "z2d_gpu_composite_timing": ([C.c_void_p, P(C.c_double)], C.c_int32),
    "z2d_gpu_batch_stats": ([C.c_void_p, P(C.c_uint64)], C.c_int32),
    "z2d_gpu_batch_counters": ([C.c_void_p, C.c_uint32, P(C.c_uint64)], C.c_int32),
    "z2d_gpu_batch_reset": ([C.c_void_p], C.c_int32),
    "z2d_gpu_batch_destroy": ([C.c_void_p], None),
    "z2d_font_load": ([C.c_void_p, C.c_uint64, P(C.c_void_p)], C.c_int32),
    "z2d_font_destroy": ([C.c_void_p], None),
    "z2d_font_info": ([C.c_void_p, P(C.c_uint32)], C.c_int32),
    "z2d_font_glyph": ([C.c_void_p, C.c_uint32, P(C.c_int32)], C.c_int32),
    "z2d_text_path": ([C.c_void_p, C.c_char_p, C.c_uint64, C.c_double, C.c_double,
                       P(abi.TextOpts), C.c_void_p, C.c_uint64, P(C.c_uint64)], C.c_int32),
    "z2d_gpu_text_show": ([C.c_void_p, P(abi.SurfaceDesc), P(abi.PatternDesc), C.c_void_p,
                           C.c_char_p, C.c_uint64, C.c_double, C.c_double, P(abi.TextOpts)],
                          C.c_int32),
    "z2d_gpu_batch_text": ([C.c_void_p, C.c_void_p, P(abi.PatternDesc), C.c_void_p, C.c_char_p,
                            C.c_uint64, C.c_double, C.c_double, P(abi.TextOpts)], C.c_int32),
    "z2d_gpu_png_bound": ([C.c_uint32, C.c_int32, C.c_int32, C.c_uint32, P(C.c_uint64)],
                          C.c_int32),
    "z2d_gpu_png_encode": ([C.c_void_p, P(abi.SurfaceDesc), C.c_uint32, C.c_void_p, C.c_uint64,
                           P(C.c_uint64)], C.c_int32),
    "z2d_gpu_png_timing": ([C.c_void_p, P(C.c_double)], C.c_int32),
    "z2d_gpu_stream_copy": ([C.c_void_p, C.c_uint64, C.c_uint32, P(C.c_double)], C.c_int32),
    "z2d_gpu_stream_read2": ([C.c_void_p, C.c_uint64, C.c_uint32, P(C.c_double)], C.c_int32),
    "z2d_stroke_polygon": ([C.c_void_p, C.c_uint64, P(abi.StrokeOpts), C.c_uint32, C.c_void_p,
                            C.c_uint64, P(C.c_uint64), P(C.c_double), P(C.c_uint32)], C.c_int32),
    "z2d_gpu_server_create": ([C.c_void_p, C.c_uint64, C.c_uint32, P(C.c_void_p)], C.c_int32),
    "z2d_gpu_server_submit": ([C.c_void_p, P(abi.RequestDesc), C.c_void_p, C.c_uint64,
                               P(C.c_uint64)], C.c_int32),
    "z2d_gpu_server_poll": ([C.c_void_p, C.c_uint64, P(C.c_int32)], C.c_int32),
    "z2d_gpu_server_wait": ([C.c_void_p, C.c_uint64, P(C.c_int32), C.c_double], C.c_int32),
    "z2d_gpu_server_wait_any": ([C.c_void_p, P(C.c_uint64), P(C.c_int32), C.c_double],
                                C.c_int32),
    "z2d_gpu_server_submit_n": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                 C.c_void_p, C.c_void_p], C.c_int32),
    "z2d_gpu_server_wait_some": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                  C.c_double], C.c_int32),
    "z2d_gpu_server_stats": ([C.c_void_p, P(C.c_uint64)], C.c_int32),
    "z2d_gpu_server_destroy": ([C.c_void_p], None),
    "z2d_gpu_composite_noalloc": ([C.c_void_p, P(abi.SurfaceDesc), C.c_int32, C.c_int32,
                                   C.c_void_p, C.c_uint64, C.c_uint32, C.c_void_p], C.c_int32),
    "z2d_gpu_composite_noalloc_max_bytes": ([], C.c_uint64),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGS)


class GpuError(RuntimeError):
    pass


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        rc = lib().z2d_gpu_create(device, C.byref(h))
        if rc != 0:
            raise GpuError(f"z2d_gpu_create failed: {rc}")
        self.h = h

    def close(self):
        if self.h:
            lib().z2d_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self):
        return lib().z2d_gpu_last_error(self.h).decode()

    def check(self, rc, what):
        if rc != 0:
            raise GpuError(f"{what} returned {rc}: {self.last_error()}")

    def sync(self):
        self.check(lib().z2d_gpu_sync(self.h), "sync")


class DeviceSurface:
    """A surface resident in HBM (z2d_gpu_surface)."""

    def __init__(self, ctx, fmt, width, height):
        self.ctx = ctx
        self.format, self.width, self.height = fmt, width, height
        self.nbytes = abi.surface_nbytes(fmt, width, height)
        h = C.c_void_p()
        ctx.check(lib().z2d_gpu_surface_create(ctx.h, fmt, width, height, C.byref(h)),
                  "surface_create")
        self.h = h

    def upload(self, arr):
        self.ctx.check(lib().z2d_gpu_surface_upload(self.h, arr.ctypes.data, arr.nbytes),
                       "surface_upload")

    def download(self, arr):
        self.ctx.check(lib().z2d_gpu_surface_download(self.h, arr.ctypes.data, arr.nbytes),
                       "surface_download")

    def desc(self):
        d = abi.SurfaceDesc()
        self.ctx.check(lib().z2d_gpu_surface_desc(self.h, C.byref(d)), "surface_desc")
        return d

    def close(self):
        if self.h:
            lib().z2d_gpu_surface_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many fills on resident surfaces rendered by one pipeline launch."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().z2d_gpu_batch_create(ctx.h, C.byref(h)), "batch_create")
        self.h = h

    def fill(self, surface, pattern, nodes, opts):
        return lib().z2d_gpu_batch_fill(self.h, surface.h, C.byref(pattern), nodes.ctypes.data,
                                        len(nodes), C.byref(opts))

    def text(self, surface, pattern, font, text, x, y, opts):
        b = text.encode("utf-8") if isinstance(text, str) else bytes(text)
        return lib().z2d_gpu_batch_text(self.h, surface.h, C.byref(pattern), font.h, b, len(b),
                                        x, y, C.byref(opts))

    def stroke(self, surface, pattern, nodes, opts):
        return lib().z2d_gpu_batch_stroke(self.h, surface.h, C.byref(pattern), nodes.ctypes.data,
                                          len(nodes), C.byref(opts))

    def draw_n(self, jobs, threads=0):
        """jobs: list of (surface, "fill" | "stroke", pattern, nodes, opts),
        appended in order as per-item fill / stroke calls would; the
        strokes' host tessellation runs on `threads` host threads (0: the
        machine's).  Returns the per-item status list."""
        n = len(jobs)
        sfc = (C.c_void_p * max(1, n))()
        items = (abi.DrawItem * max(1, n))()
        for i, (surface, kind, pat, nodes, opts) in enumerate(jobs):
            sfc[i] = surface.h
            items[i].pattern = C.pointer(pat)
            items[i].nodes = nodes.ctypes.data
            items[i].n_nodes = len(nodes)
            if kind == "stroke":
                items[i].kind = abi.DRAW_STROKE
                items[i].stroke = C.pointer(opts)
            else:
                items[i].kind = abi.DRAW_FILL
                items[i].fill = opts
        st = (C.c_int32 * max(1, n))()
        lib().z2d_gpu_batch_draw_n(self.h, sfc, items, n, threads, st)
        return list(st[:n])

    def fill_polygon(self, surface, pattern, edges, extents, aa, rule, op, precision):
        ext = (C.c_double * 4)(*extents)
        return lib().z2d_gpu_batch_fill_polygon(self.h, surface.h, C.byref(pattern),
                                                edges.ctypes.data, len(edges), ext, aa, rule, op,
                                                precision)

    def prepare(self):
        self.ctx.check(lib().z2d_gpu_batch_prepare(self.h), "batch_prepare")

    def render(self, mode=0):
        self.ctx.check(lib().z2d_gpu_batch_render(self.h, mode), "batch_render")

    def wait(self):
        self.ctx.check(lib().z2d_gpu_batch_wait(self.h), "batch_wait")

    def set_band_height(self, rows):
        self.ctx.check(lib().z2d_gpu_batch_set_band_height(self.h, rows), "set_band_height")

    def set_queue_order(self, longest_first):
        self.ctx.check(lib().z2d_gpu_batch_set_queue_order(self.h, 1 if longest_first else 0),
                       "set_queue_order")

    def autotune(self, mode=0, reps=2):
        """Pick the band height (8, 16, 32) and queue order (natural,
        longest-first) with the fastest pipeline; returns (band_h,
        longest_first, {(band_h, longest_first): mean ms})."""
        h, o = C.c_uint32(), C.c_uint32()
        ms = (C.c_double * 6)()
        self.ctx.check(lib().z2d_gpu_batch_autotune(self.h, mode, reps, C.byref(h), C.byref(o),
                                                    ms), "autotune")
        keys = [(8, 0), (16, 0), (32, 0), (8, 1), (16, 1), (32, 1)]
        return h.value, o.value, dict(zip(keys, list(ms)))

    def use_graph(self, enable=True):
        """Replay the render pipeline as one captured hipGraph."""
        self.ctx.check(lib().z2d_gpu_batch_use_graph(self.h, 1 if enable else 0), "use_graph")

    def timing(self):
        a, r = C.c_double(), C.c_double()
        lib().z2d_gpu_batch_timing(self.h, C.byref(a), C.byref(r))
        return a.value, r.value

    def stats(self):
        out = (C.c_uint64 * 4)()
        lib().z2d_gpu_batch_stats(self.h, out)
        return tuple(out)

    def counters(self, enable=True):
        """(pixel_draws, edges, tile_entries, bin_overflow, unpacked_entries,
        edge_list_capacity, 0, 0) of the last render."""
        out = (C.c_uint64 * 8)()
        self.ctx.check(lib().z2d_gpu_batch_counters(self.h, 1 if enable else 0, out), "counters")
        return tuple(out)

    def reset(self):
        lib().z2d_gpu_batch_reset(self.h)

    def close(self):
        if self.h:
            lib().z2d_gpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuBackend:
    """Replays scene calls through the C ABI (host-buffer painter calls)."""

    name = "gpu"

    def __init__(self, ctx=None):
        self.ctx = ctx or Context()

    def fill(self, sfc, pattern, nodes, opts):
        d = sfc.desc()
        rc = lib().z2d_gpu_fill(self.ctx.h, C.byref(d), C.byref(pattern), nodes.ctypes.data,
                                len(nodes), C.byref(opts))
        return rc

    def stroke(self, sfc, pattern, nodes, opts):
        d = sfc.desc()
        return lib().z2d_gpu_stroke(self.ctx.h, C.byref(d), C.byref(pattern), nodes.ctypes.data,
                                    len(nodes), C.byref(opts))

    def composite(self, sfc, dst_x, dst_y, ops, precision, out=None):
        d = sfc.desc()
        arr = (abi.CompOp * len(ops))(*ops)
        od = out.desc() if out is not None else None
        return lib().z2d_gpu_composite(self.ctx.h, C.byref(d), dst_x, dst_y, arr, len(ops),
                                       precision, C.byref(od) if od is not None else None)


class Font:
    """Font.loadBuffer (z2d_font_load): a validated copy of a TrueType font."""

    def __init__(self, data):
        self.data = bytes(data)
        h = C.c_void_p()
        rc = lib().z2d_font_load(self.data, len(self.data), C.byref(h))
        if rc != 0:
            raise GpuError(f"z2d_font_load returned {rc}")
        self.h = h

    def info(self):
        out = (C.c_uint32 * 4)()
        rc = lib().z2d_font_info(self.h, out)
        if rc != 0:
            raise GpuError(f"z2d_font_info returned {rc}")
        return dict(units_per_em=out[0], advance_width_max=out[1], number_of_hmetrics=out[2],
                    lsb_is_at_x_zero=bool(out[3] & 1), loca_long=bool(out[3] & 2),
                    cmap_full=bool(out[3] & 4))

    def glyph(self, codepoint):
        out = (C.c_int32 * 4)()
        rc = lib().z2d_font_glyph(self.h, codepoint, out)
        if rc != 0:
            raise GpuError(f"z2d_font_glyph returned {rc}")
        return dict(index=out[0], advance=out[1], lsb=out[2],
                    offset=None if out[3] < 0 else out[3])

    def text_path(self, text, x, y, opts):
        """The node array text.show hands painter.fill (z2d_text_path)."""
        b = text.encode("utf-8") if isinstance(text, str) else bytes(text)
        n = C.c_uint64()
        rc = lib().z2d_text_path(self.h, b, len(b), x, y, C.byref(opts), None, 0, C.byref(n))
        if rc not in (0, -1):
            return rc, None
        arr = np.zeros(n.value, dtype=abi.NODE_DTYPE)
        rc = lib().z2d_text_path(self.h, b, len(b), x, y, C.byref(opts), arr.ctypes.data,
                                 n.value, C.byref(n))
        return rc, arr

    def close(self):
        if getattr(self, "h", None):
            lib().z2d_font_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def text_show(ctx, desc, pattern, font, text, x, y, opts):
    """text.show on the device (z2d_gpu_text_show)."""
    b = text.encode("utf-8") if isinstance(text, str) else bytes(text)
    return lib().z2d_gpu_text_show(ctx.h, C.byref(desc), C.byref(pattern), font.h, b, len(b),
                                   x, y, C.byref(opts))


def png_bound(fmt, width, height, profile=abi.PNG_PROFILE_NONE):
    n = C.c_uint64()
    rc = lib().z2d_gpu_png_bound(fmt, width, height, profile, C.byref(n))
    if rc != 0:
        raise GpuError(f"z2d_gpu_png_bound returned {rc}")
    return n.value


def png_encode(ctx, desc, profile=abi.PNG_PROFILE_NONE):
    """writeToPNGFile into memory (z2d_gpu_png_encode): the PNG file's bytes
    for a host or device surface descriptor."""
    cap = png_bound(desc.format, desc.width, desc.height, profile)
    buf = C.create_string_buffer(cap)
    n = C.c_uint64()
    ctx.check(lib().z2d_gpu_png_encode(ctx.h, C.byref(desc), profile, buf, cap, C.byref(n)),
              "z2d_gpu_png_encode")
    return buf.raw[:n.value]


def png_timing(ctx):
    ms = C.c_double()
    ctx.check(lib().z2d_gpu_png_timing(ctx.h, C.byref(ms)), "z2d_gpu_png_timing")
    return ms.value


class Server:
    """Request loop with continuous batching under an HBM budget
    (z2d_gpu_server_*).  Each request is one surface: cleared to `clear`,
    the fills applied in order, then the optional alpha8 clip mask."""

    STATS = ("submitted", "completed", "rejected", "batches", "pixel_draws", "bytes_in_use",
             "peak_bytes", "queued")

    def __init__(self, ctx, hbm_budget, max_batch=64):
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().z2d_gpu_server_create(ctx.h, hbm_budget, max_batch, C.byref(h)),
                  "server_create")
        self.h = h
        self._pending = {}
        self._written = {}

    @staticmethod
    def prepare(fmt, width, height, fills, clear=None, clip=None, draws=None, encode=0):
        """Builds the request descriptor once (it may be submitted many
        times).  fills: list of (pattern, nodes, opts); draws (after the
        fills, in order): list of ("fill" | "stroke", pattern, nodes, opts);
        encode: 0 = surface bytes, 1 + PNG profile = a PNG file (its length
        lands in prep[-1], a c_uint64)."""
        items = (abi.FillItem * max(1, len(fills)))()
        for i, (pat, nodes, opts) in enumerate(fills):
            items[i].pattern = C.pointer(pat)
            items[i].nodes = nodes.ctypes.data
            items[i].n_nodes = len(nodes)
            items[i].opts = opts
        draws = draws or []
        ditems = (abi.DrawItem * max(1, len(draws)))()
        for i, (kind, pat, nodes, opts) in enumerate(draws):
            ditems[i].pattern = C.pointer(pat)
            ditems[i].nodes = nodes.ctypes.data
            ditems[i].n_nodes = len(nodes)
            if kind == "stroke":
                ditems[i].kind = abi.DRAW_STROKE
                ditems[i].stroke = C.pointer(opts)
            else:
                ditems[i].kind = abi.DRAW_FILL
                ditems[i].fill = opts
        rq = abi.RequestDesc()
        rq.format, rq.width, rq.height = fmt, width, height
        rq.clear = clear if clear is not None else abi.Pixel()
        rq.clip_mask = clip.ctypes.data if clip is not None else None
        rq.fills = items
        rq.n_fills = len(fills)
        rq.draws = ditems
        rq.n_draws = len(draws)
        rq.encode = encode
        written = C.c_uint64()
        rq.out_written = C.pointer(written)
        # the descriptor and what it points into
        return (rq, items, fills, clip, ditems, draws, written)

    def submit_prepared(self, prep, out=None):
        """Returns (rc, id); `out` (numpy uint8, >= surface bytes) receives
        the result and must stay alive until the request completes."""
        rid = C.c_uint64()
        w = C.c_uint64()
        prep[0].out_written = C.pointer(w)  # copied at submit: one counter per request
        rc = lib().z2d_gpu_server_submit(self.h, C.byref(prep[0]),
                                         out.ctypes.data if out is not None else None,
                                         out.nbytes if out is not None else 0, C.byref(rid))
        if rc == 0:
            self._pending[rid.value] = (out, w)  # keeps the result buffer alive
        return rc, rid.value

    def submit_many(self, preps, outs=None):
        """Submits prepared requests in one call (z2d_gpu_server_submit_n).
        Returns (rc, ids, statuses): numpy arrays, id 0 where refused.  Each
        out (numpy uint8) must stay alive until its request completes."""
        n = len(preps)
        rq = np.empty(n, np.uint64)
        op = np.zeros(n, np.uint64)
        ol = np.zeros(n, np.uint64)
        counters = []
        for i, prep in enumerate(preps):
            w = C.c_uint64()
            prep[0].out_written = C.pointer(w)  # copied at submit
            counters.append(w)
            rq[i] = C.addressof(prep[0])
            if outs is not None and outs[i] is not None:
                op[i] = outs[i].ctypes.data
                ol[i] = outs[i].nbytes
        ids = np.zeros(n, np.uint64)
        st = np.zeros(n, np.int32)
        rc = self.submit_n_raw(rq.ctypes.data, op.ctypes.data, ol.ctypes.data, n,
                               ids.ctypes.data, st.ctypes.data)
        for i in range(n):
            if ids[i]:
                self._pending[int(ids[i])] = (outs[i] if outs is not None else None, counters[i])
        return rc, ids, st

    def submit_n_raw(self, req_ptrs, out_ptrs, out_lens, n, ids, status):
        """z2d_gpu_server_submit_n on addresses (arrays of descriptor
        pointers, output pointers and lengths; ids and statuses written).
        The caller keeps every descriptor and output alive; results of
        requests submitted this way are not tracked by `written`."""
        return lib().z2d_gpu_server_submit_n(self.h, req_ptrs, out_ptrs, out_lens, n, ids, status)

    def wait_some(self, ids, status, t_done, timeout_ms=60000.0):
        """Up to len(ids) completed requests (numpy uint64 / int32 / float64
        arrays filled in place; t_done in seconds of CLOCK_MONOTONIC, the
        clock of time.monotonic()).  Returns their count, 0 on timeout."""
        k = lib().z2d_gpu_server_wait_some(self.h, ids.ctypes.data, status.ctypes.data,
                                           t_done.ctypes.data, len(ids), timeout_ms)
        if k < 0:
            raise GpuError(f"server_wait_some returned {k}")
        if self._pending:
            for i in range(k):
                self._done(int(ids[i]))
        return k

    def _done(self, rid):
        entry = self._pending.pop(rid, None)
        if entry is not None:
            self._written[rid] = entry[1].value

    def written(self, rid):
        """Bytes written to the completed request's `out` (the PNG file's
        length for an encoded request, or the size it needed)."""
        return self._written.pop(rid, None)

    def submit(self, fmt, width, height, fills, clear=None, clip=None, out=None, draws=None,
               encode=0):
        """fills: list of (pattern, nodes, opts); draws: list of
        ("fill" | "stroke", pattern, nodes, opts).  Returns (rc, id)."""
        return self.submit_prepared(
            self.prepare(fmt, width, height, fills, clear, clip, draws, encode), out)

    def wait(self, rid, timeout_ms=60000.0):
        st = C.c_int32()
        rc = lib().z2d_gpu_server_wait(self.h, rid, C.byref(st), timeout_ms)
        if rc == 1:
            self._done(rid)
            return st.value
        if rc == 0:
            raise TimeoutError(f"request {rid} not done after {timeout_ms} ms")
        raise GpuError(f"server_wait returned {rc}")

    def wait_any(self, timeout_ms=60000.0):
        """(id, status) of the next completed request, or None on timeout."""
        rid, st = C.c_uint64(), C.c_int32()
        rc = lib().z2d_gpu_server_wait_any(self.h, C.byref(rid), C.byref(st), timeout_ms)
        if rc == 1:
            self._done(rid.value)
            return rid.value, st.value
        if rc == 0:
            return None
        raise GpuError(f"server_wait_any returned {rc}")

    def stats(self):
        out = (C.c_uint64 * 8)()
        lib().z2d_gpu_server_stats(self.h, out)
        return dict(zip(self.STATS, out))

    def close(self):
        if self.h:
            lib().z2d_gpu_server_destroy(self.h)
            self.h = None
            self._pending.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def stroke_polygon(nodes, opts, surface_format):
    """Host stroke tessellation (z2d_stroke_polygon): returns (rc, edges,
    extents, aa_mode) with edges as an EDGE_DTYPE array in scaled space."""
    import numpy as np
    n = C.c_uint64()
    ext = (C.c_double * 4)()
    aa = C.c_uint32()
    rc = lib().z2d_stroke_polygon(nodes.ctypes.data, len(nodes), C.byref(opts), surface_format,
                                  None, 0, C.byref(n), ext, C.byref(aa))
    if rc:
        return rc, None, None, None
    edges = np.zeros(n.value, dtype=abi.EDGE_DTYPE)
    rc = lib().z2d_stroke_polygon(nodes.ctypes.data, len(nodes), C.byref(opts), surface_format,
                                  edges.ctypes.data, len(edges), C.byref(n), ext, C.byref(aa))
    return rc, edges, tuple(ext), aa.value
