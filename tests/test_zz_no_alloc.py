"""z2d_gpu_composite_noalloc (the compositor entry point that allocates no
device or pinned memory; its job lives in a fixed per-device scratch)
against z2d_gpu_composite through the C ABI: two calls in a row, the second
on a smaller input, each bit-exact with the allocating entry point
(SurfaceCompositor.run, src/compositor.zig:302-440).  Also its refusals:
host descriptors, and a job larger than the scratch."""
import ctypes as C

import numpy as np
import pytest

from z2d_amd import abi
from z2d_amd.scenes import SplitMix64, premultiplied_rgba, rgba_from_clamped
from gpu_util import diff_report, gpu_backend

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[no-alloc] {msg}", flush=True)


def _device_surface(ctx, rng, fmt, w, h):
    from z2d_amd.gpu import DeviceSurface
    s = DeviceSurface(ctx, fmt, w, h)
    data = rng.u8_array(abi.surface_nbytes(fmt, w, h))
    if fmt in (abi.RGBA, abi.ARGB):
        data = premultiplied_rgba(rng.next(), w * h).reshape(-1)
    s.upload(np.ascontiguousarray(data, np.uint8))
    return s


def _ops(rng, w, h, src_desc, n_stops):
    """A 3-operation batch whose job uses the scratch: a gradient source
    over the destination, a surface source through multiply, then a
    dithered pixel through dst_in."""
    keep = []
    g = abi.Gradient(abi.GRADIENT_RADIAL,
                     [w * 0.3, h * 0.4, 2.0, w * 0.6, h * 0.5, max(w, h) * 0.7],
                     [(i / max(1, n_stops - 1),
                       abi.color(0, rng.uniform(), rng.uniform(), rng.uniform(),
                                 rng.uniform(0.3, 1.0)))
                      for i in range(n_stops)])
    keep.append(g)
    o0 = abi.CompOp()
    o0.op = abi.OP["src_over"]
    o0.src.kind = abi.PARAM_GRADIENT
    o0.src.gradient = C.pointer(g.desc)
    o1 = abi.CompOp()
    o1.op = abi.OP["multiply"]
    o1.dst.kind = abi.PARAM_NONE
    o1.src.kind = abi.PARAM_SURFACE
    o1.src.surface = C.pointer(src_desc)
    o2 = abi.CompOp()
    o2.op = abi.OP["dst_in"]
    o2.src.kind = abi.PARAM_DITHER
    o2.src.dither.type = abi.DITHER_BLUE_NOISE
    o2.src.dither.source = abi.DITHER_SRC_PIXEL
    o2.src.dither.pixel = rgba_from_clamped(0.2, 0.5, 0.9, 0.75)
    o2.src.dither.scale = 8
    return (abi.CompOp * 3)(o0, o1, o2), keep


def _call_both(ctx, rng, fmt, w, h, n_stops):
    from z2d_amd.gpu import DeviceSurface, lib
    dst = _device_surface(ctx, rng, fmt, w, h)
    src = _device_surface(ctx, rng, abi.ARGB, w + 3, h + 2)
    out_ref = DeviceSurface(ctx, fmt, w, h)
    out_na = DeviceSurface(ctx, fmt, w, h)
    d_dst, d_src, d_ref, d_na = dst.desc(), src.desc(), out_ref.desc(), out_na.desc()
    ops, keep = _ops(rng, w, h, d_src, n_stops)
    rc = lib().z2d_gpu_composite(ctx.h, C.byref(d_dst), 0, 0, ops, len(ops),
                                 abi.PRECISION_INTEGER, C.byref(d_ref))
    assert rc == 0, ctx.last_error()
    rc = lib().z2d_gpu_composite_noalloc(ctx.h, C.byref(d_dst), 0, 0,
                                         C.cast(ops, C.c_void_p), len(ops),
                                         abi.PRECISION_INTEGER, C.addressof(d_na))
    assert rc == 0, ctx.last_error()
    want = np.empty(abi.surface_nbytes(fmt, w, h), np.uint8)
    got = np.empty_like(want)
    out_ref.download(want)
    out_na.download(got)
    for s in (dst, src, out_ref, out_na):
        s.close()
    return got, want, keep


def test_noalloc_matches_composite_on_two_calls():
    ctx = gpu_backend().ctx
    rng = SplitMix64(0xA110C)
    _log("call 1: 301x97 RGBA, radial gradient of 7 stops, surface, dithered pixel")
    got, want, _ = _call_both(ctx, rng, abi.RGBA, 301, 97, 7)
    msg = diff_report(got, want)
    assert not msg, f"call 1: {msg}"
    assert np.count_nonzero(want) > 0
    _log("call 2: 37x11 alpha8, 2 stops (smaller job, same scratch)")
    got, want, _ = _call_both(ctx, rng, abi.ALPHA8, 37, 11, 2)
    msg = diff_report(got, want)
    assert not msg, f"call 2: {msg}"


def test_noalloc_refusals():
    from z2d_amd.gpu import DeviceSurface, lib
    ctx = gpu_backend().ctx
    rng = SplitMix64(0xA110D)
    w, h = 16, 8
    assert lib().z2d_gpu_composite_noalloc_max_bytes() == 1 << 20
    # a host descriptor anywhere is refused
    host = np.zeros(w * h * 4, np.uint8)
    hd = abi.SurfaceDesc(abi.RGBA, w, h, 0, host.ctypes.data, host.nbytes)
    src = _device_surface(ctx, rng, abi.RGBA, w, h)
    sd = src.desc()
    o = abi.CompOp()
    o.op = abi.OP["src_over"]
    o.src.kind = abi.PARAM_SURFACE
    o.src.surface = C.pointer(sd)
    ops = (abi.CompOp * 1)(o)
    rc = lib().z2d_gpu_composite_noalloc(ctx.h, C.byref(hd), 0, 0, C.cast(ops, C.c_void_p), 1,
                                         abi.PRECISION_INTEGER, None)
    assert rc == abi.ERR_INVALID_ARGUMENT
    assert np.count_nonzero(host) == 0
    # a gradient with more stops than the scratch holds
    dst = _device_surface(ctx, rng, abi.RGBA, w, h)
    before = np.empty(w * h * 4, np.uint8)
    dst.download(before)
    dd = dst.desc()
    ops, keep = _ops(rng, w, h, sd, 60000)
    rc = lib().z2d_gpu_composite_noalloc(ctx.h, C.byref(dd), 0, 0, C.cast(ops, C.c_void_p), 3,
                                         abi.PRECISION_INTEGER, None)
    assert rc == abi.ERR_BUDGET_EXCEEDED
    assert "scratch" in ctx.last_error()
    after = np.empty_like(before)
    dst.download(after)
    assert np.array_equal(before, after)
    # and the scratch still serves the next call
    got, want, _ = _call_both(ctx, rng, abi.RGBA, 40, 9, 3)
    assert not diff_report(got, want)
    for s in (src, dst):
        s.close()
