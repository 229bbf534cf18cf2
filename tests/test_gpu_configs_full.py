"""The benchmark configurations at their full sizes on the GPU, every output
checked against the oracle:

  compositor sweep  4096x4096 RGBA8 (seed 0xC2), the 28 operators in enum
                    order out of place, then the alpha8-mask variant
                    {dst_in dst=src src=mask}, {op} (SurfaceCompositor.run,
                    src/compositor.zig:302-440), every pass bit-exact
  4K scene          3840x2160, 10,000 paths (seed 0xC3) in one batch,
                    bit-exact against the oracle's painter.fill sequence
                    (src/painter.zig:66-143)
  batched form      4096 uneven surfaces (seed 0xC4), 30% strokes, one
                    persistent launch, every surface bit-exact against the
                    host tessellator + oracle raster
  served form       requests of the batched form's distribution with clip
                    masks (seed 0xC5) submitted at once (saturation), every
                    result bit-exact against the same CPU path

The CPU side runs on a thread pool (the oracle's entry points release the
interpreter lock; compositing is per pixel, so row bands are independent)."""
import ctypes as C
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from z2d_amd import abi, scenes
from gpu_util import diff_report, gpu_backend
from oracle_backend import OracleBackend, lib as oracle_lib
from stroke_check import StrokeCheckBackend
from zscene import Surface

pytestmark = pytest.mark.gpu

WORKERS = max(1, min(16, os.cpu_count() or 1))


def _log(msg):
    print(f"[configs-full] {msg}", flush=True)


def _ctx():
    return gpu_backend().ctx


# ---------------------------------------------------------------------------
# compositor sweep
# ---------------------------------------------------------------------------
def _comp_ops(name, src_desc, mask_desc, masked):
    if masked:
        m = abi.CompOp()
        m.op = abi.OP["dst_in"]
        m.dst.kind = abi.PARAM_SURFACE
        m.dst.surface = C.pointer(src_desc)
        m.src.kind = abi.PARAM_SURFACE
        m.src.surface = C.pointer(mask_desc)
        o = abi.CompOp()
        o.op = abi.OP[name]
        return (abi.CompOp * 2)(m, o)
    o = abi.CompOp()
    o.op = abi.OP[name]
    o.src.kind = abi.PARAM_SURFACE
    o.src.surface = C.pointer(src_desc)
    return (abi.CompOp * 1)(o)


def _oracle_sweep_pass(name, masked, src, dst, mask, w, h, pool, band=128):
    """out = op(src, dst) on the oracle, in row bands on the pool."""
    out = dst.copy()

    def run(y0):
        rows = min(band, h - y0)
        o = out[y0 * w * 4:(y0 + rows) * w * 4]
        s = src[y0 * w * 4:(y0 + rows) * w * 4]
        m = mask[y0 * w:(y0 + rows) * w]
        od = abi.SurfaceDesc(abi.RGBA, w, rows, 0, o.ctypes.data, o.nbytes)
        sd = abi.SurfaceDesc(abi.RGBA, w, rows, 0, s.ctypes.data, s.nbytes)
        md = abi.SurfaceDesc(abi.ALPHA8, w, rows, 0, m.ctypes.data, m.nbytes)
        arr = _comp_ops(name, sd, md, masked)
        oracle_lib().zo_composite(C.byref(od), 0, 0, arr, len(arr), abi.PRECISION_INTEGER)

    list(pool.map(run, range(0, h, band)))
    return out


@pytest.mark.timeout(900)
def test_compositor_sweep_4096_all_operators_bit_exact():
    from z2d_amd.gpu import DeviceSurface, lib
    w = h = 4096
    _log("compositor sweep: generating 4096x4096 inputs (seed 0xC2)")
    src, dst, mask = scenes.compositor_inputs(w, h)
    ctx = _ctx()
    s_src = DeviceSurface(ctx, abi.RGBA, w, h)
    s_dst = DeviceSurface(ctx, abi.RGBA, w, h)
    s_out = DeviceSurface(ctx, abi.RGBA, w, h)
    s_mask = DeviceSurface(ctx, abi.ALPHA8, w, h)
    s_src.upload(src)
    s_dst.upload(dst)
    s_mask.upload(mask)
    d_src, d_dst, d_out, d_mask = s_src.desc(), s_dst.desc(), s_out.desc(), s_mask.desc()
    got = np.empty(w * h * 4, np.uint8)
    failures = []
    with ThreadPoolExecutor(WORKERS) as pool:
        for masked in (False, True):
            for name in abi.OPERATORS:
                arr = _comp_ops(name, d_src, d_mask, masked)
                ctx.check(lib().z2d_gpu_composite(ctx.h, C.byref(d_dst), 0, 0, arr, len(arr),
                                                  abi.PRECISION_INTEGER, C.byref(d_out)),
                          "composite")
                s_out.download(got)
                want = _oracle_sweep_pass(name, masked, src, dst, mask, w, h, pool)
                if not np.array_equal(got, want):
                    failures.append(f"{name}{' masked' if masked else ''}: "
                                    f"{diff_report(got, want)}")
                _log(f"{name}{' +mask' if masked else ''}: "
                     f"{'ok' if not failures or not failures[-1].startswith(name) else 'DIFF'}")
        # dst itself is never written (out of place)
        back = np.empty(w * h * 4, np.uint8)
        s_dst.download(back)
        assert np.array_equal(back, dst)
    for s in (s_src, s_dst, s_out, s_mask):
        s.close()
    assert not failures, "\n".join(failures)


# ---------------------------------------------------------------------------
# 4K scene
# ---------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_scene_4k_full_bit_exact():
    from z2d_amd.gpu import Batch, DeviceSurface
    _log("4K scene: generating 3840x2160, 10,000 paths (seed 0xC3)")
    sc = scenes.scene_4k()
    w, h = sc["width"], sc["height"]
    ctx = _ctx()
    surf = DeviceSurface(ctx, sc["format"], w, h)
    surf.upload(sc["buf"])
    b = Batch(ctx)
    for pat, nodes, opts in sc["draws"]:
        assert b.fill(surf, pat, nodes, opts) == 0
    b.prepare()
    b.counters(True)
    b.render(0)
    b.wait()
    cnt = b.counters(False)
    assert cnt[3] == 0, "bin overflow"
    got = np.empty(surf.nbytes, np.uint8)
    surf.download(got)
    b.close()
    surf.close()
    _log(f"4K scene: GPU done ({cnt[0]} pixel-draws), oracle rendering 10,000 paths")
    want = Surface(sc["format"], w, h)
    want.buf[: want.nbytes] = sc["buf"][: want.nbytes]
    ob = OracleBackend()
    for pat, nodes, opts in sc["draws"]:
        assert ob.fill(want, pat, nodes, opts) == 0
    msg = diff_report(got, want.data())
    assert not msg, msg
    # the render painted: white remains only where no path reached
    assert np.count_nonzero(got != 255) > w * h


# ---------------------------------------------------------------------------
# batched form
# ---------------------------------------------------------------------------
def cpu_surface(fmt, w, h, jobs, clear=None, clip=None):
    """The CPU path of one surface job: painter.fill on the oracle, strokes
    through the host tessellator and the oracle's raster, then the clip as
    Surface.composite(mask, .dst_in)."""
    s = Surface(fmt, w, h, init_px=clear)
    be = StrokeCheckBackend()
    for kind, pat, nodes, opts in jobs:
        rc = (be.stroke if kind == "stroke" else be.fill)(s, pat, nodes, opts)
        assert rc == 0, f"{kind} rc={rc}"
    if clip is not None:
        m = Surface(abi.ALPHA8, w, h)
        m.buf[: m.nbytes] = clip
        md = m.desc()
        op = abi.CompOp()
        op.op = abi.OP["dst_in"]
        op.src.kind = abi.PARAM_SURFACE
        op.src.surface = C.pointer(md)
        be.composite(s, 0, 0, [op], abi.PRECISION_INTEGER)
    return np.asarray(s.data(), np.uint8)[: s.nbytes].copy()


@pytest.mark.timeout(900)
def test_uneven_batch_4096_surfaces_bit_exact():
    from z2d_amd.gpu import Batch, DeviceSurface
    _log("batched form: generating 4096 surfaces (seed 0xC4)")
    work = scenes.uneven_surfaces(4096, 0xC4)
    ctx = _ctx()
    b = Batch(ctx)
    surfs = []
    for fmt, side, jobs, keep in work:
        s = DeviceSurface(ctx, fmt, side, side)
        surfs.append(s)
        for kind, pat, nodes, opts in jobs:
            rc = (b.stroke if kind == "stroke" else b.fill)(s, pat, nodes, opts)
            assert rc == 0, f"batch {kind} rc={rc}"
    b.prepare()
    b.counters(True)
    b.render(0)
    b.wait()
    cnt = b.counters(False)
    assert cnt[3] == 0, "bin overflow"
    _log(f"batched form: GPU done ({cnt[0]} pixel-draws), checking every surface")

    def check(i):
        fmt, side, jobs, keep = work[i]
        got = np.empty(surfs[i].nbytes, np.uint8)
        surfs[i].download(got)
        want = cpu_surface(fmt, side, side, jobs)
        msg = diff_report(got, want)
        return f"surface {i} ({abi.FORMAT_NAMES[fmt]} {side}px): {msg}" if msg else ""

    with ThreadPoolExecutor(WORKERS) as pool:
        bad = [m for m in pool.map(check, range(len(work))) if m]
    b.close()
    for s in surfs:
        s.close()
    assert not bad, "\n".join(bad[:10])


# ---------------------------------------------------------------------------
# served form
# ---------------------------------------------------------------------------
@pytest.mark.timeout(900)
def test_served_mix_saturation_bit_exact():
    from z2d_amd.gpu import Server
    n = 512
    _log(f"served form: generating {n} requests (seed 0xC5), submitting all at once")
    work = scenes.served_requests(n, 0xC5)
    srv = Server(_ctx(), 16 << 30, 64)
    subs = []
    for fmt, side, jobs, keep, clip in work:
        out = np.zeros(abi.surface_nbytes(fmt, side, side), np.uint8)
        rc, rid = srv.submit(fmt, side, side, [], clear=abi.pixel(fmt), clip=clip, out=out,
                             draws=jobs)
        assert rc == 0, f"submit rc={rc}"
        subs.append((rid, out))
    for rid, _ in subs:
        assert srv.wait(rid) == 0
    st = srv.stats()
    srv.close()
    _log(f"served form: {st['completed']} completed in {st['batches']} batches; checking")

    def check(i):
        fmt, side, jobs, keep, clip = work[i]
        want = cpu_surface(fmt, side, side, jobs, clear=abi.pixel(fmt), clip=clip)
        msg = diff_report(subs[i][1], want)
        return f"request {i} ({abi.FORMAT_NAMES[fmt]} {side}px): {msg}" if msg else ""

    with ThreadPoolExecutor(WORKERS) as pool:
        bad = [m for m in pool.map(check, range(n)) if m]
    assert st["batches"] > 1  # the load was split into several batches
    assert not bad, "\n".join(bad[:10])
