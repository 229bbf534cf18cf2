"""The request server (continuous batching under an HBM budget), checked on
the GPU through the C ABI.  Each request is one surface: initialised from a
pixel (ImageSurface.initPixel, src/surface.zig), the fills applied in order
(painter.fill, src/painter.zig:66-143), then an optional alpha8 clip mask
applied as Surface.composite(mask, .dst_in) (src/surface.zig:225-241).  The
expected surface is the oracle's rendering of the same calls; a request
renders byte-identically alone and inside a mix of other requests; a budget
smaller than the whole load still completes every request."""
import ctypes as C

import numpy as np
import pytest

from z2d_amd import abi, scenes
from gpu_util import diff_report, gpu_backend
from oracle_backend import OracleBackend
from zscene import Surface

pytestmark = pytest.mark.gpu


def _clear_pixel(rng, fmt):
    if fmt in (abi.RGBA, abi.ARGB):
        a = rng.randint(0, 255)
        return abi.pixel(fmt, rng.randint(0, a), rng.randint(0, a), rng.randint(0, a), a)
    if fmt in (abi.RGB, abi.XRGB):
        return abi.pixel(fmt, rng.randint(0, 255), rng.randint(0, 255), rng.randint(0, 255))
    return abi.pixel(fmt, a=rng.randint(0, (1 << abi.FORMAT_BITS[fmt]) - 1))


def make_request(rng, fmt=None, side_max=300, with_clip=None):
    fmt = fmt if fmt is not None else rng.randint(0, 7)
    w = rng.randint(8, side_max)
    h = rng.randint(8, side_max)
    aa = [abi.AA_NONE, abi.AA_DEFAULT, abi.AA_MULTISAMPLE_4X][rng.randint(0, 2)]
    draws, keep = scenes.random_paths(rng, rng.randint(0, 12), w, h,
                                      max_side=min(200.0, max(w, h)), aa=aa)
    clip = None
    if with_clip if with_clip is not None else rng.uniform() < 0.4:
        clip = rng.u8_array(w * h)
    return dict(fmt=fmt, w=w, h=h, clear=_clear_pixel(rng, fmt), draws=draws, keep=keep,
                clip=clip)


def oracle_render(rq):
    s = Surface(rq["fmt"], rq["w"], rq["h"], init_px=rq["clear"])
    ob = OracleBackend()
    for pat, nodes, opts in rq["draws"]:
        assert ob.fill(s, pat, nodes, opts) == 0
    if rq["clip"] is not None:
        m = Surface(abi.ALPHA8, rq["w"], rq["h"])
        m.buf[: m.nbytes] = rq["clip"]
        md = m.desc()
        op = abi.CompOp()
        op.op = abi.OP["dst_in"]
        op.src.kind = abi.PARAM_SURFACE
        op.src.surface = C.pointer(md)
        ob.composite(s, 0, 0, [op], abi.PRECISION_INTEGER)
    return np.asarray(s.data(), np.uint8)[: s.nbytes].copy()


def submit(srv, rq):
    out = np.zeros(abi.surface_nbytes(rq["fmt"], rq["w"], rq["h"]), np.uint8)
    rc, rid = srv.submit(rq["fmt"], rq["w"], rq["h"], rq["draws"], clear=rq["clear"],
                         clip=rq["clip"], out=out)
    assert rc == 0, f"submit rc={rc}"
    return rid, out


def _server(budget, max_batch=64):
    from z2d_amd.gpu import Server
    return Server(gpu_backend().ctx, budget, max_batch)


def test_server_requests_match_oracle():
    rng = scenes.SplitMix64(0x5E5)
    reqs = [make_request(rng) for _ in range(48)]
    srv = _server(256 << 20)
    try:
        subs = [submit(srv, rq) for rq in reqs]
        for i, ((rid, out), rq) in enumerate(zip(subs, reqs)):
            assert srv.wait(rid) == 0, f"request {i}"
            msg = diff_report(out, oracle_render(rq))
            assert not msg, f"request {i} fmt={rq['fmt']} {rq['w']}x{rq['h']}: {msg}"
        st = srv.stats()
        assert st["completed"] == st["submitted"] == len(reqs)
        assert st["bytes_in_use"] == 0 and st["queued"] == 0
    finally:
        srv.close()


def test_server_same_alone_and_in_mix():
    rng = scenes.SplitMix64(0xA11)
    target = make_request(rng, fmt=abi.RGBA, side_max=400, with_clip=True)
    others = [make_request(rng) for _ in range(40)]
    srv = _server(512 << 20, max_batch=16)
    try:
        rid, alone = submit(srv, target)
        assert srv.wait(rid) == 0
        subs = []
        for i, rq in enumerate(others):
            subs.append(submit(srv, rq))
            if i == 17:
                mixed_id, mixed = submit(srv, target)
        assert srv.wait(mixed_id) == 0
        for rid2, _ in subs:
            assert srv.wait(rid2) == 0
        assert not diff_report(mixed, alone)
        assert not diff_report(alone, oracle_render(target))
        assert srv.stats()["batches"] >= 2
    finally:
        srv.close()


def test_server_submit_many_and_wait_some():
    """z2d_gpu_server_submit_n: requests handed over in one call render as
    submitted one by one (the oracle's surfaces); a request that can never
    fit is refused in place (id 0, BUDGET_EXCEEDED) without stopping the
    others.  z2d_gpu_server_wait_some returns every completion once, in
    completion order, stamped on CLOCK_MONOTONIC between submit and return."""
    import time
    from z2d_amd.gpu import Server
    rng = scenes.SplitMix64(0x5B7)
    reqs = [make_request(rng, side_max=200) for _ in range(30)]
    srv = _server(64 << 20)
    try:
        preps, outs = [], []
        for i, rq in enumerate(reqs):
            if i == 11:  # larger than the whole budget
                preps.append(Server.prepare(abi.RGBA, 5000, 5000, []))
                outs.append(None)
            preps.append(Server.prepare(rq["fmt"], rq["w"], rq["h"], rq["draws"],
                                        clear=rq["clear"], clip=rq["clip"]))
            outs.append(np.zeros(abi.surface_nbytes(rq["fmt"], rq["w"], rq["h"]), np.uint8))
        t_sub = time.monotonic()
        rc, ids, st = srv.submit_many(preps, outs)
        assert rc == abi.ERR_BUDGET_EXCEEDED
        assert ids[11] == 0 and st[11] == abi.ERR_BUDGET_EXCEEDED
        live = [i for i in range(len(preps)) if i != 11]
        assert all(st[i] == 0 and ids[i] != 0 for i in live)
        c_ids = np.zeros(len(live), np.uint64)
        c_st = np.zeros(len(live), np.int32)
        c_t = np.zeros(len(live), np.float64)
        got = 0
        while got < len(live):  # several calls of at most 7 completions
            k = srv.wait_some(c_ids[got:got + 7], c_st[got:got + 7], c_t[got:got + 7], 60000.0)
            assert k > 0, "server stalled"
            got += k
        t_end = time.monotonic()
        assert sorted(c_ids.tolist()) == sorted(int(ids[i]) for i in live)
        assert not c_st.any()
        assert np.all(np.diff(c_t) >= 0) and t_sub <= c_t[0] and c_t[-1] <= t_end
        for i in live:
            rq = reqs[i if i < 11 else i - 1]
            msg = diff_report(outs[i], oracle_render(rq))
            assert not msg, f"request {i}: {msg}"
        st = srv.stats()
        assert st["completed"] == st["submitted"] == len(live)
        assert st["rejected"] == 1
    finally:
        srv.close()


def test_server_budget_admission():
    """A 1 MiB budget admits a few requests at a time; the rest wait in the
    queue and still complete.  A request that can never fit is refused."""
    rng = scenes.SplitMix64(0xB0D)
    reqs = [make_request(rng, fmt=abi.RGBA, side_max=256, with_clip=False) for _ in range(24)]
    budget = 1 << 20
    srv = _server(budget)
    try:
        big = np.zeros(600 * 600 * 4, np.uint8)
        rc, _ = srv.submit(abi.RGBA, 600, 600, [], out=big)
        assert rc == abi.ERR_BUDGET_EXCEEDED
        subs = [submit(srv, rq) for rq in reqs]
        done = {}
        while len(done) < len(subs):
            got = srv.wait_any(60000.0)
            assert got is not None, "server stalled"
            done[got[0]] = got[1]
        assert all(v == 0 for v in done.values())
        for (rid, out), rq in zip(subs, reqs):
            assert not diff_report(out, oracle_render(rq))
        st = srv.stats()
        assert st["rejected"] == 1
        assert 0 < st["peak_bytes"] <= budget
        assert st["batches"] > 1
    finally:
        srv.close()


def test_server_error_status():
    """An open path fails its request with PathNotClosed (src/painter.zig:82);
    the other requests in the batch are unaffected."""
    rng = scenes.SplitMix64(0xE77)
    good = make_request(rng, fmt=abi.RGBA, with_clip=False)
    bad = make_request(rng, fmt=abi.RGBA, with_clip=False)
    open_nodes = abi.nodes_array([(abi.MOVE_TO, 1.0, 1.0), (abi.LINE_TO, 20.0, 3.0),
                                  (abi.LINE_TO, 9.0, 15.0)])
    bad["draws"] = [(abi.solid_pattern(abi.rgba(255, 0, 0)), open_nodes, abi.fill_opts())]
    srv = _server(64 << 20)
    try:
        rid_b, _ = submit(srv, bad)
        rid_g, out = submit(srv, good)
        assert srv.wait(rid_b) == abi.ERR_PATH_NOT_CLOSED
        assert srv.wait(rid_g) == 0
        assert not diff_report(out, oracle_render(good))
    finally:
        srv.close()


def test_server_bad_curves_fail_their_request_only():
    """A curve with an infinite control point fails its request at the
    host's node walk (INVALID_ARGUMENT); a finite curve whose subdivision
    exceeds the device flattener's visit limit (an arch 1e12 px across at
    tolerance 0.001: the leftmost branch of Spline.decompose,
    src/internal/tess/Spline.zig, is 25 halvings deep, so each of the
    flattener's 16 pieces needs more than 2^20 visits) fails its request
    with UNSUPPORTED.  The other requests of the same batch render exactly
    as the oracle does."""
    rng = scenes.SplitMix64(0xBADC)
    goods = [make_request(rng, fmt=abi.RGBA, with_clip=False) for _ in range(3)]
    red = abi.solid_pattern(abi.rgba(200, 0, 0))

    def curve_req(s):
        rq = make_request(rng, fmt=abi.RGBA, with_clip=False)
        nodes = abi.nodes_array([(abi.MOVE_TO, 0.0, 0.0),
                                 (abi.CURVE_TO, 0.0, s, s, s, s, 0.0),
                                 (abi.LINE_TO, 4.0, 25.0), (abi.CLOSE_PATH,)])
        rq["draws"] = rq["draws"] + [(red, nodes, abi.fill_opts(tolerance=0.001))]
        return rq

    inf_req, huge_req = curve_req(float("inf")), curve_req(1e12)
    srv = _server(64 << 20)
    try:
        order = [goods[0], inf_req, goods[1], huge_req, goods[2]]
        subs = [submit(srv, rq) for rq in order]
        st = [srv.wait(rid, timeout_ms=120000.0) for rid, _ in subs]
        assert st[1] == abi.ERR_INVALID_ARGUMENT
        assert st[3] == abi.ERR_UNSUPPORTED
        for k in (0, 2, 4):
            assert st[k] == 0
            assert not diff_report(subs[k][1], oracle_render(order[k]))
    finally:
        srv.close()


def test_server_fills_and_strokes():
    """Requests of the batched form's distribution (30% of the paths
    stroked, dashes included) against the CPU path: the host tessellator's
    polygons and the fills rasterized by the oracle, in draw order, then the
    clip mask; also alone vs in the mix."""
    from stroke_check import StrokeCheckBackend
    rng = scenes.SplitMix64(0x57C)
    reqs = []
    for i in range(24):
        fmt = [abi.RGBA, abi.ALPHA8, abi.RGB, abi.ALPHA4][rng.randint(0, 3)]
        w, h = rng.randint(8, 260), rng.randint(8, 260)
        jobs, keep = scenes.random_jobs(rng, rng.randint(0, 14), w, h,
                                        max_side=min(200.0, max(w, h)))
        clip = rng.u8_array(w * h) if rng.uniform() < 0.4 else None
        reqs.append(dict(fmt=fmt, w=w, h=h, clear=_clear_pixel(rng, fmt), jobs=jobs, keep=keep,
                         clip=clip))

    def cpu(rq):
        s = Surface(rq["fmt"], rq["w"], rq["h"], init_px=rq["clear"])
        be = StrokeCheckBackend()
        for kind, pat, nodes, opts in rq["jobs"]:
            rc = (be.stroke if kind == "stroke" else be.fill)(s, pat, nodes, opts)
            assert rc == 0, rc
        if rq["clip"] is not None:
            m = Surface(abi.ALPHA8, rq["w"], rq["h"])
            m.buf[: m.nbytes] = rq["clip"]
            md = m.desc()
            op = abi.CompOp()
            op.op = abi.OP["dst_in"]
            op.src.kind = abi.PARAM_SURFACE
            op.src.surface = C.pointer(md)
            be.composite(s, 0, 0, [op], abi.PRECISION_INTEGER)
        return np.asarray(s.data(), np.uint8)[: s.nbytes].copy()

    srv = _server(256 << 20, max_batch=8)
    try:
        subs = []
        for rq in reqs:
            out = np.zeros(abi.surface_nbytes(rq["fmt"], rq["w"], rq["h"]), np.uint8)
            rc, rid = srv.submit(rq["fmt"], rq["w"], rq["h"], [], clear=rq["clear"],
                                 clip=rq["clip"], out=out, draws=rq["jobs"])
            assert rc == 0, rc
            subs.append((rid, out))
        for i, ((rid, out), rq) in enumerate(zip(subs, reqs)):
            assert srv.wait(rid) == 0, f"request {i}"
            msg = diff_report(out, cpu(rq))
            assert not msg, f"request {i} fmt={rq['fmt']} {rq['w']}x{rq['h']}: {msg}"
        # the last request alone
        rq = reqs[-1]
        alone = np.zeros_like(subs[-1][1])
        rc, rid = srv.submit(rq["fmt"], rq["w"], rq["h"], [], clear=rq["clear"],
                             clip=rq["clip"], out=alone, draws=rq["jobs"])
        assert rc == 0 and srv.wait(rid) == 0
        assert not diff_report(alone, subs[-1][1])
    finally:
        srv.close()


def test_server_png_responses():
    """Requests that ask for a PNG file (encode = 1 + profile): the file
    decodes to the oracle's scanline stream of the same request's raw result
    (byte for byte; the sRGB profile within 1 LSB), raw and encoded requests
    mixed in one server; a buffer smaller than the file completes with
    INVALID_ARGUMENT and reports the size needed."""
    from oracle_backend import png_scanlines
    from pngfix import png_chunks, png_stream
    rng = scenes.SplitMix64(0x96)
    reqs = []
    for i in range(18):
        fmt = [abi.RGBA, abi.RGB, abi.ALPHA8, abi.ALPHA1, abi.ARGB][rng.randint(0, 4)]
        w, h = rng.randint(8, 300), rng.randint(8, 200)
        jobs, keep = scenes.random_jobs(rng, rng.randint(0, 10), w, h,
                                        max_side=min(180.0, max(w, h)))
        prof = [abi.PNG_PROFILE_NONE, abi.PNG_PROFILE_LINEAR, abi.PNG_PROFILE_SRGB][i % 3]
        reqs.append(dict(fmt=fmt, w=w, h=h, clear=_clear_pixel(rng, fmt), jobs=jobs, keep=keep,
                         prof=prof))
    srv = _server(256 << 20, max_batch=6)
    try:
        subs = []
        for rq in reqs:
            raw = np.zeros(abi.surface_nbytes(rq["fmt"], rq["w"], rq["h"]), np.uint8)
            png = np.zeros(gpu_png_bound(rq), np.uint8)
            r1 = srv.submit(rq["fmt"], rq["w"], rq["h"], [], clear=rq["clear"], out=raw,
                            draws=rq["jobs"])
            r2 = srv.submit(rq["fmt"], rq["w"], rq["h"], [], clear=rq["clear"], out=png,
                            draws=rq["jobs"], encode=1 + rq["prof"])
            assert r1[0] == 0 and r2[0] == 0
            subs.append((r1[1], raw, r2[1], png))
        for i, ((rid1, raw, rid2, png), rq) in enumerate(zip(subs, reqs)):
            assert srv.wait(rid1) == 0 and srv.wait(rid2) == 0, f"request {i}"
            n = srv.written(rid2)
            data = png[:n].tobytes()
            assert all(ok for _, _, ok in png_chunks(data))
            want = np.frombuffer(png_scanlines(rq["fmt"], rq["w"], rq["h"], raw.tobytes(),
                                               rq["prof"]), np.uint8).astype(np.int32)
            got = np.frombuffer(png_stream(data), np.uint8).astype(np.int32)
            tol = 1 if rq["prof"] == abi.PNG_PROFILE_SRGB and rq["fmt"] <= abi.RGBA else 0
            assert got.shape == want.shape and np.abs(got - want).max(initial=0) <= tol, i
        # a buffer too small for the file
        rq = reqs[0]
        small = np.zeros(16, np.uint8)
        rc, rid = srv.submit(rq["fmt"], rq["w"], rq["h"], [], clear=rq["clear"], out=small,
                             draws=rq["jobs"], encode=1)
        assert rc == 0
        assert srv.wait(rid) == -1  # INVALID_ARGUMENT
        assert srv.written(rid) > 16
    finally:
        srv.close()


def gpu_png_bound(rq):
    from z2d_amd import gpu
    return gpu.png_bound(rq["fmt"], rq["w"], rq["h"], rq["prof"])


def test_server_resident_gradients_across_requests():
    """Gradients stay resident in HBM across the server's batches: requests
    that repeat a gradient (the same descriptor again, and an equal copy of
    it built separately) render as the oracle does, in batches after the one
    that first uploaded it and alongside gradients seen once."""
    rng = scenes.SplitMix64(0x6AD)
    base = make_request(rng, fmt=abi.RGBA, side_max=200, with_clip=False)
    stops = [(0.0, abi.color(abi.COLOR_LINEAR_RGB, 0.9, 0.1, 0.2, 1.0)),
             (0.4, abi.color(abi.COLOR_SRGB, 0.1, 0.8, 0.3, 0.7)),
             (1.0, abi.color(abi.COLOR_LINEAR_RGB, 0.2, 0.3, 0.9, 0.9))]

    def grad(kind):
        geom = ([5.0, 7.0, 150.0, 90.0] if kind == abi.GRADIENT_LINEAR
                else [60.0, 50.0, 4.0, 70.0, 60.0, 80.0])
        return abi.Gradient(kind, geom, stops, abi.INTERP_LINEAR_RGB, abi.POLAR_SHORTER,
                            abi.IDENTITY)

    reqs = []
    for i in range(6):
        keep = list(base["keep"])
        draws = []
        for j, (pat, nodes, opts) in enumerate(base["draws"]):
            # an equal gradient built again for every request
            g = grad(abi.GRADIENT_LINEAR if j % 2 == 0 else abi.GRADIENT_RADIAL)
            keep.append(g)
            draws.append((g.pattern(), nodes, opts))
        reqs.append(dict(base, draws=draws, keep=keep))
    srv = _server(64 << 20)
    try:
        for rq in reqs:  # one at a time: each in its own batch
            rid, out = submit(srv, rq)
            assert srv.wait(rid) == 0
            msg = diff_report(out, oracle_render(rq))
            assert not msg, msg
        subs = [submit(srv, rq) for rq in reqs + [make_request(rng) for _ in range(8)]]
        for (rid, out), rq in zip(subs, reqs + [None] * 8):
            assert srv.wait(rid) == 0
            if rq is not None:
                msg = diff_report(out, oracle_render(rq))
                assert not msg, msg
    finally:
        srv.close()
