"""Benchmark of the fill + composite hot path on one MI355X.

Headline workload (the configuration the metric is quoted on): one
3840x2160 RGBA8 scene of 10,000 random closed paths (70% lines / 30% cubic
Beziers, 3-64 segments, log-uniform 4-512 px boxes, 60% solid / 40%
gradient paints, 50/50 non-zero / even-odd, src_over, 4x multisample AA),
inputs resident in HBM.  One step = one pass of the hot path over the whole
scene: device flattening of the raw path nodes (Bezier subdivision), edge
build, binning, and the tile rasterize + composite of all 10,000 paths in
order.

  value          pixel-draws per second (Mpix/s): pixels with nonzero
                 coverage summed over all paths, counted on the device
  roofline       algorithmic bytes of the render stage (k_render_fast +
                 k_render: 2*bpp per pixel-draw + 32 B per band edge copy +
                 paint parameters) / its mean duration, against MI355X HBM3E
                 peak (8 TB/s)
  cpu_baseline   the C oracle (a scalar restatement of the reference
                 painter, not the Zig painter) rendering the first 4,000
                 paths of the same scene on one host core; all-core and
                 per-workload legs beside it (star latency, compositor,
                 batched surfaces)
  extra          512x512 star latency (p50/p99), 4096x4096 compositor
                 passes (GB/s), batched uneven surfaces vs one launch per
                 surface, the request server under open-loop Poisson
                 arrivals (Mpix/s, p50/p99 latency per QPS level), and the
                 PNG export of the rendered scene, and a page of text
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from z2d_amd import abi, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
METRIC = "fill+composite Mpix/s (% HBM peak); p50/p99 latency of batched render reqs"


KERNEL_SOURCES = ("z2d_amd/csrc/render.hip", "z2d_amd/csrc/render.h", "z2d_amd/csrc/dev_pixel.h",
                  "z2d_amd/csrc/composite_px.h")


def kernel_sources_hash():
    """sha256 over the render kernels' sources (the files the render stage
    is compiled from), to tie a stored profile to the code it measured."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def paint_bytes(pattern):
    if pattern.kind == abi.PATTERN_GRADIENT:
        return 96 + 20 * pattern.gradient.contents.n_stops
    return 8


def build_batch(gpu, ctx, sc, n=None):
    surf = gpu.DeviceSurface(ctx, sc["format"], sc["width"], sc["height"])
    surf.upload(sc["buf"])
    b = gpu.Batch(ctx)
    draws = sc["draws"] if n is None else sc["draws"][:n]
    for pat, nodes, opts in draws:
        rc = b.fill(surf, pat, nodes, opts)
        if rc != 0:
            raise RuntimeError(f"batch_fill rc={rc}")
    b.prepare()
    return surf, b


def count_work(b):
    b.counters(True)
    b.render(0)
    b.wait()
    c = b.counters(False)
    pix, edges, overflow = c[0], c[1], c[3]
    if overflow:
        raise RuntimeError("bin overflow while counting")
    b.last_counts = c
    return pix, edges


def bench_scene(gpu, ctx, sc, steps, warmup):
    surf, b = build_batch(gpu, ctx, sc)
    for _ in range(warmup):
        b.render(0)
        b.wait()
    pix, edges = count_work(b)
    ctx.sync()
    render_ms, pipe_ms = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.render(0)
        b.wait()
        p, r = b.timing()
        pipe_ms.append(p)
        render_ms.append(r)
    ctx.sync()
    wall = time.perf_counter() - t0
    pbytes = sum(paint_bytes(d[0]) for d in sc["draws"])
    alg_bytes = 2 * 4 * pix + 32 * edges + pbytes
    res = dict(pix=pix, edges=edges, wall=wall, render_ms=float(np.mean(render_ms)),
               pipe_ms=float(np.mean(pipe_ms)), alg_bytes=alg_bytes,
               strip_entries=int(b.last_counts[2]))
    b.close()
    surf.close()
    return res


def bench_star_latency(gpu, ctx, iters=300):
    sc = scenes.star_512()
    surf, b = build_batch(gpu, ctx, sc)
    for _ in range(10):
        b.render(0)
        b.wait()
    lat = []
    for _ in range(iters):
        t = time.perf_counter()
        b.render(0)
        b.wait()
        lat.append((time.perf_counter() - t) * 1e3)
    b.close()
    surf.close()
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


# Bytes per pixel an operator's pass has to move (out-of-place, RGBA8):
# clear writes out only; src reads src, dst reads dst; every other operator
# reads both.  The mask variant (dst_in of src by the alpha8 mask, then the
# operator) adds the mask byte wherever the operator reads src.
def _op_bytes(name, masked):
    if name == "clear":
        return 4
    if name == "dst":
        return 8
    src_side = 4 + (1 if masked else 0)
    return 4 + src_side + (0 if name == "src" else 4)


def _stream_copy(gpu, ctx, nbytes=1 << 30, iters=20, read2=False):
    """The library's reference streams (read + write GB/s; passes rotating
    over three buffer sets of nbytes each): a plain copy, or (read2) two
    reads and one write, the compositor's mix: the achievable-HBM
    references."""
    g = C.c_double()
    f = gpu.lib().z2d_gpu_stream_read2 if read2 else gpu.lib().z2d_gpu_stream_copy
    ctx.check(f(ctx.h, nbytes, iters, C.byref(g)), "stream")
    return round(g.value, 1)


def bench_compositor(gpu, ctx, w=4096, h=4096, reps=3, sets=3):
    """Compositor sweep at 4096x4096 RGBA8 (seed 0xC2): every operator in
    enum order out-of-place, out = op(src, dst), then the alpha8-mask variant
    {dst_in dst=src src=mask}, {op}.  Each operator's bytes are what it has
    to move (_op_bytes: 4 B/px for clear, 8 for src and dst, 12 otherwise,
    +1 for the mask where src is read).  The passes rotate over `sets`
    independent src/dst/out/mask sets (3 x 218 MB), so consecutive passes
    do not find their inputs in the 256 MB last-level cache; the stream
    copy rate at the same footprint is reported beside them."""
    src, dst, mask = scenes.compositor_inputs(w, h)
    surfs = []
    for _ in range(sets):
        s_src = gpu.DeviceSurface(ctx, abi.RGBA, w, h)
        s_dst = gpu.DeviceSurface(ctx, abi.RGBA, w, h)
        s_out = gpu.DeviceSurface(ctx, abi.RGBA, w, h)
        s_mask = gpu.DeviceSurface(ctx, abi.ALPHA8, w, h)
        s_src.upload(src)
        s_dst.upload(dst)
        s_mask.upload(mask)
        surfs.append((s_src, s_dst, s_out, s_mask))
    descs = [tuple(x.desc() for x in st) for st in surfs]
    f = gpu.lib().z2d_gpu_composite
    per_op, per_op_mask, dev_op, dev_mask = {}, {}, {}, {}
    tot_bytes = tot_time = tot_kern = 0.0
    timing = gpu.lib().z2d_gpu_composite_timing
    km = C.c_double(0.0)

    def op_list(name, masked, d_src, d_mask):
        if masked:
            m = abi.CompOp()
            m.op = abi.OP["dst_in"]
            m.dst.kind = abi.PARAM_SURFACE
            m.dst.surface = C.pointer(d_src)
            m.src.kind = abi.PARAM_SURFACE
            m.src.surface = C.pointer(d_mask)
            o = abi.CompOp()
            o.op = abi.OP[name]
            return (abi.CompOp * 2)(m, o)
        o = abi.CompOp()
        o.op = abi.OP[name]
        o.src.kind = abi.PARAM_SURFACE
        o.src.surface = C.pointer(d_src)
        return (abi.CompOp * 1)(o)

    for masked in (False, True):
        for name in abi.OPERATORS:
            lists = [op_list(name, masked, ds[0], ds[3]) for ds in descs]
            ds0 = descs[-1]
            arr = lists[-1]
            ctx.check(f(ctx.h, C.byref(ds0[1]), 0, 0, arr, len(arr), 0, C.byref(ds0[2])),
                      "composite")
            t = time.perf_counter()
            kern = 0.0
            n_pass = reps * sets
            for i in range(n_pass):
                ds, arr = descs[i % sets], lists[i % sets]
                ctx.check(f(ctx.h, C.byref(ds[1]), 0, 0, arr, len(arr), 0, C.byref(ds[2])),
                          "composite")
                ctx.check(timing(ctx.h, C.byref(km)), "composite timing")
                kern += km.value / 1e3
            dt = (time.perf_counter() - t) / n_pass
            kern /= n_pass
            nbytes = w * h * _op_bytes(name, masked)
            (per_op_mask if masked else per_op)[name] = round(nbytes / dt / 1e9, 1)
            (dev_mask if masked else dev_op)[name] = round(nbytes / kern / 1e9, 1)
            tot_bytes += nbytes
            tot_time += dt
            tot_kern += kern
    for st in surfs:
        for x in st:
            x.close()
    return {"GBps_all_passes": round(tot_bytes / tot_time / 1e9, 1), "per_op_GBps": per_op,
            "mask_per_op_GBps": per_op_mask, "passes": 2 * len(abi.OPERATORS),
            "note": "host-timed per call (upload of the op list, launch, sync)",
            "kernel_GBps_all_passes": round(tot_bytes / tot_kern / 1e9, 1),
            "kernel_per_op_GBps": dev_op, "kernel_mask_per_op_GBps": dev_mask,
            "rotating_sets": sets,
            "stream_copy_GBps": _stream_copy(gpu, ctx, nbytes=w * h * 4),
            "stream_read2_GBps": _stream_copy(gpu, ctx, nbytes=w * h * 4, read2=True),
            "kernel_note": "k_composite_quad device time (HIP events around the launch); "
                           "bytes each operator moves (clear 4 B/px, src/dst 8, others 12, "
                           "+1 where the mask is read) over the kernel time, passes "
                           "rotating over 3 input sets (654 MB) so no pass reads its "
                           "inputs from the last-level cache; stream_copy_GBps (1 read : "
                           "1 write) and stream_read2_GBps (2 reads : 1 write, the "
                           "operators' mix) are the library's reference streams over "
                           "buffers of the surfaces' size, rotating over 3 sets the "
                           "same way"}


def bench_uneven_batch(gpu, ctx, n_surfaces=4096, seed=0xC4):
    """Batched form: 4096 surfaces (sides log-uniform 16-1024 px, 80% RGBA8 /
    20% alpha8, cleared to 0, 1-50 paths each of the 4K scene's
    distribution; 30% of the paths stroked: width U(0.5, 8), joins and caps
    uniform, 25% dashed).  Strokes are tessellated on the host when the batch
    is built.  One persistent launch vs one launch per surface over the same
    prepared batch; the persistent launch also as a replayed hipGraph of the
    whole pipeline.  The band height is picked by the autotuner first.  The
    host build is timed both as one call per draw and through draw_n
    (strokes tessellated on host threads)."""
    surfs, items = [], []
    for fmt, side, jobs, keep in scenes.uneven_surfaces(n_surfaces, seed):
        s = gpu.DeviceSurface(ctx, fmt, side, side)
        surfs.append(s)
        surfs.append(keep)
        for kind, pat, nodes, opts in jobs:
            items.append((s, kind, pat, nodes, opts))
    # host build of the batch: one call per draw (strokes tessellated on the
    # caller's thread), then the same draws through draw_n (strokes
    # tessellated on host threads); the second batch is the one rendered
    b0 = gpu.Batch(ctx)
    t = time.perf_counter()
    for s, kind, pat, nodes, opts in items:
        rc = (b0.stroke if kind == "stroke" else b0.fill)(s, pat, nodes, opts)
        if rc:
            raise RuntimeError(f"batch {kind} rc={rc}")
    build_calls = time.perf_counter() - t
    b0.close()
    b = gpu.Batch(ctx)
    t = time.perf_counter()
    st = b.draw_n(items)
    build_n = time.perf_counter() - t
    if any(st):
        raise RuntimeError(f"batch draw_n statuses {set(st)}")
    b.prepare()
    pix, _ = count_work(b)
    res = {}
    band_h, longest, tune_ms = b.autotune(0, 2)
    res["autotuned"] = {"band_rows": band_h, "queue": "longest_first" if longest else "natural"}
    res["autotune_pipeline_ms"] = {f"{k[0]}{'_lpt' if k[1] else ''}": round(v, 3)
                                   for k, v in tune_ms.items()}
    for mode, name, graph in ((0, "batched", False), (0, "batched_graph", True),
                              (1, "per_surface", False)):
        b.use_graph(graph)
        for _ in range(2):
            b.render(mode)
            b.wait()
        t = time.perf_counter()
        for _ in range(3):
            b.render(mode)
            b.wait()
        dt = (time.perf_counter() - t) / 3
        res[name] = round(pix / dt / 1e6, 1)
    b.use_graph(False)
    res["speedup"] = round(res["batched"] / res["per_surface"], 2)
    res["speedup_graph"] = round(res["batched_graph"] / res["per_surface"], 2)
    res["host_build_ms"] = {"per_call": round(build_calls * 1e3, 1),
                            "draw_n": round(build_n * 1e3, 1),
                            "threads": min(16, os.cpu_count() or 1)}
    res["surfaces"] = n_surfaces
    res["pixel_draws"] = int(pix)
    res["stroke_share"] = 0.3
    b.close()
    return res


def serving_requests(gpu, n, seed=0xC5):
    """Request mix of the serving form: one surface job of the batched
    form's distribution each (side log-uniform 16-1024 px, 80% RGBA8 / 20%
    alpha8 cleared to 0, 1-50 paths of the 4K scene's distribution, 30% of
    them stroked), half with an alpha8 clip mask."""
    reqs = []
    for fmt, side, jobs, keep, clip in scenes.served_requests(n, seed):
        prep = gpu.Server.prepare(fmt, side, side, [], clear=abi.pixel(fmt), clip=clip,
                                  draws=jobs)
        reqs.append(dict(prep=prep, keep=keep,
                         out=np.empty(abi.surface_nbytes(fmt, side, side), np.uint8)))
    return reqs


def bench_serving(gpu, ctx, reqs, qps_levels=(250, 1000, 4000, 16000, 32000, 64000),
                  seconds=2.0, seed=0x9015, budget=16 << 30, n_sub_max=2):
    """Open-loop Poisson arrivals at each QPS level, cycling through the
    request pool, and one saturation run with 24 times the pool queued at
    t=0.  Client threads (one up to 16k QPS, n_sub_max above; arrival i by
    thread i mod their number) hand each arrival to the server at its
    scheduled time; the arrivals already due when a client wakes go over in
    one z2d_gpu_server_submit_n call, as a front end drains its socket, so
    interpreter time is spent per wake-up rather than per request.  Latency
    = the server's completion stamp (result in the caller's buffer,
    CLOCK_MONOTONIC) - scheduled arrival; throughput = pixel-draws counted
    on the device / (last completion - start).  The pool's descriptors and
    output buffers are shared by the cycled requests (results are not
    checked here; the GPU tests check the server's output).  The Python
    garbage collector runs between levels, not during one."""
    import gc
    import threading
    switch = sys.getswitchinterval()
    sys.setswitchinterval(1e-4)
    srv = gpu.Server(ctx, budget, 64)
    P = len(reqs)
    written = np.zeros(P, np.uint64)  # out_written of each pool descriptor
    for k, rq in enumerate(reqs):
        rq["prep"][0].out_written = C.cast(written.ctypes.data + 8 * k, C.POINTER(C.c_uint64))
    req_ptr = np.array([C.addressof(rq["prep"][0]) for rq in reqs], np.uint64)
    out_ptr = np.array([rq["out"].ctypes.data for rq in reqs], np.uint64)
    out_len = np.array([rq["out"].nbytes for rq in reqs], np.uint64)

    def run_level(sched, n_sub):
        """Submits arrival i at t0 + sched[i]; returns (t0, ids, done ids,
        statuses, completion times)."""
        n = len(sched)
        ids = np.zeros(n, np.uint64)
        sub_st = np.zeros(n, np.int32)
        c_ids = np.zeros(n, np.uint64)
        c_st = np.zeros(n, np.int32)
        c_t = np.zeros(n, np.float64)
        errs = []
        got = [0]

        def collect():
            while got[0] < n:
                k = srv.wait_some(c_ids[got[0]:], c_st[got[0]:], c_t[got[0]:], 30000.0)
                if k == 0:
                    return
                got[0] += k

        def submitter(k):
            mine = np.arange(k, n, n_sub)
            my_sched = sched[mine]
            my_rq, my_op, my_ol = (req_ptr[mine % P], out_ptr[mine % P], out_len[mine % P])
            my_ids = np.zeros(len(mine), np.uint64)
            my_st = np.zeros(len(mine), np.int32)
            i, m = 0, len(mine)
            while i < m:
                now = time.monotonic() - t0
                j = int(np.searchsorted(my_sched, now, side="right"))
                if j <= i:
                    dt = my_sched[i] - now
                    time.sleep(dt - 1e-4 if dt > 2e-4 else 0)
                    continue
                j = min(j, i + 256)
                rc = srv.submit_n_raw(my_rq[i:].ctypes.data, my_op[i:].ctypes.data,
                                      my_ol[i:].ctypes.data, j - i, my_ids[i:].ctypes.data,
                                      my_st[i:].ctypes.data)
                if rc:
                    errs.append(rc)
                    return
                i = j
            ids[mine] = my_ids
            sub_st[mine] = my_st

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.monotonic()
        subs = [threading.Thread(target=submitter, args=(k,)) for k in range(n_sub)]
        for t in subs:
            t.start()
        for t in subs:
            t.join()
        if errs:
            raise RuntimeError(f"server submit rc={errs[0]}")
        th.join()
        if got[0] < n or c_st.any():
            raise RuntimeError(f"serving: {got[0]}/{n} completed, statuses {set(c_st.tolist())}")
        return t0, ids, c_ids, c_t

    # untimed warm-up: one burst of the whole pool, so the per-batch buffers
    # reach their steady-state sizes before any level is measured
    run_level(np.zeros(P), 1)
    rng = np.random.default_rng(seed)
    out = {}
    levels = list(qps_levels) + ["saturation"]
    for lvl in levels:
        n = 24 * P if lvl == "saturation" else int(lvl * seconds)
        gaps = (np.zeros(n) if lvl == "saturation" else rng.exponential(1.0 / lvl, n))
        sched = np.cumsum(gaps)
        n_sub = 1 if lvl != "saturation" and lvl <= 16000 else n_sub_max
        gc.collect()
        gc.disable()
        st0 = srv.stats()
        try:
            t0, ids, c_ids, c_t = run_level(sched, n_sub)
        finally:
            gc.enable()
        st1 = srv.stats()
        order = np.argsort(ids)
        pos = order[np.searchsorted(ids[order], c_ids)]
        lat = (c_t - t0 - sched[pos]) * 1e3
        wall = float(c_t.max() - t0)
        pix = st1["pixel_draws"] - st0["pixel_draws"]
        out[str(lvl)] = dict(requests=n, Mpix_s=round(pix / wall / 1e6, 1),
                             req_s=round(n / wall, 1),
                             p50_ms=round(float(np.percentile(lat, 50)), 3),
                             p99_ms=round(float(np.percentile(lat, 99)), 3),
                             batches=int(st1["batches"] - st0["batches"]))
        log(f"serving {lvl}: {out[str(lvl)]}")
    st = srv.stats()
    out["peak_hbm_MiB"] = round(st["peak_bytes"] / 2**20, 1)
    out["hbm_budget_GiB"] = budget / 2**30
    out["rejected"] = int(st["rejected"])
    srv.close()
    sys.setswitchinterval(switch)
    return out


def bench_png(gpu, ctx, sc, reps=10):
    """PNG export of the rendered 3840x2160 scene (z2d_gpu_png_encode from a
    device surface): device time of the encode kernels, the end-to-end call
    with the file copied to host memory, and the file size.  Its CPU
    baseline is the oracle's scanline stream plus zlib level 6 on one core
    (the reference compresses with std.compress.flate at its default
    level)."""
    import zlib
    surf, b = build_batch(gpu, ctx, sc)
    b.render(0)
    b.wait()
    b.close()
    d = surf.desc()
    png = gpu.png_encode(ctx, d)  # warm-up: buffers, code objects
    dev, wall = [], []
    for _ in range(reps):
        t = time.perf_counter()
        png = gpu.png_encode(ctx, d)
        wall.append((time.perf_counter() - t) * 1e3)
        dev.append(gpu.png_timing(ctx))
    raw_len = sc["height"] * (1 + 4 * sc["width"])
    host = np.zeros(sc["width"] * sc["height"] * 4, np.uint8)
    surf.download(host)
    surf.close()
    res = dict(device_ms=round(float(np.median(dev)), 4), call_ms=round(float(np.median(wall)), 4),
               stream_bytes=raw_len, file_bytes=len(png),
               stream_GBps=round(raw_len / (np.median(dev) * 1e-3) / 1e9, 1))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_backend import png_scanlines  # the checker, timed as the baseline
    t = time.perf_counter()
    stream = png_scanlines(sc["format"], sc["width"], sc["height"], host.tobytes())
    t1 = time.perf_counter()
    z = zlib.compress(stream, 6)
    t2 = time.perf_counter()
    res["cpu_baseline"] = dict(ms=round((t2 - t) * 1e3, 1), scanlines_ms=round((t1 - t) * 1e3, 1),
                               zlib6_ms=round((t2 - t1) * 1e3, 1), file_bytes=len(z) + 57,
                               cores=1, kind="port",
                               sample="the same surface: C oracle scanlines + zlib level 6")
    return res


TEXT_WORDS = ("The quick brown fox jumps over the lázy dog; façade naïve Ærøskøbing "
              "0123456789 AVATAR Wavy typography kerning ligatures").split()


def bench_text(gpu, ctx, lines=100, reps=5):
    """A 3840x2160 page of text (Inter Regular, size 18, one line per draw,
    ~100 characters each) laid out on the host (z2d_gpu_batch_text) and
    rendered by one batch launch; the oracle rasterizing the same text
    paths on one core is its CPU baseline."""
    font_path = os.path.join(ROOT, "tests", "golden", "fonts", "Inter-Regular.ttf")
    font = gpu.Font(open(font_path, "rb").read())
    rng = scenes.SplitMix64(0x7E47)
    texts = []
    for _ in range(lines):
        words = []
        while sum(len(w) + 1 for w in words) < 100:
            words.append(TEXT_WORDS[int(rng.uniform() * len(TEXT_WORDS)) % len(TEXT_WORDS)])
        texts.append(" ".join(words))
    opts = abi.text_opts(18.0, fill=abi.fill_opts(abi.AA_DEFAULT))
    pat = abi.solid_pattern(abi.rgba(20, 20, 20, 255))
    surf = gpu.DeviceSurface(ctx, abi.RGBA, 3840, 2160)
    surf.upload(np.full(3840 * 2160 * 4, 255, np.uint8))
    b = gpu.Batch(ctx)
    t0 = time.perf_counter()
    for i, t in enumerate(texts):
        assert b.text(surf, pat, font, t, 10.0, 20.0 + 21.0 * i, opts) == 0
    t1 = time.perf_counter()
    b.prepare()
    t2 = time.perf_counter()
    b.render(0)
    b.wait()
    b.counters(True)
    b.render(0)
    b.wait()
    pix = b.counters(False)[0]
    pipe = []
    for _ in range(reps):
        b.render(0)
        b.wait()
        pipe.append(b.timing()[0])
    glyphs = sum(len(t.replace(" ", "")) for t in texts)
    nodes = 0
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_backend import OracleBackend  # the checker, timed as the baseline
    from zscene import Surface
    ref = Surface(abi.RGBA, 3840, 2160)
    ref.buf[: ref.nbytes] = 255
    paths = [font.text_path(t, 10.0, 20.0 + 21.0 * i, opts)[1] for i, t in enumerate(texts)]
    nodes = sum(len(p) for p in paths)
    ob = OracleBackend()
    t3 = time.perf_counter()
    for p in paths[:20]:
        ob.fill(ref, pat, p, opts.fill)
    cpu_s = (time.perf_counter() - t3) * lines / 20
    b.close()
    surf.close()
    ms = float(np.median(pipe))
    return dict(lines=lines, glyphs=glyphs, path_nodes=nodes, pixel_draws=int(pix),
                host_layout_ms=round((t1 - t0) * 1e3, 2), prepare_ms=round((t2 - t1) * 1e3, 2),
                device_pipeline_ms=round(ms, 4), glyphs_per_s=round(glyphs / (ms * 1e-3), 0),
                cpu_baseline=dict(ms=round(cpu_s * 1e3, 1), cores=1, kind="port",
                                  sample="oracle fill of the first 20 of the 100 line paths, "
                                         "scaled to 100"))


CPU_LABEL = "C oracle restatement, not the Zig painter"


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0))
    # the host threads one GPU's job may use (the machine's share per GPU)
    threads = max(1, min(16, avail))
    return model, os.cpu_count(), avail, threads


def _on_threads(n, fn, items):
    """fn over items on n threads (the oracle's calls release the GIL);
    returns the wall time."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(n) as ex:
        t = time.perf_counter()
        list(ex.map(fn, items))
        return time.perf_counter() - t


def cpu_baseline(gpu, ctx, sc, n_paths):
    """The CPU painter on the host cores, on bounded samples of each
    workload: the C oracle (a scalar restatement of the reference painter;
    the Zig painter itself cannot be built here), strokes through the host
    tessellator and the oracle's raster.  Single-core figures, and all-core
    ones with one job per thread (the reference harness's jobs-over-cores
    shape).  Pixel-draw counts come from the device's counters over the same
    sample."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_backend import OracleBackend, lib as oracle_lib  # the checker, timed as the baseline
    from stroke_check import StrokeCheckBackend
    from zscene import Surface
    model, ncpu, avail, threads = _cpu_info()
    ob = OracleBackend()

    # 4K scene (the headline workload): the first n_paths paths, one core;
    # all cores: one copy of a shorter prefix per thread
    _, b = build_batch(gpu, ctx, sc, n_paths)
    pix, _ = count_work(b)
    b.close()

    def scene_job(n):
        s = Surface(sc["format"], sc["width"], sc["height"])
        s.buf[: s.nbytes] = sc["buf"]
        for pat, nodes, opts in sc["draws"][:n]:
            ob.fill(s, pat, nodes, opts)

    t = time.perf_counter()
    scene_job(n_paths)
    dt = time.perf_counter() - t
    n_all = max(1, n_paths // 4)
    _, b = build_batch(gpu, ctx, sc, n_all)
    pix_all, _ = count_work(b)
    b.close()
    dt_all = _on_threads(threads, lambda i: scene_job(n_all), range(threads))

    # 512x512 star latency: one render per call
    star = scenes.star_512()
    lat = []
    for _ in range(30):
        s = Surface(star["format"], star["width"], star["height"])
        s.buf[: s.nbytes] = star["buf"]
        t0 = time.perf_counter()
        for pat, nodes, opts in star["draws"]:
            ob.fill(s, pat, nodes, opts)
        lat.append((time.perf_counter() - t0) * 1e3)

    # compositor sweep on a 1024x1024 crop of the same distribution: every
    # operator, out = op(src, dst) with out a copy of dst composited in place
    cw = ch = 1024
    src, dst, _mask = scenes.compositor_inputs(cw, ch)

    def comp_sweep(_i=0):
        so = Surface(abi.RGBA, cw, ch)
        so.buf[: so.nbytes] = src
        out = Surface(abi.RGBA, cw, ch)
        sd = so.desc()
        times = {}
        for name in abi.OPERATORS:
            out.buf[: out.nbytes] = dst
            o = abi.CompOp()
            o.op = abi.OP[name]
            o.src.kind = abi.PARAM_SURFACE
            o.src.surface = C.pointer(sd)
            od = out.desc()
            arr = (abi.CompOp * 1)(o)
            t0 = time.perf_counter()
            oracle_lib().zo_composite(C.byref(od), 0, 0, arr, 1, abi.PRECISION_INTEGER)
            times[name] = time.perf_counter() - t0
        return times

    ct = comp_sweep()
    comp_1 = {k: round(cw * ch * _op_bytes(k, False) / v / 1e9, 3) for k, v in ct.items()}
    tot_b = sum(cw * ch * _op_bytes(k, False) for k in abi.OPERATORS)
    comp_all = tot_b * threads / _on_threads(threads, comp_sweep, range(threads)) / 1e9

    # batched form: the first 256 of the 4096 uneven surfaces (30% strokes)
    work = scenes.uneven_surfaces(256, 0xC4)
    bb = gpu.Batch(ctx)
    keep = []
    for fmt, side, jobs, k in work:
        ds = gpu.DeviceSurface(ctx, fmt, side, side)
        keep.append(ds)
        for kind, pat, nodes, opts in jobs:
            rc = (bb.stroke if kind == "stroke" else bb.fill)(ds, pat, nodes, opts)
            if rc:
                raise RuntimeError(f"batch {kind} rc={rc}")
    bb.prepare()
    pix_c4, _ = count_work(bb)
    bb.close()
    for ds in keep:
        ds.close()
    sbe = StrokeCheckBackend()

    def surface_job(item):
        fmt, side, jobs, _k = item
        s = Surface(fmt, side, side)
        for kind, pat, nodes, opts in jobs:
            (sbe.stroke if kind == "stroke" else sbe.fill)(s, pat, nodes, opts)

    dt_c4 = _on_threads(1, surface_job, work)
    dt_c4_all = _on_threads(threads, surface_job, work)

    return dict(value=round(pix / dt / 1e6, 2), unit="Mpix/s", cores=1, kind="port",
                label=CPU_LABEL, cpu_model=model, host_cpus=ncpu, cpus_available=avail,
                sample=f"first {n_paths} of the 10,000 paths of the same 3840x2160 scene, "
                       f"one core, {dt:.1f} s",
                all_cores=dict(value=round(pix_all * threads / dt_all / 1e6, 2),
                               unit="Mpix/s", cores=threads,
                               sample=f"one copy of the first {n_all} paths per thread, "
                                      f"{dt_all:.1f} s wall"),
                legs={
                    "star_512_latency_ms": dict(p50=round(float(np.percentile(lat, 50)), 3),
                                                p99=round(float(np.percentile(lat, 99)), 3),
                                                cores=1, renders=len(lat)),
                    "compositor_GBps": dict(per_op=comp_1, cores=1,
                                            all_ops_all_cores=round(comp_all, 2),
                                            threads=threads,
                                            sample="1024x1024 RGBA8 inputs of the 0xC2 "
                                                   "distribution, bytes as the device "
                                                   "sweep counts them"),
                    "uneven_batch_Mpix_s": dict(one_core=round(pix_c4 / dt_c4 / 1e6, 2),
                                                all_cores=round(pix_c4 / dt_c4_all / 1e6, 2),
                                                threads=threads, surfaces=len(work),
                                                sample="first 256 of the 4096 surfaces "
                                                       "(seed 0xC4), one surface per job"),
                })


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--paths", type=int, default=10000)
    ap.add_argument("--cpu-paths", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=400,
                    help="distinct requests in the serving pool (cycled at each QPS level)")
    ap.add_argument("--comp-only", action="store_true",
                    help="only run the 4096x4096 compositor sweep (one JSON line)")
    ap.add_argument("--text-only", action="store_true",
                    help="only run the text page (one JSON line)")
    ap.add_argument("--png-only", action="store_true",
                    help="only run the PNG export of the rendered scene (one JSON line)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run the stream-copy kernel on this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py renders on one GPU only (got --gpus {args.gpus}); "
              "runs across several GPUs are not supported", file=sys.stderr)
        sys.exit(2)
    from z2d_amd import gpu
    ctx = gpu.Context()
    if args.copy_probe:
        g = C.c_double()
        ctx.check(gpu.lib().z2d_gpu_stream_copy(ctx.h, args.copy_probe << 20, 1, C.byref(g)),
                  "stream_copy")
        print(json.dumps({"stream_copy_MiB": args.copy_probe, "GBps": round(g.value, 1)}))
        return
    if args.comp_only:
        print(json.dumps({"compositor_4096": bench_compositor(gpu, ctx)}))
        return
    if args.text_only:
        print(json.dumps({"text_page_3840x2160": bench_text(gpu, ctx)}))
        return
    log("generating scene")
    sc = scenes.scene_4k(n_paths=args.paths)
    if args.png_only:
        print(json.dumps({"png_export_3840x2160": bench_png(gpu, ctx, sc)}))
        return
    log("rendering scene")
    r = bench_scene(gpu, ctx, sc, args.steps, args.warmup)
    log(f"scene: {r['render_ms']:.3f} ms render, {r['pipe_ms']:.3f} ms pipeline")
    ms = r["wall"] * 1e3 / args.steps
    value = r["pix"] * args.steps / r["wall"] / 1e6
    achieved = r["alg_bytes"] / (r["render_ms"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "scene_3840x2160_10k_paths", "width": 3840, "height": 2160,
                   "paths": args.paths, "seed": "0xC3", "aa": "multisample_4x",
                   "op": "src_over", "pixel_draws_per_step": int(r["pix"]),
                   "edges_per_step": int(r["edges"]),
                   "strip_draw_pairs_per_step": r["strip_entries"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "render stage: k_render_fast + k_render (the two strip "
                               "workers, back to back on the batch stream; HIP events "
                               "before the first and after the second)",
                     "kernel_ms": round(r["render_ms"], 4),
                     "pipeline_ms": round(r["pipe_ms"], 4)},
    }
    tj = os.path.join(ROOT, "profiles", "k_render_scene_3840x2160_10k.traffic.json")
    if os.path.exists(tj) and args.paths == 10000:
        # HBM bytes per render stage from rocprofv3 FETCH_SIZE/WRITE_SIZE
        # runs of this workload, scaled by the stream-copy calibration; the
        # profile records the hash of the kernel sources it was taken with,
        # so a number older than the kernels shows as stale
        with open(tj) as fh:
            tr = json.load(fh)
        line["roofline"]["traffic"] = tr["traffic_bytes_per_launch"]
        line["roofline"]["traffic_source"] = "profiles/k_render_scene_3840x2160_10k.md"
        line["roofline"]["traffic_kernel_sources"] = tr.get("kernel_sources_sha256")
        line["roofline"]["traffic_stale"] = tr.get("kernel_sources_sha256") != kernel_sources_hash()
        if "valu_insts_per_launch" in tr:
            # the stage is bound by VALU issue: the counted VALU instructions
            # of k_render_fast (SQ_INSTS_VALU, same profile) over this run's
            # stage time, against one wave64 VALU instruction per SIMD per 4
            # cycles on 1,024 SIMDs at 2.4 GHz
            slots = 1024 * 2.4e9 / 4 * r["render_ms"] * 1e-3
            line["roofline"]["valu_issue"] = {
                "insts": tr["valu_insts_per_launch"],
                "frac": round(tr["valu_insts_per_launch"] / slots, 4),
                "frac_traced": tr.get("valu_issue_frac_traced"),
                "source": "profiles/k_render_scene_3840x2160_10k.md"}
    if not args.no_cpu:
        log("cpu baseline")
        line["cpu_baseline"] = cpu_baseline(gpu, ctx, sc, min(args.cpu_paths, args.paths))
    if not args.no_extra:
        log("extras")
        p50, p99 = bench_star_latency(gpu, ctx)
        g = C.c_double()
        ctx.check(gpu.lib().z2d_gpu_stream_copy(ctx.h, 1 << 30, 20, C.byref(g)), "stream_copy")
        line["extra"] = {
            "stream_copy_GBps": round(g.value, 1),
            "star_512_latency_ms": {"p50": round(p50, 4), "p99": round(p99, 4)},
            "compositor_4096_GBps": bench_compositor(gpu, ctx),
            "uneven_fill_batch_Mpix_s": bench_uneven_batch(gpu, ctx),
            "serving": bench_serving(gpu, ctx, serving_requests(gpu, args.serve_requests)),
            "png_export_3840x2160": bench_png(gpu, ctx, sc),
            "text_page_3840x2160": bench_text(gpu, ctx),
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
